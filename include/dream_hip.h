/*
 * dream_hip.h -- C ABI of the MI355X (gfx950) implementation of DREAM's two
 * data-parallel hot paths:
 *
 *   SIR  the discrete-representation spatial impulse response of a
 *        rectangular element (dreamrect) and of an array of rectangular
 *        elements (dream_arr_rect);
 *   DAS  the delay-and-sum beamformer das (SAFT, TFM, RCA column/row
 *        transmit), plus linear interpolation, apodization and the adjoint
 *        for model-based imaging.
 *
 * Conventions (the reference's, so its gateways can bind these directly):
 *   - matrices are column-major: Ro is No x 3 ([x | y | z]), G is L x 3,
 *     H is nt x No (point n's SIR at H[n*nt .. n*nt+nt-1]), Y is K x n_ascans,
 *     Im is No;
 *   - units mm, us, m/s; dt in us; delays in us;
 *   - enum arguments use the reference's numbering (include/dream.h:42-102):
 *       SIRError   none=0 ignore_out_of_bounds=1 warn_out_of_bounds=2 out_of_bounds=3
 *       ErrorLevel none=0 ignore=1 warn=2 stop=3 parallel_stop=4
 *       DelayType  single=0 multiple=1
 *       FocusMet   none=0 x=1 negx=2 y=3 negy=4 xy=5 x_y=6 ud=7
 *       SteerMet   none=0 x=1 y=2 xy=3
 *       ApodMet    triangle=0 gauss=1 raised_cosine=2 hann=3 hamming=4
 *                  simply_supported=5 clamped=6 ud=7
 *       DASType    saft=0 tfm=1 rca_coltx=2 rca_rowtx=3
 *   - every function returns 0 on success and a negative DREAM_HIP_E* code
 *     on an argument or HIP runtime error; dream_hip_last_error() then holds
 *     the message (per thread).  The SIR/DAS out-of-bounds result is the
 *     separate SIRError written to *sir_err.
 *   - "_dev" entry points take device pointers for the large arrays and a
 *     hipStream_t (as void*; NULL = the default stream) and do not
 *     synchronise unless noted; the other entry points take caller-owned host memory,
 *     copy it to the device, and return after the result is back in host
 *     memory, as the reference's CPU drivers do.
 *
 * No C++ types cross this boundary.  include/dream_hip.hpp layers the
 * reference's C++ classes (ArrRect, Rect, DAS<T>) over it.
 */
#ifndef DREAM_HIP_H
#define DREAM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes. */
#define DREAM_HIP_OK            0
#define DREAM_HIP_EINVAL      (-1)   /* invalid argument                       */
#define DREAM_HIP_ERUNTIME    (-2)   /* HIP runtime error (no device, fault..)  */
#define DREAM_HIP_ENOMEM      (-3)   /* device allocation failed               */
#define DREAM_HIP_EUNSUPPORTED (-4)  /* feature not supported on this path      */
#define DREAM_HIP_ECANCELLED  (-5)   /* dream_hip_abort() was called           */

/* Output / data precision selectors. */
#define DREAM_HIP_F64 0              /* FP64 data, FP64 delays (reference T=double)         */
#define DREAM_HIP_F32_REF 1          /* FP32 data and FP32 delay arithmetic (reference T=float) */
#define DREAM_HIP_F32 2              /* FP32 data and sums, FP64 delays and sample indices    */

/* DAS sampling rules. */
#define DREAM_HIP_NEAREST 0          /* reference rule: Y[trunc(t*fs)]                     */
#define DREAM_HIP_LINEAR  1          /* linear interpolation, zero outside [0, K)         */

/* ---------------------------------------------------------------- library */

/* Message of the last failed call on this thread ("" if none). */
const char *dream_hip_last_error(void);

/* Library version string. */
const char *dream_hip_version(void);

/* Number of visible HIP devices (0 when there is none); never fails. */
int dream_hip_device_count(void);

/* Launch-shape autotuning (on by default; DREAM_HIP_AUTOTUNE=0 turns it
 * off at start-up).  When on, the first launch of a DAS forward, DAS
 * region adjoint, DAS batch or SIR batch shape class times the candidate
 * launch shapes (work-group size, split, regions per work-group, persistent
 * grid) on the caller's stream and keeps the
 * fastest, cached per process by (device architecture, shape class) and,
 * when DREAM_HIP_TUNE_CACHE names a file, in that file.  Results are the
 * same for every shape up to the summation order.  No counterpart in the
 * reference. */
void dream_hip_set_autotune(int on);

/* Cooperative cancellation: dream_hip_abort() clears the running flag; it is
 * safe to call from a signal handler or another thread.  Every entry point
 * tests the flag on entry.  Long eager calls (DAS forwards and adjoints, SIR
 * plans and the host-pointer SIR entry points) run as consecutive chunk
 * launches of about 20-30 ms of device time each and test the flag before
 * each chunk, keeping at most two chunks queued ahead of the device (the calling
 * thread waits for the device meanwhile), so a call cancelled while it runs
 * stops within about two chunks and returns DREAM_HIP_ECANCELLED with the
 * outputs of the launched chunks written.  Calls that capture into or replay
 * a hipGraph (SIR plans and DAS batches with use_graph), DAS batches and the
 * attenuated SIR test the flag on entry only.  The reference's threads test
 * `running` after every observation point (src/dream_arr_rect.cc:177-187,
 * extra_src/das.cc:180-186).  Mirrors ArrRect::abort / DAS<T>::abort /
 * set_running / is_running (include/dream_arr_rect.h:53-54, include/das.h:74-76). */
void dream_hip_abort(int signum);
void dream_hip_set_running(void);
int dream_hip_is_running(void);

/* Test hook, no counterpart in the reference: work units (pixel-channel
 * pairs or sub-element contributions) per chunk of a long call (<= 0: the
 * default, 5 x 10^10), and, when cancel_after_chunks > 0, a dream_hip_abort()
 * issued by the library itself right after that many further chunks have
 * been launched (<= 0: off).  Lets tests run small problems as many chunks
 * and cancel a call deterministically between two of them. */
void dream_hip_test_chunking(int64_t work_per_chunk, int64_t cancel_after_chunks);

/* ---------------------------------------------------------------- SIR */

/* dream_arr_rect on the GPU.  Replaces ArrRect::dream_arr_rect
 * (include/dream_arr_rect.h:40-51, src/dream_arr_rect.cc:310-433): same
 * arguments in the same order.  alpha [dB/(cm MHz)] > DBL_EPSILON adds the
 * reference's frequency-dependent attenuation (Attenuation::att,
 * src/attenuation.cc:72-140): every in-range contribution becomes a lossy
 * pulse, computed per point in the frequency domain and transformed once
 * (any nt).  Host pointers.  h (nt x No FP64)
 * is overwritten (zeroed, then accumulated), which equals the reference's
 * accumulate-into-a-zeroed-buffer contract.
 * *sir_err: for err_level stop, out_of_bounds if any contribution fell
 * outside [0, nt); for ignore/warn, that level's code if any observation
 * point's last element had an out-of-bounds contribution (the reference's
 * per-point status, src/dream_arr_rect.cc:227-239), else none. */
int dream_hip_arr_rect(double alpha, const double *Ro, int64_t No, double a, double b,
                       double dx, double dy, double dt, int64_t nt, int delay_type,
                       const double *delay, double v, double cp, int64_t num_elements,
                       const double *G, int foc_met, const double *focal, int steer_met,
                       double theta, double phi, const double *apod, int do_apod,
                       int apod_met, double apod_par, double *h, int err_level,
                       int *sir_err);

/* dreamrect on the GPU.  Replaces Rect::dreamrect (include/dreamrect.h:39-46,
 * src/dreamrect.cc:169-272); host pointers; same error rule as above with one
 * element.  Also replaces Rect::cl_dreamrect (opencl/cl_dreamrect.cc:42)
 * without its No % 64 == 0 restriction. */
int dream_hip_rect(double alpha, const double *Ro, int64_t No, double a, double b,
                   double dx, double dy, double dt, int64_t nt, int delay_type,
                   const double *delay, double v, double cp, double *h, int err_level,
                   int *sir_err);

/* Device-pointer variant of dream_hip_arr_rect.  Ro (No x 3 FP64), delay
 * (1 or No FP64) and h are DEVICE pointers; G, focal and apod are HOST
 * pointers (per-element delays and weights are evaluated on the host with
 * the C library's sin/cos/exp, as the reference does).  out_dtype is
 * DREAM_HIP_F64 or DREAM_HIP_F32 (FP32 sums, FP64 indices).  Runs on
 * `stream` and returns after synchronising it, with *sir_err and *oob_count
 * (either may be NULL) written.  (The plan API below is the asynchronous,
 * graph-capturable form.) */
int dream_hip_arr_rect_dev(const double *Ro, int64_t No, double a, double b, double dx,
                           double dy, double dt, int64_t nt, int delay_type,
                           const double *delay, double v, double cp, int64_t num_elements,
                           const double *G, int foc_met, const double *focal,
                           int steer_met, double theta, double phi, const double *apod,
                           int do_apod, int apod_met, double apod_par, void *h,
                           int out_dtype, int err_level, int *sir_err, int64_t *oob_count,
                           void *stream);

/* dream_arr_circ / dreamcirc on the GPU: the same as dream_hip_arr_rect /
 * dream_hip_rect for circular elements of radius R.  Replace
 * ArrCirc::dream_arr_circ (include/dream_arr_circ.h, src/dream_arr_circ.cc:
 * 87-420) and Circ::dreamcirc (include/dreamcirc.h, src/dreamcirc.cc:78-352):
 * the DR sum walks rows ys = -R + dy/2 (+= dy, while ys <= R) and in each
 * row xs = -sqrt(R*R - ys*ys) + dx/2 (+= dx, while xs <= sqrt(R*R - ys*ys)),
 * with the index and amplitude rules of the rectangle.  Host pointers;
 * FP64 output; alpha as for dream_hip_arr_rect. */
int dream_hip_arr_circ(double alpha, const double *Ro, int64_t No, double R, double dx,
                       double dy, double dt, int64_t nt, int delay_type, const double *delay,
                       double v, double cp, int64_t num_elements, const double *G, int foc_met,
                       const double *focal, int steer_met, double theta, double phi,
                       const double *apod, int do_apod, int apod_met, double apod_par,
                       double *h, int err_level, int *sir_err);
int dream_hip_circ(double alpha, const double *Ro, int64_t No, double R, double dx, double dy,
                   double dt, int64_t nt, int delay_type, const double *delay, double v,
                   double cp, double *h, int err_level, int *sir_err);
/* Device-pointer variant (conventions of dream_hip_arr_rect_dev). */
int dream_hip_arr_circ_dev(const double *Ro, int64_t No, double R, double dx, double dy,
                           double dt, int64_t nt, int delay_type, const double *delay, double v,
                           double cp, int64_t num_elements, const double *G, int foc_met,
                           const double *focal, int steer_met, double theta, double phi,
                           const double *apod, int do_apod, int apod_met, double apod_par,
                           void *h, int out_dtype, int err_level, int *sir_err,
                           int64_t *oob_count, void *stream);

/* dream_arr_cylind / dreamcylind on the GPU: the same as dream_hip_arr_rect /
 * dream_hip_rect for cylindrically curved elements of x-size a, y-size b
 * and curvature radius Rcurv (> 0 focused/concave, < 0 defocused/convex).
 * Replace ArrCylind::dream_arr_cylind (include/dream_arr_cylind.h,
 * src/dream_arr_cylind.cc:90-437) and Cylind::dreamcylind
 * (include/dreamcylind.h, src/dreamcylind.cc:78-353): the DR sum walks
 * angle rows phi = -phi_max + dphi/2 (+= dphi = asin(dy/|Rcurv|), while
 * phi <= phi_max) at ys = |Rcurv| sin(phi), depth zs = |Rcurv| -
 * sqrt(Rcurv^2 - ys^2) (rz = z - zs focused, z + zs defocused), and in each
 * row the rectangle's x walk; the surface element is |Rcurv| dx dphi.  With
 * b > 2|Rcurv| the reference's angles are not numbers and the response is
 * zero.  Host pointers; FP64 output; alpha as for dream_hip_arr_rect. */
int dream_hip_arr_cylind(double alpha, const double *Ro, int64_t No, double a, double b,
                         double Rcurv, double dx, double dy, double dt, int64_t nt,
                         int delay_type, const double *delay, double v, double cp,
                         int64_t num_elements, const double *G, int foc_met,
                         const double *focal, int steer_met, double theta, double phi,
                         const double *apod, int do_apod, int apod_met, double apod_par,
                         double *h, int err_level, int *sir_err);
int dream_hip_cylind(double alpha, const double *Ro, int64_t No, double a, double b, double Rcurv,
                     double dx, double dy, double dt, int64_t nt, int delay_type,
                     const double *delay, double v, double cp, double *h, int err_level,
                     int *sir_err);
/* Device-pointer variant (conventions of dream_hip_arr_rect_dev). */
int dream_hip_arr_cylind_dev(const double *Ro, int64_t No, double a, double b, double Rcurv,
                             double dx, double dy, double dt, int64_t nt, int delay_type,
                             const double *delay, double v, double cp, int64_t num_elements,
                             const double *G, int foc_met, const double *focal, int steer_met,
                             double theta, double phi, const double *apod, int do_apod,
                             int apod_met, double apod_par, void *h, int out_dtype,
                             int err_level, int *sir_err, int64_t *oob_count, void *stream);

/* dream_arr_annu on the GPU.  Replaces ArrAnnu::dream_arr_annu
 * (include/dream_arr_annu.h:40-50, src/dream_arr_annu.cc:514-634).  Gr holds
 * num_radii radii: the centre disc, then an (inner, outer) pair per ring, so
 * there are (num_radii + 1)/2 rings.  Ring n's response is the disc response
 * of radius Gr[2n] minus that of Gr[2n-1] (ring 0: disc 0, left zero when
 * Gr[0] <= DBL_EPSILON), moved later by rint(foc_n/dt) + 1 samples and cut at
 * nt, times its weight.  Focusing x / y / xy / x+y use the spherical law
 * (sqrt(r_max^2 + focal[0]^2) - sqrt(r_n^2 + focal[0]^2)) * 1000 / cp about
 * the ring's mean radius r_n; 'ud' takes focal[n] in us.  Apodization uses
 * r_n / r_max; 'ud' takes apod[n]; Hann and Hamming give 1.  Gr, focal and
 * apod are host arrays (focal and apod: num_radii values for 'ud', of which
 * the first (num_radii+1)/2 are used).  *sir_err: E_OOB if any disc of any
 * point stopped; otherwise the status of the last disc.  Host pointers for
 * Ro, delay and h; FP64 output; alpha as for dream_hip_arr_rect (every disc is
 * attenuated, then the rings are formed, as in the reference). */
int dream_hip_arr_annu(double alpha, const double *Ro, int64_t No, double dx, double dy,
                       double dt, int64_t nt, int delay_type, const double *delay, double v,
                       double cp, int64_t num_radii, const double *Gr, int foc_met,
                       const double *focal, const double *apod, int do_apod, int apod_met,
                       double apod_par, double *h, int err_level, int *sir_err);
/* Device-pointer variant (conventions of dream_hip_arr_rect_dev; Gr, focal
 * and apod stay on the host).  Allocates num_radii x No x nt doubles of
 * device scratch for the disc responses during the call. */
int dream_hip_arr_annu_dev(const double *Ro, int64_t No, double dx, double dy, double dt,
                           int64_t nt, int delay_type, const double *delay, double v, double cp,
                           int64_t num_radii, const double *Gr, int foc_met, const double *focal,
                           const double *apod, int do_apod, int apod_met, double apod_par,
                           void *h, int out_dtype, int err_level, int *sir_err,
                           int64_t *oob_count, void *stream);

/* Per-contribution sample indices of one (observation point, element) pair,
 * computed by the same device code as the SIR kernels: it_out[j] =
 * rint((t_j - delay)/dt) for the j-th sub-element in the reference's (ys
 * outer, xs inner) order.  xo,yo,zo are relative to the element centre.
 * Host pointers; returns the contribution count in *count (writes at most
 * cap).  For parity tests of the index path. */
int dream_hip_rect_indices(double xo, double yo, double zo, double a, double b, double dx,
                           double dy, double dt, double delay, double cp, int64_t *it_out,
                           int64_t cap, int64_t *count);

/* The device's FP64 reciprocal-square-root seed of each s[i] (host arrays of
 * n doubles).  The SIR and SAFT index paths certify their fast estimates
 * with margins that assume a seed relative error near 2^-24; a test hook
 * for that assumption, with no counterpart in the reference. */
int dream_hip_rsq_seed(const double *s, int64_t n, double *y);

/* Batched uneven SIR: many dream_arr_rect problems in one persistent launch.
 * Ro, delay and h are DEVICE pointers; G, focal, apod are HOST pointers. */
typedef struct DreamSirProblem {
  const double *Ro;  int64_t No;
  double a, b, dx, dy, dt;  int64_t nt;
  int32_t delay_type;  const double *delay;
  double v, cp;
  int64_t num_elements;  const double *G;
  int32_t foc_met;  const double *focal;
  int32_t steer_met;  double theta, phi;
  const double *apod;  int32_t do_apod;  int32_t apod_met;  double apod_par;
  int32_t err_level;
  void *h;                                   /* nt x No, out_dtype */
  int32_t shape;                             /* 0 rectangle (a x b), 1 disc (radius R),
                                                2 cylinder (a x b, curvature radius R) */
  double R;                                  /* disc radius / curvature radius [mm] */
} DreamSirProblem;

/* The SIR of several problems (rectangles, discs, cylindrically curved
 * elements; DreamSirProblem as for plans: Ro, delay and h DEVICE pointers)
 * with the reference's frequency-dependent attenuation when alpha
 * [dB/(cm MHz)] > DBL_EPSILON (Attenuation::att, src/attenuation.cc:72-140;
 * any nt), or without it (then as a one-shot plan).  The device-pointer
 * form of the attenuated dream_arr_rect / dream_arr_circ /
 * dream_arr_cylind; per-problem SIRError codes and out-of-bounds counts
 * (host arrays of n, may be NULL).  Synchronises `stream`. */
int dream_hip_sir_att_dev(const DreamSirProblem *problems, int32_t n, double alpha,
                          int out_dtype, int32_t *sir_err, int64_t *oob_count, void *stream);

/* Build a plan (host preparation, device tables, LPT-ordered work list) for
 * n problems; the plan owns copies of every host-side table it needs. */
int dream_hip_sir_plan_create(const DreamSirProblem *problems, int32_t n, int out_dtype,
                              void **plan);
/* Launch the plan on `stream` (captured into a hipGraph on first use when
 * use_graph != 0, replayed afterwards).  Asynchronous. */
int dream_hip_sir_plan_run(void *plan, void *stream, int use_graph);
/* After the stream has passed the run: per-problem SIRError codes and
 * out-of-bounds contribution counts (either pointer may be NULL). */
int dream_hip_sir_plan_status(void *plan, int32_t *sir_err, int64_t *oob_count);
int dream_hip_sir_plan_destroy(void *plan);

/* ---------------------------------------------------------------- DAS */

/* Create a DAS object.  Replaces DAS<T>::DAS (include/das.h:57-63); the
 * object caches device buffers and is reused while das_setup_has_changed
 * would be false (include/das.h:77-94).  precision: DREAM_HIP_F64 (T=double),
 * DREAM_HIP_F32_REF (T=float, the reference's float arithmetic) or
 * DREAM_HIP_F32 (float data, double delays).  num_r_elements == 0 forces
 * SAFT, as DAS<T>::das does (extra_src/das.cc:209-213). */
int dream_hip_das_create(int das_type, int64_t a_scan_len, int64_t No,
                         int64_t num_t_elements, int64_t num_r_elements, int precision,
                         void **handle);
int dream_hip_das_destroy(void *handle);

/* Order in which the handle visits the pixels (a permutation of 0..No-1,
 * HOST array, copied; NULL restores 0..No-1).  The result is the same for
 * any order up to the summation order; a spatially compact order makes each
 * wave's 64 pixels a compact tile whose gathers share cache lines.
 * dream_hip_das orders the pixels itself with dream_hip_pixel_order. */
int dream_hip_das_set_pixel_order(void *handle, const int32_t *perm, int64_t n);

/* Spatially compact order of No points (Morton order of the coordinates
 * quantised to 1024 cells per axis over their bounding box; inside each run
 * of 64, ordered by (z, y, x)).  Ro is No x 3 column-major FP64 (HOST); perm
 * (HOST, No entries) receives the order. */
int dream_hip_pixel_order(const double *Ro, int64_t No, int32_t *perm);

/* 1 if (das_type, a_scan_len, No, num_t, num_r, precision) differ from the
 * handle's (DAS<T>::das_setup_has_changed, include/das.h:77-94). */
int dream_hip_das_setup_has_changed(void *handle, int das_type, int64_t a_scan_len,
                                    int64_t No, int64_t num_t_elements,
                                    int64_t num_r_elements, int precision);

/* DAS<T>::das (include/das.h:67-72, extra_src/das.cc:199-298) and
 * DAS<T>::cl_das (opencl/cl_das.cc:62-159) on the GPU.  HOST pointers of the
 * handle's element type T (double for F64, float otherwise): Y (K x n_ascans),
 * Ro (No x 3), Gt (Lt x 3), Gr (Lr x 3; ignored for SAFT), delay (1 or No).
 * dt, delay and cp are T values passed as double.  Im (No) is overwritten.
 * *sir_err: the err_level code if any pixel-channel sample index fell
 * outside [0, K), else none (stop -> out_of_bounds). */
int dream_hip_das(void *handle, const void *Y, const void *Ro, const void *Gt,
                  const void *Gr, double dt, int delay_type, const void *delay, double cp,
                  void *Im, int err_level, int *sir_err);

/* Device-pointer DAS with the extensions.  Y, Ro, Gt, Gr, delay, apod_t,
 * apod_r and Im are DEVICE pointers of the handle's element type (geometry
 * and delay in double for DREAM_HIP_F32, which keeps delays in FP64).
 * interp: DREAM_HIP_NEAREST (reference) or DREAM_HIP_LINEAR.  apod_t
 * (Lt) / apod_r (Lr, unused by SAFT) may be NULL for all-ones; SAFT weights
 * A-scan n by apod_t[n], TFM/RCA weight pair (i,j) by apod_t[i]*apod_r[j].
 * n_frames >= 1 frames are beamformed in one launch: frame f reads
 * Y + f*y_frame_stride and writes Im + f*im_frame_stride (element strides).
 * Every call computes the out-of-bounds status on the device, as the
 * reference's das does (extra_src/das.cc:338-351).  With sir_err == NULL the
 * call is asynchronous and dream_hip_das_status reads the code later; with
 * sir_err it waits for its stream and writes the code.  Calls on one handle
 * must be ordered (one stream, or streams synchronised between the calls):
 * they share the handle's work buffers and status word.  Several FP32
 * frames of a two-leg (TFM with two arrays, RCA) call are first copied into
 * groups of four frames (device memory of the handle, Y's size rounded up
 * to four frames), so that one gather serves four frames. */
int dream_hip_das_dev(void *handle, const void *Y, const void *Ro, const void *Gt,
                      const void *Gr, double dt, int delay_type, const void *delay,
                      double cp, const void *apod_t, const void *apod_r, int interp,
                      int64_t n_frames, int64_t y_frame_stride, int64_t im_frame_stride,
                      void *Im, int err_level, int *sir_err, void *stream);

/* The SIRError code (at err_level) of the handle's last dream_hip_das_dev
 * call: waits for that call's stream.  No counterpart in the reference,
 * whose das returns the code synchronously; for device callers that check
 * the status after a batch of calls.  Replays of a captured call share the
 * call's status word, which then reports an out-of-bounds pair of any replay
 * since the capture. */
int dream_hip_das_status(void *handle, int err_level, int *sir_err);

/* Adjoint of dream_hip_das_dev (linear or nearest rule) for the cotangent
 * g (No per frame, DEVICE).  Outputs (DEVICE, NULL to skip; all overwritten):
 *   gY       K x n_ascans per frame (same layout and strides as Y),
 *   g_apod_t Lt, g_apod_r Lr (summed over frames),
 *   g_cp     one value of the handle's element type (summed over frames;
 *            0 for the nearest rule, whose index is piecewise constant).
 * Not available for DREAM_HIP_F32_REF handles. */
int dream_hip_das_adjoint_dev(void *handle, const void *Y, const void *Ro, const void *Gt,
                              const void *Gr, double dt, double delay, double cp,
                              const void *apod_t, const void *apod_r, int interp,
                              int64_t n_frames, int64_t y_frame_stride,
                              int64_t im_frame_stride, const void *g, void *gY,
                              void *g_apod_t, void *g_apod_r, void *g_cp, void *stream);

/* Measurement hook, no counterpart in the reference: the next n_max forward
 * calls (dream_hip_das_dev / dream_hip_das) of the handle record HIP events
 * on their stream around their kernels; dream_hip_das_kernel_times waits for
 * them and returns the number of timed calls and the summed durations (ms)
 * of the preparation kernel (the symmetrization of a reciprocal TFM call, 0
 * otherwise) and of the beamforming kernel(s), then stops timing. */
int dream_hip_das_time_kernels(void *handle, int n_max);
int dream_hip_das_kernel_times(void *handle, int *n, double *prep_ms, double *forward_ms);

/* dream_hip_das_adjoint_dev that allocates nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops).  Same arguments and results; the
 * handle must exist (creating it allocates).  Its accumulators (per-A-scan,
 * per-element and cp sums, FP64) live in a fixed 2 MiB device array per GPU
 * instead of handle memory, so n_ascans + Lt + Lr + 1 <= 262144 (larger
 * problems return DREAM_HIP_EINVAL); default launch shapes (no autotuning)
 * and one launch per kernel (cancellation is tested on entry only).  The
 * first call on a device looks the array up (hipGetSymbolAddress).  Every
 * call on one GPU shares that array, so calls that overlap on one GPU (other
 * streams, other threads) race. */
int dream_hip_das_adjoint_dev_noalloc(void *handle, const void *Y, const void *Ro,
                                      const void *Gt, const void *Gr, double dt, double delay,
                                      double cp, const void *apod_t, const void *apod_r,
                                      int interp, int64_t n_frames, int64_t y_frame_stride,
                                      int64_t im_frame_stride, const void *g, void *gY,
                                      void *g_apod_t, void *g_apod_r, void *g_cp, void *stream);

/* ---------------------------------------------------------------- das_uni */

/* DAS over implicit uniform grids: DAS_UNI<T>::DAS_UNI / cl_das_uni
 * (include/das_uni.h:50-365, opencl/das_uni_double.cl).  gt = [min_t,
 * pitch_t, max_t] (transmit elements along x, or the RCA columns), gr = [min_r,
 * pitch_r, max_r] (receive elements along x, or the RCA rows along y;
 * ignored for SAFT), Ro = the 3 x 3 column-major matrix [min_Rx dx max_Rx;
 * min_Ry dy max_Ry; min_Rz dz max_Rz] [mm].  Counts are (max - min)/pitch + 1
 * truncated; pixel n is (ny fastest, then nz, then nx).  The geometry is
 * generated on the device once per handle.  The reference kernel computes
 * in OpenCL with native_sqrt; here the arithmetic is das's (correctly rounded
 * sqrt, no contraction), with das_uni's two RCA rules: the stripe extents are
 * the [min, max] parameters, and for row transmit the delay is on the row
 * leg.  precision as dream_hip_das_create. */
int dream_hip_das_uni_create(int das_type, int64_t a_scan_len, const double *gt,
                             const double *gr, const double *Ro, int precision, void **handle);
int dream_hip_das_uni_destroy(void *handle);
/* Pixel count and element counts of a handle (any pointer may be NULL). */
int dream_hip_das_uni_shape(void *handle, int64_t *No, int64_t *num_t, int64_t *num_r);
/* DAS_UNI<T>::das_setup_has_changed: 1 if the arguments differ from the handle's. */
int dream_hip_das_uni_setup_has_changed(void *handle, int das_type, int64_t a_scan_len,
                                        const double *gt, const double *gr, const double *Ro,
                                        int precision);
/* cl_das_uni: HOST Y (K x n_ascans of the handle's element type) and Im (No);
 * scalar delay [us]; returns after Im is written. */
int dream_hip_das_uni(void *handle, const void *Y, double dt, double delay, double cp, void *Im);
/* Device form (DEVICE Y and Im, asynchronous on `stream`), with the
 * sampling rule (DREAM_HIP_NEAREST = the reference's, or DREAM_HIP_LINEAR). */
int dream_hip_das_uni_dev(void *handle, const void *Y, double dt, double delay, double cp,
                          int interp, void *Im, void *stream);

/* ---------------------------------------------------------------- convolution */

/* Convolution modes (ConvMode, include/dream.h:91-95). */
#define DREAM_HIP_CONV_EQU 0         /* Y  = A (*) B */
#define DREAM_HIP_CONV_SUM 1         /* Y += A (*) B */
#define DREAM_HIP_CONV_NEG 2         /* Y -= A (*) B */

/* FFTConvP::run (include/fftconv_p.h:39-44, extra_src/fftconv_p.cc:154-403):
 * column-wise linear convolution Y(:,n) (mode) A(:,n) (*) B(:,n), or with
 * the single filter B when B is a vector (B_M == 1 or B_N == 1).  HOST
 * pointers, FP64, column-major; Y is (A_M + B_M - 1) x A_N (read for the
 * '+=' and '-=' modes).  The reference multiplies zero-padded FFTs; this
 * evaluates the same sum directly, so results agree to rounding. */
int dream_hip_fftconv_p(double *Y, const double *A, int64_t A_M, int64_t A_N, const double *B,
                        int64_t B_M, int64_t B_N, int mode);

/* sum_fftconv (extra_src/oct_sum_fftconv.cc, add_fftconv in
 * extra_src/fftconv.cc:119-209): Y(:,n) (mode) sum_l H(:,n,l) (*) U(:,l) for an
 * M x N x L array H and a K x L matrix U.  HOST pointers, FP64; Y is
 * (M + K - 1) x N.  The reference's gateway takes only DREAM_HIP_CONV_EQU
 * (normal mode) and DREAM_HIP_CONV_SUM (in-place mode). */
int dream_hip_sum_fftconv(double *Y, const double *H, int64_t M, int64_t N, int64_t L,
                          const double *U, int64_t K, int mode);

/* Device-pointer form of both: Y(:,n) (mode) sum_{l<L} A_l(:,n) (*) B_l(:,n)
 * with A_l(:,n) at A + l*a_layer_stride + n*a_col_stride (M_A samples),
 * B_l(:,n) at B + l*b_layer_stride + n*b_col_stride (M_B samples; stride 0
 * shares a filter) and Y(:,n) at Y + n*(M_A + M_B - 1), element strides.
 * dtype: DREAM_HIP_F64 or DREAM_HIP_F32.  Asynchronous on `stream`. */
int dream_hip_conv_sum_dev(void *Y, const void *A, int64_t M_A, int64_t N, int64_t L,
                           int64_t a_col_stride, int64_t a_layer_stride, const void *B,
                           int64_t M_B, int64_t b_col_stride, int64_t b_layer_stride, int mode,
                           int dtype, void *stream);

/* Batched uneven DAS: many beamforming problems of one DAS type, precision
 * and sampling rule -- different arrays, pixel grids and A-scan lengths --
 * in one persistent forward launch.  No counterpart in the reference, whose
 * gateways beamform one problem per DAS<T>::das call (extra_src/das.cc:
 * 199-298); each problem has the meaning of one dream_hip_das_dev call with
 * one frame and a scalar delay.  All pointers are DEVICE pointers of the
 * batch's element type (geometry in double for DREAM_HIP_F32); apod_t /
 * apod_r must be given for all problems or for none; perm (No entries) is
 * the pixel visiting order or NULL. */
typedef struct DreamDasProblem {
  int64_t K, No, Lt, Lr;                     /* Lr is ignored for SAFT */
  const void *Y, *Ro, *Gt, *Gr;              /* K x n_ascans, No x 3, Lt x 3, Lr x 3 */
  double dt, delay, cp;                      /* us, us, m/s */
  const void *apod_t, *apod_r;
  const int32_t *perm;
  void *Im;                                  /* No, overwritten */
} DreamDasProblem;

/* Build a batch (device copies of the descriptors, work list of 64-pixel
 * tiles in decreasing per-tile cost).  The device arrays must stay valid
 * while the batch is used. */
int dream_hip_das_batch_create(const DreamDasProblem *problems, int32_t n, int das_type,
                               int precision, int interp, void **batch);
/* Launch on `stream` (captured into a hipGraph on first use when use_graph
 * != 0, replayed afterwards).  Asynchronous; no out-of-bounds status. */
int dream_hip_das_batch_run(void *batch, void *stream, int use_graph);
int dream_hip_das_batch_destroy(void *batch);

#ifdef __cplusplus
}
#endif

#endif /* DREAM_HIP_H */
