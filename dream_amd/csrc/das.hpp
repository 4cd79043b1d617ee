// das.hpp -- host side of the delay-and-sum (das) kernels.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace dream {

// Arguments of one DAS launch (all device pointers).
struct DasArgs {
  const void *Y;  int64_t K;
  const void *Ro;  int64_t No;
  const void *Gt;  int32_t Lt;
  const void *Gr;  int32_t Lr;
  const void *delay;  int32_t delay_multi;
  double fs_khz, inv_cp, delay_ms, cp;   // in the geometry type, passed as double
  double leg_max;                        // legs are clamped to +-leg_max (2^29 / fs)
  const void *apod_t, *apod_r;
  // RCA only: stripe extents given explicitly (sx: column x range for the
  // row leg, sy: row y range for the column leg) instead of read from the
  // first and last elements; and the delay on the receive leg (das_uni's
  // row-transmit rule) instead of the transmit leg.
  int32_t stripe_given, delay_rx;
  double sx_lo, sx_hi, sy_lo, sy_hi;
  int64_t n_frames, y_stride, im_stride;
  void *Im;
  // Forward out-of-bounds status: a pixel with an out-of-range pair sets
  // status->any_oob to status_epoch (a new value per call), so a call is
  // told from the ones before it without zeroing the word first.
  Status *status;
  uint32_t status_epoch;
  // Adjoint-only.
  const void *g;  void *gY;  void *z_t;  void *z_r;  void *z_cp;
  // Pixel visiting order (device, No entries) or null for 0..No-1, and the
  // handle's scratch for partial sums of split forward launches.
  const int32_t *perm;
  void *scratch;  size_t scratch_bytes;
  // First block (forward tile, adjoint region or pixel-adjoint tile) and
  // first frame of this launch: a long call runs as several chunk launches.
  int64_t block0, frame0, block_end;
  // TFM with one array for transmit and receive and a scalar delay: Ysym
  // holds Y[i][j] + Y[j][i] at (i, j), i < j, and Y[i][i] at (i, i) (same
  // layout and frame stride as Y).  The forward then visits the pairs i <= j
  // only, whenever the delay read on the device is zero (the pair's sample
  // index is then the same for (i, j) and (j, i): legs computed by the same
  // expression, added commutatively).
  const void *Ysym;
  int32_t recip;
  // Frames in groups of four (FP32 two-leg forwards of several frames):
  // Yg holds, per group, one 16-byte record of the four frames' samples
  // per (A-scan, sample) -- group g at Yg + g * yg_stride records, frames
  // past the last read as 0 -- so one gather serves four frames.  Null: one
  // frame per launch row.
  const void *Yg;
  int64_t yg_stride;
};

// Bytes of forward partial-sum scratch a launch of this shape uses.
size_t das_forward_scratch_bytes(int type, int prec, int64_t No, int64_t n_frames, int64_t Lt,
                                 int64_t Lr);

// Spatially compact visiting order of the pixels (Morton order of the
// quantised coordinates), so the 64 pixels of a wave form a compact tile and
// their gathers share cache lines; inside each 64-pixel tile the pixels are
// ordered by (z, y, x).  Ro is No x 3 column-major.
template <typename GT>
void pixel_order(const GT *Ro, int64_t No, int32_t *perm);

class DasHandle {
public:
  // stripes_given: the caller supplies the RCA stripe extents (das_uni), so
  // the element tables need not hold the entries the reference reads them from.
  DasHandle(int type, int64_t K, int64_t No, int64_t Lt, int64_t Lr, int precision,
            bool stripes_given = false);
  ~DasHandle();
  bool changed(int type, int64_t K, int64_t No, int64_t Lt, int64_t Lr, int precision) const;

  // Host-pointer call (DAS<T>::das semantics); returns the SIRError code.
  int run_host(const void *Y, const void *Ro, const void *Gt, const void *Gr, double dt,
               int delay_type, const void *delay, double cp, void *Im, int err_level);

  // Device-pointer forward.  The out-of-bounds status is always computed on
  // the device; the code is returned only when `want_status` (then the call
  // synchronises), else status() reads it later.
  int run_dev(const void *Y, const void *Ro, const void *Gt, const void *Gr, double dt,
              int delay_type, const void *delay, double cp, const void *apod_t,
              const void *apod_r, int interp, int64_t n_frames, int64_t y_stride,
              int64_t im_stride, void *Im, int err_level, bool want_status, hipStream_t s);

  void adjoint_dev(const void *Y, const void *Ro, const void *Gt, const void *Gr, double dt,
                   double delay, double cp, const void *apod_t, const void *apod_r, int interp,
                   int64_t n_frames, int64_t y_stride, int64_t im_stride, const void *g,
                   void *gY, void *g_apod_t, void *g_apod_r, void *g_cp, hipStream_t s,
                   double *fixed_z = nullptr, size_t fixed_z_bytes = 0);

  // The SIRError code of the last forward call at `err_level` (waits for
  // that call's stream).
  int status(int err_level);

  int type() const { return type_; }
  int precision() const { return prec_; }
  // Forward scratch for launches of n_frames frames, and its current extent.
  void ensure_scratch_for(int64_t n_frames);
  void *scratch() const { return d_scratch_; }
  size_t scratch_bytes() const { return d_scratch_bytes_; }
  void fill_args(DasArgs &a, const void *Y, const void *Ro, const void *Gt, const void *Gr,
                 double dt, int delay_type, const void *delay, double cp, const void *apod_t,
                 const void *apod_r, int64_t n_frames, int64_t y_stride, int64_t im_stride,
                 void *Im) const;

private:
  int type_;           // effective type (SAFT when Lr == 0)
  int64_t K_, No_, Lt_, Lr_;
  int prec_;
  Status *d_status_ = nullptr;
  Status *h_status_ = nullptr;   // pinned
  uint32_t epoch_ = 0;           // the last forward call's status epoch
  hipStream_t status_stream_ = nullptr;
  // Device copies for the host-pointer path (allocated on first use).
  void *d_buf_ = nullptr;
  size_t d_buf_bytes_ = 0;
  void *d_z_ = nullptr;          // adjoint channel accumulators
  size_t d_z_bytes_ = 0;
  void *d_scratch_ = nullptr;    // forward partial sums (split launches)
  size_t d_scratch_bytes_ = 0;
  void *d_sym_ = nullptr;        // reciprocal TFM: symmetrized A-scans (Ysym)
  size_t d_sym_bytes_ = 0;
  void *d_grp_ = nullptr;        // frames in groups of four (Yg)
  size_t d_grp_bytes_ = 0;
  int32_t *d_perm_ = nullptr;    // pixel order (No entries) or null
public:
  // Set (perm != null, n == No) or clear (perm == null) the pixel order; perm
  // is a host array, copied.
  void set_pixel_order(const int32_t *perm, int64_t n);
  // Kernel timing: the next n_max forward calls record HIP events on their
  // stream before the preparation kernel (the reciprocal symmetrization, if
  // any), between it and the beamforming kernel, and after that; read_timing
  // waits for them, sums both segments (ms) and stops.
  void start_timing(int n_max);
  void read_timing(int *n, double *prep_ms, double *forward_ms);
private:
  void ensure_scratch(size_t bytes);
  void group_frames(DasArgs &a, const void *Y, int64_t y_stride, int64_t n_frames,
                    hipStream_t s);
  void drop_timing();
  std::vector<hipEvent_t> tev_;  // 3 per timed call
  int t_cap_ = 0, t_n_ = 0;
};

// fs_khz, inv_cp, delay_ms (0), cp and leg_max of a launch, with the
// reference's rounding of the constants to T (extra_src/das.cc:316-318).
void das_constants(DasArgs &a, int prec, double dt, double cp);

// das_uni: DAS over implicit uniform grids.  The array elements and the
// pixels are generated on the device from [min, pitch, max] triples with the
// reference's formulas (opencl/das_uni_double.cl), then the DAS kernels run
// on them.
class DasUni {
public:
  DasUni(int type, int64_t K, const double gt[3], const double gr[3], const double Ro[9],
         int precision);
  ~DasUni();
  bool changed(int type, int64_t K, const double gt[3], const double gr[3], const double Ro[9],
               int precision) const;
  // Device pointers; Im (No) is overwritten.
  void run_dev(const void *Y, double dt, double delay, double cp, int interp, void *Im,
               hipStream_t s);
  // Host pointers (copy in, beamform, copy out).
  void run_host(const void *Y, double dt, double delay, double cp, void *Im);
  int64_t No() const { return No_; }
  int64_t Lt() const { return Lt_; }
  int64_t Lr() const { return Lr_; }
private:
  int type_, prec_;
  int64_t K_, No_ = 0, Lt_ = 0, Lr_ = 0;
  double gt_[3], gr_[3], ro_[9];
  DasHandle *h_ = nullptr;
  void *d_geom_ = nullptr;            // Ro (No x 3), Gt (Lt x 3), Gr (Lr x 3), delay
  void *d_buf_ = nullptr;             // host-path copies of Y and Im
  size_t d_buf_bytes_ = 0;
  size_t gs_ = 8;
};

// Many problems of one DAS type, precision and sampling rule in one
// persistent forward launch (device pointers; scalar delays).
class DasBatch {
public:
  DasBatch(const DreamDasProblem *problems, int32_t n, int type, int precision, int interp);
  ~DasBatch();
  void run(hipStream_t s, bool use_graph);
private:
  void choose(hipStream_t s);              // launch grid: autotuned or default
  void launch_grid(int grid, hipStream_t s);
  void launch(hipStream_t s);
  int type_, prec_, interp_;
  bool apod_ = false;
  int32_t n_probs_ = 0;
  int64_t n_items_ = 0;
  std::vector<int> grids_;                 // candidate grids, default first
  int pick_ = -1;
  DasArgs *d_args_ = nullptr;
  int64_t *d_begin_ = nullptr;
  void *d_delay_ = nullptr;
  unsigned long long *d_counter_ = nullptr;
  hipGraphExec_t graph_ = nullptr;
};

// The allocation-free adjoint's workspace: this device's copy of a fixed
// device array (its address is looked up once per device) and its size.
double *fixed_workspace(size_t *bytes);

// With a gate, a long launch runs as chunks with cancellation tests between
// them (ChunkGate); without one it is a single launch.
void launch_das_forward(const DasArgs &a, int type, int prec, int interp, bool apod,
                        hipStream_t s, ChunkGate *gate = nullptr);
void launch_das_adjoint(const DasArgs &a, int type, int prec, int interp, bool apod,
                        bool want_gy, hipStream_t s, ChunkGate *gate = nullptr);

} // namespace dream
