// dream_oracle.cc -- CPU oracle for the DREAM hot paths.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain C++ on the CPU,
// the arithmetic of the reference's FP64/FP32 CPU paths so that the HIP
// kernels in dream_amd/csrc can be checked against it.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
// product library never links, loads or falls back to it.
//
// Reference files restated (paths relative to the DREAM source tree):
//   SIR  dreamrect_serial          src/dreamrect.cc:274-335
//        Rect::dreamrect (threads) src/dreamrect.cc:79-272
//        ArrRect driver            src/dream_arr_rect.cc:90-244, 310-433
//        array helpers             src/arr_functions.cc:60-262
//        error mapping             src/dream_error.cc:40-76
//        dreamcirc / dream_arr_circ src/dreamcirc.cc:286-352, src/dream_arr_circ.cc
//        dream_arr_annu            src/dream_arr_annu.cc:186-254, 361-505, 514-634
//        dreamcylind / dream_arr_cylind src/dreamcylind.cc:275-353, 442-517,
//                                  src/dream_arr_cylind.cc:197-250
//   DAS  DAS<T>::das / smp_das     extra_src/das.cc:103-298
//        saft/tfm/rca kernels      extra_src/das.cc:306-626
// plus this project's own extensions (linear interpolation, apodization and
// the adjoint of both), for which there is no reference counterpart.
//
// Pinning (what checks each part independently of this restatement):
//   * DAS (das_f64/das_f32) and the array helpers: bit-for-bit against the
//     reference's own das.cc / arr_functions.cc compiled from source into
//     oracle/_ref (tests/test_oracle_das.py, tests/test_oracle_arr_functions.py,
//     tests/golden/*_ref.npz).
//   * dreamrect / dream_arr_rect: the reference's SIR sources need FFTW
//     (absent from this image), so they cannot be compiled here.  The SIR
//     restatement is pinned by closed forms and invariants instead
//     (tests/test_oracle_sir.py): the on-axis plateau h = cp, the
//     time-integral against 2-D quadrature of ds/(2 pi r), first-arrival
//     index, mirror symmetries, delay shift, element superposition and the
//     error-code mapping.  The last-bit operation order of the index
//     expression rint(((r*1000)/cp - delay)/dt) is fixed by reading
//     src/dreamrect.cc:303-311 and is not executable-pinned (no compiled
//     reference): "parity partially pinned" for that one detail.
//   * dream_arr_annu (no golden vectors in the reference): pinned by reduction
//     to the disc sum (one radius = dreamcirc one sample later; a ring =
//     outer disc - inner disc), the analytic on-axis ring SIR (plateau v*cp
//     between the inner and outer arrival times, window placement under 'ud'
//     delays and weights), alignment of the ring pulses at the focal depth
//     under x focusing, window values of the apodizations and the error-code
//     rule (tests/test_oracle_annu.py).
//   * Extensions (linear interpolation, apodization, adjoint): pinned by the
//     dot-product identity <A y, g> = <y, A^T g>, central finite
//     differences and reduction to the reference path (linear -> nearest at
//     integer sample times, apodization of ones -> no apodization).
//
// Build: oracle/Makefile (g++ -O3 -ffp-contract=off, no -march flags, the
// same as the reference's Release build).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

typedef int64_t idx_t;

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

// SIRError values (include/dream.h:42-47) and ErrorLevel values (:49-55).
enum { E_NONE = 0, E_IGNORE = 1, E_WARN = 2, E_OOB = 3 };
enum { L_NONE = 0, L_IGNORE = 1, L_WARN = 2, L_STOP = 3, L_PSTOP = 4 };

// DelayType (:57-60), FocusMet (:62-71), SteerMet (:73-78), ApodMet (:80-89),
// DASType (:97-102) -- same numbering as the reference enums.
enum { D_SINGLE = 0, D_MULTIPLE = 1 };
enum { F_NONE = 0, F_X, F_NEGX, F_Y, F_NEGY, F_XY, F_X_Y, F_UD };
enum { S_NONE = 0, S_X, S_Y, S_XY };
enum { A_TRIANGLE = 0, A_GAUSS, A_RAISED, A_HANN, A_HAMMING, A_SIMPLY, A_CLAMPED, A_UD };
enum { DAS_SAFT = 0, DAS_TFM, DAS_RCA_COLTX, DAS_RCA_ROWTX };

// src/dream_error.cc:40-76 -- error level -> SIRError code.
static int oob_code(int level)
{
  if (level == L_IGNORE) return E_IGNORE;
  if (level == L_WARN) return E_WARN;
  if (level == L_STOP || level == L_PSTOP) return E_OOB;
  return E_NONE;
}

// Thread count of the reference drivers: hardware_concurrency, capped by
// DREAM_NUM_THREADS and by the number of observation points, at least 1
// (src/dream_arr_rect.cc:337-354, extra_src/das.cc:220-237).  A positive
// `forced` overrides the hardware count (still capped by No).
static idx_t thread_count(idx_t No, int forced)
{
  idx_t n = (idx_t) std::thread::hardware_concurrency();
  if (forced > 0) {
    n = forced;
  } else if (const char *env = std::getenv("DREAM_NUM_THREADS")) {
    idx_t cap = (idx_t) std::stoul(env);
    if (cap < n) n = cap;
  }
  if (n > No) n = No;
  if (n < 1) n = 1;
  return n;
}

// Static contiguous block partition [t*No/T, (t+1)*No/T) and the
// reference's error aggregation: the code of the highest-numbered thread
// that saw a non-none status wins, and a single-thread run reports none
// because its status is never read back (src/dream_arr_rect.cc:405-425,
// extra_src/das.cc:270-290).
template <class Body>
static int run_partitioned(idx_t No, int forced_threads, Body body)
{
  idx_t T = thread_count(No, forced_threads);
  std::vector<int> status((size_t) T, E_NONE);
  if (T == 1) {
    body((idx_t) 0, No, status[0]);
    return E_NONE;
  }
  std::vector<std::thread> pool;
  pool.reserve((size_t) T);
  for (idx_t t = 0; t < T; t++) {
    idx_t start = t * No / T, stop = (t + 1) * No / T;
    pool.emplace_back([&, start, stop, t]() { body(start, stop, status[(size_t) t]); });
  }
  int err = E_NONE;
  for (idx_t t = 0; t < T; t++) {
    pool[(size_t) t].join();
    if (status[(size_t) t] != E_NONE) err = status[(size_t) t];
  }
  return err;
}

// ---------------------------------------------------------------------------
// Attenuation, Attenuation::att (src/attenuation.cc:72-140).  With alpha >
// DBL_EPSILON [dB/(cm MHz)] every in-range contribution adds ai times the
// lossy Green's function of its distance r [mm] and sample it instead of ai
// at it alone: the spectrum X[0] = 1, for k = 1 .. n/2
//   X[k] = exp(-r a w/(2 pi)) * exp(i(-w r a b/pi^2 - t w)),
//   w = k (2 pi/n) Fs, b = log(1/(a1 w)), a1 = dt 0.95/pi, t = it dt,
//   a = alpha/(8.686 * 10^4) [Np/(m Hz)], r and dt in SI units,
// and X[n-k] = conj(X[k]) (which for even n leaves X[n/2] conjugated), then
// FFTW's complex-to-real inverse transform of length n (only X[0 .. n/2] is
// read; the imaginary parts of X[0] and X[n/2] drop out) divided by n.  The
// inverse transform is evaluated here as the direct sum.
// ---------------------------------------------------------------------------

static double g_alpha = 0.0;

static inline bool attenuated() { return g_alpha > 2.220446049250313e-16; }

static void att_add(double r, idx_t it, double *h, double ai, idx_t n, double dt_us)
{
  thread_local std::vector<double> re, im, cs, sn, x;
  thread_local idx_t tw_n = -1;
  const idx_t half = n / 2;
  re.assign((size_t) half + 1, 0.0);
  im.assign((size_t) half + 1, 0.0);
  x.assign((size_t) n, 0.0);
  if (tw_n != n) {
    cs.resize((size_t) n);
    sn.resize((size_t) n);
    for (idx_t m = 0; m < n; m++) {
      cs[(size_t) m] = std::cos(2.0 * M_PI * (double) m / (double) n);
      sn[(size_t) m] = std::sin(2.0 * M_PI * (double) m / (double) n);
    }
    tw_n = n;
  }
  const double alpha = g_alpha / (8.686 * 10000.0);
  const double dt = dt_us * 1.0e-6;
  const double Fs = 1 / dt;
  const double dw = 2.0 * M_PI / double(n);
  const double t = double(it) * dt;
  r *= 1.0e-3;
  const double a1 = dt * 0.95 / M_PI;
  re[0] = 1.0;
  for (idx_t k = 1; k < n / 2 + 1; k++) {
    const double w = (double(k) * dw) * Fs;
    const double x1 = std::exp(-r * alpha * w / (2.0 * M_PI));
    const double b = std::log(1.0 / (a1 * w));
    const double arg = -w * r * alpha / (M_PI * M_PI) * b - t * w;
    re[(size_t) k] = x1 * std::cos(arg);
    im[(size_t) k] = (2 * k == n) ? -(x1 * std::sin(arg)) : x1 * std::sin(arg);
  }
  // Complex-to-real inverse DFT: x[j] = X0 + 2 sum_k Re(X[k] e^{+2 pi i jk/n})
  // (+ X[n/2] (-1)^j for even n), then / n.
  const idx_t kmax = (n % 2 == 0) ? half - 1 : half;
  for (idx_t j = 0; j < n; j++) {
    double acc = re[0];
    for (idx_t k = 1; k <= kmax; k++) {
      const idx_t m = (j * k) % n;
      acc += 2.0 * (re[(size_t) k] * cs[(size_t) m] - im[(size_t) k] * sn[(size_t) m]);
    }
    if (n % 2 == 0) acc += (j % 2 == 0) ? re[(size_t) half] : -re[(size_t) half];
    x[(size_t) j] = acc / double(n);
  }
  for (idx_t j = 0; j < n; j++) h[j] += ai * x[(size_t) j];
}

// Adds one in-range contribution: ai at sample it, or its attenuated pulse.
static inline void deposit(double *h, idx_t it, double ai, double r, idx_t nt, double dt)
{
  if (attenuated()) att_add(r, it, h, ai, nt, dt); else h[it] += ai;
}

extern "C" {

// Attenuation coefficient [dB/(cm MHz)] used by the SIR sums that follow
// (0: none).  Set it before a call; it is read by every thread of the call.
void dro_set_alpha(double alpha) { g_alpha = alpha; }

// ---------------------------------------------------------------------------
// Array helpers, src/arr_functions.cc:60-262.
// ---------------------------------------------------------------------------

// src/arr_functions.cc:60-84
void dro_max_dim_arr(const double *gx, const double *gy, idx_t L,
                     double *x_max, double *y_max, double *r_max)
{
  double xm = std::fabs(gx[0]);
  for (idx_t i = 1; i < L; i++) {
    double v = std::fabs(gx[i]);
    if (v > xm) xm = v;
  }
  double ym = std::fabs(gy[0]);
  for (idx_t i = 1; i < L; i++) {
    double v = std::fabs(gy[i]);
    if (v > ym) ym = v;
  }
  *x_max = xm;
  *y_max = ym;
  *r_max = std::sqrt(xm * xm + ym * ym);
}

// src/arr_functions.cc:94-148 (the enum values negx/negy fall through to
// the default branch there, giving 0).
double dro_focusing(int foc_met, double focal, double gx, double gy,
                    double x_max, double y_max, double r_max, double cp)
{
  double ext, own;
  switch (foc_met) {
  case F_X:
    ext = std::sqrt(x_max * x_max + focal * focal);
    own = std::sqrt(gx * gx + focal * focal);
    return (ext - own) * 1.0e3 / cp;
  case F_Y:
    ext = std::sqrt(y_max * y_max + focal * focal);
    own = std::sqrt(gy * gy + focal * focal);
    return (ext - own) * 1.0e3 / cp;
  case F_XY:
    ext = std::sqrt(r_max * r_max + focal * focal);
    own = std::sqrt(gx * gx + gy * gy + focal * focal);
    return (ext - own) * 1.0e3 / cp;
  case F_X_Y: {
    ext = std::sqrt(r_max * r_max + focal * focal);
    double ox = std::sqrt(gx * gx + focal * focal);
    double oy = std::sqrt(gy * gy + focal * focal);
    return (ext - (ox + oy)) * 1.0e3 / cp;
  }
  case F_UD:
    return focal;
  default:
    return 0.0;
  }
}

// src/arr_functions.cc:156-205
double dro_beamsteering(int steer_met, double theta, double phi, double gx, double gy,
                        double x_max, double y_max, double cp)
{
  const double deg2rad = M_PI / (double) 180.0;
  double sx, sy, ex, ey;
  switch (steer_met) {
  case S_X:
    sx = std::sin(theta * deg2rad);
    ex = x_max * sx;
    return (ex + gx * sx) * 1.0e3 / cp;
  case S_Y:
    sy = std::sin(phi * deg2rad);
    ey = y_max * sy;
    return (ey + gy * sy) * 1.0e3 / cp;
  case S_XY: {
    sx = std::sin(theta * deg2rad);
    ex = x_max * sx;
    double dx_ = (ex + gx * sx) * 1.0e3 / cp;
    sy = std::sin(phi * deg2rad);
    ey = y_max * sy;
    double dy_ = (ey + gy * sy) * 1.0e3 / cp;
    return dx_ + dy_;
  }
  default:
    return 0.0;
  }
}

// src/arr_functions.cc:215-262.  `prev` is returned unchanged for an
// unknown method (the reference leaves *weight untouched).
double dro_apodization(int apod_met, idx_t n, const double *apod_vec, double prev,
                       double gx, double gy, double r_max, double apod_par)
{
  double r = std::sqrt(gx * gx + gy * gy);
  switch (apod_met) {
  case A_UD:
    return apod_vec[n];
  case A_TRIANGLE:
    return 1.0 - std::fabs(r) / r_max;
  case A_GAUSS:
    return std::exp(-(apod_par * r * r) / (r_max * r_max));
  case A_RAISED:
    return apod_par + std::cos(r * M_PI / r_max);
  case A_HANN:
    return 0.5 * (1.0 - std::cos((r + r_max) * M_PI / r_max));
  case A_HAMMING: {
    double a0 = 25.0 / 46.0;
    return a0 - (1.0 - a0) * std::cos((r + r_max) * M_PI / r_max);
  }
  case A_SIMPLY:
    return 1.0 - r * r / (r_max * r_max);
  case A_CLAMPED:
    return (1.0 - r * r / (r_max * r_max)) * (1.0 - r * r / (r_max * r_max));
  default:
    return prev;
  }
}

// ---------------------------------------------------------------------------
// SIR: discrete-representation sum over a rectangle, src/dreamrect.cc:274-335.
// ---------------------------------------------------------------------------

// Sub-element centre coordinates of one side of the rectangle.  The
// reference walks `s = -w/2 + d/2; while (s <= w/2) { ...; s += d; }`
// (src/dreamrect.cc:286-331), so both the count and the coordinates are
// those of the floating-point accumulation, not of -w/2+(i+1/2)d.
idx_t dro_subgrid(double w, double d, double *out, idx_t cap)
{
  double smax = w / 2.0;
  double s = -w / 2.0 + d / 2.0;
  idx_t n = 0;
  while (s <= smax) {
    if (out && n < cap) out[n] = s;
    n++;
    s += d;
  }
  return n;
}

// One rectangle at observation point (xo,yo,zo) relative to its centre.
// Returns the status of this call: the code of the last out-of-bounds
// contribution (E_OOB stops at once), none when all landed inside.
// `oob` (optional) counts the out-of-bounds contributions.
static int rect_sum(double xo, double yo, double zo, double a, double b,
                    double dx, double dy, double dt, idx_t nt, double delay,
                    double v, double cp, double *h, int level, double weight,
                    long long *oob)
{
  int err = E_NONE;
  const double xmax = a / 2.0, ymax = b / 2.0;
  const double ds = dx * dy;
  const double rz = zo;
  for (double ys = -b / 2.0 + dy / 2.0; ys <= ymax; ys += dy) {
    const double ry = yo - ys;
    for (double xs = -a / 2.0 + dx / 2.0; xs <= xmax; xs += dx) {
      const double rx = xo - xs;
      const double r = std::sqrt(rx * rx + ry * ry + rz * rz);
      double ai = weight * v * ds / (2 * M_PI * r);
      ai /= dt;
      ai *= 1000.0;
      const double t = r * 1000.0 / cp;
      const double u = std::rint((t - delay) / dt);
      const bool inside = (u >= 0.0) && (u < (double) nt);
      if (inside) {
        deposit(h, (idx_t) u, ai, r, nt, dt);
      } else {
        if (oob) (*oob)++;
        err = oob_code(level);
        if (err == E_OOB) return err;
      }
    }
  }
  return err;
}

// Per-contribution dump for one (observation point, element) pair: the
// sample index it = rint((t - delay)/dt) (before the bounds test) and the
// amplitude ai, in the reference's (ys outer, xs inner) order.  Returns the
// number of contributions; writes at most `cap`.
idx_t dro_rect_contributions(double xo, double yo, double zo, double a, double b,
                             double dx, double dy, double dt, double delay,
                             double v, double cp, double weight,
                             int64_t *it_out, double *ai_out, idx_t cap)
{
  const double xmax = a / 2.0, ymax = b / 2.0;
  const double ds = dx * dy;
  idx_t n = 0;
  for (double ys = -b / 2.0 + dy / 2.0; ys <= ymax; ys += dy) {
    const double ry = yo - ys;
    for (double xs = -a / 2.0 + dx / 2.0; xs <= xmax; xs += dx) {
      const double rx = xo - xs;
      const double r = std::sqrt(rx * rx + ry * ry + zo * zo);
      double ai = weight * v * ds / (2 * M_PI * r);
      ai /= dt;
      ai *= 1000.0;
      const double t = r * 1000.0 / cp;
      if (n < cap) {
        if (it_out) it_out[n] = (int64_t) std::rint((t - delay) / dt);
        if (ai_out) ai_out[n] = ai;
      }
      n++;
    }
  }
  return n;
}

// Rect::dreamrect without attenuation (src/dreamrect.cc:79-272).  h is
// nt x No column-major and is accumulated into (the caller zeroes it, as
// SIRData::clear does in the bindings).
int dro_dreamrect(const double *Ro, idx_t No, double a, double b, double dx, double dy,
                  double dt, idx_t nt, int delay_type, const double *delay,
                  double v, double cp, double *h, int level, int nthreads,
                  long long *oob_total)
{
  std::atomic<bool> running(true);
  std::atomic<long long> oob_sum(0);
  int err = run_partitioned(No, nthreads, [&](idx_t start, idx_t stop, int &status) {
    long long oob = 0;
    for (idx_t n = start; n < stop; n++) {
      const double dl = (delay_type == D_SINGLE) ? delay[0] : delay[n];
      int e = rect_sum(Ro[n], Ro[n + No], Ro[n + 2 * No], a, b, dx, dy, dt, nt, dl,
                       v, cp, &h[n * nt], level, 1.0, &oob);
      if (e != E_NONE) status = e;
      if (e == E_OOB) running = false;
      if (!running) break;
    }
    oob_sum += oob;
  });
  if (oob_total) *oob_total = oob_sum.load();
  return err;
}

// ArrRect::dream_arr_rect without attenuation (src/dream_arr_rect.cc:90-244,
// 310-433).  Per observation point the status is that of the LAST element's
// rectangle sum (the element loop overwrites it, :227-239); an E_OOB status
// stops the element loop and every thread.
int dro_dream_arr_rect(const double *Ro, idx_t No, double a, double b, double dx, double dy,
                       double dt, idx_t nt, int delay_type, const double *delay,
                       double v, double cp, idx_t L, const double *G,
                       int foc_met, const double *focal, int steer_met, double theta,
                       double phi, const double *apod, int do_apod, int apod_met,
                       double apod_par, double *h, int level, int nthreads,
                       long long *oob_total)
{
  const double *gx = G, *gy = G + L, *gz = G + 2 * L;
  std::atomic<bool> running(true);
  std::atomic<long long> oob_sum(0);
  int err = run_partitioned(No, nthreads, [&](idx_t start, idx_t stop, int &status) {
    long long oob = 0;
    for (idx_t n = start; n < stop; n++) {
      const double xo = Ro[n], yo = Ro[n + No], zo = Ro[n + 2 * No];
      const double dl = (delay_type == D_SINGLE) ? delay[0] : delay[n];
      double x_max, y_max, r_max;
      dro_max_dim_arr(gx, gy, L, &x_max, &y_max, &r_max);
      int e = E_NONE;
      for (idx_t el = 0; el < L; el++) {
        double foc = (foc_met == F_UD) ? focal[el]
                   : dro_focusing(foc_met, focal[0], gx[el], gy[el], x_max, y_max, r_max, cp);
        double steer = dro_beamsteering(steer_met, theta, phi, gx[el], gy[el], x_max, y_max, cp);
        double w = do_apod ? dro_apodization(apod_met, el, apod, 1.0, gx[el], gy[el], r_max, apod_par)
                           : 1.0;
        e = rect_sum(xo - gx[el], yo - gy[el], zo - gz[el], a, b, dx, dy, dt, nt,
                     dl - foc - steer, v, cp, &h[n * nt], level, w, &oob);
        if (e == E_OOB) break;
      }
      if (e != E_NONE) status = e;
      if (e == E_OOB) running = false;
      if (!running) break;
    }
    oob_sum += oob;
  });
  if (oob_total) *oob_total = oob_sum.load();
  return err;
}

// ---------------------------------------------------------------------------
// SIR: discrete-representation sum over a circular disc, src/dreamcirc.cc:286-352.
// ---------------------------------------------------------------------------

// Rows ys = -R + dy/2, += dy while ys <= R; in row ys the x walk runs from
// -sqrt(R*R - ys*ys) + dx/2 while xs <= sqrt(R*R - ys*ys).
static int circ_sum(double xo, double yo, double zo, double R, double dx, double dy,
                    double dt, idx_t nt, double delay, double v, double cp, double *h,
                    int level, double weight, long long *oob)
{
  int err = E_NONE;
  const double ds = dx * dy;
  const double ysmin = -R, ysmax = R;
  double ys = ysmin + dy / 2.0;
  while (ys <= ysmax) {
    const double rxs = std::sqrt(R * R - ys * ys);
    const double xsmin = -rxs, xsmax = rxs;
    const double ry = yo - ys;
    double xs = xsmin + dx / 2.0;
    while (xs <= xsmax) {
      const double rx = xo - xs;
      const double r = std::sqrt(rx * rx + ry * ry + zo * zo);
      double ai = weight * v * ds / (2 * M_PI * r);
      ai /= dt;
      ai *= 1000.0;
      const double t = r * 1000.0 / cp;
      const double u = std::rint((t - delay) / dt);
      const bool inside = (u >= 0.0) && (u < (double) nt);
      if (inside) {
        deposit(h, (idx_t) u, ai, r, nt, dt);
      } else {
        if (oob) (*oob)++;
        err = oob_code(level);
        if (err == E_OOB) return err;
      }
      xs += dx;
    }
    ys += dy;
  }
  return err;
}

// Circ::dreamcirc without attenuation (src/dreamcirc.cc:78-272).
int dro_dreamcirc(const double *Ro, idx_t No, double R, double dx, double dy, double dt,
                  idx_t nt, int delay_type, const double *delay, double v, double cp, double *h,
                  int level, int nthreads, long long *oob_total)
{
  std::atomic<bool> running(true);
  std::atomic<long long> oob_sum(0);
  int err = run_partitioned(No, nthreads, [&](idx_t start, idx_t stop, int &status) {
    long long oob = 0;
    for (idx_t n = start; n < stop; n++) {
      const double dl = (delay_type == D_SINGLE) ? delay[0] : delay[n];
      int e = circ_sum(Ro[n], Ro[n + No], Ro[n + 2 * No], R, dx, dy, dt, nt, dl, v, cp,
                       &h[n * nt], level, 1.0, &oob);
      if (e != E_NONE) status = e;
      if (e == E_OOB) running = false;
      if (!running) break;
    }
    oob_sum += oob;
  });
  if (oob_total) *oob_total = oob_sum.load();
  return err;
}

// ArrCirc::dream_arr_circ without attenuation (src/dream_arr_circ.cc:87-236,
// 300-420): the element loop of dream_arr_rect with the disc sum.
int dro_dream_arr_circ(const double *Ro, idx_t No, double R, double dx, double dy, double dt,
                       idx_t nt, int delay_type, const double *delay, double v, double cp,
                       idx_t L, const double *G, int foc_met, const double *focal,
                       int steer_met, double theta, double phi, const double *apod, int do_apod,
                       int apod_met, double apod_par, double *h, int level, int nthreads,
                       long long *oob_total)
{
  const double *gx = G, *gy = G + L, *gz = G + 2 * L;
  std::atomic<bool> running(true);
  std::atomic<long long> oob_sum(0);
  int err = run_partitioned(No, nthreads, [&](idx_t start, idx_t stop, int &status) {
    long long oob = 0;
    for (idx_t n = start; n < stop; n++) {
      const double xo = Ro[n], yo = Ro[n + No], zo = Ro[n + 2 * No];
      const double dl = (delay_type == D_SINGLE) ? delay[0] : delay[n];
      double x_max, y_max, r_max;
      dro_max_dim_arr(gx, gy, L, &x_max, &y_max, &r_max);
      int e = E_NONE;
      for (idx_t el = 0; el < L; el++) {
        double foc = (foc_met == F_UD) ? focal[el]
                   : dro_focusing(foc_met, focal[0], gx[el], gy[el], x_max, y_max, r_max, cp);
        double steer = dro_beamsteering(steer_met, theta, phi, gx[el], gy[el], x_max, y_max, cp);
        double w = do_apod ? dro_apodization(apod_met, el, apod, 1.0, gx[el], gy[el], r_max, apod_par)
                           : 1.0;
        e = circ_sum(xo - gx[el], yo - gy[el], zo - gz[el], R, dx, dy, dt, nt,
                     dl - foc - steer, v, cp, &h[n * nt], level, w, &oob);
        if (e == E_OOB) break;
      }
      if (e != E_NONE) status = e;
      if (e == E_OOB) running = false;
      if (!running) break;
    }
    oob_sum += oob;
  });
  if (oob_total) *oob_total = oob_sum.load();
  return err;
}

// ---------------------------------------------------------------------------
// Cylindrically curved elements, src/dreamcylind.cc:275-353 (DR sum) with
// the surface point of cylind_f / cylind_d (:442-517; their angle checks are
// disabled in the reference, so du = 1).
// ---------------------------------------------------------------------------

// Rows are angles phi = -phi_max + dphi/2, += dphi (dphi = asin(dy/|Rc|))
// while phi <= phi_max; a row sits at ys = |Rc| sin(phi) and depth zs = |Rc|
// - sqrt(Rc^2 - ys^2) below the vertex plane (z - zs focused, z + zs
// defocused); the x walk is the rectangle's.  b > 2|Rc| makes the angles
// NaN: no rows, a zero response.
static int cylind_sum(double xo, double yo, double zo, double a, double b, double Rc,
                      double dx, double dy, double dt, idx_t nt, double delay, double v,
                      double cp, double *h, int level, double weight, long long *oob)
{
  int err = E_NONE;
  const double R = std::fabs(Rc);
  const double hb = b / 2.0;
  const double z_R = R - std::sqrt(Rc * Rc - hb * hb);
  const double span = 2.0 * std::atan(hb / (R - z_R));
  const double dphi = std::asin(dy / R);
  const double ds = R * dx * dphi;
  double ph = -span / 2.0 + dphi / 2.0;
  double ys = R * std::sin(ph);
  while (ph <= span / 2.0) {
    const double zs = R - std::sqrt(R * R - ys * ys);
    const double rz = (Rc >= 0.0) ? zo - zs : zo + zs;
    double xs = -a / 2.0 + dx / 2.0;
    while (xs <= a / 2.0) {
      const double rx = xo - xs, ry = yo - ys;
      const double r = std::sqrt(rx * rx + ry * ry + rz * rz);
      double ai = weight * v * ds * 1.0 / (2.0 * M_PI * r);
      ai /= dt;
      ai *= 1.0e3;
      const double t = r * 1.0e3 / cp;
      const double u = std::rint((t - delay) / dt);
      if (u >= 0.0 && u < (double) nt) {
        deposit(h, (idx_t) u, ai, r, nt, dt);
      } else {
        if (oob) (*oob)++;
        err = oob_code(level);
        if (err == E_OOB) return err;
      }
      xs += dx;
    }
    ph += dphi;
    ys = R * std::sin(ph);
  }
  return err;
}

// Cylind::dreamcylind without attenuation (src/dreamcylind.cc:78-273).
int dro_dreamcylind(const double *Ro, idx_t No, double a, double b, double Rc, double dx,
                    double dy, double dt, idx_t nt, int delay_type, const double *delay,
                    double v, double cp, double *h, int level, int nthreads,
                    long long *oob_total)
{
  std::atomic<bool> running(true);
  std::atomic<long long> oob_sum(0);
  int err = run_partitioned(No, nthreads, [&](idx_t start, idx_t stop, int &status) {
    long long oob = 0;
    for (idx_t n = start; n < stop; n++) {
      const double dl = (delay_type == D_SINGLE) ? delay[0] : delay[n];
      int e = cylind_sum(Ro[n], Ro[n + No], Ro[n + 2 * No], a, b, Rc, dx, dy, dt, nt, dl, v, cp,
                         &h[n * nt], level, 1.0, &oob);
      if (e != E_NONE) status = e;
      if (e == E_OOB) running = false;
      if (!running) break;
    }
    oob_sum += oob;
  });
  if (oob_total) *oob_total = oob_sum.load();
  return err;
}

// ArrCylind::dream_arr_cylind without attenuation (src/dream_arr_cylind.cc:
// 90-195, 197-250, 313-437): the element loop of dream_arr_rect with the
// curved-element sum.
int dro_dream_arr_cylind(const double *Ro, idx_t No, double a, double b, double Rc, double dx,
                         double dy, double dt, idx_t nt, int delay_type, const double *delay,
                         double v, double cp, idx_t L, const double *G, int foc_met,
                         const double *focal, int steer_met, double theta, double phi,
                         const double *apod, int do_apod, int apod_met, double apod_par,
                         double *h, int level, int nthreads, long long *oob_total)
{
  const double *gx = G, *gy = G + L, *gz = G + 2 * L;
  std::atomic<bool> running(true);
  std::atomic<long long> oob_sum(0);
  int err = run_partitioned(No, nthreads, [&](idx_t start, idx_t stop, int &status) {
    long long oob = 0;
    for (idx_t n = start; n < stop; n++) {
      const double xo = Ro[n], yo = Ro[n + No], zo = Ro[n + 2 * No];
      const double dl = (delay_type == D_SINGLE) ? delay[0] : delay[n];
      double x_max, y_max, r_max;
      dro_max_dim_arr(gx, gy, L, &x_max, &y_max, &r_max);
      int e = E_NONE;
      for (idx_t el = 0; el < L; el++) {
        double foc = (foc_met == F_UD) ? focal[el]
                   : dro_focusing(foc_met, focal[0], gx[el], gy[el], x_max, y_max, r_max, cp);
        double steer = dro_beamsteering(steer_met, theta, phi, gx[el], gy[el], x_max, y_max, cp);
        double w = do_apod ? dro_apodization(apod_met, el, apod, 1.0, gx[el], gy[el], r_max, apod_par)
                           : 1.0;
        e = cylind_sum(xo - gx[el], yo - gy[el], zo - gz[el], a, b, Rc, dx, dy, dt, nt,
                       dl - foc - steer, v, cp, &h[n * nt], level, w, &oob);
        if (e == E_OOB) break;
      }
      if (e != E_NONE) status = e;
      if (e == E_OOB) running = false;
      if (!running) break;
    }
    oob_sum += oob;
  });
  if (oob_total) *oob_total = oob_sum.load();
  return err;
}

// ---------------------------------------------------------------------------
// Annular array, src/dream_arr_annu.cc.  G holds radii: Gr[0] is the centre
// disc, then an (inner, outer) pair per ring; ring n is the outer disc 2n
// minus the inner disc 2n-1, and there are (num_radii+1)/2 rings.
// ---------------------------------------------------------------------------

// Ring radius used by focusing and apodization (:361-368): 0 for the centre
// disc, the mean of the pair's radii for the others.
void dro_annu_ring_radii(const double *Gr, idx_t num_radii, double *ring_r)
{
  const idx_t rings = (num_radii + 1) / 2;
  if (rings > 0) ring_r[0] = 0.0;
  for (idx_t n = 1; n < rings && 2 * n < num_radii; n++) ring_r[n] = (Gr[2 * n] + Gr[2 * n - 1]) / 2.0;
}

// Focusing delay in us (:376-405): the four geometric methods share one
// spherical law about the ring radius; 'ud' returns the given delay.
double dro_annu_focusing(int foc_met, double focal, double ring_r, double ring_r_max, double cp)
{
  if (foc_met == F_X || foc_met == F_Y || foc_met == F_XY || foc_met == F_X_Y) {
    const double far = std::sqrt(ring_r_max * ring_r_max + focal * focal);
    const double near = std::sqrt(ring_r * ring_r + focal * focal);
    return (far - near) * 1000 / cp;
  }
  if (foc_met == F_UD) return focal;
  return 0.0;
}

// Apodization weight of ring n (:413-447); Hann and Hamming have no annular
// form and give 1.
double dro_annu_apodization(int apod_met, idx_t n, const double *apod, double ring_r,
                            double ring_r_max, double apod_par)
{
  const double q = ring_r * ring_r / (ring_r_max * ring_r_max);
  switch (apod_met) {
  case A_UD: return apod[n];
  case A_TRIANGLE: return 1.0 - std::fabs(ring_r) / ring_r_max;
  case A_GAUSS: return std::exp(-(apod_par * ring_r * ring_r) / (ring_r_max * ring_r_max));
  case A_RAISED: return apod_par + std::cos(ring_r * M_PI / ring_r_max);
  case A_SIMPLY: return 1.0 - q;
  case A_CLAMPED: return (1.0 - q) * (1.0 - q);
  default: return 1.0;
  }
}

// ArrAnnu::dream_arr_annu without attenuation (:186-254 per point, :514-634
// threads).  Per point: every disc response through the disc sum (disc 0 is
// skipped, left zero, when Gr[0] <= DBL_EPSILON); the point's status is the
// last disc's (an E_OOB stops at once).  Each ring response is then moved
// later by rint(foc/dt) + 1 samples, cut at nt, and added with its weight
// (:459-481).  A shift that is not a number adds nothing (the reference's
// integer conversion of it is undefined).
int dro_dream_arr_annu(const double *Ro, idx_t No, double dx, double dy, double dt, idx_t nt,
                       int delay_type, const double *delay, double v, double cp,
                       idx_t num_radii, const double *Gr, int foc_met, const double *focal,
                       const double *apod, int do_apod, int apod_met, double apod_par,
                       double *h, int level, int nthreads, long long *oob_total)
{
  const idx_t rings = (num_radii + 1) / 2;
  if (rings < 1) {
    std::fill(h, h + No * nt, 0.0);
    return E_NONE;
  }
  std::vector<double> ring_r((size_t) rings), weight((size_t) rings), shift((size_t) rings);
  dro_annu_ring_radii(Gr, num_radii, ring_r.data());
  const double r_max = ring_r[(size_t) rings - 1];
  for (idx_t n = 0; n < rings; n++) {
    const double f = dro_annu_focusing(foc_met, foc_met == F_UD ? focal[n] : focal[0],
                                       ring_r[(size_t) n], r_max, cp);
    weight[(size_t) n] = do_apod ? dro_annu_apodization(apod_met, n, apod, ring_r[(size_t) n],
                                                        r_max, apod_par)
                                 : 1.0;
    shift[(size_t) n] = std::rint(f / dt) + 1.0;
  }
  std::atomic<bool> running(true);
  std::atomic<long long> oob_sum(0);
  int err = run_partitioned(No, nthreads, [&](idx_t start, idx_t stop, int &status) {
    long long oob = 0;
    std::vector<double> disc((size_t) (num_radii * nt));
    for (idx_t p = start; p < stop; p++) {
      const double dl = (delay_type == D_SINGLE) ? delay[0] : delay[p];
      std::fill(disc.begin(), disc.end(), 0.0);
      int e = E_NONE;
      for (idx_t k = 0; k < num_radii; k++) {
        if (k == 0 && Gr[0] <= 2.220446049250313e-16) continue;
        e = circ_sum(Ro[p], Ro[p + No], Ro[p + 2 * No], Gr[k], dx, dy, dt, nt, dl, v, cp,
                     &disc[(size_t) (k * nt)], level, 1.0, &oob);
        if (e == E_OOB) break;
      }
      double *hp = &h[p * nt];
      for (idx_t i = 0; i < nt; i++) hp[i] = 0.0;
      if (e != E_OOB) {
        for (idx_t n = 0; n < rings; n++) {
          const double *outer = &disc[(size_t) (2 * n * nt)];
          const double *inner = n ? &disc[(size_t) ((2 * n - 1) * nt)] : nullptr;
          const double s = shift[(size_t) n];
          for (idx_t i = 0; i < nt; i++) {
            const double j = (double) i - s;   // source sample of output sample i
            double src = 0.0;
            if (s < (double) nt && j >= 0.0 && j < (double) nt) {
              const idx_t jj = (idx_t) j;
              src = inner ? outer[jj] - inner[jj] : outer[jj];
            }
            hp[i] += weight[(size_t) n] * src;
          }
        }
      }
      if (e != E_NONE) status = e;
      if (e == E_OOB) running = false;
      if (!running) break;
    }
    oob_sum += oob;
  });
  if (oob_total) *oob_total = oob_sum.load();
  return err;
}

} // extern "C"

// ---------------------------------------------------------------------------
// DAS, extra_src/das.cc:306-626.  T is the arithmetic type (float or double);
// the three constants are double expressions rounded to T, as the
// reference's `T Fs_khz = (1.0/dt)*1000.0;` lines are (:316-318).
// ---------------------------------------------------------------------------

template <class T>
struct DasConst {
  T fs_khz, inv_cp, delay_ms;
  DasConst(T dt, T cp, T delay)
    : fs_khz((T) ((1.0 / (double) dt) * 1000.0)),
      inv_cp((T) (1.0 / (double) cp)),
      delay_ms((T) ((double) delay / 1000.0)) {}
};

// Truncation toward zero of the sample position, as dream_idx_type(x) does;
// non-finite or huge values are reported out of range.
template <class T>
static inline bool trunc_index(T x, idx_t K, idx_t *k)
{
  if (!(x > (T) -9.0e18 && x < (T) 9.0e18)) return false;
  idx_t kk = (idx_t) x;
  *k = kk;
  return kk >= 0 && kk < K;
}

template <class T>
static inline T dist3(T dx, T dy, T dz) { return std::sqrt(dx * dx + dy * dy + dz * dz); }

// Stripe-clamped lateral distance used by the RCA kernels
// (extra_src/das.cc:477-485, 494-502).
template <class T>
static inline T stripe_gap(T p, T lo, T hi)
{
  if (p >= lo && p <= hi) return (T) 0.0;
  return (p < lo) ? (lo - p) : (hi - p);
}

template <class T>
static int das_point(int type, const T *Y, idx_t K, const T *Gt, idx_t Lt, const T *Gr,
                     idx_t Lr, T xo, T yo, T zo, T dt, T delay, T cp, T &im, int level)
{
  int err = E_NONE;
  const DasConst<T> c(dt, cp, delay);
  T acc = (T) 0.0;
  const T *yp = Y;
  idx_t k;
  auto visit = [&](T t_dp) -> bool {   // returns false to bail out (stop)
    if (trunc_index<T>(t_dp * c.fs_khz, K, &k)) {
      acc += yp[k];
    } else {
      err = oob_code(level);
      if (err == E_OOB) return false;
    }
    yp += K;
    return true;
  };
  if (type == DAS_SAFT) {                                   // :306-357
    for (idx_t n = 0; n < Lt; n++) {
      T t = dist3<T>(Gt[n] - xo, Gt[n + Lt] - yo, Gt[n + 2 * Lt] - zo) * c.inv_cp;
      t += c.delay_ms;
      T t_dp = (T) (2.0 * t);
      if (!visit(t_dp)) { im = acc; return err; }
    }
  } else if (type == DAS_TFM) {                             // :365-427
    for (idx_t i = 0; i < Lt; i++) {
      T tt = dist3<T>(Gt[i] - xo, Gt[i + Lt] - yo, Gt[i + 2 * Lt] - zo) * c.inv_cp;
      tt += c.delay_ms;
      for (idx_t j = 0; j < Lr; j++) {
        T tr = dist3<T>(Gr[j] - xo, Gr[j + Lr] - yo, Gr[j + 2 * Lr] - zo) * c.inv_cp;
        if (!visit(tt + tr)) { im = acc; return err; }
      }
    }
  } else {
    // RCA: Gt = column electrodes (along x), Gr = row electrodes (along y).
    // The stripe extents are read as the reference reads them (:464-467,
    // :560-563): rows' y at offsets Lc and 2 Lc - 1, columns' x at 0 and
    // Lr - 1 -- the first/last y of the rows and x of the columns for square
    // arrays; for non-square ones the same (other) entries the reference uses.
    const T *Gc = Gt, *Gw = Gr;
    const idx_t Lc = Lt, Lw = Lr;
    const T y_lo = Gw[0 + Lc], y_hi = Gw[Lc - 1 + Lc];
    const T x_lo = Gc[0], x_hi = Gc[Lw - 1];
    const T gy_c = stripe_gap<T>(yo, y_lo, y_hi);
    const T gx_r = stripe_gap<T>(xo, x_lo, x_hi);
    if (type == DAS_RCA_COLTX) {                            // :437-529
      for (idx_t ic = 0; ic < Lc; ic++) {
        T tc = dist3<T>(Gc[ic] - xo, gy_c, Gc[ic + 2 * Lc] - zo) * c.inv_cp;
        tc += c.delay_ms;
        for (idx_t ir = 0; ir < Lw; ir++) {
          T tr = dist3<T>(gx_r, Gw[ir + Lw] - yo, Gw[ir + 2 * Lw] - zo) * c.inv_cp;
          if (!visit(tc + tr)) { im = acc; return err; }
        }
      }
    } else {                                                // :533-626
      for (idx_t ir = 0; ir < Lw; ir++) {
        T tr = dist3<T>(gx_r, Gw[ir + Lw] - yo, Gw[ir + 2 * Lw] - zo) * c.inv_cp;
        for (idx_t ic = 0; ic < Lc; ic++) {
          T tc = dist3<T>(Gc[ic] - xo, gy_c, Gc[ic + 2 * Lc] - zo) * c.inv_cp;
          tc += c.delay_ms;
          if (!visit(tc + tr)) { im = acc; return err; }
        }
      }
    }
  }
  im = acc;
  return err;
}

// DAS<T>::das (extra_src/das.cc:199-298): forces SAFT when Lr == 0, threads
// over pixels, per-pixel delay when delay_type is multiple.
template <class T>
static int das_run(int type, const T *Y, idx_t K, const T *Ro, idx_t No, const T *Gt,
                   idx_t Lt, const T *Gr, idx_t Lr, T dt, int delay_type, const T *delay,
                   T cp, T *Im, int level, int nthreads)
{
  if (Lr == 0) type = DAS_SAFT;
  std::atomic<bool> running(true);
  return run_partitioned(No, nthreads, [&](idx_t start, idx_t stop, int &status) {
    for (idx_t no = start; no < stop; no++) {
      T dl = (delay_type == D_SINGLE) ? delay[0] : delay[no];
      int e = das_point<T>(type, Y, K, Gt, Lt, Gr, Lr, Ro[no], Ro[no + No], Ro[no + 2 * No],
                           dt, dl, cp, Im[no], level);
      if (e != E_NONE) status = e;
      if (e == E_OOB) running = false;
      if (!running) break;
    }
  });
}

extern "C" {

int dro_das_f64(int type, const double *Y, idx_t K, const double *Ro, idx_t No,
                const double *Gt, idx_t Lt, const double *Gr, idx_t Lr, double dt,
                int delay_type, const double *delay, double cp, double *Im, int level,
                int nthreads)
{
  return das_run<double>(type, Y, K, Ro, No, Gt, Lt, Gr, Lr, dt, delay_type, delay, cp, Im,
                         level, nthreads);
}

int dro_das_f32(int type, const float *Y, idx_t K, const float *Ro, idx_t No,
                const float *Gt, idx_t Lt, const float *Gr, idx_t Lr, float dt,
                int delay_type, const float *delay, float cp, float *Im, int level,
                int nthreads)
{
  return das_run<float>(type, Y, K, Ro, No, Gt, Lt, Gr, Lr, dt, delay_type, delay, cp, Im,
                        level, nthreads);
}

} // extern "C"

// ---------------------------------------------------------------------------
// Extension: DAS with linear interpolation and apodization, and its adjoint
// (no reference counterpart).  FP64 throughout.
//
//   Im[p] = sum_{i,j} w_i v_j * s_ij(u_ij)
//   u_ij  = (t_i + t_j) * Fs_khz       (SAFT: u = 2 (t + delay_ms) Fs_khz)
//   t_i   = dist_i * (1/cp) + delay_ms (transmit leg; receive leg has no delay)
//   interp 0: s = Y[trunc(u)] when 0 <= trunc(u) < K (the reference rule)
//   interp 1: k = floor(u), f = u - k,
//             s = (1-f) Yz[k] + f Yz[k+1], Yz = Y inside [0,K), 0 outside.
// SAFT uses one weight per A-scan (w), TFM/RCA the product of the transmit
// weight w_i and the receive weight v_j.
// ---------------------------------------------------------------------------

namespace {

struct Geo {
  int type;
  const double *Gt, *Gr;
  idx_t Lt, Lr;
  double fs, cp, delay_ms, inv_cp;
};

// Leg distances for pixel (xo,yo,zo): transmit legs dt_[i], receive legs
// dr_[j] (SAFT: only dt_).
static void legs(const Geo &g, double xo, double yo, double zo, double *dtx, double *drx)
{
  if (g.type == DAS_SAFT || g.type == DAS_TFM) {
    for (idx_t i = 0; i < g.Lt; i++)
      dtx[i] = dist3<double>(g.Gt[i] - xo, g.Gt[i + g.Lt] - yo, g.Gt[i + 2 * g.Lt] - zo);
    if (g.type == DAS_TFM)
      for (idx_t j = 0; j < g.Lr; j++)
        drx[j] = dist3<double>(g.Gr[j] - xo, g.Gr[j + g.Lr] - yo, g.Gr[j + 2 * g.Lr] - zo);
    return;
  }
  const double y_lo = g.Gr[g.Lt], y_hi = g.Gr[g.Lt - 1 + g.Lt];
  const double x_lo = g.Gt[0], x_hi = g.Gt[g.Lr - 1];
  const double gy_c = stripe_gap<double>(yo, y_lo, y_hi);
  const double gx_r = stripe_gap<double>(xo, x_lo, x_hi);
  for (idx_t i = 0; i < g.Lt; i++)
    dtx[i] = dist3<double>(g.Gt[i] - xo, gy_c, g.Gt[i + 2 * g.Lt] - zo);
  for (idx_t j = 0; j < g.Lr; j++)
    drx[j] = dist3<double>(gx_r, g.Gr[j + g.Lr] - yo, g.Gr[j + 2 * g.Lr] - zo);
}

// A-scan index of the pair (i = transmit, j = receive).
static inline idx_t ascan(const Geo &g, idx_t i, idx_t j)
{
  if (g.type == DAS_SAFT) return i;
  if (g.type == DAS_RCA_ROWTX) return j * g.Lt + i;   // rows outer, A-scan r*Lc + c
  return i * g.Lr + j;                                 // tfm, rca_coltx
}

static inline double yz(const double *y, idx_t K, idx_t k) { return (k >= 0 && k < K) ? y[k] : 0.0; }

} // namespace

extern "C" {

// Forward of the extension.  apod_t (length Lt) and apod_r (length Lr, unused
// for SAFT) may be null for all-ones.  Returns 1 if any pair's nearest-lower
// sample index fell outside [0,K), else 0.
int dro_das_ext_f64(int type, int interp, const double *Y, idx_t K, const double *Ro, idx_t No,
                    const double *Gt, idx_t Lt, const double *Gr, idx_t Lr, double dt,
                    double delay, double cp, const double *apod_t, const double *apod_r,
                    double *Im, int nthreads)
{
  if (Lr == 0) type = DAS_SAFT;
  Geo g{type, Gt, Gr, Lt, Lr, (1.0 / dt) * 1000.0, cp, delay / 1000.0, 1.0 / cp};
  std::atomic<int> any_oob(0);
  idx_t T = thread_count(No, nthreads);
  std::vector<std::thread> pool;
  for (idx_t th = 0; th < T; th++) {
    pool.emplace_back([&, th]() {
      std::vector<double> dtx((size_t) Lt), drx((size_t) (Lr > 0 ? Lr : 1));
      int oob = 0;
      for (idx_t p = th * No / T; p < (th + 1) * No / T; p++) {
        legs(g, Ro[p], Ro[p + No], Ro[p + 2 * No], dtx.data(), drx.data());
        double acc = 0.0;
        idx_t nj = (type == DAS_SAFT) ? 1 : Lr;
        for (idx_t i = 0; i < Lt; i++) {
          double ti = dtx[(size_t) i] * g.inv_cp + g.delay_ms;
          double wi = apod_t ? apod_t[i] : 1.0;
          for (idx_t j = 0; j < nj; j++) {
            double u, w = wi;
            if (type == DAS_SAFT) {
              u = 2.0 * ti * g.fs;
            } else {
              u = (ti + drx[(size_t) j] * g.inv_cp) * g.fs;
              w *= apod_r ? apod_r[j] : 1.0;
            }
            const double *y = Y + ascan(g, i, j) * K;
            double s;
            if (interp == 0) {
              idx_t k;
              if (trunc_index<double>(u, K, &k)) s = y[k]; else { s = 0.0; oob = 1; }
            } else {
              double fk = std::floor(u);
              if (!(fk > -9.0e18 && fk < 9.0e18)) fk = -2.0;
              idx_t k = (idx_t) fk;
              double f = u - fk;
              if (k < 0 || k >= K) oob = 1;
              s = (1.0 - f) * yz(y, K, k) + f * yz(y, K, k + 1);
            }
            acc += w * s;
          }
        }
        Im[p] = acc;
      }
      if (oob) any_oob = 1;
    });
  }
  for (auto &t : pool) t.join();
  return any_oob.load();
}

// Adjoint of dro_das_ext_f64 with respect to Y, the apodization weights and
// cp, for the cotangent g (length No).  Outputs (any may be null):
//   gY      K x n_ascans,  gY = A^T g
//   g_apod_t Lt,  g_apod_r Lr,  g_cp scalar.
// In interp 0 the sample index is piecewise constant in cp, so g_cp = 0.
void dro_das_ext_adj_f64(int type, int interp, const double *Y, idx_t K, const double *Ro,
                         idx_t No, const double *Gt, idx_t Lt, const double *Gr, idx_t Lr,
                         double dt, double delay, double cp, const double *apod_t,
                         const double *apod_r, const double *g, double *gY, double *g_apod_t,
                         double *g_apod_r, double *g_cp, int nthreads)
{
  if (Lr == 0) type = DAS_SAFT;
  Geo geo{type, Gt, Gr, Lt, Lr, (1.0 / dt) * 1000.0, cp, delay / 1000.0, 1.0 / cp};
  const idx_t nj = (type == DAS_SAFT) ? 1 : Lr;
  const idx_t n_ascans = (type == DAS_SAFT) ? Lt : Lt * Lr;
  idx_t T = thread_count(No, nthreads);
  struct Part { std::vector<double> gy, gat, gar; double gcp = 0.0; };
  std::vector<Part> parts((size_t) T);
  std::vector<std::thread> pool;
  for (idx_t th = 0; th < T; th++) {
    pool.emplace_back([&, th]() {
      Part &P = parts[(size_t) th];
      if (gY) P.gy.assign((size_t) (n_ascans * K), 0.0);
      P.gat.assign((size_t) Lt, 0.0);
      P.gar.assign((size_t) (Lr > 0 ? Lr : 1), 0.0);
      std::vector<double> dtx((size_t) Lt), drx((size_t) (Lr > 0 ? Lr : 1));
      for (idx_t p = th * No / T; p < (th + 1) * No / T; p++) {
        const double gp = g[p];
        legs(geo, Ro[p], Ro[p + No], Ro[p + 2 * No], dtx.data(), drx.data());
        for (idx_t i = 0; i < Lt; i++) {
          double ti = dtx[(size_t) i] * geo.inv_cp + geo.delay_ms;
          double wi = apod_t ? apod_t[i] : 1.0;
          for (idx_t j = 0; j < nj; j++) {
            double u, du_dcp, wj = 1.0;
            if (type == DAS_SAFT) {
              u = 2.0 * ti * geo.fs;
              du_dcp = -2.0 * dtx[(size_t) i] / (cp * cp) * geo.fs;
            } else {
              u = (ti + drx[(size_t) j] * geo.inv_cp) * geo.fs;
              du_dcp = -(dtx[(size_t) i] + drx[(size_t) j]) / (cp * cp) * geo.fs;
              wj = apod_r ? apod_r[j] : 1.0;
            }
            const idx_t a = ascan(geo, i, j);
            const double *y = Y + a * K;
            double *gy = gY ? P.gy.data() + a * K : nullptr;
            double s, slope = 0.0;
            if (interp == 0) {
              idx_t k;
              if (trunc_index<double>(u, K, &k)) {
                s = y[k];
                if (gy) gy[k] += wi * wj * gp;
              } else {
                s = 0.0;
              }
            } else {
              double fk = std::floor(u);
              idx_t k = (idx_t) fk;
              double f = u - fk;
              double y0 = yz(y, K, k), y1 = yz(y, K, k + 1);
              s = (1.0 - f) * y0 + f * y1;
              slope = y1 - y0;
              if (gy) {
                if (k >= 0 && k < K) gy[k] += wi * wj * (1.0 - f) * gp;
                if (k + 1 >= 0 && k + 1 < K) gy[k + 1] += wi * wj * f * gp;
              }
            }
            P.gat[(size_t) i] += gp * wj * s;
            if (type != DAS_SAFT) P.gar[(size_t) j] += gp * wi * s;
            P.gcp += gp * wi * wj * slope * du_dcp;
          }
        }
      }
    });
  }
  for (auto &t : pool) t.join();
  if (gY) std::memset(gY, 0, sizeof(double) * (size_t) (n_ascans * K));
  if (g_apod_t) std::memset(g_apod_t, 0, sizeof(double) * (size_t) Lt);
  if (g_apod_r && Lr > 0) std::memset(g_apod_r, 0, sizeof(double) * (size_t) Lr);
  double gcp = 0.0;
  for (auto &P : parts) {
    if (gY) for (size_t q = 0; q < P.gy.size(); q++) gY[q] += P.gy[q];
    if (g_apod_t) for (idx_t i = 0; i < Lt; i++) g_apod_t[i] += P.gat[(size_t) i];
    if (g_apod_r && type != DAS_SAFT) for (idx_t j = 0; j < Lr; j++) g_apod_r[j] += P.gar[(size_t) j];
    gcp += P.gcp;
  }
  if (g_cp) *g_cp = gcp;
}

} // extern "C"
