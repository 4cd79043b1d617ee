// abi.hip -- the extern "C" boundary of libdream_hip.so (include/dream_hip.h).
//
// Every entry point converts C++ exceptions into DREAM_HIP_E* codes and keeps
// the message for dream_hip_last_error(); no C++ type crosses the boundary.

#include <atomic>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "conv.hpp"
#include "das.hpp"
#include "sir.hpp"
#include "tune.hpp"

namespace dream {

namespace {
thread_local std::string g_last_error;
std::atomic<bool> g_running(true);
thread_local hipStream_t g_stream = nullptr;
} // namespace

void hip_fail(hipError_t e, const char *what, const char *file, int line)
{
  char buf[512];
  std::snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int) e,
                hipGetErrorString(e), what, file, line);
  throw Error(e == hipErrorOutOfMemory ? DREAM_HIP_ENOMEM : DREAM_HIP_ERUNTIME, buf);
}

bool cancelled() { return !g_running.load(std::memory_order_relaxed); }

namespace {
std::atomic<int64_t> g_chunk_work(0);          // 0: the default
std::atomic<int64_t> g_cancel_after(0);        // > 0: abort after that many chunks
} // namespace

int64_t chunk_work()
{
  const int64_t w = g_chunk_work.load(std::memory_order_relaxed);
  return w > 0 ? w : (int64_t) 50000000000ll;
}

// A gate costs nothing until a call is actually cut into chunks: the capture
// test and the two events are set up at the first chunk.
ChunkGate::ChunkGate(hipStream_t s) : s_(s) {}

ChunkGate::~ChunkGate()
{
  for (auto &e : ev_)
    if (e) (void) hipEventDestroy(e);
}

void ChunkGate::before()
{
  if (n_ == 0 && !ev_[0] && !capturing_) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    DREAM_HIP_CHECK(hipStreamIsCapturing(s_, &st));
    capturing_ = st != hipStreamCaptureStatusNone;
    if (!capturing_)
      for (auto &e : ev_) DREAM_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  if (capturing_) return;
  if (n_ >= 2) DREAM_HIP_CHECK(hipEventSynchronize(ev_[n_ & 1]));
  if (cancelled()) throw Error(DREAM_HIP_ECANCELLED, "cancelled between work chunks");
}

void ChunkGate::after()
{
  if (capturing_) return;
  DREAM_HIP_CHECK(hipEventRecord(ev_[n_ & 1], s_));
  n_++;
  int64_t c = g_cancel_after.load();
  while (c > 0 && !g_cancel_after.compare_exchange_weak(c, c - 1)) {
  }
  if (c == 1) g_running.store(false);
}

hipStream_t lib_stream()
{
  if (!g_stream) DREAM_HIP_CHECK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  return g_stream;
}

template <class F>
int guarded(F f)
{
  try {
    g_last_error.clear();
    return f();
  } catch (const Error &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_last_error = "host allocation failed";
    return DREAM_HIP_ENOMEM;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return DREAM_HIP_ERUNTIME;
  }
}

// Device copy of a host array of doubles (freed by the owner).
struct DevBuf {
  void *p = nullptr;
  DevBuf() = default;
  DevBuf(const void *src, size_t bytes, hipStream_t s)
  {
    if (!bytes) return;
    if (hipMalloc(&p, bytes) != hipSuccess) throw Error(DREAM_HIP_ENOMEM, "device allocation failed");
    DREAM_HIP_CHECK(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, s));
  }
  explicit DevBuf(size_t bytes)
  {
    if (bytes && hipMalloc(&p, bytes) != hipSuccess)
      throw Error(DREAM_HIP_ENOMEM, "device allocation failed");
  }
  ~DevBuf() { if (p) (void) hipFree(p); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
};

SirProblem make_problem(const double *Ro, int64_t No, double a, double b, double dx, double dy,
                        double dt, int64_t nt, int delay_type, const double *delay, double v,
                        double cp, int64_t L, const double *G, int foc_met, const double *focal,
                        int steer_met, double theta, double phi, const double *apod,
                        int do_apod, int apod_met, double apod_par, void *h, int err_level);

// A disc problem: radius R in place of the rectangle's sides.
SirProblem make_circ_problem(const double *Ro, int64_t No, double R, double dx, double dy,
                             double dt, int64_t nt, int delay_type, const double *delay, double v,
                             double cp, int64_t L, const double *G, int foc_met,
                             const double *focal, int steer_met, double theta, double phi,
                             const double *apod, int do_apod, int apod_met, double apod_par,
                             void *h, int err_level);

SirProblem make_problem(const double *Ro, int64_t No, double a, double b, double dx, double dy,
                        double dt, int64_t nt, int delay_type, const double *delay, double v,
                        double cp, int64_t L, const double *G, int foc_met, const double *focal,
                        int steer_met, double theta, double phi, const double *apod,
                        int do_apod, int apod_met, double apod_par, void *h, int err_level)
{
  require(L >= 1 && G, "dream_arr_rect: need a grid G with at least one element");
  require(foc_met >= 0 && foc_met <= 7, "dream_arr_rect: unknown focusing method");
  require(steer_met >= 0 && steer_met <= 3, "dream_arr_rect: unknown steering method");
  require(!do_apod || (apod_met >= 0 && apod_met <= 7), "dream_arr_rect: unknown apodization");
  require(err_level >= 0 && err_level <= 4, "dream_arr_rect: unknown error level");
  SirProblem p;
  p.Ro = Ro;
  p.No = No;
  p.a = a; p.b = b; p.dx = dx; p.dy = dy; p.dt = dt; p.nt = nt;
  p.delay_type = delay_type;
  p.delay = delay;
  p.v = v;
  p.cp = cp;
  p.G.assign(G, G + 3 * L);
  p.foc_met = foc_met;
  if (foc_met == 7) {
    require(focal != nullptr, "dream_arr_rect: 'ud' focusing needs L focal delays");
    p.focal.assign(focal, focal + L);
  } else {
    p.focal.assign(1, focal ? focal[0] : 0.0);
  }
  p.steer_met = steer_met;
  p.theta = theta;
  p.phi = phi;
  p.do_apod = do_apod;
  p.apod_met = apod_met;
  p.apod_par = apod_par;
  if (do_apod && apod_met == 7) {
    require(apod != nullptr, "dream_arr_rect: 'ud' apodization needs L weights");
    p.apod.assign(apod, apod + L);
  }
  p.err_level = err_level;
  p.h = h;
  return p;
}

SirProblem make_circ_problem(const double *Ro, int64_t No, double R, double dx, double dy,
                             double dt, int64_t nt, int delay_type, const double *delay, double v,
                             double cp, int64_t L, const double *G, int foc_met,
                             const double *focal, int steer_met, double theta, double phi,
                             const double *apod, int do_apod, int apod_met, double apod_par,
                             void *h, int err_level)
{
  SirProblem p = make_problem(Ro, No, 0.0, 0.0, dx, dy, dt, nt, delay_type, delay, v, cp, L, G,
                              foc_met, focal, steer_met, theta, phi, apod, do_apod, apod_met,
                              apod_par, h, err_level);
  p.shape = SIR_CIRC;
  p.R = R;
  return p;
}

// A cylindrically curved problem: sides a, b and curvature radius Rcurv.
SirProblem make_cylind_problem(const double *Ro, int64_t No, double a, double b, double Rcurv,
                               double dx, double dy, double dt, int64_t nt, int delay_type,
                               const double *delay, double v, double cp, int64_t L,
                               const double *G, int foc_met, const double *focal, int steer_met,
                               double theta, double phi, const double *apod, int do_apod,
                               int apod_met, double apod_par, void *h, int err_level)
{
  SirProblem p = make_problem(Ro, No, a, b, dx, dy, dt, nt, delay_type, delay, v, cp, L, G,
                              foc_met, focal, steer_met, theta, phi, apod, do_apod, apod_met,
                              apod_par, h, err_level);
  p.shape = SIR_CYLIND;
  p.R = Rcurv;
  return p;
}

// One SIR problem on the device: run a single-problem plan, then read its status.
void run_single(const SirProblem &p, int out_dtype, hipStream_t s, int *sir_err,
                int64_t *oob_count)
{
  std::vector<SirProblem> probs{p};
  SirPlan plan(probs, out_dtype);
  plan.run(s, false);
  int32_t code = 0;
  int64_t oob = 0;
  plan.status(&code, &oob);   // synchronises `s`
  if (sir_err) *sir_err = code;
  if (oob_count) *oob_count = oob;
}

// Host-pointer SIR of one problem (FP64 output): copy in, run, copy out.
int run_host_sir(SirProblem p, double alpha, const double *Ro, const double *delay, double *h,
                 int *sir_err)
{
  require(p.No >= 0 && p.nt >= 1, "SIR: need No >= 0 and nt >= 1");
  require(p.No == 0 || (Ro && delay && h), "SIR: null argument");
  hipStream_t s = lib_stream();
  const int64_t nd = p.delay_type ? p.No : 1;
  DevBuf dRo(Ro, (size_t) p.No * 3 * 8, s), dD(delay, (size_t) nd * 8, s);
  DevBuf dH((size_t) (p.No * p.nt) * 8);
  p.Ro = (const double *) dRo.p;
  p.delay = (const double *) dD.p;
  p.h = dH.p;
  int code = 0;
  if (alpha > 2.220446049250313e-16) {
    int32_t c = 0;
    run_sir_att(std::vector<SirProblem>{p}, alpha, DREAM_HIP_F64, s, &c, nullptr);
    code = c;
  } else {
    run_single(p, DREAM_HIP_F64, s, &code, nullptr);
  }
  if (p.No) DREAM_HIP_CHECK(hipMemcpy(h, dH.p, (size_t) (p.No * p.nt) * 8, hipMemcpyDeviceToHost));
  if (sir_err) *sir_err = code;
  return DREAM_HIP_OK;
}

} // namespace dream

using namespace dream;

extern "C" {

const char *dream_hip_last_error(void) { return g_last_error.c_str(); }

const char *dream_hip_version(void) { return "dream_hip 0.1 (gfx950)"; }

int dream_hip_device_count(void)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void dream_hip_set_autotune(int on) { set_autotune(on != 0); }

void dream_hip_test_chunking(int64_t work_per_chunk, int64_t cancel_after_chunks)
{
  g_chunk_work.store(work_per_chunk > 0 ? work_per_chunk : 0);
  g_cancel_after.store(cancel_after_chunks > 0 ? cancel_after_chunks : 0);
}

void dream_hip_abort(int) { g_running.store(false); }
void dream_hip_set_running(void) { g_running.store(true); }
int dream_hip_is_running(void) { return g_running.load() ? 1 : 0; }

// ------------------------------------------------------------------ SIR

int dream_hip_arr_rect_dev(const double *Ro, int64_t No, double a, double b, double dx,
                           double dy, double dt, int64_t nt, int delay_type,
                           const double *delay, double v, double cp, int64_t num_elements,
                           const double *G, int foc_met, const double *focal, int steer_met,
                           double theta, double phi, const double *apod, int do_apod,
                           int apod_met, double apod_par, void *h, int out_dtype, int err_level,
                           int *sir_err, int64_t *oob_count, void *stream)
{
  return guarded([&]() {
    hipStream_t s = (hipStream_t) stream;   // NULL: the default stream
    std::vector<SirProblem> probs;
    probs.push_back(make_problem(Ro, No, a, b, dx, dy, dt, nt, delay_type, delay, v, cp,
                                 num_elements, G, foc_met, focal, steer_met, theta, phi, apod,
                                 do_apod, apod_met, apod_par, h, err_level));
    SirPlan plan(probs, out_dtype);
    plan.run(s, false);
    int32_t code = 0;
    int64_t oob = 0;
    plan.status(&code, &oob);   // synchronises `s`
    if (sir_err) *sir_err = code;
    if (oob_count) *oob_count = oob;
    return DREAM_HIP_OK;
  });
}

int dream_hip_arr_rect(double alpha, const double *Ro, int64_t No, double a, double b,
                       double dx, double dy, double dt, int64_t nt, int delay_type,
                       const double *delay, double v, double cp, int64_t num_elements,
                       const double *G, int foc_met, const double *focal, int steer_met,
                       double theta, double phi, const double *apod, int do_apod, int apod_met,
                       double apod_par, double *h, int err_level, int *sir_err)
{
  return guarded([&]() {
    require(No == 0 || (Ro && delay && h), "dream_arr_rect: null argument");
    return run_host_sir(make_problem(nullptr, No, a, b, dx, dy, dt, nt, delay_type, nullptr, v,
                                     cp, num_elements, G, foc_met, focal, steer_met, theta, phi,
                                     apod, do_apod, apod_met, apod_par, nullptr, err_level),
                        alpha, Ro, delay, h, sir_err);
  });
}

int dream_hip_rect(double alpha, const double *Ro, int64_t No, double a, double b, double dx,
                   double dy, double dt, int64_t nt, int delay_type, const double *delay,
                   double v, double cp, double *h, int err_level, int *sir_err)
{
  const double G[3] = {0.0, 0.0, 0.0};
  const double focal = 0.0;
  return dream_hip_arr_rect(alpha, Ro, No, a, b, dx, dy, dt, nt, delay_type, delay, v, cp, 1, G,
                            0, &focal, 0, 0.0, 0.0, nullptr, 0, 0, 0.0, h, err_level, sir_err);
}

int dream_hip_arr_circ_dev(const double *Ro, int64_t No, double R, double dx, double dy,
                           double dt, int64_t nt, int delay_type, const double *delay, double v,
                           double cp, int64_t num_elements, const double *G, int foc_met,
                           const double *focal, int steer_met, double theta, double phi,
                           const double *apod, int do_apod, int apod_met, double apod_par,
                           void *h, int out_dtype, int err_level, int *sir_err,
                           int64_t *oob_count, void *stream)
{
  return guarded([&]() {
    run_single(make_circ_problem(Ro, No, R, dx, dy, dt, nt, delay_type, delay, v, cp,
                                 num_elements, G, foc_met, focal, steer_met, theta, phi, apod,
                                 do_apod, apod_met, apod_par, h, err_level),
               out_dtype, (hipStream_t) stream, sir_err, oob_count);
    return DREAM_HIP_OK;
  });
}

int dream_hip_arr_circ(double alpha, const double *Ro, int64_t No, double R, double dx,
                       double dy, double dt, int64_t nt, int delay_type, const double *delay,
                       double v, double cp, int64_t num_elements, const double *G, int foc_met,
                       const double *focal, int steer_met, double theta, double phi,
                       const double *apod, int do_apod, int apod_met, double apod_par,
                       double *h, int err_level, int *sir_err)
{
  return guarded([&]() {
    return run_host_sir(make_circ_problem(nullptr, No, R, dx, dy, dt, nt, delay_type, nullptr, v,
                                          cp, num_elements, G, foc_met, focal, steer_met, theta,
                                          phi, apod, do_apod, apod_met, apod_par, nullptr,
                                          err_level),
                        alpha, Ro, delay, h, sir_err);
  });
}

int dream_hip_circ(double alpha, const double *Ro, int64_t No, double R, double dx, double dy,
                   double dt, int64_t nt, int delay_type, const double *delay, double v,
                   double cp, double *h, int err_level, int *sir_err)
{
  const double G[3] = {0.0, 0.0, 0.0};
  const double focal = 0.0;
  return dream_hip_arr_circ(alpha, Ro, No, R, dx, dy, dt, nt, delay_type, delay, v, cp, 1, G, 0,
                            &focal, 0, 0.0, 0.0, nullptr, 0, 0, 0.0, h, err_level, sir_err);
}

int dream_hip_arr_cylind_dev(const double *Ro, int64_t No, double a, double b, double Rcurv,
                             double dx, double dy, double dt, int64_t nt, int delay_type,
                             const double *delay, double v, double cp, int64_t num_elements,
                             const double *G, int foc_met, const double *focal, int steer_met,
                             double theta, double phi, const double *apod, int do_apod,
                             int apod_met, double apod_par, void *h, int out_dtype,
                             int err_level, int *sir_err, int64_t *oob_count, void *stream)
{
  return guarded([&]() {
    run_single(make_cylind_problem(Ro, No, a, b, Rcurv, dx, dy, dt, nt, delay_type, delay, v, cp,
                                   num_elements, G, foc_met, focal, steer_met, theta, phi, apod,
                                   do_apod, apod_met, apod_par, h, err_level),
               out_dtype, (hipStream_t) stream, sir_err, oob_count);
    return DREAM_HIP_OK;
  });
}

int dream_hip_arr_cylind(double alpha, const double *Ro, int64_t No, double a, double b,
                         double Rcurv, double dx, double dy, double dt, int64_t nt,
                         int delay_type, const double *delay, double v, double cp,
                         int64_t num_elements, const double *G, int foc_met,
                         const double *focal, int steer_met, double theta, double phi,
                         const double *apod, int do_apod, int apod_met, double apod_par,
                         double *h, int err_level, int *sir_err)
{
  return guarded([&]() {
    return run_host_sir(make_cylind_problem(nullptr, No, a, b, Rcurv, dx, dy, dt, nt, delay_type,
                                            nullptr, v, cp, num_elements, G, foc_met, focal,
                                            steer_met, theta, phi, apod, do_apod, apod_met,
                                            apod_par, nullptr, err_level),
                        alpha, Ro, delay, h, sir_err);
  });
}

int dream_hip_cylind(double alpha, const double *Ro, int64_t No, double a, double b, double Rcurv,
                     double dx, double dy, double dt, int64_t nt, int delay_type,
                     const double *delay, double v, double cp, double *h, int err_level,
                     int *sir_err)
{
  const double G[3] = {0.0, 0.0, 0.0};
  const double focal = 0.0;
  return dream_hip_arr_cylind(alpha, Ro, No, a, b, Rcurv, dx, dy, dt, nt, delay_type, delay, v,
                              cp, 1, G, 0, &focal, 0, 0.0, 0.0, nullptr, 0, 0, 0.0, h, err_level,
                              sir_err);
}

int dream_hip_arr_annu_dev(const double *Ro, int64_t No, double dx, double dy, double dt,
                           int64_t nt, int delay_type, const double *delay, double v, double cp,
                           int64_t num_radii, const double *Gr, int foc_met, const double *focal,
                           const double *apod, int do_apod, int apod_met, double apod_par,
                           void *h, int out_dtype, int err_level, int *sir_err,
                           int64_t *oob_count, void *stream)
{
  return guarded([&]() {
    require(err_level >= 0 && err_level <= 4, "dream_arr_annu: unknown error level");
    require(No == 0 || (Ro && delay && h), "dream_arr_annu: null argument");
    AnnuProblem a;
    a.Ro = Ro;  a.No = No;
    a.dx = dx;  a.dy = dy;  a.dt = dt;  a.nt = nt;
    a.delay_type = delay_type;  a.delay = delay;
    a.v = v;  a.cp = cp;
    a.num_radii = num_radii;  a.Gr = Gr;
    a.foc_met = foc_met;  a.focal = focal;
    a.apod = apod;  a.do_apod = do_apod;  a.apod_met = apod_met;  a.apod_par = apod_par;
    a.err_level = err_level;
    a.h = h;
    run_arr_annu(a, out_dtype, (hipStream_t) stream, sir_err, oob_count);
    return DREAM_HIP_OK;
  });
}

int dream_hip_arr_annu(double alpha, const double *Ro, int64_t No, double dx, double dy,
                       double dt, int64_t nt, int delay_type, const double *delay, double v,
                       double cp, int64_t num_radii, const double *Gr, int foc_met,
                       const double *focal, const double *apod, int do_apod, int apod_met,
                       double apod_par, double *h, int err_level, int *sir_err)
{
  return guarded([&]() {
    require(No >= 0 && nt >= 1, "dream_arr_annu: need No >= 0 and nt >= 1");
    require(No == 0 || (Ro && delay && h), "dream_arr_annu: null argument");
    hipStream_t s = lib_stream();
    const int64_t nd = delay_type ? No : 1;
    DevBuf dRo(Ro, (size_t) No * 3 * 8, s), dD(delay, (size_t) nd * 8, s);
    DevBuf dH((size_t) No * nt * 8);
    require(err_level >= 0 && err_level <= 4, "dream_arr_annu: unknown error level");
    AnnuProblem an;
    an.Ro = (const double *) dRo.p;  an.No = No;
    an.dx = dx;  an.dy = dy;  an.dt = dt;  an.nt = nt;
    an.delay_type = delay_type;  an.delay = (const double *) dD.p;
    an.v = v;  an.cp = cp;
    an.num_radii = num_radii;  an.Gr = Gr;
    an.foc_met = foc_met;  an.focal = focal;
    an.apod = apod;  an.do_apod = do_apod;  an.apod_met = apod_met;  an.apod_par = apod_par;
    an.err_level = err_level;
    an.alpha = alpha;
    an.h = dH.p;
    int code = 0;
    run_arr_annu(an, DREAM_HIP_F64, s, &code, nullptr);
    if (No) DREAM_HIP_CHECK(hipMemcpy(h, dH.p, (size_t) No * nt * 8, hipMemcpyDeviceToHost));
    if (sir_err) *sir_err = code;
    return DREAM_HIP_OK;
  });
}

int dream_hip_rect_indices(double xo, double yo, double zo, double a, double b, double dx,
                           double dy, double dt, double delay, double cp, int64_t *it_out,
                           int64_t cap, int64_t *count)
{
  return guarded([&]() {
    require(it_out || cap == 0, "rect_indices: null output");
    require(count != nullptr, "rect_indices: null count");
    rect_indices(xo, yo, zo, a, b, dx, dy, dt, delay, cp, it_out, cap, count);
    return DREAM_HIP_OK;
  });
}

int dream_hip_rsq_seed(const double *s, int64_t n, double *y)
{
  return guarded([&]() {
    require(n >= 0, "rsq_seed: negative count");
    require(n == 0 || (s && y), "rsq_seed: null array");
    rsq_seed(s, n, y);
    return DREAM_HIP_OK;
  });
}

namespace {
std::vector<SirProblem> sir_problems(const DreamSirProblem *problems, int32_t n)
{
  std::vector<SirProblem> probs;
  probs.reserve((size_t) n);
  for (int32_t i = 0; i < n; i++) {
    const DreamSirProblem &q = problems[i];
    require(q.shape == SIR_RECT || q.shape == SIR_CIRC || q.shape == SIR_CYLIND,
            "SIR: unknown shape");
    probs.push_back(make_problem(q.Ro, q.No, q.a, q.b, q.dx, q.dy, q.dt, q.nt, q.delay_type,
                                 q.delay, q.v, q.cp, q.num_elements, q.G, q.foc_met, q.focal,
                                 q.steer_met, q.theta, q.phi, q.apod, q.do_apod, q.apod_met,
                                 q.apod_par, q.h, q.err_level));
    probs.back().shape = q.shape;
    probs.back().R = q.R;
  }
  return probs;
}
} // namespace

int dream_hip_sir_att_dev(const DreamSirProblem *problems, int32_t n, double alpha,
                          int out_dtype, int32_t *sir_err, int64_t *oob_count, void *stream)
{
  return guarded([&]() {
    require((problems || n == 0) && n >= 0, "sir_att_dev: bad arguments");
    if (cancelled()) throw Error(DREAM_HIP_ECANCELLED, "SIR cancelled");
    hipStream_t s = (hipStream_t) stream;   // NULL: the default stream
    const std::vector<SirProblem> probs = sir_problems(problems, n);
    if (alpha > 2.220446049250313e-16) {          // DBL_EPSILON
      run_sir_att(probs, alpha, out_dtype, s, sir_err, oob_count);
    } else {
      SirPlan plan(probs, out_dtype);
      plan.run(s, false);
      plan.status(sir_err, oob_count);
    }
    return DREAM_HIP_OK;
  });
}

int dream_hip_sir_plan_create(const DreamSirProblem *problems, int32_t n, int out_dtype,
                              void **plan)
{
  return guarded([&]() {
    require(plan != nullptr && (problems || n == 0) && n >= 0, "sir_plan_create: bad arguments");
    std::vector<SirProblem> probs;
    probs.reserve((size_t) n);
    for (int32_t i = 0; i < n; i++) {
      const DreamSirProblem &q = problems[i];
      require(q.shape == SIR_RECT || q.shape == SIR_CIRC || q.shape == SIR_CYLIND,
              "sir_plan_create: unknown shape");
      probs.push_back(make_problem(q.Ro, q.No, q.a, q.b, q.dx, q.dy, q.dt, q.nt, q.delay_type,
                                   q.delay, q.v, q.cp, q.num_elements, q.G, q.foc_met, q.focal,
                                   q.steer_met, q.theta, q.phi, q.apod, q.do_apod, q.apod_met,
                                   q.apod_par, q.h, q.err_level));
      probs.back().shape = q.shape;
      probs.back().R = q.R;
    }
    *plan = new SirPlan(probs, out_dtype);
    return DREAM_HIP_OK;
  });
}

int dream_hip_sir_plan_run(void *plan, void *stream, int use_graph)
{
  return guarded([&]() {
    require(plan != nullptr, "sir_plan_run: null plan");
    static_cast<SirPlan *>(plan)->run((hipStream_t) stream, use_graph != 0);
    return DREAM_HIP_OK;
  });
}

int dream_hip_sir_plan_status(void *plan, int32_t *sir_err, int64_t *oob_count)
{
  return guarded([&]() {
    require(plan != nullptr, "sir_plan_status: null plan");
    static_cast<SirPlan *>(plan)->status(sir_err, oob_count);
    return DREAM_HIP_OK;
  });
}

int dream_hip_sir_plan_destroy(void *plan)
{
  return guarded([&]() {
    delete static_cast<SirPlan *>(plan);
    return DREAM_HIP_OK;
  });
}

// ------------------------------------------------------------------ DAS

int dream_hip_das_create(int das_type, int64_t a_scan_len, int64_t No, int64_t num_t_elements,
                         int64_t num_r_elements, int precision, void **handle)
{
  return guarded([&]() {
    require(handle != nullptr, "das_create: null handle pointer");
    *handle = new DasHandle(das_type, a_scan_len, No, num_t_elements, num_r_elements, precision);
    return DREAM_HIP_OK;
  });
}

int dream_hip_das_destroy(void *handle)
{
  return guarded([&]() {
    delete static_cast<DasHandle *>(handle);
    return DREAM_HIP_OK;
  });
}

int dream_hip_das_setup_has_changed(void *handle, int das_type, int64_t a_scan_len, int64_t No,
                                    int64_t num_t_elements, int64_t num_r_elements,
                                    int precision)
{
  if (!handle) return 1;
  return static_cast<DasHandle *>(handle)->changed(das_type, a_scan_len, No, num_t_elements,
                                                   num_r_elements, precision)
             ? 1
             : 0;
}

int dream_hip_das_set_pixel_order(void *handle, const int32_t *perm, int64_t n)
{
  return guarded([&]() {
    require(handle != nullptr, "das: null handle");
    static_cast<DasHandle *>(handle)->set_pixel_order(perm, n);
    return DREAM_HIP_OK;
  });
}

int dream_hip_pixel_order(const double *Ro, int64_t No, int32_t *perm)
{
  return guarded([&]() {
    require(No >= 0 && (No == 0 || (Ro && perm)), "pixel_order: bad arguments");
    pixel_order(Ro, No, perm);
    return DREAM_HIP_OK;
  });
}

int dream_hip_das(void *handle, const void *Y, const void *Ro, const void *Gt, const void *Gr,
                  double dt, int delay_type, const void *delay, double cp, void *Im,
                  int err_level, int *sir_err)
{
  return guarded([&]() {
    require(handle != nullptr, "das: null handle");
    require(err_level >= 0 && err_level <= 4, "das: unknown error level");
    int code = static_cast<DasHandle *>(handle)->run_host(Y, Ro, Gt, Gr, dt, delay_type, delay, cp,
                                                          Im, err_level);
    if (sir_err) *sir_err = code;
    return DREAM_HIP_OK;
  });
}

int dream_hip_das_dev(void *handle, const void *Y, const void *Ro, const void *Gt, const void *Gr,
                      double dt, int delay_type, const void *delay, double cp,
                      const void *apod_t, const void *apod_r, int interp, int64_t n_frames,
                      int64_t y_frame_stride, int64_t im_frame_stride, void *Im, int err_level,
                      int *sir_err, void *stream)
{
  return guarded([&]() {
    require(handle != nullptr, "das: null handle");
    require(err_level >= 0 && err_level <= 4, "das: unknown error level");
    hipStream_t s = (hipStream_t) stream;   // NULL: the default stream
    int code = static_cast<DasHandle *>(handle)->run_dev(
        Y, Ro, Gt, Gr, dt, delay_type, delay, cp, apod_t, apod_r, interp, n_frames,
        y_frame_stride, im_frame_stride, Im, err_level, sir_err != nullptr, s);
    if (sir_err) *sir_err = code;
    return DREAM_HIP_OK;
  });
}

int dream_hip_das_status(void *handle, int err_level, int *sir_err)
{
  return guarded([&]() {
    require(handle != nullptr && sir_err != nullptr, "das status: null handle or code");
    require(err_level >= 0 && err_level <= 4, "das: unknown error level");
    *sir_err = static_cast<DasHandle *>(handle)->status(err_level);
    return DREAM_HIP_OK;
  });
}

int dream_hip_das_adjoint_dev(void *handle, const void *Y, const void *Ro, const void *Gt,
                              const void *Gr, double dt, double delay, double cp,
                              const void *apod_t, const void *apod_r, int interp,
                              int64_t n_frames, int64_t y_frame_stride, int64_t im_frame_stride,
                              const void *g, void *gY, void *g_apod_t, void *g_apod_r,
                              void *g_cp, void *stream)
{
  return guarded([&]() {
    require(handle != nullptr, "das adjoint: null handle");
    hipStream_t s = (hipStream_t) stream;   // NULL: the default stream
    static_cast<DasHandle *>(handle)->adjoint_dev(Y, Ro, Gt, Gr, dt, delay, cp, apod_t, apod_r,
                                                  interp, n_frames, y_frame_stride,
                                                  im_frame_stride, g, gY, g_apod_t, g_apod_r,
                                                  g_cp, s);
    return DREAM_HIP_OK;
  });
}

int dream_hip_das_time_kernels(void *handle, int n_max)
{
  return guarded([&]() {
    require(handle != nullptr, "das timing: null handle");
    static_cast<DasHandle *>(handle)->start_timing(n_max);
    return DREAM_HIP_OK;
  });
}

int dream_hip_das_kernel_times(void *handle, int *n, double *prep_ms, double *forward_ms)
{
  return guarded([&]() {
    require(handle != nullptr, "das timing: null handle");
    static_cast<DasHandle *>(handle)->read_timing(n, prep_ms, forward_ms);
    return DREAM_HIP_OK;
  });
}

int dream_hip_das_adjoint_dev_noalloc(void *handle, const void *Y, const void *Ro,
                                      const void *Gt, const void *Gr, double dt, double delay,
                                      double cp, const void *apod_t, const void *apod_r,
                                      int interp, int64_t n_frames, int64_t y_frame_stride,
                                      int64_t im_frame_stride, const void *g, void *gY,
                                      void *g_apod_t, void *g_apod_r, void *g_cp, void *stream)
{
  return guarded([&]() {
    require(handle != nullptr, "das adjoint: null handle");
    hipStream_t s = (hipStream_t) stream;   // NULL: the default stream
    size_t bytes = 0;
    double *z = fixed_workspace(&bytes);
    static_cast<DasHandle *>(handle)->adjoint_dev(Y, Ro, Gt, Gr, dt, delay, cp, apod_t, apod_r,
                                                  interp, n_frames, y_frame_stride,
                                                  im_frame_stride, g, gY, g_apod_t, g_apod_r,
                                                  g_cp, s, z, bytes);
    return DREAM_HIP_OK;
  });
}

// ------------------------------------------------------------------ das_uni

int dream_hip_das_uni_create(int das_type, int64_t a_scan_len, const double *gt,
                             const double *gr, const double *Ro, int precision, void **handle)
{
  return guarded([&]() {
    require(handle && gt && Ro && (das_type == 0 || gr), "das_uni_create: null argument");
    require(a_scan_len >= 1 && a_scan_len < (1ll << 26),
            "das_uni: A-scan length out of range (1 .. 2^26)");
    *handle = new DasUni(das_type, a_scan_len, gt, das_type == 0 ? gt : gr, Ro, precision);
    return DREAM_HIP_OK;
  });
}

int dream_hip_das_uni_destroy(void *handle)
{
  return guarded([&]() {
    delete static_cast<DasUni *>(handle);
    return DREAM_HIP_OK;
  });
}

int dream_hip_das_uni_shape(void *handle, int64_t *No, int64_t *num_t, int64_t *num_r)
{
  return guarded([&]() {
    require(handle != nullptr, "das_uni: null handle");
    const DasUni *u = static_cast<const DasUni *>(handle);
    if (No) *No = u->No();
    if (num_t) *num_t = u->Lt();
    if (num_r) *num_r = u->Lr();
    return DREAM_HIP_OK;
  });
}

int dream_hip_das_uni_setup_has_changed(void *handle, int das_type, int64_t a_scan_len,
                                        const double *gt, const double *gr, const double *Ro,
                                        int precision)
{
  if (!handle || !gt || !Ro || (das_type != 0 && !gr)) return 1;
  return static_cast<DasUni *>(handle)->changed(das_type, a_scan_len, gt, das_type == 0 ? gt : gr,
                                                Ro, precision) ? 1 : 0;
}

int dream_hip_das_uni(void *handle, const void *Y, double dt, double delay, double cp, void *Im)
{
  return guarded([&]() {
    require(handle && Y && Im, "das_uni: null argument");
    static_cast<DasUni *>(handle)->run_host(Y, dt, delay, cp, Im);
    return DREAM_HIP_OK;
  });
}

int dream_hip_das_uni_dev(void *handle, const void *Y, double dt, double delay, double cp,
                          int interp, void *Im, void *stream)
{
  return guarded([&]() {
    require(handle && Y && Im, "das_uni: null argument");
    static_cast<DasUni *>(handle)->run_dev(Y, dt, delay, cp, interp, Im, (hipStream_t) stream);
    return DREAM_HIP_OK;
  });
}

// ------------------------------------------------------------------ convolution

int dream_hip_conv_sum_dev(void *Y, const void *A, int64_t M_A, int64_t N, int64_t L,
                           int64_t a_col_stride, int64_t a_layer_stride, const void *B,
                           int64_t M_B, int64_t b_col_stride, int64_t b_layer_stride, int mode,
                           int dtype, void *stream)
{
  return guarded([&]() {
    require(N == 0 || (Y && A && B), "conv: null device pointer");
    require(dtype == DREAM_HIP_F64 || dtype == DREAM_HIP_F32, "conv: dtype must be F64 or F32");
    hipStream_t s = (hipStream_t) stream;
    if (dtype == DREAM_HIP_F64)
      launch_conv_sum<double>((double *) Y, (const double *) A, M_A, N, L, a_col_stride,
                              a_layer_stride, (const double *) B, M_B, b_col_stride,
                              b_layer_stride, mode, s);
    else
      launch_conv_sum<float>((float *) Y, (const float *) A, M_A, N, L, a_col_stride,
                             a_layer_stride, (const float *) B, M_B, b_col_stride,
                             b_layer_stride, mode, s);
    return DREAM_HIP_OK;
  });
}

namespace {
// Host-pointer convolution: copy in, convolve, copy out (synchronous).
int host_conv_sum(double *Y, const double *A, int64_t MA, int64_t N, int64_t L, int64_t a_col,
                  int64_t a_layer, int64_t a_total, const double *B, int64_t MB, int64_t b_col,
                  int64_t b_layer, int64_t b_total, int mode)
{
  require(MA >= 1 && MB >= 1 && N >= 0 && L >= 1, "conv: bad sizes");
  require(mode == CONV_EQU || mode == CONV_SUM || mode == CONV_NEG, "conv: unknown mode");
  const int64_t MY = MA + MB - 1;
  if (N == 0) return DREAM_HIP_OK;
  require(Y && A && B, "conv: null argument");
  hipStream_t s = lib_stream();
  DevBuf dA(A, (size_t) a_total * 8, s), dB(B, (size_t) b_total * 8, s);
  DevBuf dY = mode == CONV_EQU ? DevBuf((size_t) (MY * N) * 8) : DevBuf(Y, (size_t) (MY * N) * 8, s);
  launch_conv_sum<double>((double *) dY.p, (const double *) dA.p, MA, N, L, a_col, a_layer,
                          (const double *) dB.p, MB, b_col, b_layer, mode, s);
  DREAM_HIP_CHECK(hipMemcpyAsync(Y, dY.p, (size_t) (MY * N) * 8, hipMemcpyDeviceToHost, s));
  DREAM_HIP_CHECK(hipStreamSynchronize(s));
  return DREAM_HIP_OK;
}
} // namespace

int dream_hip_fftconv_p(double *Y, const double *A, int64_t A_M, int64_t A_N, const double *B,
                        int64_t B_M, int64_t B_N, int mode)
{
  return guarded([&]() {
    require(A_M >= 1 && A_N >= 0 && B_M >= 1 && B_N >= 1, "fftconv_p: bad sizes");
    // A vector B is one filter for every column (extra_src/fftconv_p.cc:170-173).
    const bool vec = B_M == 1 || B_N == 1;
    require(vec || B_N == A_N,
            "fftconv_p: B must be a vector or a matrix with the same number of columns as A");
    const int64_t MB = vec ? B_M * B_N : B_M;
    return host_conv_sum(Y, A, A_M, A_N, 1, A_M, 0, A_M * A_N, B, MB, vec ? 0 : B_M, 0,
                         vec ? MB : B_M * B_N, mode);
  });
}

int dream_hip_sum_fftconv(double *Y, const double *H, int64_t M, int64_t N, int64_t L,
                          const double *U, int64_t K, int mode)
{
  return guarded([&]() {
    require(M >= 1 && N >= 0 && L >= 1 && K >= 1, "sum_fftconv: bad sizes");
    return host_conv_sum(Y, H, M, N, L, M, M * N, M * N * L, U, K, 0, K, K * L, mode);
  });
}

int dream_hip_das_batch_create(const DreamDasProblem *problems, int32_t n, int das_type,
                               int precision, int interp, void **batch)
{
  return guarded([&]() {
    require(batch != nullptr && (problems || n == 0) && n >= 0, "das_batch_create: bad arguments");
    *batch = new DasBatch(problems, n, das_type, precision, interp);
    return DREAM_HIP_OK;
  });
}

int dream_hip_das_batch_run(void *batch, void *stream, int use_graph)
{
  return guarded([&]() {
    require(batch != nullptr, "das_batch_run: null batch");
    static_cast<DasBatch *>(batch)->run((hipStream_t) stream, use_graph != 0);
    return DREAM_HIP_OK;
  });
}

int dream_hip_das_batch_destroy(void *batch)
{
  return guarded([&]() {
    delete static_cast<DasBatch *>(batch);
    return DREAM_HIP_OK;
  });
}

} // extern "C"
