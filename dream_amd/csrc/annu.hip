// annu.hip -- annular arrays (dream_arr_annu) on gfx950.
//
// Reference: ArrAnnu::dream_arr_annu_serial (src/dream_arr_annu.cc:186-254)
// and its helpers annular_disc_radii (:361-368), focusing_annular
// (:376-405), apodization_annular (:413-447), superpos_annular (:459-481)
// and resp_annular (:489-505).  The radii vector Gr holds the centre disc's
// radius, then an (inner, outer) pair per ring; ring n's response is the
// disc response of radius Gr[2n] minus that of Gr[2n-1] (ring 0: disc 0).
//
// Mapping.  Every disc is one SIR problem of the disc kernel (shape
// SIR_CIRC, one element at the origin, no focusing or apodization), and all
// discs of a call run in one persistent SirPlan launch, writing FP64
// responses to a scratch block of num_radii x No x nt doubles.  One
// elementwise kernel then forms the array response of every (point, sample):
//     h[i] = sum_n w_n * ring_n[i - s_n]   (0 where i - s_n is outside [0, nt))
// with s_n = rint(foc_n / dt) + 1, ring by ring in the reference's order, so
// that for equal disc responses the sums round as the reference's do.  The
// disc sums dominate the cost (pi R^2 / (dx dy) sub-elements per point and
// disc); the combine pass reads num_radii doubles per output sample.

#include "sir.hpp"

#include <cmath>
#include <vector>

namespace dream {

namespace {

constexpr int kCombineThreads = 256;

// One ring of the superposition: disc slots of the outer and inner disc (-1:
// a zero response), weight and sample shift (a double so that a shift that
// is not a number or out of range adds nothing).
struct AnnuRing {
  int32_t outer, inner;
  double weight, shift;
};

typedef const __attribute__((address_space(4))) AnnuRing cring;

template <typename TO>
__global__ void __launch_bounds__(kCombineThreads)
annu_combine_kernel(const double *__restrict__ disc, int64_t nt, int64_t No, int rings,
                    const AnnuRing *__restrict__ tab, TO *__restrict__ h)
{
  const int64_t idx = (int64_t) blockIdx.x * kCombineThreads + threadIdx.x;
  if (idx >= No * nt) return;
  const int64_t p = idx / nt;
  const int64_t i = idx - p * nt;
  const int64_t slot = No * nt;
  const double ntd = (double) nt;
  cring *ring = (cring *) tab;
  double acc = 0.0;
  for (int n = 0; n < rings; n++) {
    const int32_t outer = ring[n].outer, inner = ring[n].inner;
    const double w = ring[n].weight, s = ring[n].shift;
    const double j = (double) i - s;
    double src = 0.0;
    if (s < ntd && j >= 0.0 && j < ntd) {
      const int64_t k = p * nt + (int64_t) j;
      const double o = outer >= 0 ? disc[outer * slot + k] : 0.0;
      src = inner >= 0 ? o - disc[inner * slot + k] : o;
    }
    acc += w * src;
  }
  h[idx] = (TO) acc;
}

// FocusMet values (include/dream.h:62-71).
enum { F_NONE = 0, F_X = 1, F_Y = 3, F_XY = 5, F_X_Y = 6, F_UD = 7 };
// ApodMet values (:80-89).
enum { A_TRIANGLE = 0, A_GAUSS, A_RAISED, A_HANN, A_HAMMING, A_SIMPLY, A_CLAMPED, A_UD };

double ring_focus(int foc_met, double focal, double r, double r_max, double cp)
{
  if (foc_met == F_X || foc_met == F_Y || foc_met == F_XY || foc_met == F_X_Y) {
    const double far = std::sqrt(r_max * r_max + focal * focal);
    const double near = std::sqrt(r * r + focal * focal);
    return (far - near) * 1000 / cp;
  }
  return foc_met == F_UD ? focal : 0.0;
}

double ring_weight(int apod_met, int64_t n, const double *apod, double r, double r_max,
                   double par)
{
  const double q = r * r / (r_max * r_max);
  switch (apod_met) {
  case A_UD: return apod[n];
  case A_TRIANGLE: return 1.0 - std::fabs(r) / r_max;
  case A_GAUSS: return std::exp(-(par * r * r) / (r_max * r_max));
  case A_RAISED: return par + std::cos(r * M_PI / r_max);
  case A_SIMPLY: return 1.0 - q;
  case A_CLAMPED: return (1.0 - q) * (1.0 - q);
  default: return 1.0;   // Hann and Hamming have no annular form
  }
}

} // namespace

void run_arr_annu(const AnnuProblem &a, int out_dtype, hipStream_t s, int *sir_err,
                  int64_t *oob_count)
{
  require(a.num_radii >= 1 && a.Gr, "dream_arr_annu: need at least one radius");
  require(a.nt >= 1 && a.No >= 0, "dream_arr_annu: need No >= 0 and nt >= 1");
  require(a.foc_met >= 0 && a.foc_met <= 7, "dream_arr_annu: unknown focusing method");
  require(!a.do_apod || (a.apod_met >= 0 && a.apod_met <= 7),
          "dream_arr_annu: unknown apodization");
  require(a.foc_met != F_UD || a.focal, "dream_arr_annu: 'ud' focusing needs the ring delays");
  require(!(a.do_apod && a.apod_met == A_UD) || a.apod,
          "dream_arr_annu: 'ud' apodization needs the ring weights");
  require(out_dtype == DREAM_HIP_F64 || out_dtype == DREAM_HIP_F32,
          "dream_arr_annu: output type must be f32 or f64");
  const int64_t R = a.num_radii, rings = (R + 1) / 2;

  // Ring radii, weights and shifts (host, once per call).
  std::vector<double> rr((size_t) rings, 0.0);
  for (int64_t n = 1; n < rings; n++) rr[(size_t) n] = (a.Gr[2 * n] + a.Gr[2 * n - 1]) / 2.0;
  const double r_max = rr[(size_t) rings - 1];
  const bool skip0 = a.Gr[0] <= 2.220446049250313e-16;
  std::vector<AnnuRing> tab((size_t) rings);
  for (int64_t n = 0; n < rings; n++) {
    const double f = ring_focus(a.foc_met, a.foc_met == F_UD ? a.focal[n] : (a.focal ? a.focal[0] : 0.0),
                                rr[(size_t) n], r_max, a.cp);
    AnnuRing &t = tab[(size_t) n];
    t.outer = (n == 0 && skip0) ? -1 : (int32_t) (2 * n);
    t.inner = n == 0 ? -1 : (int32_t) (2 * n - 1);
    t.weight = a.do_apod ? ring_weight(a.apod_met, n, a.apod, rr[(size_t) n], r_max, a.apod_par)
                         : 1.0;
    t.shift = std::rint(f / a.dt) + 1.0;
  }
  if (sir_err) *sir_err = E_NONE;
  if (oob_count) *oob_count = 0;
  if (a.No == 0) return;

  // Disc responses: one SIR problem per radius (disc 0 skipped when it has
  // no area, as the reference leaves it zero).
  const size_t slot = (size_t) a.No * (size_t) a.nt;
  double *disc = nullptr;
  if (hipMalloc(&disc, slot * (size_t) R * sizeof(double)) != hipSuccess)
    throw Error(DREAM_HIP_ENOMEM, "dream_arr_annu: device allocation failed");
  AnnuRing *d_tab = nullptr;
  try {
    if (hipMalloc(&d_tab, sizeof(AnnuRing) * (size_t) rings) != hipSuccess)
      throw Error(DREAM_HIP_ENOMEM, "dream_arr_annu: device allocation failed");
    DREAM_HIP_CHECK(hipMemcpyAsync(d_tab, tab.data(), sizeof(AnnuRing) * (size_t) rings,
                                   hipMemcpyHostToDevice, s));
    const double G0[3] = {0.0, 0.0, 0.0};
    std::vector<SirProblem> probs;
    for (int64_t k = 0; k < R; k++) {
      if (k == 0 && skip0) continue;
      SirProblem p;
      p.Ro = a.Ro;  p.No = a.No;
      p.shape = SIR_CIRC;  p.R = a.Gr[k];
      p.dx = a.dx;  p.dy = a.dy;  p.dt = a.dt;  p.nt = a.nt;
      p.delay_type = a.delay_type;  p.delay = a.delay;
      p.v = a.v;  p.cp = a.cp;
      p.G.assign(G0, G0 + 3);
      p.foc_met = 0;  p.focal.assign(1, 0.0);
      p.err_level = a.err_level;
      p.h = disc + (size_t) k * slot;
      probs.push_back(p);
    }
    int code = E_NONE;
    int64_t oob = 0;
    if (!probs.empty()) {
      std::vector<int32_t> codes(probs.size());
      std::vector<int64_t> oobs(probs.size());
      if (a.alpha > 2.220446049250313e-16) {
        run_sir_att(probs, a.alpha, DREAM_HIP_F64, s, codes.data(), oobs.data());
      } else {
        SirPlan plan(probs, DREAM_HIP_F64);
        plan.run(s, false);
        plan.status(codes.data(), oobs.data());   // synchronises `s`
      }
      // A point's status is its last disc's; a stop ends the call at once.
      for (size_t q = 0; q < probs.size(); q++) {
        oob += oobs[q];
        if (codes[q] == E_OOB) code = E_OOB;
      }
      if (code != E_OOB) code = codes.back();   // the last disc is radius R - 1
    }
    if (code != E_OOB) {
      const int64_t total = a.No * a.nt;
      const int64_t blocks = (total + kCombineThreads - 1) / kCombineThreads;
      require(blocks <= 0x7fffffffLL, "dream_arr_annu: problem too large");
      if (out_dtype == DREAM_HIP_F64)
        annu_combine_kernel<double><<<(unsigned) blocks, kCombineThreads, 0, s>>>(
            disc, a.nt, a.No, (int) rings, d_tab, (double *) a.h);
      else
        annu_combine_kernel<float><<<(unsigned) blocks, kCombineThreads, 0, s>>>(
            disc, a.nt, a.No, (int) rings, d_tab, (float *) a.h);
      DREAM_HIP_CHECK(hipGetLastError());
    }
    DREAM_HIP_CHECK(hipStreamSynchronize(s));
    if (sir_err) *sir_err = code;
    if (oob_count) *oob_count = oob;
  } catch (...) {
    (void) hipStreamSynchronize(s);
    if (d_tab) (void) hipFree(d_tab);
    (void) hipFree(disc);
    throw;
  }
  (void) hipFree(d_tab);
  (void) hipFree(disc);
}

} // namespace dream
