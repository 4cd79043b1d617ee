// att.hip -- SIR with frequency-dependent attenuation (alpha > 0) on gfx950.
//
// Reference: Attenuation::att (src/attenuation.cc:72-140), called from the
// DR sums for every in-range contribution (e.g. src/dreamrect.cc:341-395).
// Each contribution of amplitude ai, distance r and sample it adds ai times
// the inverse real FFT (length nt, divided by nt) of the lossy Green's
// function spectrum
//     X[0] = 1,  X[k] = exp(-r a w_k/(2 pi)) exp(i(-w_k r a b_k/pi^2 - t w_k)),
//     w_k = k (2 pi/nt) Fs,  b_k = log(1/(a1 w_k)),  a1 = dt 0.95/pi,  t = it dt
// (r, dt in SI units; a = alpha/(8.686 10^4) in Np/(m Hz)), k = 1 .. nt/2.
//
// The transform is linear, so the sum over contributions is taken in the
// frequency domain: one work-group per (point, chunk of contributions)
// stages 256 contributions at a time (geometry, index and amplitude exactly
// as the DR sums compute them) and every thread accumulates the spectrum
// bins it owns, evaluating X[k] with the reference's expressions; a second
// kernel turns each point's spectrum into its nt samples by the direct
// complex-to-real sum (the imaginary parts of X[0] and X[nt/2] drop out, as
// in FFTW's c2r transform).  Cost: nt/2 complex exponentials per
// contribution, as the reference pays, without its per-contribution FFT.
// A work-group holds at most kBinBlock bins; longer responses split each
// point's bins over several work-groups (any nt).

#include "sir.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace dream {

namespace {

constexpr int kThreads = 256;

typedef const __attribute__((address_space(1))) double gdouble;

// One problem on the device.
struct AttDev {
  const double *Ro;  int64_t No;
  const double *delay;  int32_t delay_multi;  int32_t L;
  const double *sub;    // 3 x S: sub-element x | y | z offset (rz = z - zoff)
  int64_t S;            // sub-elements per element
  const double *el;     // 6 x L: gx gy gz foc steer ((w v) ds)
  const double *freq;   // 2 x K: w_k | b_k  (K = nt/2 + 1; entry 0 unused)
  const double *tw;     // 2 x nt: cos(2 pi m/nt) | sin(2 pi m/nt)
  int64_t nt;
  double cp, dt, dt_s, alpha;   // alpha in Np/(m Hz)
  double *spec;         // No x K x 2 (re, im), zeroed before the launch
  void *h;  int32_t out_f64;  int32_t chunks;
  int32_t bin_blocks;   // the K bins in blocks of kBinBlock, one work-group each
  Status *status;
  int64_t item_begin;
};

// Spectrum bins one work-group accumulates in LDS (two FP64 accumulators
// each, beside the 256-contribution staging): nt up to 2 (kBinBlock - 1) in
// one block, longer ones over several.
constexpr int64_t kBinBlock = 3584;

__device__ __forceinline__ int find_att_problem(const AttDev *P, int n, int64_t item)
{
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (P[mid].item_begin <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ unsigned wave_sum_u(unsigned v)
{
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Spectrum bins [k0, k1) of one (point, chunk): sum over the chunk's
// contributions of ai * X[k].  A work item is (point, chunk, bin block).
__global__ void __launch_bounds__(kThreads)
att_spectrum_kernel(const AttDev *__restrict__ probs, int n_probs)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int64_t item = blockIdx.x;
  const int pi = find_att_problem(probs, n_probs, item);
  const AttDev &P = probs[pi];
  const int64_t local = item - P.item_begin;
  const int64_t per_point = (int64_t) P.chunks * P.bin_blocks;
  const int64_t n = local / per_point;
  const int64_t rest = local - n * per_point;
  const int c = (int) (rest / P.bin_blocks);
  const int bb = (int) (rest - (int64_t) c * P.bin_blocks);
  const int64_t K = P.nt / 2 + 1;
  const int64_t k0 = (int64_t) bb * kBinBlock, k1 = std::min<int64_t>(K, k0 + kBinBlock);
  const int64_t nk = k1 - k0;
  const int64_t total = (int64_t) P.L * P.S;
  const int64_t c0 = (int64_t) c * total / P.chunks, c1 = (int64_t) (c + 1) * total / P.chunks;

  double *acc_re = reinterpret_cast<double *>(smem);        // nk
  double *acc_im = acc_re + nk;                             // nk
  double *st_r = acc_im + nk;                               // kThreads: r [m]
  double *st_t = st_r + kThreads;                           //           it * dt [s]
  double *st_a = st_t + kThreads;                           //           ai (0: skipped)
  for (int64_t q = threadIdx.x; q < nk; q += kThreads) { acc_re[q] = 0.0; acc_im[q] = 0.0; }

  gdouble *Ro = (gdouble *) P.Ro;
  const double xo = Ro[n], yo = Ro[n + P.No], zo = Ro[n + 2 * P.No];
  const double dlay = P.delay_multi ? ((gdouble *) P.delay)[n] : ((gdouble *) P.delay)[0];
  gdouble *el = (gdouble *) P.el;
  gdouble *sub = (gdouble *) P.sub;
  gdouble *wk = (gdouble *) P.freq;
  gdouble *bk = wk + K;
  const int L = P.L;
  const double ntd = (double) P.nt;
  const double alpha = P.alpha;
  unsigned oob = 0;
  bool oob_last = false;

  for (int64_t cb = c0; cb < c1; cb += kThreads) {
    const int nb = (int) std::min<int64_t>(kThreads, c1 - cb);
    __syncthreads();                        // the previous batch is consumed
    if ((int) threadIdx.x < nb) {
      const int64_t f = cb + threadIdx.x;
      const int e = (int) (f / P.S);
      const int64_t s = f - (int64_t) e * P.S;
      const double xe = xo - el[e], ye = yo - el[L + e], ze = zo - el[2 * L + e];
      const double d = (dlay - el[3 * L + e]) - el[4 * L + e];
      const double rx = xe - sub[s];
      const double ry = ye - sub[P.S + s];
      const double rz = ze - sub[2 * P.S + s];
      const double r = __builtin_sqrt((rx * rx + ry * ry) + rz * rz);
      double ai = el[5 * L + e] / (2 * M_PI * r);
      ai /= P.dt;
      ai *= 1000.0;
      const double t = (r * 1000.0) / P.cp;
      const double u = __builtin_rint((t - d) / P.dt);
      if (u >= 0.0 && u < ntd) {
        st_r[threadIdx.x] = r * 1.0e-3;
        st_t[threadIdx.x] = u * P.dt_s;
        st_a[threadIdx.x] = ai;
      } else {
        st_a[threadIdx.x] = 0.0;
        oob++;
        oob_last |= (e == L - 1);
      }
    }
    __syncthreads();
    for (int64_t q = threadIdx.x; q < nk; q += kThreads) {
      const int64_t k = k0 + q;
      double re = acc_re[q], im = acc_im[q];
      if (k == 0) {
        for (int j = 0; j < nb; j++) re += st_a[j];     // X[0] = 1
      } else {
        const double w = wk[k], b = bk[k];
        for (int j = 0; j < nb; j++) {
          const double ai = st_a[j];
          if (ai == 0.0) continue;
          const double r = st_r[j], tt = st_t[j];
          const double x1 = exp(-r * alpha * w / (2.0 * M_PI));
          const double arg = -w * r * alpha / (M_PI * M_PI) * b - tt * w;
          double sn, cs;
          sincos(arg, &sn, &cs);
          re += ai * (x1 * cs);
          im += ai * (x1 * sn);
        }
      }
      acc_re[q] = re;
      acc_im[q] = im;
    }
  }
  __syncthreads();
  double *spec = P.spec + n * K * 2;
  for (int64_t q = threadIdx.x; q < nk; q += kThreads) {
    const int64_t k = k0 + q;
    if (P.chunks == 1) {
      spec[2 * k] = acc_re[q];
      spec[2 * k + 1] = acc_im[q];
    } else {
      unsafeAtomicAdd(spec + 2 * k, acc_re[q]);
      unsafeAtomicAdd(spec + 2 * k + 1, acc_im[q]);
    }
  }
  // Every bin block sees the same contributions: the first one counts them.
  const unsigned wsum = bb == 0 ? wave_sum_u(oob) : 0u;
  const bool any_last = __any(oob_last);
  if ((threadIdx.x & 63) == 0 && wsum) {
    atomicOr(&P.status->any_oob, 1u);
    if (any_last) atomicOr(&P.status->last_elem_oob, 1u);
    atomicAdd(&P.status->oob, (unsigned long long) wsum);
  }
}

// Samples of one point from its spectrum: the complex-to-real inverse DFT
//   x[j] = (X0 + 2 sum_{k=1}^{kmax} (Re X_k cos(2 pi jk/n) - Im X_k sin(2 pi jk/n))
//          + [n even] Re X_{n/2} (-1)^j) / n.
__global__ void __launch_bounds__(kThreads)
att_c2r_kernel(const AttDev *__restrict__ probs, int n_probs, const int64_t *point_begin)
{
  const int64_t g = blockIdx.x;
  int lo = 0, hi = n_probs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (point_begin[mid] <= g) lo = mid; else hi = mid - 1;
  }
  const AttDev &P = probs[lo];
  const int64_t n = g - point_begin[lo];
  const int64_t nt = P.nt, K = nt / 2 + 1;
  const int64_t kmax = (nt % 2 == 0) ? K - 2 : K - 1;
  gdouble *spec = (gdouble *) (P.spec + n * K * 2);
  gdouble *cs = (gdouble *) P.tw;
  gdouble *sn = cs + nt;
  const double inv = (double) nt;
  for (int64_t j = threadIdx.x; j < nt; j += kThreads) {
    double acc = spec[0];
    int64_t m = 0;                          // (j * k) mod nt
    for (int64_t k = 1; k <= kmax; k++) {
      m += j;
      if (m >= nt) m -= nt;
      acc += 2.0 * (spec[2 * k] * cs[m] - spec[2 * k + 1] * sn[m]);
    }
    if (nt % 2 == 0) acc += (j % 2 == 0) ? spec[2 * (K - 1)] : -spec[2 * (K - 1)];
    const double x = acc / inv;
    if (P.out_f64) reinterpret_cast<double *>(P.h)[n * nt + j] = x;
    else reinterpret_cast<float *>(P.h)[n * nt + j] = (float) x;
  }
}

__global__ void att_zero_kernel(double *p, int64_t n)
{
  for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t) gridDim.x * blockDim.x)
    p[i] = 0.0;
}

__global__ void att_zero_status(Status *st, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) { st[i].any_oob = 0u; st[i].last_elem_oob = 0u; st[i].oob = 0ull; }
}

// Sub-element offsets of one element in the reference's walk order.
void sub_elements(const SirProblem &p, std::vector<double> &x, std::vector<double> &y,
                  std::vector<double> &z)
{
  x.clear(); y.clear(); z.clear();
  if (p.shape == SIR_CIRC) {
    std::vector<double> ys, x0, cnt;
    disc_rows(p.R, p.dx, p.dy, ys, x0, cnt);
    for (size_t r = 0; r < ys.size(); r++) {
      double xs = x0[r];
      for (int64_t i = 0; i < (int64_t) cnt[r]; i++) {
        x.push_back(xs);
        y.push_back(ys[r]);
        z.push_back(0.0);
        xs += p.dx;
      }
    }
    return;
  }
  std::vector<double> xs = subgrid(p.a, p.dx), ys, zr;
  if (p.shape == SIR_CYLIND) {
    cylinder_rows(p.b, p.R, p.dy, ys, zr);
  } else {
    ys = subgrid(p.b, p.dy);
    zr.assign(ys.size(), 0.0);
  }
  for (size_t r = 0; r < ys.size(); r++)
    for (double v : xs) {
      x.push_back(v);
      y.push_back(ys[r]);
      z.push_back(zr[r]);
    }
}

} // namespace

void run_sir_att(const std::vector<SirProblem> &problems, double alpha_db, int out_dtype,
                 hipStream_t s, int32_t *codes, int64_t *oobs)
{
  require(out_dtype == DREAM_HIP_F64 || out_dtype == DREAM_HIP_F32,
          "SIR: output precision must be DREAM_HIP_F64 or DREAM_HIP_F32");
  const size_t P = problems.size();
  if (P == 0) return;
  int dev = 0, cus = 256;
  DREAM_HIP_CHECK(hipGetDevice(&dev));
  DREAM_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t target = 8 * (int64_t) std::max(cus, 1);
  const double alpha = alpha_db / (8.686 * 10000.0);

  std::vector<double> tables;
  std::vector<AttDev> dev_p(P);
  std::vector<size_t> off_sub(P), off_el(P), off_freq(P), off_tw(P);
  std::vector<int64_t> spec_off(P), point_begin(P);
  int64_t spec_total = 0, points = 0, items = 0;
  size_t max_K = 1;
  for (size_t i = 0; i < P; i++) {
    const SirProblem &p = problems[i];
    const int64_t L = (int64_t) (p.G.size() / 3);
    require(L >= 1 && p.nt >= 1 && p.No >= 0, "SIR: need elements, nt >= 1 and No >= 0");
    require(p.dx > 0 && p.dy > 0 && p.dt > 0, "SIR: dx, dy and dt must be positive");
    require(p.No == 0 || (p.Ro && p.delay && p.h), "SIR: null Ro, delay or h");
    std::vector<double> x, y, z;
    sub_elements(p, x, y, z);
    const int64_t S = (int64_t) x.size();
    off_sub[i] = tables.size();
    tables.insert(tables.end(), x.begin(), x.end());
    tables.insert(tables.end(), y.begin(), y.end());
    tables.insert(tables.end(), z.begin(), z.end());
    std::vector<double> el = element_table(p);
    off_el[i] = tables.size();
    tables.insert(tables.end(), el.begin(), el.end());
    // Frequency tables with the reference's expressions (src/attenuation.cc:84-110).
    const int64_t nt = p.nt, K = nt / 2 + 1;
    const double dt_s = p.dt * 1.0e-6, Fs = 1 / dt_s, dw = 2.0 * M_PI / double(nt);
    const double a1 = dt_s * 0.95 / M_PI;
    off_freq[i] = tables.size();
    std::vector<double> w((size_t) K, 0.0), b((size_t) K, 0.0);
    for (int64_t k = 1; k < K; k++) {
      w[(size_t) k] = (double(k) * dw) * Fs;
      b[(size_t) k] = std::log(1.0 / (a1 * w[(size_t) k]));
    }
    tables.insert(tables.end(), w.begin(), w.end());
    tables.insert(tables.end(), b.begin(), b.end());
    off_tw[i] = tables.size();
    for (int64_t m = 0; m < nt; m++) tables.push_back(std::cos(2.0 * M_PI * (double) m / (double) nt));
    for (int64_t m = 0; m < nt; m++) tables.push_back(std::sin(2.0 * M_PI * (double) m / (double) nt));

    AttDev &d = dev_p[i];
    std::memset(&d, 0, sizeof d);
    d.Ro = p.Ro;  d.No = p.No;
    d.delay = p.delay;  d.delay_multi = p.delay_type != 0;
    d.L = (int32_t) L;  d.S = S;
    d.nt = nt;  d.cp = p.cp;  d.dt = p.dt;  d.dt_s = dt_s;  d.alpha = alpha;
    d.h = p.h;  d.out_f64 = out_dtype == DREAM_HIP_F64;
    const int64_t total = L * S;
    int64_t chunks = 1;
    if (p.No > 0 && p.No < target) chunks = (target + p.No - 1) / p.No;
    chunks = std::max<int64_t>(1, std::min<int64_t>(chunks, (total + kThreads - 1) / kThreads));
    require(chunks < (1ll << 31), "SIR: too many chunks");
    d.chunks = (int32_t) chunks;
    d.bin_blocks = (int32_t) ((K + kBinBlock - 1) / kBinBlock);
    d.item_begin = items;
    items += (S > 0) ? p.No * chunks * d.bin_blocks : 0;
    spec_off[i] = spec_total;
    spec_total += p.No * K * 2;
    point_begin[i] = points;
    points += p.No;
    max_K = std::max<size_t>(max_K, (size_t) std::min<int64_t>(K, kBinBlock));
  }
  const size_t lds = (2 * max_K + 3 * kThreads) * sizeof(double);

  double *d_tables = nullptr, *d_spec = nullptr;
  AttDev *d_probs = nullptr;
  Status *d_status = nullptr;
  int64_t *d_pb = nullptr;
  auto release = [&]() {
    (void) hipFree(d_tables);  (void) hipFree(d_spec);  (void) hipFree(d_probs);
    (void) hipFree(d_status);  (void) hipFree(d_pb);
  };
  try {
    DREAM_HIP_CHECK(hipMalloc(&d_tables, std::max<size_t>(tables.size(), 1) * sizeof(double)));
    DREAM_HIP_CHECK(hipMalloc(&d_spec, (size_t) std::max<int64_t>(spec_total, 1) * sizeof(double)));
    DREAM_HIP_CHECK(hipMalloc(&d_probs, P * sizeof(AttDev)));
    DREAM_HIP_CHECK(hipMalloc(&d_status, P * sizeof(Status)));
    DREAM_HIP_CHECK(hipMalloc(&d_pb, P * sizeof(int64_t)));
    for (size_t i = 0; i < P; i++) {
      AttDev &d = dev_p[i];
      d.sub = d_tables + off_sub[i];
      d.el = d_tables + off_el[i];
      d.freq = d_tables + off_freq[i];
      d.tw = d_tables + off_tw[i];
      d.spec = d_spec + spec_off[i];
      d.status = d_status + i;
    }
    DREAM_HIP_CHECK(hipMemcpyAsync(d_tables, tables.data(), tables.size() * sizeof(double),
                                   hipMemcpyHostToDevice, s));
    DREAM_HIP_CHECK(hipMemcpyAsync(d_probs, dev_p.data(), P * sizeof(AttDev),
                                   hipMemcpyHostToDevice, s));
    DREAM_HIP_CHECK(hipMemcpyAsync(d_pb, point_begin.data(), P * sizeof(int64_t),
                                   hipMemcpyHostToDevice, s));
    att_zero_status<<<(unsigned) ((P + 255) / 256), 256, 0, s>>>(d_status, (int) P);
    if (spec_total) {
      const int blocks = (int) std::min<int64_t>((spec_total + 255) / 256, 4096);
      att_zero_kernel<<<blocks, 256, 0, s>>>(d_spec, spec_total);
    }
    DREAM_HIP_CHECK(hipGetLastError());
    if (items > 0) {
      require(items < (1ll << 31), "SIR with attenuation: too many work items");
      hipLaunchKernelGGL(att_spectrum_kernel, dim3((unsigned) items), dim3(kThreads), lds, s,
                         d_probs, (int) P);
      DREAM_HIP_CHECK(hipGetLastError());
    }
    if (points > 0) {
      require(points < (1ll << 31), "SIR with attenuation: too many points");
      hipLaunchKernelGGL(att_c2r_kernel, dim3((unsigned) points), dim3(kThreads), 0, s, d_probs,
                         (int) P, d_pb);
      DREAM_HIP_CHECK(hipGetLastError());
    }
    std::vector<Status> st(P);
    DREAM_HIP_CHECK(hipMemcpyAsync(st.data(), d_status, P * sizeof(Status),
                                   hipMemcpyDeviceToHost, s));
    DREAM_HIP_CHECK(hipStreamSynchronize(s));
    for (size_t i = 0; i < P; i++) {
      const int lv = problems[i].err_level;
      int code = E_NONE;
      if (lv == L_STOP || lv == L_PSTOP) {
        if (st[i].any_oob) code = E_OOB;
      } else if (st[i].last_elem_oob) {
        code = level_code(lv);
      }
      if (codes) codes[i] = code;
      if (oobs) oobs[i] = (int64_t) st[i].oob;
    }
  } catch (...) {
    (void) hipStreamSynchronize(s);
    release();
    throw;
  }
  release();
}

} // namespace dream
