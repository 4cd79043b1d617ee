// conv.hip -- batched linear convolution (fftconv_p, sum_fftconv) on gfx950.
//
// Reference: FFTConvP::run (extra_src/fftconv_p.cc:154-403, per column
// fftconv, extra_src/fftconv.cc:32-106) and sum_fftconv (add_fftconv,
// extra_src/fftconv.cc:119-209; extra_src/oct_sum_fftconv.cc).  Both are
// linear convolutions computed by the reference through zero-padded FFTs of
// length M_A + M_B - 1; the result is the same sum
//     y[k] = sum_l sum_i b_l[i] a_l[k - i],   k = 0 .. M_A + M_B - 2,
// which is evaluated here directly in the data type (no FFT round trip).
//
// Mapping.  One work-group per (output column, tile of kTile outputs).  The
// filter b is consumed in chunks of kChunk taps; for each chunk the taps and
// the window of a they meet are staged in LDS, and each thread accumulates
// kR outputs spaced kThreads apart, so consecutive lanes read consecutive
// LDS words (no bank conflicts) and the taps are broadcast.

#include "conv.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

namespace dream {

namespace {

constexpr int kThreads = 256;
constexpr int kR = 4;
constexpr int kTile = kThreads * kR;      // outputs per work-group
constexpr int kChunk = 256;               // taps staged per step

template <typename T>
__global__ void __launch_bounds__(kThreads)
conv_sum_kernel(T *__restrict__ Y, const T *__restrict__ A, int64_t MA, int64_t L,
                int64_t a_col, int64_t a_layer, const T *__restrict__ B, int64_t MB,
                int64_t b_col, int64_t b_layer, int mode)
{
  __shared__ T sb[kChunk];
  __shared__ T sa[kTile + kChunk];
  const int64_t n = blockIdx.y;
  const int64_t k0 = (int64_t) blockIdx.x * kTile;
  const int64_t MY = MA + MB - 1;
  const int t = threadIdx.x;
  T acc[kR];
#pragma unroll
  for (int r = 0; r < kR; r++) acc[r] = T(0);

  // Taps i that meet an output of this tile: k - i in [0, MA).
  const int64_t i_lo = std::max<int64_t>(0, k0 - MA + 1);
  const int64_t i_hi = std::min<int64_t>(MB - 1, k0 + kTile - 1);
  for (int64_t l = 0; l < L; l++) {
    const T *a = A + l * a_layer + n * a_col;
    const T *b = B + l * b_layer + n * b_col;
    for (int64_t i0 = i_lo; i0 <= i_hi; i0 += kChunk) {
      __syncthreads();                            // the previous chunk is consumed
      for (int q = t; q < kChunk; q += kThreads) {
        const int64_t i = i0 + q;
        sb[q] = i <= i_hi ? b[i] : T(0);
      }
      // a window: j = k - i for k in [k0, k0 + kTile), i in [i0, i0 + kChunk).
      const int64_t j0 = k0 - i0 - (kChunk - 1);
      for (int q = t; q < kTile + kChunk - 1; q += kThreads) {
        const int64_t j = j0 + q;
        sa[q] = (j >= 0 && j < MA) ? a[j] : T(0);
      }
      __syncthreads();
      const int cnt = (int) std::min<int64_t>(kChunk, i_hi - i0 + 1);
      // Output k0 + r*kThreads + t reads sa[r*kThreads + t - ii + kChunk - 1].
      const int base = t + kChunk - 1;
      int ii = 0;
      for (; ii + 4 <= cnt; ii += 4) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const T bv = sb[ii + u];
#pragma unroll
          for (int r = 0; r < kR; r++) acc[r] += bv * sa[r * kThreads + base - (ii + u)];
        }
      }
      for (; ii < cnt; ii++) {
        const T bv = sb[ii];
#pragma unroll
        for (int r = 0; r < kR; r++) acc[r] += bv * sa[r * kThreads + base - ii];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kR; r++) {
    const int64_t k = k0 + (int64_t) r * kThreads + t;
    if (k >= MY) continue;
    T *y = Y + n * MY + k;
    if (mode == CONV_SUM) *y += acc[r];
    else if (mode == CONV_NEG) *y -= acc[r];
    else *y = acc[r];
  }
}

// ------------------------------------------------------------------ FFT path
//
// For long filters the same sums go through FFTs of a power-of-two length
// NF >= M_A + M_B - 1 (zero padding makes the cyclic convolution linear, as
// in the reference's zero-padded FFTs).  One work-group per pair of output
// columns holds NF complex samples in LDS: column 2p is the real part, 2p+1
// the imaginary part, so one complex transform serves two real columns.
// Forward transforms are radix-2 decimation in frequency (natural order in,
// bit-reversed out), the product is formed in that order, and the inverse is
// decimation in time (bit-reversed in, natural out): no permutation passes.
// A filter shared by the columns is transformed once per layer by the same
// code (fft_filter_kernel); per-column filters are transformed in the
// work-group, two at a time, and separated by conjugate symmetry, as are the
// two data columns then: C = A1 B1 + i A2 B2.  Layers add in the frequency
// domain before the one inverse transform.  Twiddles exp(-2 pi i j / NF),
// j < NF/2, come from a host table (correctly rounded cos and sin).

template <typename T>
struct Cx {
  T x, y;
};
template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b)
{
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ Cx<T> cmulc(Cx<T> a, Cx<T> b)            // a * conj(b)
{
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}

__device__ __forceinline__ unsigned bitrev(unsigned v, int bits)
{
  return __brev(v) >> (32 - bits);
}

// In place, natural order in, bit-reversed order out.
template <typename T>
__device__ void fft_dif(Cx<T> *X, int NF, const Cx<T> *__restrict__ tw)
{
  for (int h = NF >> 1, ts = 1; h >= 1; h >>= 1, ts <<= 1) {
    for (int b = threadIdx.x; b < (NF >> 1); b += blockDim.x) {
      const int j = b & (h - 1);
      const int i0 = ((b - j) << 1) + j, i1 = i0 + h;
      const Cx<T> u = X[i0], v = X[i1];
      X[i0] = {u.x + v.x, u.y + v.y};
      X[i1] = cmul<T>({u.x - v.x, u.y - v.y}, tw[j * ts]);
    }
    __syncthreads();
  }
}

// In place, bit-reversed order in, natural order out; the inverse transform
// without the 1/NF factor.
template <typename T>
__device__ void ifft_dit(Cx<T> *X, int NF, const Cx<T> *__restrict__ tw)
{
  for (int h = 1, ts = NF >> 1; h < NF; h <<= 1, ts >>= 1) {
    for (int b = threadIdx.x; b < (NF >> 1); b += blockDim.x) {
      const int j = b & (h - 1);
      const int i0 = ((b - j) << 1) + j, i1 = i0 + h;
      const Cx<T> u = X[i0], t = cmulc<T>(X[i1], tw[j * ts]);
      X[i0] = {u.x + t.x, u.y + t.y};
      X[i1] = {u.x - t.x, u.y - t.y};
    }
    __syncthreads();
  }
}

// Load two real columns (b: may be null) of length M into X, zero-padded.
template <typename T>
__device__ void load_pair(Cx<T> *X, int NF, const T *__restrict__ a, const T *__restrict__ b,
                          int64_t M)
{
  for (int k = threadIdx.x; k < NF; k += blockDim.x)
    X[k] = {k < M ? a[k] : T(0), (b && k < M) ? b[k] : T(0)};
  __syncthreads();
}

// Spectrum of filter l (shared by the columns): FBR[l] = FFT(b_l), bit-reversed order.
template <typename T>
__global__ void __launch_bounds__(kThreads)
fft_filter_kernel(const T *__restrict__ B, int64_t MB, int64_t b_layer, int NF,
                  const Cx<T> *__restrict__ tw, Cx<T> *__restrict__ FBR)
{
  extern __shared__ __align__(16) unsigned char smem[];
  Cx<T> *X = reinterpret_cast<Cx<T> *>(smem);
  const int64_t l = blockIdx.x;
  load_pair<T>(X, NF, B + l * b_layer, nullptr, MB);
  fft_dif<T>(X, NF, tw);
  for (int k = threadIdx.x; k < NF; k += blockDim.x) FBR[l * NF + k] = X[k];
}

// Split W = FFT(p + i q) of two real sequences (bit-reversed storage):
// P[k] = (W[k] + conj W[-k]) / 2, Q[k] = (W[k] - conj W[-k]) / (2i), and
// form C[k] (+)= P[k] * U[k] + i Q[k] * V[k] at k and at -k.
template <typename T>
__device__ __forceinline__ void pair_terms(const Cx<T> *W, int NF, int bits, int k, Cx<T> &P,
                                           Cx<T> &Q)
{
  const Cx<T> w = W[bitrev((unsigned) k, bits)];
  const Cx<T> v = W[bitrev((unsigned) ((NF - k) & (NF - 1)), bits)];
  P = {(w.x + v.x) * T(0.5), (w.y - v.y) * T(0.5)};
  Q = {(w.y + v.y) * T(0.5), (v.x - w.x) * T(0.5)};
}

// One work-group per pair of output columns (2p, 2p + 1).
//   shared: the layers' filter spectra FBR (null: per-column filters, then
//   transformed here).  LDS: X (NF), then S (NF) when L > 1 or the filters
//   are per column, then Wb (NF) for per-column filters.
template <typename T>
__global__ void __launch_bounds__(kThreads)
conv_fft_kernel(T *__restrict__ Y, const T *__restrict__ A, int64_t MA, int64_t N, int64_t L,
                int64_t a_col, int64_t a_layer, const T *__restrict__ B, int64_t MB,
                int64_t b_col, int64_t b_layer, int mode, int NF, int bits,
                const Cx<T> *__restrict__ tw, const Cx<T> *__restrict__ FBR)
{
  extern __shared__ __align__(16) unsigned char smem[];
  Cx<T> *X = reinterpret_cast<Cx<T> *>(smem);
  const bool per_col = FBR == nullptr;
  const bool sep = per_col || L > 1;             // accumulate in S
  Cx<T> *S = X + NF;
  Cx<T> *Wb = S + NF;
  const int64_t n1 = 2 * (int64_t) blockIdx.x, n2 = n1 + 1;
  const bool has2 = n2 < N;
  if (sep)
    for (int k = threadIdx.x; k < NF; k += blockDim.x) S[k] = {T(0), T(0)};
  for (int64_t l = 0; l < L; l++) {
    const T *a1 = A + l * a_layer + n1 * a_col;
    load_pair<T>(X, NF, a1, has2 ? a1 + a_col : nullptr, MA);
    fft_dif<T>(X, NF, tw);
    if (!per_col) {
      // One real filter for both columns: (A1 + i A2) B = A1 B + i A2 B.
      const Cx<T> *F = FBR + l * NF;
      for (int k = threadIdx.x; k < NF; k += blockDim.x) {
        const Cx<T> c = cmul<T>(X[k], F[k]);
        if (sep) S[k] = {S[k].x + c.x, S[k].y + c.y};
        else X[k] = c;
      }
      __syncthreads();
      continue;
    }
    const T *b1 = B + l * b_layer + n1 * b_col;
    load_pair<T>(Wb, NF, b1, has2 ? b1 + b_col : nullptr, MB);
    fft_dif<T>(Wb, NF, tw);
    // C[k] = A1[k] B1[k] + i A2[k] B2[k] for k and -k together (one thread).
    for (int k = threadIdx.x; k <= NF / 2; k += blockDim.x) {
      Cx<T> A1, A2, B1, B2;
      pair_terms<T>(X, NF, bits, k, A1, A2);
      pair_terms<T>(Wb, NF, bits, k, B1, B2);
      const Cx<T> p = cmul<T>(A1, B1), q = cmul<T>(A2, B2);
      const unsigned rk = bitrev((unsigned) k, bits);
      S[rk] = {S[rk].x + p.x - q.y, S[rk].y + p.y + q.x};
      const int km = (NF - k) & (NF - 1);
      if (km != k) {
        // At -k: conj(A1 B1) + i conj(A2 B2) (both columns real).
        const unsigned rm = bitrev((unsigned) km, bits);
        S[rm] = {S[rm].x + p.x + q.y, S[rm].y - p.y + q.x};
      }
    }
    __syncthreads();
  }
  Cx<T> *R = sep ? S : X;
  ifft_dit<T>(R, NF, tw);
  const int64_t MY = MA + MB - 1;
  const T inv = T(1) / T(NF);
  for (int64_t k = threadIdx.x; k < MY; k += blockDim.x) {
    const T v1 = R[k].x * inv, v2 = R[k].y * inv;
    T *y1 = Y + n1 * MY + k;
    if (mode == CONV_SUM) *y1 += v1; else if (mode == CONV_NEG) *y1 -= v1; else *y1 = v1;
    if (has2) {
      T *y2 = y1 + MY;
      if (mode == CONV_SUM) *y2 += v2; else if (mode == CONV_NEG) *y2 -= v2; else *y2 = v2;
    }
  }
}

int max_lds_bytes()
{
  static int bytes = 0;
  if (!bytes) {
    int dev = 0, v = 0;
    DREAM_HIP_CHECK(hipGetDevice(&dev));
    DREAM_HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    bytes = v > 0 ? v : 65536;
  }
  return bytes;
}

// DREAM_HIP_CONV = direct | fft forces a path (for comparisons and tests);
// otherwise the FFT path is taken when it fits in LDS and the direct sum
// would cost several times more arithmetic.
template <typename T>
int fft_length_if_chosen(int64_t MA, int64_t MB, int64_t L, bool per_col)
{
  const char *e = std::getenv("DREAM_HIP_CONV");
  const bool force_direct = e && std::strcmp(e, "direct") == 0;
  const bool force_fft = e && std::strcmp(e, "fft") == 0;
  if (force_direct) return 0;
  const int64_t MY = MA + MB - 1;
  int64_t NF = 2;
  while (NF < MY) NF <<= 1;
  const int bufs = per_col ? 3 : (L > 1 ? 2 : 1);
  if ((size_t) bufs * NF * sizeof(Cx<T>) > (size_t) max_lds_bytes() || NF > (1 << 20)) return 0;
  if (force_fft) return (int) NF;
  int bits = 0;
  while ((1ll << bits) < NF) bits++;
  // Per column: direct 2 MA MB L flops; FFT (two columns per transform)
  // about 5 NF log2 NF per transform, (L + 1) transforms per pair (2 L + 1
  // with per-column filters).
  const double direct = 2.0 * (double) MA * (double) MB * (double) L;
  const double fft = 5.0 * (double) NF * bits * (per_col ? 2 * L + 1 : L + 1) / 2.0;
  return direct > 4.0 * fft ? (int) NF : 0;
}

// Device twiddle table exp(-2 pi i j / NF), j < NF/2, per (device, NF,
// type), built once and kept for the life of the process.
template <typename T>
const Cx<T> *twiddles(int NF)
{
  static std::mutex mu;
  static std::map<std::pair<int, int>, void *> cache;
  int dev = 0;
  DREAM_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  void *&p = cache[{dev, NF}];
  if (!p) {
    std::vector<Cx<T>> h((size_t) NF / 2);
    for (int j = 0; j < NF / 2; j++) {
      const double ang = -2.0 * M_PI * (double) j / (double) NF;
      h[(size_t) j] = {(T) std::cos(ang), (T) std::sin(ang)};
    }
    void *d = nullptr;
    DREAM_HIP_CHECK(hipMalloc(&d, h.size() * sizeof(Cx<T>)));
    DREAM_HIP_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(Cx<T>), hipMemcpyHostToDevice));
    p = d;
  }
  return static_cast<const Cx<T> *>(p);
}

template <typename T>
void launch_conv_fft(T *Y, const T *A, int64_t MA, int64_t N, int64_t L, int64_t a_col,
                     int64_t a_layer, const T *B, int64_t MB, int64_t b_col, int64_t b_layer,
                     int mode, int NF, hipStream_t s)
{
  int bits = 0;
  while ((1 << bits) < NF) bits++;
  const Cx<T> *tw = twiddles<T>(NF);
  const bool per_col = b_col != 0;
  Cx<T> *F = nullptr;
  if (!per_col) {
    // The layers' filter spectra, stream-ordered scratch.
    DREAM_HIP_CHECK(hipMallocAsync((void **) &F, (size_t) L * NF * sizeof(Cx<T>), s));
    hipLaunchKernelGGL(fft_filter_kernel<T>, dim3((unsigned) L), dim3(kThreads),
                       (size_t) NF * sizeof(Cx<T>), s, B, MB, b_layer, NF, tw, F);
    DREAM_HIP_CHECK(hipGetLastError());
  }
  const int bufs = per_col ? 3 : (L > 1 ? 2 : 1);
  const unsigned pairs = (unsigned) ((N + 1) / 2);
  hipLaunchKernelGGL(conv_fft_kernel<T>, dim3(pairs), dim3(kThreads),
                     (size_t) bufs * NF * sizeof(Cx<T>), s, Y, A, MA, N, L, a_col, a_layer, B,
                     MB, b_col, b_layer, mode, NF, bits, tw, F);
  const hipError_t e = hipGetLastError();
  if (F) (void) hipFreeAsync(F, s);
  DREAM_HIP_CHECK(e);
}

} // namespace

template <typename T>
void launch_conv_sum(T *Y, const T *A, int64_t MA, int64_t N, int64_t L, int64_t a_col,
                     int64_t a_layer, const T *B, int64_t MB, int64_t b_col, int64_t b_layer,
                     int mode, hipStream_t s)
{
  require(MA >= 1 && MB >= 1 && N >= 0 && L >= 1, "conv: need M_A, M_B, L >= 1 and N >= 0");
  require(mode == CONV_EQU || mode == CONV_SUM || mode == CONV_NEG, "conv: unknown mode");
  if (N == 0) return;
  require(N < (1ll << 31), "conv: too many columns");
  const int NF = fft_length_if_chosen<T>(MA, MB, L, b_col != 0);
  if (NF) {
    launch_conv_fft<T>(Y, A, MA, N, L, a_col, a_layer, B, MB, b_col, b_layer, mode, NF, s);
    return;
  }
  const int64_t MY = MA + MB - 1;
  const int64_t tiles = (MY + kTile - 1) / kTile;
  require(tiles < (1ll << 31), "conv: output too long");
  hipLaunchKernelGGL(conv_sum_kernel<T>, dim3((unsigned) tiles, (unsigned) N), dim3(kThreads), 0,
                     s, Y, A, MA, L, a_col, a_layer, B, MB, b_col, b_layer, mode);
  DREAM_HIP_CHECK(hipGetLastError());
}

template void launch_conv_sum<double>(double *, const double *, int64_t, int64_t, int64_t,
                                      int64_t, int64_t, const double *, int64_t, int64_t,
                                      int64_t, int, hipStream_t);
template void launch_conv_sum<float>(float *, const float *, int64_t, int64_t, int64_t, int64_t,
                                     int64_t, const float *, int64_t, int64_t, int64_t, int,
                                     hipStream_t);

} // namespace dream
