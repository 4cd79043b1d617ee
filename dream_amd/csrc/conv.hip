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
