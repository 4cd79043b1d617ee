// das.hip -- delay-and-sum beamforming (SAFT, TFM, RCA) and its adjoint on
// gfx950.
//
// Reference: DAS<T>::das_{saft,tfm,rca_serial_coltx,rca_serial_rowtx}_serial
// (extra_src/das.cc:306-626).  Per pixel (xo,yo,zo) and A-scan:
//   leg(g)  = sqrt((gx-xo)^2 + (gy-yo)^2 + (gz-zo)^2) * (1/cp)      [ms]
//   SAFT    k = trunc(2 (leg_n + delay_ms) Fs_khz)            A-scan n
//   TFM     k = trunc(((leg_i + delay_ms) + leg_j) Fs_khz)    A-scan i*Lr + j
//   RCA     column leg with the stripe-clamped y distance, row leg with the
//           stripe-clamped x distance; A-scan c*Lr + r (column transmit) or
//           r*Lc + c (row transmit); the column leg carries the delay.
//   Im[p]  += Y[k] for 0 <= k < K.
// The geometry type GT (double, or float for the reference's float path)
// computes every delay with the reference's operations in the reference's
// order (no contraction, correctly rounded sqrt), so the sample indices are
// the reference's bit for bit; the data type DT carries samples and sums.
//
// Forward mapping.  One lane per pixel, 64 pixels per wave, 4 independent
// waves per work-group.  All lanes of a wave visit the same A-scan at the
// same time, so neighbouring pixels gather neighbouring samples of one
// A-scan (coalesced along the sample axis).  For two-leg modes the receive
// legs of a chunk of RC receivers live in registers (RC sqrt per pixel per
// chunk), the transmit leg is recomputed once per (transmit, chunk), and the
// inner RC pairs cost one add, one multiply, one truncation and one gather.
//
// Extensions (no reference counterpart): linear interpolation with zero
// outside [0, K), per-element apodization, and the adjoint with respect to
// Y (scatter into per-work-group LDS A-scan accumulators), the apodization
// weights and cp.

#include <algorithm>
#include <cmath>
#include <cstring>

#include "das.hpp"

namespace dream {

namespace {

constexpr int kThreads = 256;
constexpr int kRC = 16;          // receive legs held in registers per chunk

template <typename GT>
__device__ __forceinline__ GT leg3(GT dx, GT dy, GT dz)
{
  return __builtin_sqrt(dx * dx + dy * dy + dz * dz);
}
template <>
__device__ __forceinline__ float leg3<float>(float dx, float dy, float dz)
{
  return __builtin_sqrtf(dx * dx + dy * dy + dz * dz);
}

// Stripe-clamped lateral distance (extra_src/das.cc:477-485).
template <typename GT>
__device__ __forceinline__ GT stripe_gap(GT p, GT lo, GT hi)
{
  if (p >= lo && p <= hi) return GT(0);
  return (p < lo) ? (lo - p) : (hi - p);
}

// delay[us] -> ms in GT, as (T)(delay/1000.0) of extra_src/das.cc:318.
template <typename GT>
__device__ __forceinline__ GT delay_ms_of(GT d)
{
  return (GT) ((double) d / 1000.0);
}

__device__ __forceinline__ float wave_sum_f(float v)
{
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_sum_d(double v)
{
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One pair's sample: nearest-lower (reference) or linear.  Accumulates
// w*value into acc, returns whether the nearest-lower index was in range.
template <typename DT, typename GT, int INTERP>
__device__ __forceinline__ bool sample(const DT *__restrict__ y, int K, GT tdp, GT fs, DT w,
                                       DT &acc)
{
  if (INTERP == DREAM_HIP_NEAREST) {
    const GT x = tdp * fs;
    if (x > GT(-1) && x < (GT) K) {
      acc += w * y[(int) x];
      return true;
    }
    return false;
  } else {
    const GT u = tdp * fs;
    const GT fk = floor(u);
    if (fk > GT(-2) && fk < (GT) K) {
      const int k = (int) fk;
      const DT f = (DT) (u - fk);
      const DT y0 = k >= 0 ? y[k] : DT(0);
      const DT y1 = k + 1 < K ? y[k + 1] : DT(0);
      acc += w * (y0 + f * (y1 - y0));
    }
    return fk >= GT(0) && fk < (GT) K;
  }
}

// Value and slope (y[k+1]-y[k]) of one pair for the adjoint; k via the
// chosen rule, f the interpolation fraction (0 for nearest).
template <typename DT, typename GT, int INTERP>
__device__ __forceinline__ void sample2(const DT *__restrict__ y, int K, GT u, DT &val, DT &slope,
                                        int &k, DT &f)
{
  if (INTERP == DREAM_HIP_NEAREST) {
    f = DT(0);
    slope = DT(0);
    if (u > GT(-1) && u < (GT) K) {
      k = (int) u;
      val = y[k];
    } else {
      k = -2;
      val = DT(0);
    }
  } else {
    const GT fk = floor(u);
    if (fk > GT(-2) && fk < (GT) K) {
      k = (int) fk;
      f = (DT) (u - fk);
      const DT y0 = k >= 0 ? y[k] : DT(0);
      const DT y1 = k + 1 < K ? y[k + 1] : DT(0);
      slope = y1 - y0;
      val = y0 + f * slope;
    } else {
      k = -2;
      f = DT(0);
      val = DT(0);
      slope = DT(0);
    }
  }
}

struct PixelFrame {
  int64_t p;       // pixel index in Ro / Im
  bool valid;
};

__device__ __forceinline__ int64_t pixel_of(const DasArgs &A, const int32_t *perm, int64_t idx)
{
  return perm ? (int64_t) perm[idx] : idx;
}

template <typename GT>
struct Stripes {
  GT y_lo, y_hi, x_lo, x_hi;
};

// --------------------------------------------------------------------- forward

template <typename DT, typename GT, int TYPE, int INTERP, bool APOD>
__global__ void __launch_bounds__(kThreads)
das_fwd_kernel(DasArgs A, const int32_t *__restrict__ perm)
{
  const int64_t idx = (int64_t) blockIdx.x * kThreads + threadIdx.x;
  const int64_t frame = blockIdx.y;
  const bool valid = idx < A.No;
  const int64_t p = valid ? pixel_of(A, perm, idx) : 0;
  const GT *Ro = (const GT *) A.Ro;
  const GT xo = Ro[p], yo = Ro[p + A.No], zo = Ro[p + 2 * A.No];
  const GT *gdel = (const GT *) A.delay;
  const GT dl = A.delay_multi ? delay_ms_of<GT>(gdel[p]) : (GT) A.delay_ms;
  const GT fs = (GT) A.fs_khz, inv_cp = (GT) A.inv_cp;
  const DT *Y = (const DT *) A.Y + frame * A.y_stride;
  const int K = (int) A.K;
  const GT *Gt = (const GT *) A.Gt;
  const GT *Gr = (const GT *) A.Gr;
  const DT *at = (const DT *) A.apod_t;
  const DT *ar = (const DT *) A.apod_r;
  const int Lt = A.Lt, Lr = A.Lr;
  DT acc = DT(0);
  bool oob = false;

  if (TYPE == DAS_SAFT) {
    for (int n = 0; n < Lt; n++) {
      GT t = leg3<GT>(Gt[n] - xo, Gt[n + Lt] - yo, Gt[n + 2 * Lt] - zo) * inv_cp;
      t += dl;
      const GT tdp = t + t;                       // (T)(2.0*t), exact
      const DT w = APOD ? at[n] : DT(1);
      oob |= !sample<DT, GT, INTERP>(Y + (int64_t) n * K, K, tdp, fs, w, acc);
    }
  } else {
    GT gy_c = GT(0), gx_r = GT(0);
    if (TYPE == DAS_RCA_COLTX || TYPE == DAS_RCA_ROWTX) {
      gy_c = stripe_gap<GT>(yo, Gr[Lr], Gr[2 * Lr - 1]);
      gx_r = stripe_gap<GT>(xo, Gt[0], Gt[Lt - 1]);
    }
    for (int j0 = 0; j0 < Lr; j0 += kRC) {
      GT tr[kRC];
#pragma unroll
      for (int jj = 0; jj < kRC; jj++) {
        const int j = min(j0 + jj, Lr - 1);
        if (TYPE == DAS_TFM)
          tr[jj] = leg3<GT>(Gr[j] - xo, Gr[j + Lr] - yo, Gr[j + 2 * Lr] - zo) * inv_cp;
        else
          tr[jj] = leg3<GT>(gx_r, Gr[j + Lr] - yo, Gr[j + 2 * Lr] - zo) * inv_cp;
      }
      const int nj = min(kRC, Lr - j0);
      for (int i = 0; i < Lt; i++) {
        GT tt;
        if (TYPE == DAS_TFM)
          tt = leg3<GT>(Gt[i] - xo, Gt[i + Lt] - yo, Gt[i + 2 * Lt] - zo) * inv_cp;
        else
          tt = leg3<GT>(Gt[i] - xo, gy_c, Gt[i + 2 * Lt] - zo) * inv_cp;
        tt += dl;
        const DT wi = APOD ? at[i] : DT(1);
        DT part = DT(0);
        if (nj == kRC) {
#pragma unroll
          for (int jj = 0; jj < kRC; jj++) {
            const int j = j0 + jj;
            const int64_t a = (TYPE == DAS_RCA_ROWTX) ? (int64_t) j * Lt + i : (int64_t) i * Lr + j;
            const DT w = APOD ? wi * ar[j] : DT(1);
            oob |= !sample<DT, GT, INTERP>(Y + a * K, K, tt + tr[jj], fs, w, part);
          }
        } else {
          for (int jj = 0; jj < nj; jj++) {
            const int j = j0 + jj;
            const int64_t a = (TYPE == DAS_RCA_ROWTX) ? (int64_t) j * Lt + i : (int64_t) i * Lr + j;
            const DT w = APOD ? wi * ar[j] : DT(1);
            GT trj = tr[0];
#pragma unroll
            for (int q = 1; q < kRC; q++) if (q == jj) trj = tr[q];
            oob |= !sample<DT, GT, INTERP>(Y + a * K, K, tt + trj, fs, w, part);
          }
        }
        acc += part;
      }
    }
  }
  if (valid) ((DT *) A.Im)[frame * A.im_stride + p] = acc;
  oob = oob && valid;
  if (__any(oob) && (threadIdx.x & 63) == 0) atomicOr(&A.status->any_oob, 1u);
}

// --------------------------------------------------------------------- adjoint

// Channel-major adjoint: one work-group per (pixel block, A-scan, frame).
// Scatters the A-scan's Y-gradient into LDS, then adds the touched window
// into gY; accumulates sum_p g_p*value (z_t[a]) and sum_p g_p*w*slope*du/dcp
// (z_cp) for the A-scan.
template <typename DT, typename GT, int TYPE, int INTERP, bool APOD>
__global__ void __launch_bounds__(kThreads)
das_adj_channel_kernel(DasArgs A, const int32_t *__restrict__ perm, int64_t pix_per_block,
                       int blocks_per_channel)
{
  extern __shared__ __align__(16) unsigned char smem[];
  DT *gacc = reinterpret_cast<DT *>(smem);
  __shared__ int s_kmin, s_kmax;
  __shared__ double s_red[2][kThreads / 64];
  const int K = (int) A.K;
  const int64_t a = blockIdx.y;                 // A-scan
  const int64_t frame = blockIdx.z;
  const int blk = blockIdx.x;
  const int Lt = A.Lt, Lr = A.Lr;
  int i, j = 0;
  if (TYPE == DAS_SAFT) { i = (int) a; }
  else if (TYPE == DAS_RCA_ROWTX) { j = (int) (a / Lt); i = (int) (a - (int64_t) j * Lt); }
  else { i = (int) (a / Lr); j = (int) (a - (int64_t) i * Lr); }
  const bool want_gy = A.gY != nullptr;

  if (want_gy) for (int k = threadIdx.x; k < K; k += kThreads) gacc[k] = DT(0);
  if (threadIdx.x == 0) { s_kmin = K; s_kmax = -1; }
  __syncthreads();

  const GT *Ro = (const GT *) A.Ro;
  const GT *Gt = (const GT *) A.Gt;
  const GT *Gr = (const GT *) A.Gr;
  const GT fs = (GT) A.fs_khz, inv_cp = (GT) A.inv_cp, dl = (GT) A.delay_ms, cp = (GT) A.cp;
  const DT *y = (const DT *) A.Y + frame * A.y_stride + a * K;
  const DT *g = (const DT *) A.g + frame * A.im_stride;
  const DT *at = (const DT *) A.apod_t;
  const DT *ar = (const DT *) A.apod_r;
  const DT w = APOD ? (TYPE == DAS_SAFT ? at[i] : at[i] * ar[j]) : DT(1);
  GT y_lo = 0, y_hi = 0, x_lo = 0, x_hi = 0;
  if (TYPE == DAS_RCA_COLTX || TYPE == DAS_RCA_ROWTX) {
    y_lo = Gr[Lr]; y_hi = Gr[2 * Lr - 1]; x_lo = Gt[0]; x_hi = Gt[Lt - 1];
  }
  const int64_t p0 = (int64_t) blk * pix_per_block;
  const int64_t p1 = min(p0 + pix_per_block, A.No);
  double z = 0.0, zc = 0.0;
  int kmin = K, kmax = -1;
  for (int64_t idx = p0 + threadIdx.x; idx < p1; idx += kThreads) {
    const int64_t p = pixel_of(A, perm, idx);
    const GT xo = Ro[p], yo = Ro[p + A.No], zo = Ro[p + 2 * A.No];
    const DT gp = g[p];
    GT u, tsum;   // tsum: total travel time without the delay, for du/dcp
    if (TYPE == DAS_SAFT) {
      GT t = leg3<GT>(Gt[i] - xo, Gt[i + Lt] - yo, Gt[i + 2 * Lt] - zo) * inv_cp;
      tsum = t + t;
      t += dl;
      u = (t + t) * fs;
    } else {
      GT tt, tr;
      if (TYPE == DAS_TFM) {
        tt = leg3<GT>(Gt[i] - xo, Gt[i + Lt] - yo, Gt[i + 2 * Lt] - zo) * inv_cp;
        tr = leg3<GT>(Gr[j] - xo, Gr[j + Lr] - yo, Gr[j + 2 * Lr] - zo) * inv_cp;
      } else {
        tt = leg3<GT>(Gt[i] - xo, stripe_gap<GT>(yo, y_lo, y_hi), Gt[i + 2 * Lt] - zo) * inv_cp;
        tr = leg3<GT>(stripe_gap<GT>(xo, x_lo, x_hi), Gr[j + Lr] - yo, Gr[j + 2 * Lr] - zo) * inv_cp;
      }
      tsum = tt + tr;
      tt += dl;
      u = (tt + tr) * fs;
    }
    DT val, slope, f;
    int k;
    sample2<DT, GT, INTERP>(y, K, u, val, slope, k, f);
    z += (double) (gp * val);
    // d u / d cp = -(travel time without delay) * fs / cp
    zc += (double) (gp * w * slope) * (-(double) tsum * (double) fs / (double) cp);
    if (want_gy && k > -2) {
      const DT gw = gp * w;
      if (INTERP == DREAM_HIP_NEAREST) {
        atomicAdd(&gacc[k], gw);
        kmin = min(kmin, k);
        kmax = max(kmax, k);
      } else {
        if (k >= 0) atomicAdd(&gacc[k], gw * (DT(1) - f));
        if (k + 1 < K) atomicAdd(&gacc[k + 1], gw * f);
        kmin = min(kmin, max(k, 0));
        kmax = max(kmax, min(k + 1, K - 1));
      }
    }
  }
  // Block reductions.
  z = wave_sum_d(z);
  zc = wave_sum_d(zc);
  if ((threadIdx.x & 63) == 0) {
    s_red[0][threadIdx.x >> 6] = z;
    s_red[1][threadIdx.x >> 6] = zc;
  }
  if (want_gy) {
    atomicMin(&s_kmin, kmin);
    atomicMax(&s_kmax, kmax);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double zs = 0.0, cs = 0.0;
    for (int q = 0; q < kThreads / 64; q++) { zs += s_red[0][q]; cs += s_red[1][q]; }
    atomicAdd(&((double *) A.z_t)[a], zs);
    atomicAdd((double *) A.z_cp, cs);
  }
  if (want_gy) {
    DT *gy = (DT *) A.gY + frame * A.y_stride + a * K;
    if (blocks_per_channel == 1) {
      for (int k = threadIdx.x; k < K; k += kThreads) gy[k] = gacc[k];
    } else {
      for (int k = s_kmin + threadIdx.x; k <= s_kmax; k += kThreads) {
        const DT v = gacc[k];
        if (v != DT(0)) unsafeAtomicAdd(&gy[k], v);
      }
    }
  }
}

// Pixel-major adjoint for two-leg modes without gY: accumulates
// z_t[i] = sum_p g_p sum_j w_j v(p,i,j), z_r[j] = sum_p g_p sum_i w_i v(p,i,j)
// and z_cp, in LDS per work-group, flushed once.  Grid-stride over
// (frame, 256-pixel tile).
template <typename DT, typename GT, int TYPE, int INTERP, bool APOD>
__global__ void __launch_bounds__(kThreads)
das_adj_pixel_kernel(DasArgs A, const int32_t *__restrict__ perm, int64_t n_tiles)
{
  extern __shared__ __align__(16) unsigned char smem[];
  double *zt = reinterpret_cast<double *>(smem);
  double *zr = zt + A.Lt;
  __shared__ double s_cp[kThreads / 64];
  const int Lt = A.Lt, Lr = A.Lr;
  for (int q = threadIdx.x; q < Lt + Lr; q += kThreads) zt[q] = 0.0;
  __syncthreads();
  const GT *Ro = (const GT *) A.Ro;
  const GT *Gt = (const GT *) A.Gt;
  const GT *Gr = (const GT *) A.Gr;
  const GT fs = (GT) A.fs_khz, inv_cp = (GT) A.inv_cp, dl = (GT) A.delay_ms;
  const double cpd = A.cp;
  const DT *at = (const DT *) A.apod_t;
  const DT *ar = (const DT *) A.apod_r;
  const int K = (int) A.K;
  const int64_t tiles_per_frame = (A.No + kThreads - 1) / kThreads;
  const GT y_lo = Gr[Lr], y_hi = Gr[2 * Lr - 1], x_lo = Gt[0], x_hi = Gt[Lt - 1];
  double zcp = 0.0;
  const bool lane0 = (threadIdx.x & 63) == 0;

  for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    const int64_t frame = tile / tiles_per_frame;
    const int64_t idx = (tile - frame * tiles_per_frame) * kThreads + threadIdx.x;
    const bool valid = idx < A.No;
    const int64_t p = valid ? pixel_of(A, perm, idx) : 0;
    const GT xo = Ro[p], yo = Ro[p + A.No], zo = Ro[p + 2 * A.No];
    const DT gp = valid ? ((const DT *) A.g)[frame * A.im_stride + p] : DT(0);
    const DT *Y = (const DT *) A.Y + frame * A.y_stride;
    GT gy_c = GT(0), gx_r = GT(0);
    if (TYPE != DAS_TFM) {
      gy_c = stripe_gap<GT>(yo, y_lo, y_hi);
      gx_r = stripe_gap<GT>(xo, x_lo, x_hi);
    }
    for (int j0 = 0; j0 < Lr; j0 += kRC) {
      GT tr[kRC];
      DT zj[kRC];
#pragma unroll
      for (int jj = 0; jj < kRC; jj++) {
        const int j = min(j0 + jj, Lr - 1);
        if (TYPE == DAS_TFM)
          tr[jj] = leg3<GT>(Gr[j] - xo, Gr[j + Lr] - yo, Gr[j + 2 * Lr] - zo) * inv_cp;
        else
          tr[jj] = leg3<GT>(gx_r, Gr[j + Lr] - yo, Gr[j + 2 * Lr] - zo) * inv_cp;
        zj[jj] = DT(0);
      }
      const int nj = min(kRC, Lr - j0);
      for (int i = 0; i < Lt; i++) {
        GT tt;
        if (TYPE == DAS_TFM)
          tt = leg3<GT>(Gt[i] - xo, Gt[i + Lt] - yo, Gt[i + 2 * Lt] - zo) * inv_cp;
        else
          tt = leg3<GT>(Gt[i] - xo, gy_c, Gt[i + 2 * Lt] - zo) * inv_cp;
        const GT tt0 = tt;
        tt += dl;
        const DT wi = APOD ? at[i] : DT(1);
        DT zi = DT(0);
        DT zc = DT(0);
#pragma unroll
        for (int jj = 0; jj < kRC; jj++) {
          if (jj < nj) {
            const int j = j0 + jj;
            const int64_t a = (TYPE == DAS_RCA_ROWTX) ? (int64_t) j * Lt + i : (int64_t) i * Lr + j;
            const DT wj = APOD ? ar[j] : DT(1);
            DT val, slope, f;
            int k;
            sample2<DT, GT, INTERP>(Y + a * K, K, (tt + tr[jj]) * fs, val, slope, k, f);
            zi += wj * val;
            zj[jj] += wi * val;
            if (INTERP == DREAM_HIP_LINEAR) zc += wi * wj * slope * (DT) (tt0 + tr[jj]);
          }
        }
        const float zsum = wave_sum_f((float) (gp * zi));
        if (lane0 && zsum != 0.0f) atomicAdd(&zt[i], (double) zsum);
        if (INTERP == DREAM_HIP_LINEAR) zcp += (double) (gp * zc);
      }
#pragma unroll
      for (int jj = 0; jj < kRC; jj++) {
        if (jj < nj) {
          const float zsum = wave_sum_f((float) (gp * zj[jj]));
          if (lane0 && zsum != 0.0f) atomicAdd(&zr[j0 + jj], (double) zsum);
        }
      }
    }
  }
  zcp = wave_sum_d(zcp) * (-(double) A.fs_khz / cpd);
  if (lane0) s_cp[threadIdx.x >> 6] = zcp;
  __syncthreads();
  for (int q = threadIdx.x; q < Lt + Lr; q += kThreads) {
    const double v = zt[q];
    if (v != 0.0) atomicAdd(q < Lt ? &((double *) A.z_t)[q] : &((double *) A.z_r)[q - Lt], v);
  }
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < kThreads / 64; q++) s += s_cp[q];
    atomicAdd((double *) A.z_cp, s);
  }
}

// From per-A-scan sums Z[a] (channel kernel) to the apodization gradients:
// SAFT g_t[n] = Z[n]; two-leg g_t[i] = sum_j w_j Z[i,j], g_r[j] = sum_i w_i Z[i,j].
template <typename DT>
__global__ void adj_finalize_channel(int type, int Lt, int Lr, const double *Z, const DT *at,
                                     const DT *ar, const double *zcp, DT *g_t, DT *g_r, DT *g_cp)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (type == DAS_SAFT) {
    if (q < Lt && g_t) g_t[q] = (DT) Z[q];
  } else {
    auto Zij = [&](int i, int j) {
      return Z[type == DAS_RCA_ROWTX ? (int64_t) j * Lt + i : (int64_t) i * Lr + j];
    };
    if (q < Lt && g_t) {
      double s = 0.0;
      for (int j = 0; j < Lr; j++) s += (ar ? (double) ar[j] : 1.0) * Zij(q, j);
      g_t[q] = (DT) s;
    }
    if (q < Lr && g_r) {
      double s = 0.0;
      for (int i = 0; i < Lt; i++) s += (at ? (double) at[i] : 1.0) * Zij(i, q);
      g_r[q] = (DT) s;
    }
  }
  if (q == 0 && g_cp) *g_cp = (DT) *zcp;
}

template <typename DT>
__global__ void adj_finalize_pixel(int Lt, int Lr, const double *zt, const double *zr,
                                   const double *zcp, DT *g_t, DT *g_r, DT *g_cp)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < Lt && g_t) g_t[q] = (DT) zt[q];
  if (q < Lr && g_r) g_r[q] = (DT) zr[q];
  if (q == 0 && g_cp) *g_cp = (DT) *zcp;
}

__global__ void zero_words(unsigned int *p, int64_t n)
{
  for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t) gridDim.x * blockDim.x)
    p[i] = 0u;
}

void zero_async(void *p, size_t bytes, hipStream_t s)
{
  if (!bytes || !p) return;
  const int64_t words = (int64_t) (bytes / 4);
  const int blocks = (int) std::min<int64_t>((words + 255) / 256, 8192);
  zero_words<<<blocks, 256, 0, s>>>(reinterpret_cast<unsigned int *>(p), words);
  DREAM_HIP_CHECK(hipGetLastError());
}

int cu_count()
{
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    DREAM_HIP_CHECK(hipGetDevice(&dev));
    DREAM_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (cus <= 0) cus = 256;
  }
  return cus;
}

// ------------------------------------------------------------------ dispatch

template <typename DT, typename GT, int TYPE, int INTERP, bool APOD>
void fwd_launch(const DasArgs &a, const int32_t *perm, hipStream_t s)
{
  dim3 grid((unsigned) ((a.No + kThreads - 1) / kThreads), (unsigned) a.n_frames);
  hipLaunchKernelGGL((das_fwd_kernel<DT, GT, TYPE, INTERP, APOD>), grid, dim3(kThreads), 0, s,
                     a, perm);
}

template <typename DT, typename GT, int TYPE>
void fwd_dispatch_interp(const DasArgs &a, const int32_t *perm, int interp, bool apod,
                         hipStream_t s)
{
  if (interp == DREAM_HIP_NEAREST) {
    if (apod) fwd_launch<DT, GT, TYPE, DREAM_HIP_NEAREST, true>(a, perm, s);
    else fwd_launch<DT, GT, TYPE, DREAM_HIP_NEAREST, false>(a, perm, s);
  } else {
    if (apod) fwd_launch<DT, GT, TYPE, DREAM_HIP_LINEAR, true>(a, perm, s);
    else fwd_launch<DT, GT, TYPE, DREAM_HIP_LINEAR, false>(a, perm, s);
  }
}

template <typename DT, typename GT>
void fwd_dispatch_type(const DasArgs &a, const int32_t *perm, int type, int interp, bool apod,
                       hipStream_t s)
{
  switch (type) {
  case DAS_SAFT: fwd_dispatch_interp<DT, GT, DAS_SAFT>(a, perm, interp, apod, s); break;
  case DAS_TFM: fwd_dispatch_interp<DT, GT, DAS_TFM>(a, perm, interp, apod, s); break;
  case DAS_RCA_COLTX: fwd_dispatch_interp<DT, GT, DAS_RCA_COLTX>(a, perm, interp, apod, s); break;
  default: fwd_dispatch_interp<DT, GT, DAS_RCA_ROWTX>(a, perm, interp, apod, s); break;
  }
}

template <typename DT, typename GT, int TYPE, int INTERP, bool APOD>
void adj_launch(const DasArgs &a, const int32_t *perm, bool want_gy, hipStream_t s)
{
  const int64_t n_ascans = (TYPE == DAS_SAFT) ? a.Lt : (int64_t) a.Lt * a.Lr;
  if (TYPE == DAS_SAFT || want_gy) {
    // Channel-major: enough blocks per A-scan to fill the device.
    const int64_t target = (int64_t) cu_count() * 8;
    int64_t bpc = std::max<int64_t>(1, target / std::max<int64_t>(n_ascans * a.n_frames, 1));
    bpc = std::min<int64_t>(bpc, std::max<int64_t>(1, (a.No + kThreads - 1) / kThreads));
    const int64_t ppb = (a.No + bpc - 1) / bpc;
    bpc = (a.No + ppb - 1) / std::max<int64_t>(ppb, 1);
    if (bpc < 1) bpc = 1;
    dim3 grid((unsigned) bpc, (unsigned) n_ascans, (unsigned) a.n_frames);
    const size_t lds = want_gy ? (size_t) a.K * sizeof(DT) : 0;
    if (want_gy && bpc > 1) zero_async(a.gY, (size_t) (a.n_frames * a.y_stride) * sizeof(DT), s);
    hipLaunchKernelGGL((das_adj_channel_kernel<DT, GT, TYPE, INTERP, APOD>), grid, dim3(kThreads),
                       lds, s, a, perm, ppb, (int) bpc);
  } else {
    const int64_t tiles = ((a.No + kThreads - 1) / kThreads) * a.n_frames;
    const int grid = (int) std::min<int64_t>(tiles, (int64_t) cu_count() * 4);
    const size_t lds = (size_t) (a.Lt + a.Lr) * sizeof(double);
    hipLaunchKernelGGL((das_adj_pixel_kernel<DT, GT, TYPE, INTERP, APOD>), dim3(grid),
                       dim3(kThreads), lds, s, a, perm, tiles);
  }
}

template <typename DT, typename GT, int TYPE>
void adj_dispatch_interp(const DasArgs &a, const int32_t *perm, int interp, bool apod,
                         bool want_gy, hipStream_t s)
{
  if (interp == DREAM_HIP_NEAREST) {
    if (apod) adj_launch<DT, GT, TYPE, DREAM_HIP_NEAREST, true>(a, perm, want_gy, s);
    else adj_launch<DT, GT, TYPE, DREAM_HIP_NEAREST, false>(a, perm, want_gy, s);
  } else {
    if (apod) adj_launch<DT, GT, TYPE, DREAM_HIP_LINEAR, true>(a, perm, want_gy, s);
    else adj_launch<DT, GT, TYPE, DREAM_HIP_LINEAR, false>(a, perm, want_gy, s);
  }
}

template <typename DT, typename GT>
void adj_dispatch_type(const DasArgs &a, const int32_t *perm, int type, int interp, bool apod,
                       bool want_gy, hipStream_t s)
{
  switch (type) {
  case DAS_SAFT: adj_dispatch_interp<DT, GT, DAS_SAFT>(a, perm, interp, apod, want_gy, s); break;
  case DAS_TFM: adj_dispatch_interp<DT, GT, DAS_TFM>(a, perm, interp, apod, want_gy, s); break;
  case DAS_RCA_COLTX:
    adj_dispatch_interp<DT, GT, DAS_RCA_COLTX>(a, perm, interp, apod, want_gy, s);
    break;
  default: adj_dispatch_interp<DT, GT, DAS_RCA_ROWTX>(a, perm, interp, apod, want_gy, s); break;
  }
}

size_t elem_size(int prec) { return prec == DREAM_HIP_F64 ? 8 : 4; }
size_t geom_size(int prec) { return prec == DREAM_HIP_F32_REF ? 4 : 8; }

} // namespace

void launch_das_forward(const DasArgs &a, int type, int prec, int interp, bool apod,
                        hipStream_t s)
{
  const int32_t *perm = nullptr;
  if (prec == DREAM_HIP_F64) fwd_dispatch_type<double, double>(a, perm, type, interp, apod, s);
  else if (prec == DREAM_HIP_F32) fwd_dispatch_type<float, double>(a, perm, type, interp, apod, s);
  else fwd_dispatch_type<float, float>(a, perm, type, interp, apod, s);
  DREAM_HIP_CHECK(hipGetLastError());
}

void launch_das_adjoint(const DasArgs &a, int type, int prec, int interp, bool apod,
                        bool want_gy, hipStream_t s)
{
  const int32_t *perm = nullptr;
  if (prec == DREAM_HIP_F64) adj_dispatch_type<double, double>(a, perm, type, interp, apod, want_gy, s);
  else if (prec == DREAM_HIP_F32) adj_dispatch_type<float, double>(a, perm, type, interp, apod, want_gy, s);
  else throw Error(DREAM_HIP_EINVAL, "the adjoint is not available for DREAM_HIP_F32_REF handles");
  DREAM_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ handle

DasHandle::DasHandle(int type, int64_t K, int64_t No, int64_t Lt, int64_t Lr, int precision)
  : type_(Lr == 0 ? DAS_SAFT : type), K_(K), No_(No), Lt_(Lt), Lr_(Lr), prec_(precision)
{
  require(type >= DAS_SAFT && type <= DAS_RCA_ROWTX, "das: unknown DAS type");
  require(precision == DREAM_HIP_F64 || precision == DREAM_HIP_F32_REF ||
              precision == DREAM_HIP_F32,
          "das: unknown precision");
  require(K >= 1 && K < (1ll << 31), "das: A-scan length out of range");
  require(No >= 0 && Lt >= 1 && Lr >= 0, "das: bad array sizes");
  const int64_t n_ascans = type_ == DAS_SAFT ? Lt : Lt * Lr;
  require(n_ascans * K < (1ll << 40), "das: data too large");
  if (type_ != DAS_SAFT) require(Lr >= 1, "das: TFM/RCA need receive elements");
  DREAM_HIP_CHECK(hipMalloc(&d_status_, sizeof(Status)));
  DREAM_HIP_CHECK(hipHostMalloc(&h_status_, sizeof(Status)));
}

DasHandle::~DasHandle()
{
  (void) hipFree(d_status_);
  (void) hipHostFree(h_status_);
  (void) hipFree(d_buf_);
  (void) hipFree(d_z_);
}

bool DasHandle::changed(int type, int64_t K, int64_t No, int64_t Lt, int64_t Lr,
                        int precision) const
{
  const int eff = Lr == 0 ? DAS_SAFT : type;
  return eff != type_ || K != K_ || No != No_ || Lt != Lt_ || Lr != Lr_ || precision != prec_;
}

void DasHandle::fill_args(DasArgs &a, const void *Y, const void *Ro, const void *Gt,
                          const void *Gr, double dt, int delay_type, const void *delay,
                          double cp, const void *apod_t, const void *apod_r, int64_t n_frames,
                          int64_t y_stride, int64_t im_stride, void *Im) const
{
  std::memset(&a, 0, sizeof a);
  a.Y = Y;
  a.K = K_;
  a.Ro = Ro;
  a.No = No_;
  a.Gt = Gt;
  a.Lt = (int32_t) Lt_;
  a.Gr = Gr;
  a.Lr = type_ == DAS_SAFT ? 0 : (int32_t) Lr_;
  a.delay = delay;
  a.delay_multi = delay_type != 0;
  // The reference's constants: double expressions rounded to T
  // (extra_src/das.cc:316-318).
  if (prec_ == DREAM_HIP_F32_REF) {
    a.fs_khz = (double) (float) ((1.0 / dt) * 1000.0);
    a.inv_cp = (double) (float) (1.0 / cp);
    a.delay_ms = 0.0;
    a.cp = cp;
  } else {
    a.fs_khz = (1.0 / dt) * 1000.0;
    a.inv_cp = 1.0 / cp;
    a.cp = cp;
  }
  a.apod_t = apod_t;
  a.apod_r = apod_r;
  a.n_frames = n_frames;
  a.y_stride = y_stride;
  a.im_stride = im_stride;
  a.Im = Im;
  a.status = d_status_;
}

int DasHandle::run_dev(const void *Y, const void *Ro, const void *Gt, const void *Gr, double dt,
                       int delay_type, const void *delay, double cp, const void *apod_t,
                       const void *apod_r, int interp, int64_t n_frames, int64_t y_stride,
                       int64_t im_stride, void *Im, int err_level, bool want_status,
                       hipStream_t s)
{
  if (cancelled()) throw Error(DREAM_HIP_ECANCELLED, "das cancelled");
  require(interp == DREAM_HIP_NEAREST || interp == DREAM_HIP_LINEAR, "das: unknown interpolation");
  require(n_frames >= 1, "das: n_frames must be >= 1");
  const int64_t n_ascans = type_ == DAS_SAFT ? Lt_ : Lt_ * Lr_;
  if (y_stride <= 0) y_stride = n_ascans * K_;
  if (im_stride <= 0) im_stride = No_;
  require(No_ == 0 || (Y && Ro && Gt && delay && Im), "das: null device pointer");
  require(type_ == DAS_SAFT || Gr, "das: null receive array");
  const bool apod = apod_t || apod_r;
  require(!apod || apod_t, "das: apod_r given without apod_t");
  require(!apod || type_ == DAS_SAFT || apod_r, "das: TFM/RCA apodization needs apod_r");
  DasArgs a;
  fill_args(a, Y, Ro, Gt, Gr, dt, delay_type, delay, cp, apod_t, apod_r, n_frames, y_stride,
            im_stride, Im);
  if (!a.delay_multi && No_ > 0) {
    // The scalar delay is read here once so its ms value is exact host math.
    double d;
    if (geom_size(prec_) == 8) {
      DREAM_HIP_CHECK(hipMemcpyAsync(&d, delay, 8, hipMemcpyDeviceToHost, s));
      DREAM_HIP_CHECK(hipStreamSynchronize(s));
    } else {
      float f;
      DREAM_HIP_CHECK(hipMemcpyAsync(&f, delay, 4, hipMemcpyDeviceToHost, s));
      DREAM_HIP_CHECK(hipStreamSynchronize(s));
      d = f;
    }
    a.delay_ms = prec_ == DREAM_HIP_F32_REF ? (double) (float) (d / 1000.0) : d / 1000.0;
  }
  zero_async(d_status_, sizeof(Status), s);
  if (No_ > 0) launch_das_forward(a, type_, prec_, interp, apod, s);
  if (!want_status) return E_NONE;
  DREAM_HIP_CHECK(hipMemcpyAsync(h_status_, d_status_, sizeof(Status), hipMemcpyDeviceToHost, s));
  DREAM_HIP_CHECK(hipStreamSynchronize(s));
  if (!h_status_->any_oob) return E_NONE;
  return level_code(err_level);
}

int DasHandle::run_host(const void *Y, const void *Ro, const void *Gt, const void *Gr, double dt,
                        int delay_type, const void *delay, double cp, void *Im, int err_level)
{
  const size_t es = elem_size(prec_);
  const size_t gs = geom_size(prec_);
  const int64_t n_ascans = type_ == DAS_SAFT ? Lt_ : Lt_ * Lr_;
  const int64_t nd = delay_type ? No_ : 1;
  // The host arrays are in T (float for both FP32 precisions); the FP64
  // delay path of DREAM_HIP_F32 takes them widened to double.
  const size_t in_gs = prec_ == DREAM_HIP_F64 ? 8 : 4;
  const size_t bY = (size_t) (n_ascans * K_) * es;
  const size_t bRo = (size_t) No_ * 3 * gs, bGt = (size_t) Lt_ * 3 * gs;
  const size_t bGr = (size_t) (type_ == DAS_SAFT ? 0 : Lr_ * 3) * gs;
  const size_t bD = (size_t) nd * gs, bIm = (size_t) No_ * es;
  auto al = [](size_t b) { return (b + 255) & ~(size_t) 255; };
  const size_t total = al(bY) + al(bRo) + al(bGt) + al(bGr) + al(bD) + al(bIm);
  if (total > d_buf_bytes_) {
    (void) hipFree(d_buf_);
    d_buf_ = nullptr;
    d_buf_bytes_ = 0;
    if (hipMalloc(&d_buf_, total) != hipSuccess)
      throw Error(DREAM_HIP_ENOMEM, "das: device allocation failed");
    d_buf_bytes_ = total;
  }
  char *base = (char *) d_buf_;
  char *dY = base, *dRo = dY + al(bY), *dGt = dRo + al(bRo), *dGr = dGt + al(bGt);
  char *dD = dGr + al(bGr), *dIm = dD + al(bD);
  hipStream_t s = lib_stream();
  DREAM_HIP_CHECK(hipMemcpyAsync(dY, Y, bY, hipMemcpyHostToDevice, s));
  auto put_geom = [&](char *dst, const void *src, int64_t n) {
    if (!n) return;
    if (gs == in_gs) {
      DREAM_HIP_CHECK(hipMemcpyAsync(dst, src, (size_t) n * gs, hipMemcpyHostToDevice, s));
    } else {
      std::vector<double> w((size_t) n);
      const float *f = (const float *) src;
      for (int64_t q = 0; q < n; q++) w[(size_t) q] = f[q];
      DREAM_HIP_CHECK(hipMemcpy(dst, w.data(), (size_t) n * 8, hipMemcpyHostToDevice));
    }
  };
  put_geom(dRo, Ro, No_ * 3);
  put_geom(dGt, Gt, Lt_ * 3);
  if (bGr) put_geom(dGr, Gr, Lr_ * 3);
  put_geom(dD, delay, nd);
  int code = run_dev(dY, dRo, dGt, bGr ? dGr : nullptr, dt, delay_type, dD, cp, nullptr, nullptr,
                     DREAM_HIP_NEAREST, 1, 0, 0, dIm, err_level, true, s);
  DREAM_HIP_CHECK(hipMemcpyAsync(Im, dIm, bIm, hipMemcpyDeviceToHost, s));
  DREAM_HIP_CHECK(hipStreamSynchronize(s));
  return code;
}

void DasHandle::adjoint_dev(const void *Y, const void *Ro, const void *Gt, const void *Gr,
                            double dt, double delay, double cp, const void *apod_t,
                            const void *apod_r, int interp, int64_t n_frames, int64_t y_stride,
                            int64_t im_stride, const void *g, void *gY, void *g_apod_t,
                            void *g_apod_r, void *g_cp, hipStream_t s)
{
  if (cancelled()) throw Error(DREAM_HIP_ECANCELLED, "das adjoint cancelled");
  require(prec_ != DREAM_HIP_F32_REF, "das adjoint: not available for DREAM_HIP_F32_REF handles");
  require(interp == DREAM_HIP_NEAREST || interp == DREAM_HIP_LINEAR, "das: unknown interpolation");
  require(n_frames >= 1, "das: n_frames must be >= 1");
  require(Y && Ro && Gt && g, "das adjoint: null device pointer");
  require(type_ == DAS_SAFT || Gr, "das adjoint: null receive array");
  const bool apod = apod_t || apod_r;
  require(!apod || (apod_t && (type_ == DAS_SAFT || apod_r)),
          "das adjoint: apodization needs apod_t (and apod_r for TFM/RCA)");
  const int64_t n_ascans = type_ == DAS_SAFT ? Lt_ : Lt_ * Lr_;
  if (y_stride <= 0) y_stride = n_ascans * K_;
  if (im_stride <= 0) im_stride = No_;
  require(!gY || (size_t) K_ * elem_size(prec_) <= 65536,
          "das adjoint: the Y-gradient needs K * sizeof(T) <= 64 KiB (one A-scan in LDS)");
  require(type_ == DAS_SAFT || gY || (size_t) (Lt_ + Lr_) * 8 <= 65536,
          "das adjoint: too many elements for the apodization accumulators");
  // Accumulators: Z per A-scan (channel kernel) or z_t/z_r (pixel kernel), z_cp.
  const size_t zbytes = (size_t) (n_ascans + Lt_ + Lr_ + 1) * sizeof(double);
  if (zbytes > d_z_bytes_) {
    (void) hipFree(d_z_);
    d_z_ = nullptr;
    d_z_bytes_ = 0;
    if (hipMalloc(&d_z_, zbytes) != hipSuccess)
      throw Error(DREAM_HIP_ENOMEM, "das adjoint: device allocation failed");
    d_z_bytes_ = zbytes;
  }
  double *Z = (double *) d_z_;
  double *zt = Z + n_ascans, *zr = zt + Lt_, *zcp = zr + Lr_;
  zero_async(d_z_, zbytes, s);
  DasArgs a;
  fill_args(a, Y, Ro, Gt, Gr, dt, 0, nullptr, cp, apod_t, apod_r, n_frames, y_stride, im_stride,
            nullptr);
  a.delay_ms = delay / 1000.0;
  a.g = g;
  a.gY = gY;
  a.z_cp = zcp;
  const bool channel = type_ == DAS_SAFT || gY != nullptr;
  a.z_t = channel ? (void *) Z : (void *) zt;
  a.z_r = zr;
  if (No_ > 0) launch_das_adjoint(a, type_, prec_, interp, apod, gY != nullptr, s);
  else if (gY) zero_async(gY, (size_t) (n_frames * y_stride) * elem_size(prec_), s);
  const int n = (int) std::max<int64_t>(std::max<int64_t>(Lt_, Lr_), 1);
  const int blocks = (n + 255) / 256;
  if (prec_ == DREAM_HIP_F64) {
    if (channel)
      hipLaunchKernelGGL(adj_finalize_channel<double>, dim3(blocks), dim3(256), 0, s, type_,
                         (int) Lt_, (int) Lr_, Z, (const double *) apod_t, (const double *) apod_r,
                         zcp, (double *) g_apod_t, (double *) g_apod_r, (double *) g_cp);
    else
      hipLaunchKernelGGL(adj_finalize_pixel<double>, dim3(blocks), dim3(256), 0, s, (int) Lt_,
                         (int) Lr_, zt, zr, zcp, (double *) g_apod_t, (double *) g_apod_r,
                         (double *) g_cp);
  } else {
    if (channel)
      hipLaunchKernelGGL(adj_finalize_channel<float>, dim3(blocks), dim3(256), 0, s, type_,
                         (int) Lt_, (int) Lr_, Z, (const float *) apod_t, (const float *) apod_r,
                         zcp, (float *) g_apod_t, (float *) g_apod_r, (float *) g_cp);
    else
      hipLaunchKernelGGL(adj_finalize_pixel<float>, dim3(blocks), dim3(256), 0, s, (int) Lt_,
                         (int) Lr_, zt, zr, zcp, (float *) g_apod_t, (float *) g_apod_r,
                         (float *) g_cp);
  }
  DREAM_HIP_CHECK(hipGetLastError());
}

} // namespace dream
