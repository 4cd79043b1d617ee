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
#include <atomic>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <utility>
#include <vector>

#include "das.hpp"
#include "tune.hpp"

// The allocation-free adjoint's workspace (dream_hip_das_adjoint_dev_noalloc):
// per-A-scan sums, per-element sums and the cp sum, FP64, 2 MiB.  Each GPU
// the library runs on holds its own copy until the process ends; host code
// uses only the per-device address fixed_workspace() looks up, never this
// name.
static __device__ __align__(256) double scratch[(size_t) 1 << 18];

namespace dream {

namespace {

constexpr int kThreads = 256;
constexpr int kRC = 16;          // receive legs held in registers per chunk

// Correctly rounded sqrt(s) for s in [2^-767, inf): the hardware rsqrt
// seed, one Goldschmidt step for the root g and half-reciprocal h, then two
// Newton corrections of g from the residual s - g^2 -- the operations of
// the compiler's IEEE square root, which only adds a scaling of arguments
// below 2^-767 (so that the residual stays normal) and passes 0 and inf
// through.  Five operations fewer per root.
__device__ __forceinline__ double sqrt_normal(double s)
{
  const double y = __builtin_amdgcn_rsq(s);
  double g = s * y, h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, s);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, s);
  return __builtin_fma(d, h, g);
}

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

// leg3 for code that the whole wave runs together (the two-leg forward's
// legs): the same value, by sqrt_normal where it applies.  Zero, tiny,
// infinite or NaN squared distances (a pixel on an element, or non-finite
// input) take the compiler's square root; the test is wave-wide, so the
// common case runs no branch body.
template <typename GT>
__device__ __forceinline__ GT leg3_wave(GT dx, GT dy, GT dz)
{
  if constexpr (std::is_same<GT, double>::value) {
    const double s = dx * dx + dy * dy + dz * dz;
    double r = sqrt_normal(s);
    const bool odd = !(s >= 0x1p-767 && s < __builtin_inf());
    if (__builtin_expect(__any(odd), 0))
      if (odd) r = __builtin_sqrt(s);
    return r;
  } else {
    return leg3<GT>(dx, dy, dz);
  }
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
// Wave sum in the data type: FP64 plans reduce in FP64 all the way.
__device__ __forceinline__ float wave_sum_t(float v) { return wave_sum_f(v); }
__device__ __forceinline__ double wave_sum_t(double v) { return wave_sum_d(v); }

__device__ __forceinline__ int64_t pixel_of(const int32_t *perm, int64_t idx)
{
  return perm ? (int64_t) perm[idx] : idx;
}

// --------------------------------------------------------------------- gathers

// A-scan samples are read with raw buffer loads over one frame of Y: an
// offset past the frame returns 0 in hardware without touching memory, so
// every gather of an unrolled block is issued unconditionally and the block
// keeps all its loads in flight (one wait per block instead of one per pair).
constexpr unsigned kNoSample = 0xFFFFFFF0u;   // byte offset beyond any frame

__device__ __forceinline__ __amdgpu_buffer_rsrc_t frame_rsrc(const void *base, int64_t bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short) 0, (int) bytes,
                                           0x00020000);
}

// The forward reads one A-scan at a time through a structured view of the
// frame: K records of sizeof(DT) bytes, with the A-scan's byte offset a K
// sizeof(DT) (< 4 GiB, the frame limit) in the scalar offset, which the
// hardware adds after its range check.  The check (record index < K) is
// then the bounds test of the sample index itself: an index outside [0, K),
// negative ones included (|k| <= 2^28, so they wrap to >= 2^32 - 2^28 > K),
// reads 0 without touching memory, and a gather costs no more than the
// conversion of its position (a raw view cannot do this: its range check
// covers the scalar offset too).  kNoRecord marks a pair that is not live.
constexpr unsigned kNoRecord = 0xFFFFFFFFu;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t ascan_rsrc(const void *frame, int K, int elem)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(frame), (short) elem, K,
                                           0x00020000);
}

extern "C" __device__ float das_struct_load_f32(__amdgpu_buffer_rsrc_t r, int index, int voffset,
                                                int soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.f32");
extern "C" __device__ double das_struct_load_f64(__amdgpu_buffer_rsrc_t r, int index, int voffset,
                                                 int soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.f64");

// Four frames' samples of one (A-scan, sample) record (frame groups).
typedef float frames4 __attribute__((ext_vector_type(4)));
extern "C" __device__ frames4 das_struct_load_v4f32(__amdgpu_buffer_rsrc_t r, int index,
                                                    int voffset, int soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v4f32");

// Sample `index` of the A-scan at byte offset `soff` (0 outside [0, K)).
template <typename DT>
__device__ __forceinline__ DT sload(__amdgpu_buffer_rsrc_t r, unsigned index, unsigned soff)
{
  if constexpr (std::is_same<DT, float>::value)
    return das_struct_load_f32(r, (int) index, 0, (int) soff, 0);
  else if constexpr (std::is_same<DT, frames4>::value)
    return das_struct_load_v4f32(r, (int) index, 0, (int) soff, 0);
  else
    return das_struct_load_f64(r, (int) index, 0, (int) soff, 0);
}

// What a lane sums: one frame's sample (DT), or four frames' (FR = 4).
template <typename DT, int FR>
using lane_t = typename std::conditional<FR == 4, frames4, DT>::type;

template <typename DT>
__device__ __forceinline__ DT bload(__amdgpu_buffer_rsrc_t r, unsigned byte_off);
template <>
__device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, unsigned byte_off)
{
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int) byte_off, 0, 0));
}
template <>
__device__ __forceinline__ double bload<double>(__amdgpu_buffer_rsrc_t r, unsigned byte_off)
{
  auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int) byte_off, 0, 0);
  return __hiloint2double((int) v[1], (int) v[0]);
}

// Legs are clamped to |t| <= leg_max = 2^26 / fs before use, so every sample
// position u = (t_t + t_r) * fs lies in [-2^27, 2^27] and converts to int
// exactly as trunc/floor would; a NaN leg becomes the clamp value and lands
// out of bounds, as in the reference.  Only legs that are out of bounds
// anyway are changed.
template <typename GT>
__device__ __forceinline__ GT clamp_leg(GT t, GT tmax)
{
  return fmin(fmax(t, -tmax), tmax);
}

// SAFT sample position u = 2 (sqrt(s) / cp + delay) fs, with the rounding
// of the IEEE chain (extra_src/das.cc:329-335).  FAST (FP64 geometry only):
// s by two FMAs and sqrt(s) from the rsqrt seed and one Goldschmidt step,
// relative error below 2^-47.5 (the seed is accurate to 2^-24.2 on gfx950),
// then u = (r / cp + delay) (2 fs) (doubling is exact, so this is the
// chain's own rounding of its operands).  Carried through, the fast u
// differs from the IEEE one by less than 2^-46 (|u| + |2 fs delay| + 1).
// It is kept when it lies farther than 2^-29 |u| + `bnd` from an integer,
// `bnd` = 2^-40 (|2 fs delay| + 1) per pixel: then trunc and floor of the
// two agree, and |u| < 2^28, so the integer conversion is exact with no
// clamp.  Otherwise (about one pair in 10^5, and for s = 0 or non-finite
// s) the IEEE chain is recomputed, clamped as everywhere else.  The kept
// value moves the linear-interpolation weight by at most 2^-40 |u|, so FAST
// is only used where that is far below the output's precision
// (nearest-lower sampling, or FP32 data).  `tleg` returns the leg time
// sqrt(s) / cp.
template <typename GT, bool FAST>
__device__ __forceinline__ GT saft_pos(GT dx, GT dy, GT dz, GT inv_cp, GT dl, GT tmax, GT fs,
                                       GT fs2, GT bnd, GT &tleg)
{
  if (FAST) {
    const double s = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
    double r, rinv;
    sqrt_rsqrt<1>(s, r, rinv);
    tleg = r * inv_cp;
    const double u = (tleg + dl) * fs2;
    if (__builtin_fabs(u - __builtin_rint(u)) > __builtin_fma(__builtin_fabs(u), 0x1p-29, bnd))
      return u;
  }
  tleg = leg3<GT>(dx, dy, dz) * inv_cp;
  const GT t = tleg + dl;
  return clamp_leg<GT>(t + t, tmax) * fs;     // (T)(2.0*t) * fs, 2t exact
}

template <typename DT, typename GT, int INTERP>
constexpr bool kSaftFast = std::is_same<GT, double>::value &&
                           (INTERP == DREAM_HIP_NEAREST || sizeof(DT) == 4);

// Forward gathers go through the frame's structured view (ascan_rsrc): the
// record index is the sample index k itself and the A-scan's offset is
// scalar, so the hardware range check does the bounds test.  Positions
// arrive with |u| < 2^28 (legs clamped to leg_max = 2^26 / fs, or a
// certified SAFT position), so the integer conversion is exact.  A pair
// that is not `live` (padding past the last element of a block, or the
// lower half of a diagonal reciprocal block) gets kNoRecord and reads 0
// through the same unconditional load, so the loads of a block never wait
// on a branch; with ALL_LIVE (a block without such pairs) the select is
// left out.  Indices (and the linear weight) come first; callers issue a
// block's loads together and combine after.
template <typename DT, typename GT, int INTERP, bool ALL_LIVE>
__device__ __forceinline__ void gather_index(GT u, bool live, unsigned &i0, unsigned &i1, DT &f,
                                             int &k)
{
  if (INTERP == DREAM_HIP_NEAREST) {
    k = (int) u;
    i0 = (ALL_LIVE || live) ? (unsigned) k : kNoRecord;
  } else {
    const GT fk = floor(u);
    k = (int) fk;
    i0 = (ALL_LIVE || live) ? (unsigned) k : kNoRecord;
    i1 = (ALL_LIVE || live) ? (unsigned) (k + 1) : kNoRecord;
    f = (DT) (u - fk);
  }
}

// --------------------------------------------------------------------- forward

// Work of a 64-pixel tile is cut into jobs: (receive chunk of kRC, transmit
// block of kTB) for two-leg modes, blocks of kSB elements for SAFT.  The 4
// waves of a work-group share the tile and take jobs round-robin; `split`
// work-groups (gridDim.y) share a tile when the tiles alone cannot fill the
// device, writing partial sums that das_reduce_kernel adds in fixed order.
constexpr int kTB = 16;    // transmit elements per two-leg job
constexpr int kSB = 8;     // SAFT elements per job (one unrolled block)

constexpr int kMaxWaves = 8;     // waves per forward work-group (64 .. 512 threads)

// The sum over the A-scans of jobs job0, job0 + jstride, ... of the
// pixel at visiting position idx (one lane per pixel, 64 per wave; every
// lane of the wave runs the same jobs).  `oob` is set when a pair of the
// pixel fell outside [0, K) (only tracked when STATUS).
template <typename DT, typename GT, int TYPE, int INTERP, bool APOD, bool STATUS,
          bool RECIP = false, int FR = 1>
__device__ __forceinline__ lane_t<DT, FR> fwd_tile_sum(const DasArgs &A, const int32_t *__restrict__ perm,
                                           int64_t idx, int64_t frame, int job0, int jstride,
                                           bool &oob)
{
  const bool valid = idx < A.No;
  const int64_t p = valid ? pixel_of(perm, idx) : 0;
  const GT *Ro = (const GT *) A.Ro;
  const GT xo = Ro[p], yo = Ro[p + A.No], zo = Ro[p + 2 * A.No];
  const GT *gdel = (const GT *) A.delay;
  const GT dl = delay_ms_of<GT>(gdel[A.delay_multi ? p : 0]);
  const GT fs = (GT) A.fs_khz, inv_cp = (GT) A.inv_cp;
  const GT tmax = (GT) A.leg_max;
  const int K = (int) A.K;
  const unsigned Ku = (unsigned) K;
  // Reciprocal TFM (see DasArgs::Ysym; the host sets it only for a scalar
  // delay): the test reads the delay at one address and is made explicitly
  // wave-uniform, so the job loop below stays scalar.
  bool recip = false;
  if (RECIP && TYPE == DAS_TFM)
    recip = __builtin_amdgcn_readfirstlane((int) (delay_ms_of<GT>(gdel[0]) == GT(0))) != 0;
  // One frame, or (FR = 4) frame group `frame` of the grouped copy.
  using VT = lane_t<DT, FR>;
  const VT *Yf = FR == 4 ? (const VT *) A.Yg + frame * A.yg_stride
                         : (const VT *) (recip ? A.Ysym : A.Y) + frame * A.y_stride;
  const __amdgpu_buffer_rsrc_t rY = ascan_rsrc(Yf, K, (int) sizeof(VT));
  const unsigned KB = Ku * (unsigned) sizeof(VT);          // bytes per A-scan
  const GT *Gt = (const GT *) A.Gt;
  const GT *Gr = (const GT *) A.Gr;
  const DT *at = (const DT *) A.apod_t;
  const DT *ar = (const DT *) A.apod_r;
  const int Lt = A.Lt, Lr = A.Lr;
  VT acc = VT{};

  if (TYPE == DAS_SAFT) {
    const GT bnd = (GT) (0x1p-40 * (fabs(2.0 * (double) fs * (double) dl) + 1.0));
    const GT fs2 = fs + fs;
    const int J = (Lt + kSB - 1) / kSB;
    for (int job = job0; job < J; job += jstride) {
      const int n0 = job * kSB;
      GT u[kSB];
      bool any_in = false;
#pragma unroll
      for (int q = 0; q < kSB; q++) {
        const int n = min(n0 + q, Lt - 1);
        GT tleg;
        u[q] = saft_pos<GT, kSaftFast<DT, GT, INTERP>>(Gt[n] - xo, Gt[n + Lt] - yo,
                                                      Gt[n + 2 * Lt] - zo, inv_cp, dl, tmax,
                                                      fs, fs2, bnd, tleg);
        any_in |= (n0 + q < Lt) && u[q] > GT(-1) && u[q] < (GT) K;
      }
      if (!__any(any_in && valid)) {               // the whole block is out of range
        if (STATUS) oob = true;
        continue;
      }
      // All loads of the block are issued before any is used.
      unsigned o0[kSB], o1[kSB];
      DT f[kSB];
      VT y0[kSB], y1[kSB];
#pragma unroll
      for (int q = 0; q < kSB; q++) {
        const bool live = n0 + q < Lt;
        int k;
        gather_index<DT, GT, INTERP, false>(u[q], live, o0[q], o1[q], f[q], k);
        if (STATUS) oob |= live && (unsigned) k >= Ku;
      }
#pragma unroll
      for (int q = 0; q < kSB; q++) {
        const unsigned soff = (unsigned) min(n0 + q, Lt - 1) * KB;
        y0[q] = sload<VT>(rY, o0[q], soff);
        if (INTERP != DREAM_HIP_NEAREST) y1[q] = sload<VT>(rY, o1[q], soff);
      }
      VT part = VT{};
#pragma unroll
      for (int q = 0; q < kSB; q++) {
        const VT v = INTERP == DREAM_HIP_NEAREST ? y0[q] : y0[q] + f[q] * (y1[q] - y0[q]);
        part += APOD ? at[min(n0 + q, Lt - 1)] * v : v;
      }
      acc += part;
    }
  } else {
    GT gy_c = GT(0), gx_r = GT(0);
    if (TYPE == DAS_RCA_COLTX || TYPE == DAS_RCA_ROWTX) {
      const bool given = A.stripe_given != 0;
      gy_c = stripe_gap<GT>(yo, given ? (GT) A.sy_lo : Gr[Lt], given ? (GT) A.sy_hi : Gr[2 * Lt - 1]);
      gx_r = stripe_gap<GT>(xo, given ? (GT) A.sx_lo : Gt[0], given ? (GT) A.sx_hi : Gt[Lr - 1]);
    }
    const bool delay_rx = A.delay_rx != 0;
    const int nchunks = (Lr + kRC - 1) / kRC;
    // Reciprocal TFM visits the blocks b <= c only (Lt == Lr, kTB == kRC).
    const int J = recip ? nchunks * (nchunks + 1) / 2 : nchunks * ((Lt + kTB - 1) / kTB);
    for (int job = job0; job < J; job += jstride) {
      int c = job % nchunks, b = job / nchunks;
      if (recip) {
        b = 0;
        c = job;
        while (c >= nchunks - b) {
          c -= nchunks - b;
          b++;
        }
        c += b;
      }
      const bool diag = recip && b == c;     // in a diagonal block only j >= i
      const int j0 = c * kRC, i0 = b * kTB, i1 = min(i0 + kTB, Lt);
      const bool full = j0 + kRC <= Lr;
      GT tr[kRC];
#pragma unroll
      for (int jj = 0; jj < kRC; jj++) {
        const int j = min(j0 + jj, Lr - 1);
        if (TYPE == DAS_TFM)
          tr[jj] = leg3_wave<GT>(Gr[j] - xo, Gr[j + Lr] - yo, Gr[j + 2 * Lr] - zo) * inv_cp;
        else
          tr[jj] = leg3_wave<GT>(gx_r, Gr[j + Lr] - yo, Gr[j + 2 * Lr] - zo) * inv_cp;
        if (delay_rx) tr[jj] += dl;
        tr[jj] = clamp_leg<GT>(tr[jj], tmax);
      }
      // Extremes of the chunk's receive legs: a transmit row whose sample
      // positions (tt + tr) * fs all lie outside (-1, K) for every lane of
      // the wave issues no gathers (rounding is monotone, so the bound is exact).
      GT trmin = tr[0], trmax = tr[0];
#pragma unroll
      for (int jj = 1; jj < kRC; jj++) {
        trmin = fmin(trmin, tr[jj]);       // clamped duplicates of the last
        trmax = fmax(trmax, tr[jj]);       // receiver do not widen the range
      }
      // Rows of the job; ALL_LIVE (every pair of the block is live: a full
      // chunk off the diagonal) leaves the per-pair live select out.
      auto rows = [&](auto all_live) {
        constexpr bool ALL_LIVE = decltype(all_live)::value;
        for (int i = i0; i < i1; i++) {
          GT tt;
          if (TYPE == DAS_TFM)
            tt = leg3_wave<GT>(Gt[i] - xo, Gt[i + Lt] - yo, Gt[i + 2 * Lt] - zo) * inv_cp;
          else
            tt = leg3_wave<GT>(Gt[i] - xo, gy_c, Gt[i + 2 * Lt] - zo) * inv_cp;
          if (!delay_rx) tt += dl;
          tt = clamp_leg<GT>(tt, tmax);
          // The row's sample positions lie in [lo, hi] (rounding is monotone,
          // so the extreme receive legs give the extreme positions).  That
          // makes the out-of-bounds status a per-row test: a pair is out when
          // its index trunc(u) (nearest; floor(u) when interpolating) falls
          // outside [0, K).
          const GT lo = (tt + trmin) * fs, hi = (tt + trmax) * fs;
          const bool row_out = !valid || !(lo < (GT) K) || !(hi > GT(-1));
          if (STATUS)
            oob |= !(INTERP == DREAM_HIP_NEAREST ? lo > GT(-1) : lo >= GT(0)) || !(hi < (GT) K);
          if (__all(row_out)) continue;
          const DT wi = APOD ? at[i] : DT(1);
          // Receivers in groups of kG (the whole chunk for nearest-lower
          // sampling, half of it for linear interpolation, half that again
          // for frame groups): every load of a group is issued before any
          // is used.  Receivers past Lr (in the
          // last chunk) read 0.
          constexpr int kG = (INTERP == DREAM_HIP_NEAREST ? kRC : kRC / 2) / (FR == 4 ? 2 : 1);
          VT part = VT{};
#pragma unroll
          for (int g0 = 0; g0 < kRC; g0 += kG) {
            unsigned o0[kG], o1[kG];
            DT f[kG];
            VT y0[kG], y1[kG];
#pragma unroll
            for (int jj = 0; jj < kG; jj++) {
              const bool live = (full || j0 + g0 + jj < Lr) && (!diag || j0 + g0 + jj >= i);
              int k;
              gather_index<DT, GT, INTERP, ALL_LIVE>((tt + tr[g0 + jj]) * fs, live, o0[jj], o1[jj],
                                                     f[jj], k);
            }
#pragma unroll
            for (int jj = 0; jj < kG; jj++) {
              const int j = min(j0 + g0 + jj, Lr - 1);
              const int a = (TYPE == DAS_RCA_ROWTX) ? j * Lt + i : i * Lr + j;
              const unsigned soff = (unsigned) a * KB;
              y0[jj] = sload<VT>(rY, o0[jj], soff);
              if (INTERP != DREAM_HIP_NEAREST) y1[jj] = sload<VT>(rY, o1[jj], soff);
            }
#pragma unroll
            for (int jj = 0; jj < kG; jj++) {
              const VT v =
                  INTERP == DREAM_HIP_NEAREST ? y0[jj] : y0[jj] + f[jj] * (y1[jj] - y0[jj]);
              part += APOD ? ar[min(j0 + g0 + jj, Lr - 1)] * v : v;
            }
          }
          acc += APOD ? wi * part : part;
        }
      };
      if (full && !diag) {
        rows(std::true_type{});
        continue;
      }
      if constexpr (RECIP && TYPE == DAS_TFM && INTERP == DREAM_HIP_NEAREST && !APOD) {
        if (diag && full) {
          // A full diagonal block of the reciprocal path: row i0 + il is
          // receiver j0 + il of the chunk (one array, the same expression,
          // a zero delay: the same leg bit for bit), so the rows take their
          // transmit legs from tr[] and, unrolled, issue only their live
          // pairs jj >= il.  The sums are the generic rows' (their dead
          // pairs add zeros).
#pragma unroll
          for (int il = 0; il < kRC; il++) {
            const GT tt = tr[il];
            const GT lo = (tt + trmin) * fs, hi = (tt + trmax) * fs;
            const bool row_out = !valid || !(lo < (GT) K) || !(hi > GT(-1));
            if (STATUS) oob |= !(lo > GT(-1)) || !(hi < (GT) K);
            if (__all(row_out)) continue;
            const unsigned abase = (unsigned) ((i0 + il) * Lr + j0) * KB;
            unsigned o0[kRC];
            VT y0[kRC];
#pragma unroll
            for (int jj = il; jj < kRC; jj++) {
              unsigned o1;
              DT f;
              int k;
              gather_index<DT, GT, INTERP, true>((tt + tr[jj]) * fs, true, o0[jj], o1, f, k);
            }
#pragma unroll
            for (int jj = il; jj < kRC; jj++) y0[jj] = sload<VT>(rY, o0[jj], abase + jj * KB);
            VT part = VT{};
#pragma unroll
            for (int jj = il; jj < kRC; jj++) part += y0[jj];
            acc += part;
          }
          continue;
        }
      }
      rows(std::false_type{});
    }
  }
  return acc;
}

template <typename DT, typename GT, int TYPE, int INTERP, bool APOD, bool STATUS,
          bool RECIP = false, int FR = 1>
__global__ void __launch_bounds__(64 * kMaxWaves)
das_fwd_kernel(DasArgs A, const int32_t *__restrict__ perm, int split, DT *__restrict__ partial)
{
  using VT = lane_t<DT, FR>;
  __shared__ VT red[kMaxWaves][64];
  const int nwaves = blockDim.x >> 6;
  // The wave index is uniform across the wave; say so, so that job, element
  // and A-scan indices live in SGPRs.
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t idx = ((int64_t) blockIdx.x + A.block0) * 64 + lane;
  const int sidx = blockIdx.y;
  const int64_t frame = (int64_t) blockIdx.z + A.frame0;
  const bool valid = idx < A.No;
  const int64_t p = valid ? pixel_of(perm, idx) : 0;
  bool oob = false;
  const VT acc = fwd_tile_sum<DT, GT, TYPE, INTERP, APOD, STATUS, RECIP, FR>(
      A, perm, idx, frame, sidx * nwaves + wave, nwaves * split, oob);
  // Sum the waves' partial sums in a fixed order.
  red[wave][lane] = acc;
  __syncthreads();
  if (wave == 0 && valid) {
    VT sum = red[0][lane];
    for (int w = 1; w < nwaves; w++) sum += red[w][lane];
    if constexpr (FR == 1) {
      if (split == 1) ((DT *) A.Im)[frame * A.im_stride + p] = sum;
      else partial[((int64_t) sidx * A.n_frames + frame) * A.No + p] = sum;
    } else {
      // Group `frame` holds frames FR * frame .. FR * frame + FR - 1.
#pragma unroll
      for (int q = 0; q < FR; q++) {
        const int64_t fr = FR * frame + q;
        if (fr < A.n_frames) {
          if (split == 1) ((DT *) A.Im)[fr * A.im_stride + p] = sum[q];
          else partial[((int64_t) sidx * A.n_frames + fr) * A.No + p] = sum[q];
        }
      }
    }
  }
  if (STATUS) {
    // Every writer of a call stores the same value, the call's epoch, so a
    // plain store does it: it stays in the writing XCD's L2 until the
    // kernel ends, where an atomic (or a device-scope load) from thousands
    // of waves would queue at one memory address -- up to a fifth of a
    // short launch when the waves of an out-of-range image end together.
    // Calls of one handle are stream-ordered, so a later call never sees
    // an earlier call's store land after its own.
    oob = oob && valid;
    if (__any(oob) && lane == 0) A.status->any_oob = A.status_epoch;
  }
}

// Im[f][p] = sum over s of partial[s][f][p], s ascending (split launches),
// for the pixels at visiting-order indices [idx0, idx0 + n_idx) of frames
// [frame0, frame0 + nf): the pixels of one chunk of the forward launch.
template <typename DT>
__global__ void das_reduce_kernel(const DT *__restrict__ partial, int split, int64_t No,
                                  int64_t n_frames, int64_t frame0, int64_t nf, int64_t idx0,
                                  int64_t n_idx, const int32_t *__restrict__ perm,
                                  int64_t im_stride, DT *__restrict__ Im)
{
  const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_idx * nf) return;
  const int64_t f = frame0 + q / n_idx, idx = idx0 + q % n_idx;
  if (idx >= No) return;
  const int64_t p = pixel_of(perm, idx);
  const int64_t at = f * No + p, stride = n_frames * No;
  // Same order of additions as a plain loop; the loads of four partial
  // sums are issued together.
  DT sum;
  int s;
  if (split >= 4) {
    const DT a = partial[at], b = partial[stride + at];
    const DT c = partial[2 * stride + at], d = partial[3 * stride + at];
    sum = a;
    sum += b;
    sum += c;
    sum += d;
    s = 4;
  } else {
    sum = partial[at];
    s = 1;
  }
  for (; s + 4 <= split; s += 4) {
    const DT a = partial[s * stride + at], b = partial[(s + 1) * stride + at];
    const DT c = partial[(s + 2) * stride + at], d = partial[(s + 3) * stride + at];
    sum += a;
    sum += b;
    sum += c;
    sum += d;
  }
  for (; s < split; s++) sum += partial[s * stride + at];
  Im[f * im_stride + p] = sum;
}

// The same sums when a launch covers every pixel of its frames and the
// image rows are contiguous (im_stride == No): out[i] = sum over s of
// partial[s * stride + i], four consecutive values per thread, no pixel
// order to look up (n4 groups of four; partial and out offset by the
// launch's first frame, aligned to four values).
template <typename DT>
__global__ void das_reduce_flat_kernel(const DT *__restrict__ partial, int split, int64_t stride,
                                       int64_t n4, DT *__restrict__ out)
{
  typedef DT v4 __attribute__((ext_vector_type(4)));
  const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n4) return;
  const v4 *P = (const v4 *) partial + i;
  const int64_t st = stride / 4;
  v4 sum = P[0];
  int t = 1;
  for (; t + 3 <= split; t += 3) {
    const v4 a = P[t * st], b = P[(t + 1) * st], c = P[(t + 2) * st];
    sum += a;
    sum += b;
    sum += c;
  }
  for (; t < split; t++) sum += P[t * st];
  ((v4 *) out)[i] = sum;
}

// Ysym[f][i L + j] = Y[f][i L + j] + Y[f][j L + i] (i < j), Y[f][i L + i]
// (i == j), for the reciprocal TFM forward; nothing when the scalar delay
// read on the device is not zero (the forward then reads Y).  One
// work-group per pair i <= j (blockIdx.x runs over the L (L + 1) / 2 pairs
// column by column) and frame.  16-byte accesses when K and the frame
// stride keep every A-scan 16-byte aligned, two per thread and operand
// issued before either is used.
template <typename DT, typename GT>
__global__ void __launch_bounds__(256)
das_symmetrize_kernel(const DT *__restrict__ Y, DT *__restrict__ Ys, int L, int64_t K,
                      int64_t y_stride, const GT *__restrict__ delay)
{
  if (delay_ms_of<GT>(delay[0]) != GT(0)) return;
  // Pair t = j (j + 1) / 2 + i, 0 <= i <= j < L.
  const int t = (int) blockIdx.x;
  int j = (int) ((sqrtf(8.0f * (float) t + 1.0f) - 1.0f) * 0.5f);
  while ((j + 1) * (j + 2) / 2 <= t) j++;
  while (j * (j + 1) / 2 > t) j--;
  const int i = t - j * (j + 1) / 2;
  const int64_t f = blockIdx.z;
  const DT *a = Y + f * y_stride + ((int64_t) i * L + j) * K;
  const DT *b = Y + f * y_stride + ((int64_t) j * L + i) * K;
  DT *o = Ys + f * y_stride + ((int64_t) i * L + j) * K;
  constexpr int V = 16 / sizeof(DT);
  const bool vec = (K % V) == 0 && (y_stride % V) == 0 &&
                   ((reinterpret_cast<uintptr_t>(Y) | reinterpret_cast<uintptr_t>(Ys)) & 15) == 0;
  const bool diag = i == j;
  const int64_t step = blockDim.x;
  if (vec) {
    using V4 = typename std::conditional<sizeof(DT) == 4, float4, double2>::type;
    const V4 *av = reinterpret_cast<const V4 *>(a);
    const V4 *bv = reinterpret_cast<const V4 *>(b);
    V4 *ov = reinterpret_cast<V4 *>(o);
    const int64_t n = K / V;
    for (int64_t k = threadIdx.x; k < n; k += 2 * step) {
      const bool two = k + step < n;
      V4 x0 = av[k], x1 = two ? av[k + step] : V4{};
      if (!diag) {
        const V4 y0 = bv[k], y1 = two ? bv[k + step] : V4{};
        x0 = x0 + y0;
        x1 = x1 + y1;
      }
      ov[k] = x0;
      if (two) ov[k + step] = x1;
    }
    return;
  }
  if (diag) {
    for (int64_t k = threadIdx.x; k < K; k += step) o[k] = a[k];
  } else {
    for (int64_t k = threadIdx.x; k < K; k += step) o[k] = a[k] + b[k];
  }
}

// Yg[g][e] = the samples e (of the n_ascans K of one frame) of frames
// 4g .. 4g+3 as one 16-byte record, 0 for frames past the last: the frame
// groups of a multi-frame FP32 two-leg forward.  Reads are coalesced per
// frame, writes per record.
__global__ void __launch_bounds__(256)
das_group_frames_kernel(const float *__restrict__ Y, int64_t y_stride, int64_t E,
                        int64_t n_frames, frames4 *__restrict__ Yg)
{
  const int64_t g = blockIdx.y;
  const int64_t step = (int64_t) gridDim.x * blockDim.x;
  for (int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; e < E; e += step) {
    frames4 v;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int64_t f = 4 * g + q;
      v[q] = f < n_frames ? Y[f * y_stride + e] : 0.0f;
    }
    Yg[g * E + e] = v;
  }
}

// ---------------------------------------------------------------- uneven batch

__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v)
{
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned) v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned) (v >> 32));
  return ((unsigned long long) hi << 32) | lo;
}

// Many problems of one type in one persistent launch.  A work item is
// (problem, 64-pixel tile); items are numbered problem by problem, problems
// in decreasing per-tile cost, and each WAVE pulls items from counter[0] and
// runs every job of its tile, so uneven problems keep all waves busy until
// the list drains.  The last wave to leave resets both counters, so the
// launch can be captured into a hipGraph and replayed.
template <typename DT, typename GT, int TYPE, int INTERP, bool APOD>
__global__ void __launch_bounds__(kThreads)
das_batch_kernel(const DasArgs *__restrict__ args, const int64_t *__restrict__ item_begin,
                 int n_probs, int64_t n_items, unsigned long long *counter)
{
  const int lane = threadIdx.x & 63;
  const unsigned long long total_waves = (unsigned long long) gridDim.x * (blockDim.x >> 6);
  for (;;) {
    unsigned long long got = 0;
    if (lane == 0) got = atomicAdd(&counter[0], 1ull);
    const int64_t item = (int64_t) uniform_u64(__shfl(got, 0, 64));
    if (item >= n_items) break;
    int lo = 0, hi = n_probs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (item_begin[mid] <= item) lo = mid; else hi = mid - 1;
    }
    const DasArgs &A = args[lo];
    const int64_t idx = (item - item_begin[lo]) * 64 + lane;
    bool oob = false;
    const DT acc = fwd_tile_sum<DT, GT, TYPE, INTERP, APOD, false>(A, A.perm, idx, 0, 0, 1, oob);
    if (idx < A.No) ((DT *) A.Im)[pixel_of(A.perm, idx)] = acc;
  }
  if (lane == 0) {
    __threadfence();
    const unsigned long long done = atomicAdd(&counter[1], 1ull);
    if (done == total_waves - 1) {
      atomicExch(&counter[0], 0ull);
      atomicExch(&counter[1], 0ull);
    }
  }
}

// Both taps of one pair for the adjoint: y0 = Y[k], y1 = Y[k+1] (0 outside
// [0, K)), the fraction f, and k (-2 when neither tap exists).  The nearest
// rule uses k = trunc(u), y1 = y0 and f = 0.  Branch-free: loads are issued
// unconditionally through the frame's buffer resource.
template <typename DT, typename GT, int INTERP>
__device__ __forceinline__ void fetch_taps(__amdgpu_buffer_rsrc_t r, unsigned abase, int K, GT u,
                                           bool live, DT &y0, DT &y1, DT &f, int &k)
{
  if (INTERP == DREAM_HIP_NEAREST) {
    const int kk = (int) u;                      // u in [-2^27, 2^27]
    const bool inb = live && (unsigned) kk < (unsigned) K;
    k = inb ? kk : -2;
    y0 = bload<DT>(r, inb ? (abase + (unsigned) kk) * (unsigned) sizeof(DT) : kNoSample);
    y1 = y0;
    f = DT(0);
  } else {
    const GT fk = floor(u);
    const int kk = (int) fk;
    const bool t0 = live && (unsigned) kk < (unsigned) K;
    const bool t1 = live && (unsigned) (kk + 1) < (unsigned) K;
    y0 = bload<DT>(r, t0 ? (abase + (unsigned) kk) * (unsigned) sizeof(DT) : kNoSample);
    y1 = bload<DT>(r, t1 ? (abase + (unsigned) (kk + 1)) * (unsigned) sizeof(DT) : kNoSample);
    f = (t0 || t1) ? (DT) (u - fk) : DT(0);
    k = (t0 || t1) ? kk : -2;
  }
}

// ------------------------------------------------------------ region adjoint

// Region-major adjoint: one work-group per (region of kRegion consecutive
// pixels in visiting order, group of A-scans, frame).  Each lane keeps its
// kRPU pixels (coordinates and g) in registers for the whole group, so the
// pixel data are read once per group rather than once per A-scan.  The
// Y-gradient of every A-scan of the group is scattered into an LDS window of
// kWin samples whose start comes from the distance range between the
// element(s) and the region's bounding box (exact interval bound plus a
// 2-sample margin); a contribution that falls outside its window (regions
// with a large extent) is added to gY directly.  Windows are added into the
// zeroed gY once at the end.  Per-A-scan sums of g*value (for the
// apodization gradients) stay in registers and are reduced once.
constexpr int kRPU = 4;                         // pixels per lane
constexpr int kRegion = kThreads * kRPU;        // pixels per region
constexpr int kSaftGroup = 8;                   // SAFT A-scans per work-group
constexpr int kGI = 2, kGJ = 4;                 // two-leg group: transmits x receives
constexpr int kWin = 512;                       // LDS window (samples) per A-scan

template <typename T>
__device__ __forceinline__ void lds_atomic_add(T *p, T v)
{
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Range of |c - p| for p in [lo, hi] (c a point coordinate).
template <typename GT>
__device__ __forceinline__ void gap_range(GT c, GT lo, GT hi, GT &mn, GT &mx)
{
  const GT a = fabs(c - lo), b = fabs(c - hi);
  mn = (c >= lo && c <= hi) ? GT(0) : fmin(a, b);
  mx = fmax(a, b);
}

// Range of the stripe gap |stripe_gap(p, lo, hi)| for p in [pmin, pmax].
template <typename GT>
__device__ __forceinline__ void stripe_range(GT pmin, GT pmax, GT lo, GT hi, GT &mn, GT &mx)
{
  mn = (pmax < lo) ? lo - pmax : ((pmin > hi) ? pmin - hi : GT(0));
  mx = fmax(fabs(stripe_gap<GT>(pmin, lo, hi)), fabs(stripe_gap<GT>(pmax, lo, hi)));
}

__device__ __forceinline__ int window_start(double umin)
{
  if (!(umin > -4.0e8)) umin = -4.0e8;          // also catches NaN
  if (umin > 4.0e8) umin = 4.0e8;
  return (int) floor(umin) - 2;
}

template <typename DT, typename GT, int TYPE, int INTERP, bool APOD>
__global__ void __launch_bounds__(kThreads)
das_adj_region_kernel(DasArgs A, const int32_t *__restrict__ perm, int n_gj, int subs)
{
  constexpr int NA = TYPE == DAS_SAFT ? kSaftGroup : kGI * kGJ;
  extern __shared__ __align__(16) unsigned char smem[];
  // NA x kWin FP64 accumulators (only when gY is wanted): gfx950's LDS FP64
  // atomic add is about 20 times faster than its FP32 one.
  double *win = reinterpret_cast<double *>(smem);
  __shared__ GT s_box[6][kThreads / 64];
  __shared__ int s_w0[NA];
  __shared__ double s_z[NA][kThreads / 64];
  __shared__ double s_zc[kThreads / 64];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t region = (int64_t) blockIdx.x + A.block0;
  const int grp = blockIdx.y;
  const int64_t frame = (int64_t) blockIdx.z + A.frame0;
  const int Lt = A.Lt, Lr = A.Lr;
  const int K = (int) A.K;
  const bool want_gy = A.gY != nullptr;
  // The group's A-scans: SAFT elements n0.. ; two-leg transmits i0.. x receives j0..
  int n0 = 0, i0 = 0, j0 = 0;
  if (TYPE == DAS_SAFT) n0 = grp * kSaftGroup;
  else { i0 = (grp / n_gj) * kGI; j0 = (grp % n_gj) * kGJ; }

  const GT *Ro = (const GT *) A.Ro;
  const GT *Gt = (const GT *) A.Gt;
  const GT *Gr = (const GT *) A.Gr;
  const GT fs = (GT) A.fs_khz, inv_cp = (GT) A.inv_cp, dl = (GT) A.delay_ms;
  const GT tmax = (GT) A.leg_max;
  const GT bnd = (GT) (0x1p-40 * (fabs(2.0 * (double) fs * (double) dl) + 1.0));
  const GT fs2 = fs + fs;
  const DT *g = (const DT *) A.g + frame * A.im_stride;
  const DT *at = (const DT *) A.apod_t;
  const DT *ar = (const DT *) A.apod_r;
  GT y_lo = 0, y_hi = 0, x_lo = 0, x_hi = 0;
  if (TYPE == DAS_RCA_COLTX || TYPE == DAS_RCA_ROWTX) {
    y_lo = Gr[Lt]; y_hi = Gr[2 * Lt - 1]; x_lo = Gt[0]; x_hi = Gt[Lr - 1];
  }

  // The work-group covers `subs` consecutive regions, one after the other,
  // with one set of windows (flushed once) spanning all of them.
  // Lanes take their 64-pixel block transposed (lane l -> position
  // 8 (l mod 8) + l / 8): the visiting order runs along x inside a block,
  // and consecutive lanes a row apart scatter into different LDS bins.
  const int lane_pos = (int) ((threadIdx.x & ~63u) | ((threadIdx.x & 7u) << 3) | ((threadIdx.x >> 3) & 7u));
  const int64_t base0 = region * (int64_t) subs * kRegion;
  // Bounding box of the work-group's pixels (only needed for the LDS windows).
  if (want_gy) {
    GT bmin[3] = {INFINITY, INFINITY, INFINITY}, bmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t idx = base0 + lane_pos; idx < base0 + (int64_t) subs * kRegion && idx < A.No;
         idx += kThreads) {
      const int64_t p = pixel_of(perm, idx);
      const GT x = Ro[p], y = Ro[p + A.No], z = Ro[p + 2 * A.No];
      bmin[0] = fmin(bmin[0], x); bmax[0] = fmax(bmax[0], x);
      bmin[1] = fmin(bmin[1], y); bmax[1] = fmax(bmax[1], y);
      bmin[2] = fmin(bmin[2], z); bmax[2] = fmax(bmax[2], z);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
      GT mn = bmin[c], mx = bmax[c];
      for (int off = 32; off > 0; off >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, off, 64));
        mx = fmax(mx, __shfl_xor(mx, off, 64));
      }
      if (lane == 0) { s_box[c][wave] = mn; s_box[3 + c][wave] = mx; }
    }
    for (int s = threadIdx.x; s < NA * kWin; s += kThreads) win[s] = 0.0;
    __syncthreads();
    if (threadIdx.x < NA) {
      GT lo[3], hi[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        lo[c] = s_box[c][0]; hi[c] = s_box[3 + c][0];
        for (int w = 1; w < kThreads / 64; w++) {
          lo[c] = fmin(lo[c], s_box[c][w]); hi[c] = fmax(hi[c], s_box[3 + c][w]);
        }
      }
      // Distance ranges of the legs to the box, then the sample range.
      double umin;
      if (TYPE == DAS_SAFT) {
        const int n = min(n0 + (int) threadIdx.x, Lt - 1);
        GT m0, m1, m2, x0, x1, x2;
        gap_range<GT>(Gt[n], lo[0], hi[0], m0, x0);
        gap_range<GT>(Gt[n + Lt], lo[1], hi[1], m1, x1);
        gap_range<GT>(Gt[n + 2 * Lt], lo[2], hi[2], m2, x2);
        const double dmin = sqrt((double) (m0 * m0 + m1 * m1 + m2 * m2));
        umin = 2.0 * (dmin * (double) inv_cp + (double) dl) * (double) fs;
      } else {
        const int i = min(i0 + (int) threadIdx.x / kGJ, Lt - 1);
        const int j = min(j0 + (int) threadIdx.x % kGJ, Lr - 1);
        GT m0, m1, m2, x0, x1, x2;
        if (TYPE == DAS_TFM) {
          gap_range<GT>(Gt[i], lo[0], hi[0], m0, x0);
          gap_range<GT>(Gt[i + Lt], lo[1], hi[1], m1, x1);
        } else {
          gap_range<GT>(Gt[i], lo[0], hi[0], m0, x0);
          stripe_range<GT>(lo[1], hi[1], y_lo, y_hi, m1, x1);
        }
        gap_range<GT>(Gt[i + 2 * Lt], lo[2], hi[2], m2, x2);
        const double dt_min = sqrt((double) (m0 * m0 + m1 * m1 + m2 * m2));
        if (TYPE == DAS_TFM) {
          gap_range<GT>(Gr[j], lo[0], hi[0], m0, x0);
        } else {
          stripe_range<GT>(lo[0], hi[0], x_lo, x_hi, m0, x0);
        }
        gap_range<GT>(Gr[j + Lr], lo[1], hi[1], m1, x1);
        gap_range<GT>(Gr[j + 2 * Lr], lo[2], hi[2], m2, x2);
        const double dr_min = sqrt((double) (m0 * m0 + m1 * m1 + m2 * m2));
        umin = ((dt_min * (double) inv_cp + (double) dl) + dr_min * (double) inv_cp) * (double) fs;
      }
      s_w0[threadIdx.x] = window_start(umin);
    }
  }
  __syncthreads();

  const __amdgpu_buffer_rsrc_t ry =
      frame_rsrc((const DT *) A.Y + frame * A.y_stride, A.y_stride * (int64_t) sizeof(DT));
  DT *gy_frame = want_gy ? (DT *) A.gY + frame * A.y_stride : nullptr;
  DT zsum[NA];
  DT zc = DT(0);
#pragma unroll
  for (int s = 0; s < NA; s++) zsum[s] = DT(0);

  for (int sub = 0; sub < subs; sub++) {
    // The lane's pixels in this region.
    GT xo[kRPU], yo[kRPU], zo[kRPU];
    DT gp[kRPU];
    bool live[kRPU];
    const int64_t base = base0 + (int64_t) sub * kRegion;
#pragma unroll
    for (int q = 0; q < kRPU; q++) {
      const int64_t idx = base + (int64_t) q * kThreads + lane_pos;
      live[q] = idx < A.No;
      const int64_t p = pixel_of(perm, live[q] ? idx : 0);
      xo[q] = Ro[p]; yo[q] = Ro[p + A.No]; zo[q] = Ro[p + 2 * A.No];
      gp[q] = live[q] ? g[p] : DT(0);
    }

    // Two-leg: receive legs of the group for the lane's pixels.
    GT tr[TYPE == DAS_SAFT ? 1 : kGJ][kRPU];
    GT gx_r[kRPU], gy_c[kRPU];
    if (TYPE != DAS_SAFT) {
#pragma unroll
      for (int q = 0; q < kRPU; q++) {
        gx_r[q] = GT(0); gy_c[q] = GT(0);
        if (TYPE != DAS_TFM) {
          gy_c[q] = stripe_gap<GT>(yo[q], y_lo, y_hi);
          gx_r[q] = stripe_gap<GT>(xo[q], x_lo, x_hi);
        }
#pragma unroll
        for (int jj = 0; jj < kGJ; jj++) {
          const int j = min(j0 + jj, Lr - 1);
          if (TYPE == DAS_TFM)
            tr[jj][q] = leg3<GT>(Gr[j] - xo[q], Gr[j + Lr] - yo[q], Gr[j + 2 * Lr] - zo[q]) * inv_cp;
          else
            tr[jj][q] = leg3<GT>(gx_r[q], Gr[j + Lr] - yo[q], Gr[j + 2 * Lr] - zo[q]) * inv_cp;
        }
      }
    }

#pragma unroll
    for (int s = 0; s < NA; s++) {
      int a, ii = 0, jj = 0;
      bool ok;
      DT w;
      if (TYPE == DAS_SAFT) {
        const int n = n0 + s;
        ok = n < Lt;
        a = min(n, Lt - 1);
        w = APOD ? at[a] : DT(1);
      } else {
        ii = s / kGJ; jj = s % kGJ;
        const int i = i0 + ii, j = j0 + jj;
        ok = i < Lt && j < Lr;
        const int ic = min(i, Lt - 1), jc = min(j, Lr - 1);
        a = (TYPE == DAS_RCA_ROWTX) ? jc * Lt + ic : ic * Lr + jc;
        w = APOD ? at[ic] * ar[jc] : DT(1);
      }
      if (!ok) continue;                           // uniform
      const unsigned abase = (unsigned) a * (unsigned) K;
      const int w0 = want_gy ? s_w0[s] : 0;
      GT u[kRPU], tsum[kRPU];
#pragma unroll
      for (int q = 0; q < kRPU; q++) {
        if (TYPE == DAS_SAFT) {
          GT tleg;
          u[q] = saft_pos<GT, kSaftFast<DT, GT, INTERP>>(Gt[a] - xo[q], Gt[a + Lt] - yo[q],
                                                        Gt[a + 2 * Lt] - zo[q], inv_cp, dl, tmax,
                                                        fs, fs2, bnd, tleg);
          tsum[q] = tleg + tleg;
        } else {
          const int i = min(i0 + ii, Lt - 1);
          GT tt;
          if (TYPE == DAS_TFM)
            tt = leg3<GT>(Gt[i] - xo[q], Gt[i + Lt] - yo[q], Gt[i + 2 * Lt] - zo[q]) * inv_cp;
          else
            tt = leg3<GT>(Gt[i] - xo[q], gy_c[q], Gt[i + 2 * Lt] - zo[q]) * inv_cp;
          tsum[q] = tt + tr[jj][q];
          tt += dl;
          u[q] = (clamp_leg<GT>(tt, tmax) + clamp_leg<GT>(tr[jj][q], tmax)) * fs;
        }
      }
      DT y0[kRPU], y1[kRPU], f[kRPU];
      int k[kRPU];
#pragma unroll
      for (int q = 0; q < kRPU; q++)
        fetch_taps<DT, GT, INTERP>(ry, abase, K, u[q], live[q], y0[q], y1[q], f[q], k[q]);
#pragma unroll
      for (int q = 0; q < kRPU; q++) {
        const DT slope = INTERP == DREAM_HIP_LINEAR ? y1[q] - y0[q] : DT(0);
        zsum[s] += gp[q] * (y0[q] + f[q] * slope);
        if (INTERP == DREAM_HIP_LINEAR) zc += gp[q] * w * slope * (DT) tsum[q];
        if (want_gy && k[q] > -2) {
          const DT gw = gp[q] * w;
          if (INTERP == DREAM_HIP_NEAREST) {
            const unsigned l = (unsigned) (k[q] - w0);
            if (l < (unsigned) kWin) lds_atomic_add(&win[s * kWin + l], (double) gw);
            else unsafeAtomicAdd(&gy_frame[abase + k[q]], gw);
          } else {
            const unsigned l = (unsigned) (k[q] - w0);
            const DT v0 = gw * (DT(1) - f[q]), v1 = gw * f[q];
            if (l + 1u < (unsigned) kWin) {         // both taps in the window
              lds_atomic_add(&win[s * kWin + l], (double) v0);
              lds_atomic_add(&win[s * kWin + l + 1], (double) v1);
            } else {
              if (k[q] >= 0) unsafeAtomicAdd(&gy_frame[abase + k[q]], v0);
              if (k[q] + 1 < K) unsafeAtomicAdd(&gy_frame[abase + k[q] + 1], v1);
            }
          }
        }
      }
    }

  }  // sub-regions

  // Per-A-scan sums and the cp sum: wave, then work-group, then one atomic.
#pragma unroll
  for (int s = 0; s < NA; s++) {
    const double v = wave_sum_d((double) zsum[s]);
    if (lane == 0) s_z[s][wave] = v;
  }
  {
    const double v = wave_sum_d((double) zc);
    if (lane == 0) s_zc[wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < NA) {
    const int s = threadIdx.x;
    int a = -1;
    if (TYPE == DAS_SAFT) { if (n0 + s < Lt) a = n0 + s; }
    else {
      const int i = i0 + s / kGJ, j = j0 + s % kGJ;
      if (i < Lt && j < Lr) a = (TYPE == DAS_RCA_ROWTX) ? j * Lt + i : i * Lr + j;
    }
    double v = 0.0;
    for (int w = 0; w < kThreads / 64; w++) v += s_z[s][w];
    if (a >= 0 && v != 0.0) atomicAdd(&((double *) A.z_t)[a], v);
  }
  if (threadIdx.x == 0 && INTERP == DREAM_HIP_LINEAR) {
    double v = 0.0;
    for (int w = 0; w < kThreads / 64; w++) v += s_zc[w];
    v *= -(double) A.fs_khz / A.cp;             // du/dcp = -(travel time) fs / cp
    if (v != 0.0) atomicAdd((double *) A.z_cp, v);
  }
  // Windows into gY.
  if (want_gy) {
    for (int q = threadIdx.x; q < NA * kWin; q += kThreads) {
      const int s = q / kWin;
      const DT v = (DT) win[q];
      if (v == DT(0)) continue;
      int a;
      if (TYPE == DAS_SAFT) a = n0 + s;
      else {
        const int i = i0 + s / kGJ, j = j0 + s % kGJ;
        a = (TYPE == DAS_RCA_ROWTX) ? j * Lt + i : i * Lr + j;
      }
      const int kk = s_w0[s] + (q - s * kWin);
      if (kk < 0 || kk >= K) continue;           // a linear tap just outside the A-scan
      unsafeAtomicAdd(&gy_frame[(int64_t) a * K + kk], v);
    }
  }
}

// Pixel-major adjoint for two-leg modes without gY: accumulates
// z_t[i] = sum_p g_p sum_j w_j v(p,i,j), z_r[j] = sum_p g_p sum_i w_i v(p,i,j)
// and z_cp, in LDS per work-group, flushed once.  Grid-stride over
// (frame, 256-pixel tile); gathers as in the forward kernel.  FR = 4: over
// (frame group, tile), each gather bringing four frames' samples, which
// the lane weights with its four g values (nearest-lower sampling).
template <typename DT, typename GT, int TYPE, int INTERP, bool APOD, int FR = 1>
__global__ void __launch_bounds__(kThreads)
das_adj_pixel_kernel(DasArgs A, const int32_t *__restrict__ perm, int64_t n_tiles)
{
  static_assert(FR == 1 || (INTERP == DREAM_HIP_NEAREST && std::is_same<DT, float>::value),
                "frame groups: FP32, nearest-lower sampling");
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
  const GT y_lo = Gr[Lt], y_hi = Gr[2 * Lt - 1], x_lo = Gt[0], x_hi = Gt[Lr - 1];
  const GT tmax = (GT) A.leg_max;
  double zcp = 0.0;
  const bool lane0 = (threadIdx.x & 63) == 0;

  for (int64_t tile = A.block0 + blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    const int64_t frame = tile / tiles_per_frame;
    const int64_t idx = (tile - frame * tiles_per_frame) * kThreads + threadIdx.x;
    const bool valid = idx < A.No;
    const int64_t p = valid ? pixel_of(perm, idx) : 0;
    const GT xo = Ro[p], yo = Ro[p + A.No], zo = Ro[p + 2 * A.No];
    // `frame` is the frame group when FR = 4: its g values in gp4, and gp
    // (the weight of the per-row sums) 1.
    const DT *gv = (const DT *) A.g;
    DT gp = DT(0);
    frames4 gp4 = frames4{};
    if constexpr (FR == 4) {
      gp = DT(1);
#pragma unroll
      for (int q = 0; q < 4; q++)
        gp4[q] = valid && 4 * frame + q < A.n_frames ? gv[(4 * frame + q) * A.im_stride + p] : 0.0f;
    } else {
      gp = valid ? gv[frame * A.im_stride + p] : DT(0);
    }
    const DT *Yf = (const DT *) A.Y + (FR == 4 ? 0 : frame * A.y_stride);
    const __amdgpu_buffer_rsrc_t ry =
        FR == 4 ? ascan_rsrc((const frames4 *) A.Yg + frame * A.yg_stride, K, 16)
                : ascan_rsrc(Yf, K, (int) sizeof(DT));
    const __amdgpu_buffer_rsrc_t rf = frame_rsrc(Yf, A.y_stride * (int64_t) sizeof(DT));
    const unsigned KB = (unsigned) K * (unsigned) (FR == 4 ? 16 : sizeof(DT));
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
      for (int i = 0; i < Lt; i++) {
        GT tt;
        if (TYPE == DAS_TFM)
          tt = leg3<GT>(Gt[i] - xo, Gt[i + Lt] - yo, Gt[i + 2 * Lt] - zo) * inv_cp;
        else
          tt = leg3<GT>(Gt[i] - xo, gy_c, Gt[i + 2 * Lt] - zo) * inv_cp;
        const GT tt0 = tt;
        tt += dl;
        const DT wi = APOD ? at[i] : DT(1);
        // Nearest-lower taps through the structured view of the frame (0
        // outside [0, K) and for pairs that are not live); linear ones
        // through the frame view with their own bounds tests (fewer
        // registers there).
        DT y0[kRC], y1[kRC], f[kRC];
#pragma unroll
        for (int jj = 0; jj < kRC; jj++) {
          const int j = min(j0 + jj, Lr - 1);
          const int a = (TYPE == DAS_RCA_ROWTX) ? j * Lt + i : i * Lr + j;
          const GT u = (clamp_leg<GT>(tt, tmax) + clamp_leg<GT>(tr[jj], tmax)) * fs;
          const bool live = valid && j0 + jj < Lr;
          if (INTERP == DREAM_HIP_NEAREST) {
            unsigned i0, i1;
            int kk;
            gather_index<DT, GT, INTERP, false>(u, live, i0, i1, f[jj], kk);
            if constexpr (FR == 4) {
              const frames4 v = sload<frames4>(ry, i0, (unsigned) a * KB);
              y0[jj] = gp4[0] * v[0] + gp4[1] * v[1] + gp4[2] * v[2] + gp4[3] * v[3];
            } else {
              y0[jj] = sload<DT>(ry, i0, (unsigned) a * KB);
            }
            y1[jj] = y0[jj];
            f[jj] = DT(0);
          } else {
            int kk;
            fetch_taps<DT, GT, INTERP>(rf, (unsigned) a * (unsigned) K, K, u, live, y0[jj],
                                       y1[jj], f[jj], kk);
          }
        }
        DT zi = DT(0), zc = DT(0);
#pragma unroll
        for (int jj = 0; jj < kRC; jj++) {
          const DT wj = APOD ? ar[min(j0 + jj, Lr - 1)] : DT(1);
          const DT slope = INTERP == DREAM_HIP_LINEAR ? y1[jj] - y0[jj] : DT(0);
          const DT val = y0[jj] + f[jj] * slope;
          zi += wj * val;
          zj[jj] += wi * val;
          if (INTERP == DREAM_HIP_LINEAR) zc += wi * wj * slope * (DT) (tt0 + tr[jj]);
        }
        const DT zsum = wave_sum_t(gp * zi);
        if (lane0 && zsum != DT(0)) atomicAdd(&zt[i], (double) zsum);
        if (INTERP == DREAM_HIP_LINEAR) zcp += (double) (gp * zc);
      }
#pragma unroll
      for (int jj = 0; jj < kRC; jj++) {
        if (j0 + jj < Lr) {
          const DT zsum = wave_sum_t(gp * zj[jj]);
          if (lane0 && zsum != DT(0)) atomicAdd(&zr[j0 + jj], (double) zsum);
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

// From per-A-scan sums Z[a] (region kernel) to the apodization gradients:
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

// The reference reads the RCA stripe extents at rows[Lt], rows[2 Lt - 1]
// (y column of the rows, offset by the column count) and columns[Lr - 1]
// (extra_src/das.cc:464-467, :560-563); those entries exist in the 3 x L
// tables only when 2 Lt <= 3 Lr and Lr <= 3 Lt.
bool rca_extents_in_bounds(int64_t Lt, int64_t Lr)
{
  return 2 * Lt <= 3 * Lr && Lr <= 3 * Lt;
}

// DREAM_HIP_RECIPROCAL=0 switches the reciprocal TFM forward off (the full
// pair sum runs), for comparisons.
// Frame groups (DasArgs::Yg) for multi-frame FP32 two-leg forwards;
// DREAM_HIP_FRAME_GROUPS=0 gives one frame per launch row (for comparisons).
bool frame_groups_enabled()
{
  static const bool on = [] {
    const char *e = std::getenv("DREAM_HIP_FRAME_GROUPS");
    return !(e && e[0] == '0');
  }();
  return on;
}

bool recip_enabled()
{
  static const bool on = [] {
    const char *e = std::getenv("DREAM_HIP_RECIPROCAL");
    return !(e && e[0] == '0');
  }();
  return on;
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

// Launch shape of the forward kernel: `waves` waves per work-group share a
// 64-pixel tile; `split` work-groups share a tile (partial sums + reduce
// kernel) only when tiles x kMaxWaves waves cannot fill the device.
struct FwdShape {
  int waves, split;
};

// Jobs per 64-pixel tile.
int64_t forward_jobs(int type, int64_t Lt, int64_t Lr, bool recip)
{
  if (type == DAS_SAFT) return (Lt + kSB - 1) / kSB;
  const int64_t nc = (Lr + kRC - 1) / kRC;
  return recip ? nc * (nc + 1) / 2 : nc * ((Lt + kTB - 1) / kTB);
}

FwdShape forward_shape(int type, int64_t No, int64_t n_frames, int64_t Lt, int64_t Lr,
                       bool recip = false)
{
  const int64_t tiles = ((No + 63) / 64) * n_frames;
  const int64_t target = (int64_t) cu_count() * 32;      // 8 waves per SIMD
  const int64_t J = forward_jobs(type, Lt, Lr, recip);
  FwdShape sh{4, 1};
  if (tiles == 0) return sh;                              // no pixels: nothing launches
  if (tiles * 4 < target) sh.waves = kMaxWaves;
  sh.waves = (int) std::max<int64_t>(1, std::min<int64_t>(sh.waves, J));
  const int64_t waves_total = tiles * sh.waves;
  if (waves_total * 2 <= target) {
    const int64_t max_split = std::max<int64_t>(1, (J + sh.waves - 1) / sh.waves);
    sh.split = (int) std::min<int64_t>(std::min<int64_t>((target + waves_total - 1) / waves_total,
                                                         max_split), 64);
  }
  return sh;
}

// Candidate launch shapes for the autotuner; the heuristic shape first.
std::vector<FwdShape> forward_candidates(int type, int64_t No, int64_t n_frames, int64_t Lt,
                                         int64_t Lr, bool recip = false)
{
  const FwdShape h = forward_shape(type, No, n_frames, Lt, Lr, recip);
  const int64_t J = forward_jobs(type, Lt, Lr, recip);
  std::vector<FwdShape> c{h};
  const int splits[3] = {1, h.split, 2 * h.split};
  // 2, 4, 8 waves, and the wave counts that deal J jobs evenly (ceil(J / k)).
  std::vector<int> waves{2, 4, 8};
  for (int64_t k = 1; k <= 4; k++) {
    const int64_t w = (J + k - 1) / k;
    if (w >= 2 && w <= kMaxWaves && std::find(waves.begin(), waves.end(), (int) w) == waves.end())
      waves.push_back((int) w);
  }
  auto add = [&](int w, int sp) {
    if (w > std::max<int64_t>(J, 1) || sp > 64) return;
    if ((int64_t) w * sp > std::max<int64_t>(J, 1) * 2) return;     // most waves idle
    // Split partial sums (up to 8 bytes each) beyond 4 GiB are not worth
    // the memory.
    if (sp > 1 && (double) sp * (double) n_frames * (double) No * 8.0 > 4294967296.0) return;
    for (const FwdShape &x : c)
      if (x.waves == w && x.split == sp) return;
    c.push_back(FwdShape{w, sp});
  };
  for (int w : waves)
    for (int sp : splits) add(w, sp);
  // Small work-groups and many per tile: fewer jobs per wave and more,
  // smaller work-groups (measured faster for the FP64 and reference-float
  // TFM images and the eight-frame RCA call).
  add(1, 8);
  add(1, 16);
  add(2, 4);
  // One or two waves per work-group with as many work-groups per tile as
  // deal the jobs evenly, k jobs per wave.
  for (int64_t k = 1; k <= 3; k++) {
    const int64_t per = (J + k - 1) / k;
    for (int w = 1; w <= 2; w++)
      if (per % w == 0) add(w, (int) (per / w));
  }
  return c;
}

template <typename DT, typename GT, int TYPE, int INTERP, bool APOD>
void fwd_launch(const DasArgs &a, hipStream_t s, ChunkGate *gate);

// Pixel-channel pairs per pixel of one frame.
inline int64_t pairs_per_pixel(int type, int64_t Lt, int64_t Lr)
{
  return type == DAS_SAFT ? Lt : Lt * Lr;
}

// Runs the (blocks x frames) grid of a launch as chunks of about
// chunk_work() pairs, `work_per_block` pairs each block: whole frames per
// chunk when a frame is small, else block ranges within a frame.  `launch`
// takes (block0, n_blocks, frame0, n_frames); a gate tests cancellation
// between chunks (none when the call is one chunk).
template <class F>
void run_chunked(int64_t blocks, int64_t frames, int64_t work_per_block, ChunkGate *gate,
                 F launch)
{
  const int64_t per_chunk = std::max<int64_t>(1, chunk_work() / std::max<int64_t>(work_per_block, 1));
  if (!gate || per_chunk >= blocks * frames) {
    launch((int64_t) 0, blocks, (int64_t) 0, frames);
    return;
  }
  if (per_chunk >= blocks) {
    const int64_t fpc = per_chunk / blocks;
    for (int64_t f0 = 0; f0 < frames; f0 += fpc) {
      gate->before();
      launch((int64_t) 0, blocks, f0, std::min(fpc, frames - f0));
      gate->after();
    }
    return;
  }
  for (int64_t f0 = 0; f0 < frames; f0++)
    for (int64_t b0 = 0; b0 < blocks; b0 += per_chunk) {
      gate->before();
      launch(b0, std::min(per_chunk, blocks - b0), f0, (int64_t) 1);
      gate->after();
    }
}

template <typename DT, typename GT, int TYPE, int INTERP, bool APOD>
void fwd_launch_shape(const DasArgs &a0, FwdShape sh, hipStream_t s, ChunkGate *gate = nullptr)
{
  if (sh.split > 1 && (size_t) sh.split * a0.n_frames * a0.No * sizeof(DT) > a0.scratch_bytes)
    sh.split = 1;
  dim3 block((unsigned) (64 * sh.waves));
  const int64_t tiles = (a0.No + 63) / 64;
  // Launch rows: frames, or groups of four frames (a0.Yg, FP32 two-leg).
  constexpr bool kGroups = std::is_same<DT, float>::value && TYPE != DAS_SAFT;
  const bool groups = kGroups && a0.Yg != nullptr;
  const int64_t n_rows = groups ? (a0.n_frames + 3) / 4 : a0.n_frames;
  // A chunk is a range of tiles of some launch rows; with split partial
  // sums its pixels are reduced right after it, so every launched chunk's
  // pixels are written when a later chunk is cancelled.
  auto launch = [&](int64_t b0, int64_t nb, int64_t f0, int64_t nf) {
    DasArgs a = a0;
    a.block0 = b0;
    a.frame0 = f0;
    dim3 grid((unsigned) nb, (unsigned) sh.split, (unsigned) nf);
    // The reciprocal TFM kernel is its own instance (no apodization), so the
    // other kernels keep their registers.
    constexpr bool kRecip = TYPE == DAS_TFM && !APOD;
    if (groups) {
      if constexpr (kGroups) {
        if (a.status)
          hipLaunchKernelGGL((das_fwd_kernel<DT, GT, TYPE, INTERP, APOD, true, false, 4>), grid,
                             block, 0, s, a, a.perm, sh.split, (DT *) a.scratch);
        else
          hipLaunchKernelGGL((das_fwd_kernel<DT, GT, TYPE, INTERP, APOD, false, false, 4>), grid,
                             block, 0, s, a, a.perm, sh.split, (DT *) a.scratch);
      }
    } else if (kRecip && a.recip) {
      if (a.status)
        hipLaunchKernelGGL((das_fwd_kernel<DT, GT, TYPE, INTERP, APOD, true, kRecip>), grid,
                           block, 0, s, a, a.perm, sh.split, (DT *) a.scratch);
      else
        hipLaunchKernelGGL((das_fwd_kernel<DT, GT, TYPE, INTERP, APOD, false, kRecip>), grid,
                           block, 0, s, a, a.perm, sh.split, (DT *) a.scratch);
    } else if (a.status) {
      hipLaunchKernelGGL((das_fwd_kernel<DT, GT, TYPE, INTERP, APOD, true>), grid, block, 0, s,
                         a, a.perm, sh.split, (DT *) a.scratch);
    } else {
      hipLaunchKernelGGL((das_fwd_kernel<DT, GT, TYPE, INTERP, APOD, false>), grid, block, 0, s,
                         a, a.perm, sh.split, (DT *) a.scratch);
    }
    DREAM_HIP_CHECK(hipGetLastError());
    if (sh.split > 1) {
      const int64_t n_idx = std::min<int64_t>(nb * 64, a.No - b0 * 64);
      const int64_t fr0 = groups ? 4 * f0 : f0;
      const int64_t frn = groups ? std::min<int64_t>(4 * (f0 + nf), a.n_frames) - fr0 : nf;
      const int64_t n = n_idx * frn;
      DT *out = (DT *) a.Im + fr0 * a.No;
      if (b0 == 0 && n_idx == a.No && a.im_stride == a.No && a.No % 4 == 0 &&
          (uintptr_t) out % (4 * sizeof(DT)) == 0) {
        hipLaunchKernelGGL(das_reduce_flat_kernel<DT>, dim3((unsigned) ((n / 4 + 255) / 256)),
                           dim3(256), 0, s, (const DT *) a.scratch + fr0 * a.No, sh.split,
                           a.n_frames * a.No, n / 4, out);
      } else {
        hipLaunchKernelGGL(das_reduce_kernel<DT>, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0,
                           s, (const DT *) a.scratch, sh.split, a.No, a.n_frames, fr0, frn, b0 * 64,
                           n_idx, a.perm, a.im_stride, (DT *) a.Im);
      }
      DREAM_HIP_CHECK(hipGetLastError());
    }
  };
  run_chunked(tiles, n_rows, 64 * pairs_per_pixel(TYPE, a0.Lt, a0.Lr) * (groups ? 4 : 1), gate,
              launch);
}

// The heuristic shape, or the autotuner's pick for this shape class (with
// or without the out-of-bounds status).
template <typename DT, typename GT, int TYPE, int INTERP, bool APOD>
void fwd_launch(const DasArgs &a, hipStream_t s, ChunkGate *gate)
{
  const std::vector<FwdShape> cand =
      forward_candidates(TYPE, a.No, a.n_frames, a.Lt, a.Lr, a.recip != 0);
  int pick = 0;
  if (cand.size() > 1) {
    char key[192];
    std::snprintf(key, sizeof key, "das_fwd|%d|%d|%d|%d|%d|%lld|%d|%d|%lld|%lld|%d|%d|%d", TYPE,
                  (int) sizeof(DT), (int) sizeof(GT), INTERP, (int) APOD, (long long) a.No, a.Lt,
                  a.Lr, (long long) a.K, (long long) a.n_frames, (int) a.recip,
                  (int) (a.status != nullptr), (int) (a.Yg != nullptr));
    const int t = autotune(key, (int) cand.size(),
                           [&](int i, hipStream_t st) { fwd_launch_shape<DT, GT, TYPE, INTERP, APOD>(a, cand[(size_t) i], st); },
                           s);
    if (t >= 0) pick = t;
  }
  fwd_launch_shape<DT, GT, TYPE, INTERP, APOD>(a, cand[(size_t) pick], s, gate);
}

template <typename DT, typename GT, int TYPE>
void fwd_dispatch_interp(const DasArgs &a, int interp, bool apod, hipStream_t s, ChunkGate *g)
{
  if (interp == DREAM_HIP_NEAREST) {
    if (apod) fwd_launch<DT, GT, TYPE, DREAM_HIP_NEAREST, true>(a, s, g);
    else fwd_launch<DT, GT, TYPE, DREAM_HIP_NEAREST, false>(a, s, g);
  } else {
    if (apod) fwd_launch<DT, GT, TYPE, DREAM_HIP_LINEAR, true>(a, s, g);
    else fwd_launch<DT, GT, TYPE, DREAM_HIP_LINEAR, false>(a, s, g);
  }
}

template <typename DT, typename GT>
void fwd_dispatch_type(const DasArgs &a, int type, int interp, bool apod, hipStream_t s,
                       ChunkGate *g)
{
  switch (type) {
  case DAS_SAFT: fwd_dispatch_interp<DT, GT, DAS_SAFT>(a, interp, apod, s, g); break;
  case DAS_TFM: fwd_dispatch_interp<DT, GT, DAS_TFM>(a, interp, apod, s, g); break;
  case DAS_RCA_COLTX: fwd_dispatch_interp<DT, GT, DAS_RCA_COLTX>(a, interp, apod, s, g); break;
  default: fwd_dispatch_interp<DT, GT, DAS_RCA_ROWTX>(a, interp, apod, s, g); break;
  }
}

template <typename DT, typename GT, int TYPE, int INTERP, bool APOD>
void adj_launch(const DasArgs &a, const int32_t *perm, bool want_gy, hipStream_t s,
                ChunkGate *gate)
{
  if (TYPE == DAS_SAFT || want_gy) {
    // Region-major: (pixel region, A-scan group, frame) work-groups.
    // A work-group takes `subs` consecutive regions (one window flush for
    // all of them).  Candidates 1, 2, 4 while the grid keeps at least four
    // work-groups per CU; the autotuner picks (the windows of a larger
    // group cover a wider delay range, which coarse pixel grids can
    // overflow into direct gY atomics), the first is the default.
    const int64_t regions1 = (a.No + kRegion - 1) / kRegion;
    int n_gj = 1;
    int64_t groups;
    if (TYPE == DAS_SAFT) {
      groups = (a.Lt + kSaftGroup - 1) / kSaftGroup;
    } else {
      n_gj = (a.Lr + kGJ - 1) / kGJ;
      groups = (int64_t) ((a.Lt + kGI - 1) / kGI) * n_gj;
    }
    // Largest first (the default), then smaller; a fixed array, so the
    // launch takes no host memory.
    int subs_cand[3], n_subs = 0;
    for (int sb = 4; sb >= 2; sb /= 2)
      if ((regions1 + sb - 1) / sb * groups * a.n_frames >= 4 * (int64_t) cu_count())
        subs_cand[n_subs++] = sb;
    subs_cand[n_subs++] = 1;
    const int na = TYPE == DAS_SAFT ? kSaftGroup : kGI * kGJ;
    const size_t lds = want_gy ? (size_t) na * kWin * sizeof(double) : 0;
    const int64_t n_z = TYPE == DAS_SAFT ? a.Lt : (int64_t) a.Lt * a.Lr;
    auto run = [&](int subs, bool clear_z, hipStream_t st, ChunkGate *g) {
      if (clear_z) {
        zero_async(a.z_t, (size_t) n_z * sizeof(double), st);
        zero_async(a.z_cp, sizeof(double), st);
      }
      if (want_gy) zero_async(a.gY, (size_t) (a.n_frames * a.y_stride) * sizeof(DT), st);
      const int64_t regions = (regions1 + subs - 1) / subs;
      const int64_t per_region = (int64_t) subs * kRegion * pairs_per_pixel(TYPE, a.Lt, a.Lr);
      run_chunked(regions, a.n_frames, per_region, g,
                  [&](int64_t b0, int64_t nb, int64_t f0, int64_t nf) {
                    DasArgs c = a;
                    c.block0 = b0;
                    c.frame0 = f0;
                    dim3 grid((unsigned) nb, (unsigned) groups, (unsigned) nf);
                    hipLaunchKernelGGL((das_adj_region_kernel<DT, GT, TYPE, INTERP, APOD>), grid,
                                       dim3(kThreads), lds, st, c, perm, n_gj, subs);
                    DREAM_HIP_CHECK(hipGetLastError());
                  });
    };
    int pick = 0;
    if (n_subs > 1 && !autotune_suspended()) {
      char key[192];
      std::snprintf(key, sizeof key, "das_adj|%d|%d|%d|%d|%d|%d|%lld|%d|%d|%lld|%lld", TYPE,
                    (int) sizeof(DT), (int) sizeof(GT), INTERP, (int) APOD, (int) want_gy,
                    (long long) a.No, a.Lt, a.Lr, (long long) a.K, (long long) a.n_frames);
      const int before = autotune_cache_size();
      const int t = autotune(key, n_subs,
                             [&](int i, hipStream_t st) { run(subs_cand[i], true, st, nullptr); },
                             s);
      if (t >= 0) pick = t;
      // Candidates ran here: start the accumulators again.
      if (autotune_cache_size() != before) {
        zero_async(a.z_t, (size_t) n_z * sizeof(double), s);
        zero_async(a.z_cp, sizeof(double), s);
      }
    }
    run(subs_cand[pick], false, s, gate);
  } else {
    // Grid-stride over (frame or frame group, 256-pixel tile); a chunk is a
    // range of tiles.
    constexpr bool kGroups = std::is_same<DT, float>::value && INTERP == DREAM_HIP_NEAREST;
    const bool groups = kGroups && a.Yg != nullptr;
    const int64_t rows = groups ? (a.n_frames + 3) / 4 : a.n_frames;
    const int64_t tiles = ((a.No + kThreads - 1) / kThreads) * rows;
    const size_t lds = (size_t) (a.Lt + a.Lr) * sizeof(double);
    run_chunked(tiles, 1,
                (int64_t) kThreads * pairs_per_pixel(TYPE, a.Lt, a.Lr) * (groups ? 4 : 1), gate,
                [&](int64_t b0, int64_t nb, int64_t, int64_t) {
                  DasArgs c = a;
                  c.block0 = b0;
                  const int grid = (int) std::min<int64_t>(nb, (int64_t) cu_count() * 4);
                  if (groups) {
                    if constexpr (kGroups)
                      hipLaunchKernelGGL((das_adj_pixel_kernel<DT, GT, TYPE, INTERP, APOD, 4>),
                                         dim3(grid), dim3(kThreads), lds, s, c, perm, b0 + nb);
                  } else {
                    hipLaunchKernelGGL((das_adj_pixel_kernel<DT, GT, TYPE, INTERP, APOD>),
                                       dim3(grid), dim3(kThreads), lds, s, c, perm, b0 + nb);
                  }
                  DREAM_HIP_CHECK(hipGetLastError());
                });
  }
}

template <typename DT, typename GT, int TYPE>
void adj_dispatch_interp(const DasArgs &a, const int32_t *perm, int interp, bool apod,
                         bool want_gy, hipStream_t s, ChunkGate *g)
{
  if (interp == DREAM_HIP_NEAREST) {
    if (apod) adj_launch<DT, GT, TYPE, DREAM_HIP_NEAREST, true>(a, perm, want_gy, s, g);
    else adj_launch<DT, GT, TYPE, DREAM_HIP_NEAREST, false>(a, perm, want_gy, s, g);
  } else {
    if (apod) adj_launch<DT, GT, TYPE, DREAM_HIP_LINEAR, true>(a, perm, want_gy, s, g);
    else adj_launch<DT, GT, TYPE, DREAM_HIP_LINEAR, false>(a, perm, want_gy, s, g);
  }
}

template <typename DT, typename GT>
void adj_dispatch_type(const DasArgs &a, const int32_t *perm, int type, int interp, bool apod,
                       bool want_gy, hipStream_t s, ChunkGate *g)
{
  switch (type) {
  case DAS_SAFT: adj_dispatch_interp<DT, GT, DAS_SAFT>(a, perm, interp, apod, want_gy, s, g); break;
  case DAS_TFM: adj_dispatch_interp<DT, GT, DAS_TFM>(a, perm, interp, apod, want_gy, s, g); break;
  case DAS_RCA_COLTX:
    adj_dispatch_interp<DT, GT, DAS_RCA_COLTX>(a, perm, interp, apod, want_gy, s, g);
    break;
  default: adj_dispatch_interp<DT, GT, DAS_RCA_ROWTX>(a, perm, interp, apod, want_gy, s, g); break;
  }
}

template <typename DT, typename GT, int TYPE, int INTERP, bool APOD>
void batch_launch(const DasArgs *args, const int64_t *begin, int n, int64_t items,
                  unsigned long long *counter, int grid, hipStream_t s)
{
  hipLaunchKernelGGL((das_batch_kernel<DT, GT, TYPE, INTERP, APOD>), dim3(grid), dim3(kThreads), 0,
                     s, args, begin, n, items, counter);
}

template <typename DT, typename GT, int TYPE>
void batch_dispatch_interp(int interp, bool apod, const DasArgs *args, const int64_t *begin, int n,
                           int64_t items, unsigned long long *counter, int grid, hipStream_t s)
{
  if (interp == DREAM_HIP_NEAREST) {
    if (apod) batch_launch<DT, GT, TYPE, DREAM_HIP_NEAREST, true>(args, begin, n, items, counter, grid, s);
    else batch_launch<DT, GT, TYPE, DREAM_HIP_NEAREST, false>(args, begin, n, items, counter, grid, s);
  } else {
    if (apod) batch_launch<DT, GT, TYPE, DREAM_HIP_LINEAR, true>(args, begin, n, items, counter, grid, s);
    else batch_launch<DT, GT, TYPE, DREAM_HIP_LINEAR, false>(args, begin, n, items, counter, grid, s);
  }
}

template <typename DT, typename GT>
void batch_dispatch_type(int type, int interp, bool apod, const DasArgs *args,
                         const int64_t *begin, int n, int64_t items,
                         unsigned long long *counter, int grid, hipStream_t s)
{
  switch (type) {
  case DAS_SAFT:
    batch_dispatch_interp<DT, GT, DAS_SAFT>(interp, apod, args, begin, n, items, counter, grid, s);
    break;
  case DAS_TFM:
    batch_dispatch_interp<DT, GT, DAS_TFM>(interp, apod, args, begin, n, items, counter, grid, s);
    break;
  case DAS_RCA_COLTX:
    batch_dispatch_interp<DT, GT, DAS_RCA_COLTX>(interp, apod, args, begin, n, items, counter, grid, s);
    break;
  default:
    batch_dispatch_interp<DT, GT, DAS_RCA_ROWTX>(interp, apod, args, begin, n, items, counter, grid, s);
    break;
  }
}

// Pixels of a uniform grid in das_uni's linear order (y fastest, then z,
// then x; opencl/das_uni_double.cl:32-47): x = min_x + ((T) nx) * dx etc.
template <typename GT>
__global__ void uni_grid_kernel(GT *Ro, int64_t No, int64_t Nx, int64_t Ny, int64_t Nz, GT min_x,
                                GT dx, GT min_y, GT dy, GT min_z, GT dz)
{
  const int64_t no = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (no >= No) return;
  const int64_t yq = no / Ny;
  const int64_t ny = no - yq * Ny;
  const int64_t zq = yq / Nz;
  const int64_t nz = yq - zq * Nz;
  const int64_t xq = zq / Nx;
  const int64_t nx = zq - xq * Nx;
  Ro[no] = min_x + ((GT) nx) * dx;
  Ro[no + No] = min_y + ((GT) ny) * dy;
  Ro[no + 2 * No] = min_z + ((GT) nz) * dz;
}

size_t elem_size(int prec) { return prec == DREAM_HIP_F64 ? 8 : 4; }
size_t geom_size(int prec) { return prec == DREAM_HIP_F32_REF ? 4 : 8; }

} // namespace

size_t das_forward_scratch_bytes(int type, int prec, int64_t No, int64_t n_frames, int64_t Lt,
                                 int64_t Lr)
{
  int split = 1;
  for (bool recip : {false, true})
    for (const FwdShape &sh : forward_candidates(Lr == 0 ? DAS_SAFT : type, No, n_frames, Lt, Lr,
                                                 recip))
      split = std::max(split, sh.split);
  return split > 1 ? (size_t) split * (size_t) n_frames * (size_t) No * elem_size(prec) : 0;
}

template <typename GT>
void pixel_order(const GT *Ro, int64_t No, int32_t *perm)
{
  double lo[3], hi[3];
  for (int c = 0; c < 3; c++) {
    lo[c] = hi[c] = No ? (double) Ro[c * No] : 0.0;
    for (int64_t n = 1; n < No; n++) {
      const double v = (double) Ro[n + c * No];
      lo[c] = std::min(lo[c], v);
      hi[c] = std::max(hi[c], v);
    }
  }
  // Spread 10-bit cell indices to every third bit and interleave (z, y, x).
  auto spread = [](uint32_t v) {
    uint64_t x = v & 0x3ff;
    x = (x | (x << 16)) & 0x030000FFull;
    x = (x | (x << 8)) & 0x0300F00Full;
    x = (x | (x << 4)) & 0x030C30C3ull;
    x = (x | (x << 2)) & 0x09249249ull;
    return x;
  };
  std::vector<std::pair<uint64_t, int32_t>> key((size_t) No);
  for (int64_t n = 0; n < No; n++) {
    uint64_t code = 0;
    for (int c = 0; c < 3; c++) {
      const double span = hi[c] - lo[c];
      const uint32_t q = span > 0 ? (uint32_t) std::min(1023.0, ((double) Ro[n + c * No] - lo[c]) / span * 1023.0) : 0u;
      code |= spread(q) << c;
    }
    key[(size_t) n] = {code, (int32_t) n};
  }
  std::stable_sort(key.begin(), key.end(),
                   [](const std::pair<uint64_t, int32_t> &x, const std::pair<uint64_t, int32_t> &y) {
                     return x.first < y.first;
                   });
  for (int64_t n = 0; n < No; n++) perm[n] = key[(size_t) n].second;
  // Within each 64-pixel block (one wave), order by (z, y, x): the 4-lane
  // quads, whose gathers the L1 serves together, then run along x, where the
  // sample index changes more slowly than along z for arrays in the z = 0
  // plane (measured on MI355X: C2 TFM 2.5 % and C5 RCA 13 % faster than
  // Morton quads).
  for (int64_t b = 0; b + 64 <= No; b += 64) {
    std::stable_sort(perm + b, perm + b + 64, [&](int32_t u, int32_t v) {
      for (int c = 2; c >= 0; c--) {
        const double a = (double) Ro[u + c * No], d = (double) Ro[v + c * No];
        if (a != d) return a < d;
      }
      return false;
    });
  }
}

template void pixel_order<double>(const double *, int64_t, int32_t *);
template void pixel_order<float>(const float *, int64_t, int32_t *);

void launch_das_forward(const DasArgs &a, int type, int prec, int interp, bool apod,
                        hipStream_t s, ChunkGate *gate)
{
  if (prec == DREAM_HIP_F64) fwd_dispatch_type<double, double>(a, type, interp, apod, s, gate);
  else if (prec == DREAM_HIP_F32) fwd_dispatch_type<float, double>(a, type, interp, apod, s, gate);
  else fwd_dispatch_type<float, float>(a, type, interp, apod, s, gate);
  DREAM_HIP_CHECK(hipGetLastError());
}

void launch_das_adjoint(const DasArgs &a, int type, int prec, int interp, bool apod,
                        bool want_gy, hipStream_t s, ChunkGate *gate)
{
  const int32_t *perm = a.perm;
  if (prec == DREAM_HIP_F64)
    adj_dispatch_type<double, double>(a, perm, type, interp, apod, want_gy, s, gate);
  else if (prec == DREAM_HIP_F32)
    adj_dispatch_type<float, double>(a, perm, type, interp, apod, want_gy, s, gate);
  else throw Error(DREAM_HIP_EINVAL, "the adjoint is not available for DREAM_HIP_F32_REF handles");
  DREAM_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ handle

void das_constants(DasArgs &a, int prec, double dt, double cp)
{
  // The reference's constants: double expressions rounded to T
  // (extra_src/das.cc:316-318).
  if (prec == DREAM_HIP_F32_REF) {
    a.fs_khz = (double) (float) ((1.0 / dt) * 1000.0);
    a.inv_cp = (double) (float) (1.0 / cp);
  } else {
    a.fs_khz = (1.0 / dt) * 1000.0;
    a.inv_cp = 1.0 / cp;
  }
  a.delay_ms = 0.0;
  a.cp = cp;
  a.leg_max = std::ldexp(1.0, 26) / a.fs_khz;
}

DasHandle::DasHandle(int type, int64_t K, int64_t No, int64_t Lt, int64_t Lr, int precision,
                     bool stripes_given)
  : type_(Lr == 0 ? DAS_SAFT : type), K_(K), No_(No), Lt_(Lt), Lr_(Lr), prec_(precision)
{
  require(type >= DAS_SAFT && type <= DAS_RCA_ROWTX, "das: unknown DAS type");
  require(precision == DREAM_HIP_F64 || precision == DREAM_HIP_F32_REF ||
              precision == DREAM_HIP_F32,
          "das: unknown precision");
  require(K >= 1 && K < (1ll << 26), "das: A-scan length out of range (1 .. 2^26)");
  require(No >= 0 && Lt >= 1 && Lr >= 0, "das: bad array sizes");
  const int64_t n_ascans = type_ == DAS_SAFT ? Lt : Lt * Lr;
  require(n_ascans * K < (1ll << 40), "das: data too large");
  if (type_ != DAS_SAFT) require(Lr >= 1, "das: TFM/RCA need receive elements");
  if ((type_ == DAS_RCA_COLTX || type_ == DAS_RCA_ROWTX) && !stripes_given)
    require(rca_extents_in_bounds(Lt, Lr),
          "das: RCA stripe extents are read at rows[Lt .. 2 Lt - 1] and columns[Lr - 1] "
          "(extra_src/das.cc:464-467); need 2 Lt <= 3 Lr and Lr <= 3 Lt");
  DREAM_HIP_CHECK(hipMalloc(&d_status_, sizeof(Status)));
  DREAM_HIP_CHECK(hipMemset(d_status_, 0, sizeof(Status)));
  DREAM_HIP_CHECK(hipHostMalloc(&h_status_, sizeof(Status)));
}

DasHandle::~DasHandle()
{
  drop_timing();
  (void) hipFree(d_sym_);
  (void) hipFree(d_grp_);
  (void) hipFree(d_status_);
  (void) hipHostFree(h_status_);
  (void) hipFree(d_buf_);
  (void) hipFree(d_z_);
  (void) hipFree(d_scratch_);
  (void) hipFree(d_perm_);
}

void DasHandle::ensure_scratch_for(int64_t n_frames)
{
  ensure_scratch(das_forward_scratch_bytes(type_, prec_, No_, n_frames, Lt_, Lr_));
}

void DasHandle::drop_timing()
{
  for (hipEvent_t e : tev_) (void) hipEventDestroy(e);
  tev_.clear();
  t_cap_ = t_n_ = 0;
}

void DasHandle::start_timing(int n_max)
{
  require(n_max >= 1 && n_max <= 100000, "das timing: n_max out of range (1 .. 100000)");
  drop_timing();
  tev_.assign((size_t) 3 * n_max, nullptr);
  for (hipEvent_t &e : tev_) DREAM_HIP_CHECK(hipEventCreate(&e));
  t_cap_ = n_max;
}

void DasHandle::read_timing(int *n, double *prep_ms, double *forward_ms)
{
  double p = 0.0, f = 0.0;
  for (int i = 0; i < t_n_; i++) {
    float a = 0.0f, b = 0.0f;
    DREAM_HIP_CHECK(hipEventSynchronize(tev_[(size_t) 3 * i + 2]));
    DREAM_HIP_CHECK(hipEventElapsedTime(&a, tev_[(size_t) 3 * i], tev_[(size_t) 3 * i + 1]));
    DREAM_HIP_CHECK(hipEventElapsedTime(&b, tev_[(size_t) 3 * i + 1], tev_[(size_t) 3 * i + 2]));
    p += a;
    f += b;
  }
  if (n) *n = t_n_;
  if (prep_ms) *prep_ms = p;
  if (forward_ms) *forward_ms = f;
  drop_timing();
}

void DasHandle::ensure_scratch(size_t bytes)
{
  if (bytes <= d_scratch_bytes_) return;
  (void) hipFree(d_scratch_);
  d_scratch_ = nullptr;
  d_scratch_bytes_ = 0;
  if (hipMalloc(&d_scratch_, bytes) != hipSuccess)
    throw Error(DREAM_HIP_ENOMEM, "das: device allocation failed (partial sums)");
  d_scratch_bytes_ = bytes;
}

void DasHandle::set_pixel_order(const int32_t *perm, int64_t n)
{
  if (!perm) {
    (void) hipFree(d_perm_);
    d_perm_ = nullptr;
    return;
  }
  require(n == No_, "das: the pixel order must have one entry per pixel");
  std::vector<char> seen((size_t) No_, 0);
  for (int64_t q = 0; q < n; q++) {
    require(perm[q] >= 0 && perm[q] < No_ && !seen[(size_t) perm[q]],
            "das: the pixel order must be a permutation of 0..No-1");
    seen[(size_t) perm[q]] = 1;
  }
  if (!d_perm_) DREAM_HIP_CHECK(hipMalloc(&d_perm_, (size_t) std::max<int64_t>(No_, 1) * 4));
  if (n) DREAM_HIP_CHECK(hipMemcpy(d_perm_, perm, (size_t) n * 4, hipMemcpyHostToDevice));
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
  das_constants(a, prec_, dt, cp);
  a.apod_t = apod_t;
  a.apod_r = apod_r;
  a.n_frames = n_frames;
  a.y_stride = y_stride;
  a.im_stride = im_stride;
  a.Im = Im;
  a.status = d_status_;
  a.status_epoch = epoch_;
  a.perm = d_perm_;
  a.scratch = d_scratch_;
  a.scratch_bytes = d_scratch_bytes_;
}

// Several FP32 frames of a two-leg array: the frames in groups of four
// (a.Yg), so that one gather serves four frames -- a copy of Y in the
// handle, made on stream s; one frame per launch row if it cannot be had.
void DasHandle::group_frames(DasArgs &a, const void *Y, int64_t y_stride, int64_t n_frames,
                             hipStream_t s)
{
  const int64_t E = (type_ == DAS_SAFT ? Lt_ : Lt_ * Lr_) * K_;
  if (prec_ != DREAM_HIP_F32 || type_ == DAS_SAFT || n_frames < 2 || !frame_groups_enabled() ||
      (uint64_t) E * 16 >= 0xFFFFFFF0ull)
    return;
  const int64_t G = (n_frames + 3) / 4;
  const size_t bytes = (size_t) (G * E) * 16;
  if (bytes > d_grp_bytes_) {
    // No allocation while the stream is being captured into a graph: the
    // captured call reads one frame per gather.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return;
    (void) hipFree(d_grp_);
    d_grp_ = nullptr;
    d_grp_bytes_ = 0;
    if (hipMalloc(&d_grp_, bytes) == hipSuccess) d_grp_bytes_ = bytes;
    else (void) hipGetLastError();
  }
  if (!d_grp_) return;
  const unsigned bx = (unsigned) std::min<int64_t>((E + 255) / 256, (int64_t) 1 << 20);
  hipLaunchKernelGGL(das_group_frames_kernel, dim3(bx, (unsigned) G), dim3(256), 0, s,
                     (const float *) Y, y_stride, E, n_frames, (frames4 *) d_grp_);
  DREAM_HIP_CHECK(hipGetLastError());
  a.Yg = d_grp_;
  a.yg_stride = E;
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
  require(y_stride >= n_ascans * K_, "das: Y frames must not overlap (stride >= K x A-scans)");
  require((uint64_t) y_stride * elem_size(prec_) < 0xFFFFFFF0ull,
          "das: one frame of Y must be smaller than 4 GiB");
  ensure_scratch_for(n_frames);
  DasArgs a;
  fill_args(a, Y, Ro, Gt, Gr, dt, delay_type, delay, cp, apod_t, apod_r, n_frames, y_stride,
            im_stride, Im);
  // A new epoch per call (the word is cleared once per 2^32 calls).
  if (++epoch_ == 0) {
    zero_async(d_status_, sizeof(Status), s);
    epoch_ = 1;
  }
  a.status_epoch = epoch_;
  status_stream_ = s;
  const bool timed = t_n_ < t_cap_ && No_ > 0;
  if (timed) DREAM_HIP_CHECK(hipEventRecord(tev_[(size_t) 3 * t_n_], s));
  // Reciprocal TFM: one array for both sides (the same pointer), one scalar
  // delay, no apodization.  The pairs (i, j) and (j, i) then share their
  // sample index whenever the delay is zero, which the symmetrize kernel and
  // the forward test on the device; Ysym holds the pair sums.
  if (No_ > 0 && type_ == DAS_TFM && Gt == Gr && Lt_ == Lr_ && delay_type == 0 && !apod &&
      recip_enabled()) {
    const size_t bytes = (size_t) (n_frames * y_stride) * elem_size(prec_);
    if (bytes > d_sym_bytes_) {
      (void) hipFree(d_sym_);
      d_sym_ = nullptr;
      d_sym_bytes_ = 0;
      if (hipMalloc(&d_sym_, bytes) != hipSuccess)
        throw Error(DREAM_HIP_ENOMEM, "das: device allocation failed (symmetrized A-scans)");
      d_sym_bytes_ = bytes;
    }
    const dim3 grid((unsigned) (Lt_ * (Lt_ + 1) / 2), 1u, (unsigned) n_frames);
    if (prec_ == DREAM_HIP_F64)
      hipLaunchKernelGGL((das_symmetrize_kernel<double, double>), grid, dim3(256), 0, s,
                         (const double *) Y, (double *) d_sym_, (int) Lt_, K_, y_stride,
                         (const double *) delay);
    else if (prec_ == DREAM_HIP_F32)
      hipLaunchKernelGGL((das_symmetrize_kernel<float, double>), grid, dim3(256), 0, s,
                         (const float *) Y, (float *) d_sym_, (int) Lt_, K_, y_stride,
                         (const double *) delay);
    else
      hipLaunchKernelGGL((das_symmetrize_kernel<float, float>), grid, dim3(256), 0, s,
                         (const float *) Y, (float *) d_sym_, (int) Lt_, K_, y_stride,
                         (const float *) delay);
    DREAM_HIP_CHECK(hipGetLastError());
    a.Ysym = d_sym_;
    a.recip = 1;
  }
  if (No_ > 0 && !a.recip) group_frames(a, Y, y_stride, n_frames, s);
  if (timed) DREAM_HIP_CHECK(hipEventRecord(tev_[(size_t) 3 * t_n_ + 1], s));
  ChunkGate gate(s);
  if (No_ > 0) launch_das_forward(a, type_, prec_, interp, apod, s, &gate);
  if (timed) DREAM_HIP_CHECK(hipEventRecord(tev_[(size_t) 3 * t_n_++ + 2], s));
  if (!want_status) return E_NONE;
  return status(err_level);
}

int DasHandle::status(int err_level)
{
  if (epoch_ == 0) return E_NONE;
  hipStream_t s = status_stream_;
  DREAM_HIP_CHECK(hipMemcpyAsync(h_status_, d_status_, sizeof(Status), hipMemcpyDeviceToHost, s));
  DREAM_HIP_CHECK(hipStreamSynchronize(s));
  if (h_status_->any_oob != epoch_) return E_NONE;
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
  // Visit the pixels in a spatially compact order.
  {
    std::vector<int32_t> perm((size_t) No_);
    if (prec_ == DREAM_HIP_F64) pixel_order((const double *) Ro, No_, perm.data());
    else pixel_order((const float *) Ro, No_, perm.data());
    set_pixel_order(perm.data(), No_);
  }
  put_geom(dRo, Ro, No_ * 3);
  put_geom(dGt, Gt, Lt_ * 3);
  // One array for both sides (equal bytes): pass one device copy for both,
  // which lets a TFM pair (i, j) with (j, i).
  const bool same = bGr && Lr_ == Lt_ &&
                    std::memcmp(Gt, Gr, (size_t) Lt_ * 3 * in_gs) == 0;
  if (bGr && !same) put_geom(dGr, Gr, Lr_ * 3);
  put_geom(dD, delay, nd);
  int code = run_dev(dY, dRo, dGt, bGr ? (same ? dGt : dGr) : nullptr, dt, delay_type, dD, cp,
                     nullptr, nullptr, DREAM_HIP_NEAREST, 1, 0, 0, dIm, err_level, true, s);
  DREAM_HIP_CHECK(hipMemcpyAsync(Im, dIm, bIm, hipMemcpyDeviceToHost, s));
  DREAM_HIP_CHECK(hipStreamSynchronize(s));
  return code;
}

void DasHandle::adjoint_dev(const void *Y, const void *Ro, const void *Gt, const void *Gr,
                            double dt, double delay, double cp, const void *apod_t,
                            const void *apod_r, int interp, int64_t n_frames, int64_t y_stride,
                            int64_t im_stride, const void *g, void *gY, void *g_apod_t,
                            void *g_apod_r, void *g_cp, hipStream_t s, double *fixed_z,
                            size_t fixed_z_bytes)
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
  require(y_stride >= n_ascans * K_,
          "das adjoint: Y frames must not overlap (stride >= K x A-scans)");
  require((uint64_t) y_stride * elem_size(prec_) < 0xFFFFFFF0ull,
          "das adjoint: one frame of Y must be smaller than 4 GiB");
  require(type_ == DAS_SAFT || gY || (size_t) (Lt_ + Lr_) * 8 <= 65536,
          "das adjoint: too many elements for the apodization accumulators");
  // Accumulators: Z per A-scan (region kernel) or z_t/z_r (pixel kernel), z_cp.
  const size_t zbytes = (size_t) (n_ascans + Lt_ + Lr_ + 1) * sizeof(double);
  double *Z;
  if (fixed_z) {
    // Allocation-free call: the caller's fixed workspace, default launch
    // shapes, one launch per kernel (no chunk events).
    require(zbytes <= fixed_z_bytes,
            "das adjoint: n_ascans + Lt + Lr + 1 exceeds the fixed workspace (262144 doubles)");
    Z = fixed_z;
  } else {
    if (zbytes > d_z_bytes_) {
      (void) hipFree(d_z_);
      d_z_ = nullptr;
      d_z_bytes_ = 0;
      if (hipMalloc(&d_z_, zbytes) != hipSuccess)
        throw Error(DREAM_HIP_ENOMEM, "das adjoint: device allocation failed");
      d_z_bytes_ = zbytes;
    }
    Z = (double *) d_z_;
  }
  double *zt = Z + n_ascans, *zr = zt + Lt_, *zcp = zr + Lr_;
  zero_async(Z, zbytes, s);
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
  // The pixel kernel's nearest-lower taps read frames in groups of four
  // (not in the allocation-free call).
  if (No_ > 0 && !channel && !fixed_z && interp == DREAM_HIP_NEAREST)
    group_frames(a, Y, y_stride, n_frames, s);
  ChunkGate gate(s);
  if (No_ > 0) {
    if (fixed_z) {
      SuspendAutotune no_tuning;
      launch_das_adjoint(a, type_, prec_, interp, apod, gY != nullptr, s, nullptr);
    } else {
      launch_das_adjoint(a, type_, prec_, interp, apod, gY != nullptr, s, &gate);
    }
  }
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

double *fixed_workspace(size_t *bytes)
{
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void *addr[kMaxDevices] = {};
  int dev = 0;
  DREAM_HIP_CHECK(hipGetDevice(&dev));
  require(dev >= 0 && dev < kMaxDevices, "fixed workspace: device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      throw Error(DREAM_HIP_ERUNTIME, "fixed workspace: hipGetSymbolAddress failed");
    addr[dev] = p;
  }
  *bytes = sizeof(scratch);
  return static_cast<double *>(addr[dev]);
}

// ------------------------------------------------------------------ das_uni

namespace {

// Element / grid count (max - min)/pitch + 1 truncated, in T as the
// reference computes it (include/das_uni.h:70-78, extra_src/oct_das_uni.cc:280-290).
template <typename T>
int64_t uni_count(double mn, double pitch, double mx)
{
  return (int64_t) ((double) (((T) mx - (T) mn) / (T) pitch) + 1.0);
}

template <typename GT>
void uni_tables(int type, const double *gt, const double *gr, int64_t Lt, int64_t Lr,
                std::vector<char> &out)
{
  // Transmit (SAFT/TFM) or column (RCA) elements along x; receive (TFM)
  // elements along x, or row (RCA) elements along y; all at z = 0.
  const int64_t Lrr = type == DAS_SAFT ? 0 : Lr;
  std::vector<GT> g((size_t) (3 * (Lt + Lrr)), GT(0));
  for (int64_t n = 0; n < Lt; n++) g[(size_t) n] = (GT) gt[0] + ((GT) n) * (GT) gt[1];
  GT *r = g.data() + 3 * Lt;
  for (int64_t n = 0; n < Lrr; n++) {
    const GT v = (GT) gr[0] + ((GT) n) * (GT) gr[1];
    if (type == DAS_TFM) r[n] = v; else r[n + Lrr] = v;
  }
  out.assign((const char *) g.data(), (const char *) g.data() + g.size() * sizeof(GT));
}

} // namespace

DasUni::DasUni(int type, int64_t K, const double gt[3], const double gr[3], const double Ro[9],
               int precision)
  : type_(type), prec_(precision), K_(K)
{
  require(type >= DAS_SAFT && type <= DAS_RCA_ROWTX, "das_uni: unknown DAS type");
  require(precision == DREAM_HIP_F64 || precision == DREAM_HIP_F32_REF ||
              precision == DREAM_HIP_F32,
          "das_uni: unknown precision");
  std::memcpy(gt_, gt, sizeof gt_);
  std::memcpy(gr_, type == DAS_SAFT ? gt : gr, sizeof gr_);
  std::memcpy(ro_, Ro, sizeof ro_);
  const bool f = precision == DREAM_HIP_F32_REF;
  gs_ = f ? 4 : 8;
  auto count = [&](double mn, double p, double mx) {
    require(p > 0 && mx >= mn, "das_uni: need pitch > 0 and max >= min");
    return f ? uni_count<float>(mn, p, mx) : uni_count<double>(mn, p, mx);
  };
  Lt_ = count(gt[0], gt[1], gt[2]);
  Lr_ = type == DAS_SAFT ? 0 : count(gr[0], gr[1], gr[2]);
  const int64_t Nx = count(Ro[0], Ro[3], Ro[6]), Ny = count(Ro[1], Ro[4], Ro[7]),
                Nz = count(Ro[2], Ro[5], Ro[8]);
  require(Nx >= 1 && Ny >= 1 && Nz >= 1 && Lt_ >= 1, "das_uni: empty grid");
  No_ = Nx * Ny * Nz;
  h_ = new DasHandle(type, K, No_, Lt_, Lr_, precision, true);
  // Device geometry: Ro, then the element tables, then the delay.
  std::vector<char> tabs;
  if (f) uni_tables<float>(type, gt, gr, Lt_, Lr_, tabs);
  else uni_tables<double>(type, gt, gr, Lt_, Lr_, tabs);
  const size_t ro_bytes = (size_t) No_ * 3 * gs_;
  DREAM_HIP_CHECK(hipMalloc(&d_geom_, ro_bytes + tabs.size() + 16));
  char *base = (char *) d_geom_;
  DREAM_HIP_CHECK(hipMemcpy(base + ro_bytes, tabs.data(), tabs.size(), hipMemcpyHostToDevice));
  const int blocks = (int) ((No_ + 255) / 256);
  if (f)
    hipLaunchKernelGGL(uni_grid_kernel<float>, dim3(blocks), dim3(256), 0, lib_stream(),
                       (float *) base, No_, Nx, Ny, Nz, (float) Ro[0], (float) Ro[3],
                       (float) Ro[1], (float) Ro[4], (float) Ro[2], (float) Ro[5]);
  else
    hipLaunchKernelGGL(uni_grid_kernel<double>, dim3(blocks), dim3(256), 0, lib_stream(),
                       (double *) base, No_, Nx, Ny, Nz, Ro[0], Ro[3], Ro[1], Ro[4], Ro[2], Ro[5]);
  DREAM_HIP_CHECK(hipGetLastError());
  DREAM_HIP_CHECK(hipStreamSynchronize(lib_stream()));
  // Visit the pixels in Morton order of their grid indices.
  std::vector<double> idx((size_t) No_ * 3);
  for (int64_t no = 0; no < No_; no++) {
    const int64_t yq = no / Ny, zq = yq / Nz;
    idx[(size_t) no] = (double) (zq % Nx);
    idx[(size_t) (no + No_)] = (double) (no % Ny);
    idx[(size_t) (no + 2 * No_)] = (double) (yq % Nz);
  }
  std::vector<int32_t> perm((size_t) No_);
  pixel_order<double>(idx.data(), No_, perm.data());
  h_->set_pixel_order(perm.data(), No_);
}

DasUni::~DasUni()
{
  delete h_;
  (void) hipFree(d_geom_);
  (void) hipFree(d_buf_);
}

bool DasUni::changed(int type, int64_t K, const double gt[3], const double gr[3],
                     const double Ro[9], int precision) const
{
  if (type != type_ || K != K_ || precision != prec_) return true;
  for (int i = 0; i < 3; i++)
    if (gt[i] != gt_[i] || (type != DAS_SAFT && gr[i] != gr_[i])) return true;
  for (int i = 0; i < 9; i++)
    if (Ro[i] != ro_[i]) return true;
  return false;
}

void DasUni::run_dev(const void *Y, double dt, double delay, double cp, int interp, void *Im,
                     hipStream_t s)
{
  if (cancelled()) throw Error(DREAM_HIP_ECANCELLED, "das_uni cancelled");
  char *base = (char *) d_geom_;
  const size_t ro_bytes = (size_t) No_ * 3 * gs_;
  const void *Gt = base + ro_bytes;
  const void *Gr = base + ro_bytes + (size_t) Lt_ * 3 * gs_;
  void *dd = base + ro_bytes + (size_t) (Lt_ + (type_ == DAS_SAFT ? 0 : Lr_)) * 3 * gs_;
  // The scalar delay in the geometry type, stream-ordered before the launch.
  if (gs_ == 8) {
    DREAM_HIP_CHECK(hipMemcpyAsync(dd, &delay, 8, hipMemcpyHostToDevice, s));
  } else {
    const float df = (float) delay;
    DREAM_HIP_CHECK(hipMemcpyAsync(dd, &df, 4, hipMemcpyHostToDevice, s));
  }
  DasArgs a;
  h_->fill_args(a, Y, base, Gt, type_ == DAS_SAFT ? nullptr : Gr, dt, 0, dd, cp, nullptr,
                nullptr, 1, 0, 0, Im);
  if (type_ == DAS_RCA_COLTX || type_ == DAS_RCA_ROWTX) {
    // The stripe extents are the [min, max] parameters (opencl/das_uni_double.cl:197-200).
    a.stripe_given = 1;
    a.sx_lo = gs_ == 8 ? gt_[0] : (double) (float) gt_[0];
    a.sx_hi = gs_ == 8 ? gt_[2] : (double) (float) gt_[2];
    a.sy_lo = gs_ == 8 ? gr_[0] : (double) (float) gr_[0];
    a.sy_hi = gs_ == 8 ? gr_[2] : (double) (float) gr_[2];
    // Row transmit: the delay is on the row (receive-index) leg
    // (opencl/das_uni_double.cl:344-346).
    a.delay_rx = type_ == DAS_RCA_ROWTX;
  }
  a.y_stride = (type_ == DAS_SAFT ? Lt_ : Lt_ * Lr_) * K_;
  require((uint64_t) a.y_stride * (prec_ == DREAM_HIP_F64 ? 8 : 4) < 0xFFFFFFF0ull,
          "das_uni: Y must be smaller than 4 GiB");
  a.im_stride = No_;
  a.status = nullptr;
  h_->ensure_scratch_for(1);
  a.scratch = h_->scratch();
  a.scratch_bytes = h_->scratch_bytes();
  require(interp == DREAM_HIP_NEAREST || interp == DREAM_HIP_LINEAR, "das_uni: unknown interpolation");
  ChunkGate gate(s);
  launch_das_forward(a, type_, prec_, interp, false, s, &gate);
}

void DasUni::run_host(const void *Y, double dt, double delay, double cp, void *Im)
{
  const size_t es = prec_ == DREAM_HIP_F64 ? 8 : 4;
  const size_t bY = (size_t) ((type_ == DAS_SAFT ? Lt_ : Lt_ * Lr_) * K_) * es;
  const size_t bIm = (size_t) No_ * es;
  const size_t total = ((bY + 255) & ~(size_t) 255) + bIm;
  if (total > d_buf_bytes_) {
    (void) hipFree(d_buf_);
    d_buf_ = nullptr;
    d_buf_bytes_ = 0;
    if (hipMalloc(&d_buf_, total) != hipSuccess)
      throw Error(DREAM_HIP_ENOMEM, "das_uni: device allocation failed");
    d_buf_bytes_ = total;
  }
  char *dY = (char *) d_buf_, *dIm = dY + ((bY + 255) & ~(size_t) 255);
  hipStream_t s = lib_stream();
  DREAM_HIP_CHECK(hipMemcpyAsync(dY, Y, bY, hipMemcpyHostToDevice, s));
  run_dev(dY, dt, delay, cp, DREAM_HIP_NEAREST, dIm, s);
  DREAM_HIP_CHECK(hipMemcpyAsync(Im, dIm, bIm, hipMemcpyDeviceToHost, s));
  DREAM_HIP_CHECK(hipStreamSynchronize(s));
}

// ------------------------------------------------------------------ batch

DasBatch::DasBatch(const DreamDasProblem *probs, int32_t n, int type, int precision, int interp)
  : type_(type), prec_(precision), interp_(interp)
{
  require(type >= DAS_SAFT && type <= DAS_RCA_ROWTX, "das batch: unknown DAS type");
  require(precision == DREAM_HIP_F64 || precision == DREAM_HIP_F32_REF ||
              precision == DREAM_HIP_F32,
          "das batch: unknown precision");
  require(interp == DREAM_HIP_NEAREST || interp == DREAM_HIP_LINEAR,
          "das batch: unknown interpolation");
  require(n >= 0 && (n == 0 || probs), "das batch: bad problem list");
  const size_t es = elem_size(precision), gs = geom_size(precision);
  std::vector<DasArgs> args((size_t) n);
  std::vector<double> cost((size_t) n);
  for (int32_t i = 0; i < n; i++) {
    const DreamDasProblem &p = probs[i];
    const bool saft = type == DAS_SAFT;
    require(p.K >= 1 && p.K < (1ll << 26), "das batch: A-scan length out of range (1 .. 2^26)");
    require(p.No >= 0 && p.Lt >= 1 && (saft || p.Lr >= 1), "das batch: bad array sizes");
    require((type != DAS_RCA_COLTX && type != DAS_RCA_ROWTX) || rca_extents_in_bounds(p.Lt, p.Lr),
            "das batch: RCA needs 2 Lt <= 3 Lr and Lr <= 3 Lt (stripe extents)");
    const int64_t n_ascans = saft ? p.Lt : p.Lt * p.Lr;
    require((uint64_t) (n_ascans * p.K) * es < 0xFFFFFFF0ull,
            "das batch: Y of one problem must be smaller than 4 GiB");
    require(p.No == 0 || (p.Y && p.Ro && p.Gt && p.Im && (saft || p.Gr)),
            "das batch: null device pointer");
    const bool apod = p.apod_t != nullptr;
    require(!apod || saft || p.apod_r, "das batch: TFM/RCA apodization needs apod_r");
    if (i == 0) apod_ = apod;
    require(apod == apod_, "das batch: all problems or none must be apodized");
    DasArgs &a = args[(size_t) i];
    std::memset(&a, 0, sizeof a);
    a.Y = p.Y;
    a.K = p.K;
    a.Ro = p.Ro;
    a.No = p.No;
    a.Gt = p.Gt;
    a.Lt = (int32_t) p.Lt;
    a.Gr = saft ? nullptr : p.Gr;
    a.Lr = saft ? 0 : (int32_t) p.Lr;
    das_constants(a, precision, p.dt, p.cp);
    a.apod_t = p.apod_t;
    a.apod_r = p.apod_r;
    a.n_frames = 1;
    a.y_stride = n_ascans * p.K;
    a.im_stride = p.No;
    a.Im = p.Im;
    a.perm = p.perm;
    cost[(size_t) i] = (double) n_ascans;
  }
  // Problems in decreasing per-tile cost; items numbered problem by problem.
  std::vector<int32_t> order((size_t) n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t x, int32_t y) { return cost[(size_t) x] > cost[(size_t) y]; });
  std::vector<DasArgs> sorted;
  std::vector<int64_t> begin;
  std::vector<double> delays;
  int64_t item = 0;
  for (int32_t k = 0; k < n; k++) {
    const int32_t i = order[(size_t) k];
    const int64_t tiles = (probs[i].No + 63) / 64;
    if (tiles == 0) continue;
    sorted.push_back(args[(size_t) i]);
    begin.push_back(item);
    delays.push_back(probs[i].delay);
    item += tiles;
  }
  n_probs_ = (int32_t) sorted.size();
  n_items_ = item;
  const size_t nb = std::max<size_t>(sorted.size(), 1);
  DREAM_HIP_CHECK(hipMalloc(&d_args_, nb * sizeof(DasArgs)));
  DREAM_HIP_CHECK(hipMalloc(&d_begin_, nb * sizeof(int64_t)));
  DREAM_HIP_CHECK(hipMalloc(&d_delay_, nb * sizeof(double)));
  DREAM_HIP_CHECK(hipMalloc(&d_counter_, 2 * sizeof(unsigned long long)));
  DREAM_HIP_CHECK(hipMemset(d_counter_, 0, 2 * sizeof(unsigned long long)));
  // Scalar delays in the geometry type (float for the reference-float path).
  if (gs == 8) {
    if (!delays.empty())
      DREAM_HIP_CHECK(hipMemcpy(d_delay_, delays.data(), delays.size() * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(delays.begin(), delays.end());
    if (!f.empty())
      DREAM_HIP_CHECK(hipMemcpy(d_delay_, f.data(), f.size() * 4, hipMemcpyHostToDevice));
  }
  for (size_t k = 0; k < sorted.size(); k++) {
    sorted[k].delay = (const char *) d_delay_ + k * gs;
    sorted[k].delay_multi = 0;
  }
  if (!sorted.empty()) {
    DREAM_HIP_CHECK(hipMemcpy(d_args_, sorted.data(), sorted.size() * sizeof(DasArgs),
                              hipMemcpyHostToDevice));
    DREAM_HIP_CHECK(hipMemcpy(d_begin_, begin.data(), begin.size() * sizeof(int64_t),
                              hipMemcpyHostToDevice));
  }
  // Candidate grids: 2, 4 or 8 waves per SIMD, never more waves than items;
  // the first is the default when the autotuner does not run.
  for (int per_simd : {8, 4, 2}) {
    const int64_t waves = std::min<int64_t>((int64_t) cu_count() * 4 * per_simd, n_items_);
    const int g = (int) std::max<int64_t>(1, (waves + 3) / 4);
    if (std::find(grids_.begin(), grids_.end(), g) == grids_.end()) grids_.push_back(g);
  }
}

DasBatch::~DasBatch()
{
  if (graph_) (void) hipGraphExecDestroy(graph_);
  (void) hipFree(d_args_);
  (void) hipFree(d_begin_);
  (void) hipFree(d_delay_);
  (void) hipFree(d_counter_);
}

void DasBatch::launch_grid(int grid, hipStream_t s)
{
  if (n_items_ == 0) return;
  if (prec_ == DREAM_HIP_F64)
    batch_dispatch_type<double, double>(type_, interp_, apod_, d_args_, d_begin_, n_probs_,
                                        n_items_, d_counter_, grid, s);
  else if (prec_ == DREAM_HIP_F32)
    batch_dispatch_type<float, double>(type_, interp_, apod_, d_args_, d_begin_, n_probs_,
                                       n_items_, d_counter_, grid, s);
  else
    batch_dispatch_type<float, float>(type_, interp_, apod_, d_args_, d_begin_, n_probs_,
                                      n_items_, d_counter_, grid, s);
  DREAM_HIP_CHECK(hipGetLastError());
}

void DasBatch::launch(hipStream_t s) { launch_grid(grids_[(size_t) pick_], s); }

void DasBatch::choose(hipStream_t s)
{
  if (pick_ >= 0) return;
  char key[160];
  std::snprintf(key, sizeof key, "das_batch|%d|%d|%d|%d|%d|%lld", type_, prec_, interp_,
                (int) apod_, n_probs_, (long long) n_items_);
  const int t = n_items_ ? autotune(key, (int) grids_.size(),
                                    [&](int i, hipStream_t st) { launch_grid(grids_[(size_t) i], st); }, s)
                         : -1;
  pick_ = t >= 0 ? t : 0;
}

void DasBatch::run(hipStream_t s, bool use_graph)
{
  if (cancelled()) throw Error(DREAM_HIP_ECANCELLED, "das batch cancelled");
  choose(s);
  if (!use_graph) {
    launch(s);
    return;
  }
  if (!graph_) {
    // Capture on the library's own stream; the graph runs on the caller's.
    hipStream_t cs = lib_stream();
    hipGraph_t g = nullptr;
    DREAM_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    launch(cs);
    DREAM_HIP_CHECK(hipStreamEndCapture(cs, &g));
    DREAM_HIP_CHECK(hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0));
    DREAM_HIP_CHECK(hipGraphDestroy(g));
  }
  DREAM_HIP_CHECK(hipGraphLaunch(graph_, s));
}

} // namespace dream
