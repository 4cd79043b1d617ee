// sir.hip -- discrete-representation spatial impulse response of arrays of
// rectangular elements (dream_arr_rect, dreamrect) on gfx950.
//
// Reference algorithm: src/dreamrect.cc:274-335 (the DR sum) driven by
// src/dream_arr_rect.cc:192-244 (per-element delay and weight).  For every
// observation point, element and sub-element centre (xs, ys):
//   r  = sqrt((rx*rx + ry*ry) + rz*rz)
//   ai = ((w v) ds) / (2 pi r) / dt * 1000
//   it = rint((((r*1000)/cp) - d)/dt),  d = (delay - foc) - steer
//   h[it] += ai when 0 <= it < nt.
//
// Work decomposition.  A work item is one observation point and a chunk of
// its (element, row) pairs; the whole batch of problems is one list of work
// items in cost order, served either one work-group per item or by a
// persistent grid pulling items from an atomic counter.  Each 256-thread
// work-group accumulates into an LDS histogram of nt bins and writes the
// column once (or adds its bins into a zeroed column when a point is split
// over several items).  Each lane walks one row of sub-elements, keeping a
// run-length (bin, sum) in registers and touching LDS only when the bin
// changes.  Rows are handed to lanes in a strided order so the 64 lanes of
// a wave see rows spread over the whole element and hit different bins.
// The LDS atomics of those flushes bound the kernel, and along x the bin
// changes about every other step when the element lies to the side of the
// point; for flat rectangles a point may instead walk columns (y inner, in
// row segments), where |ry| <= b/2 keeps the bin for several steps.  Each
// point takes the walk its run-length model prices lower.
//
// Sample index.  The index must equal the reference's FP64 sequence bit for
// bit.  A fast estimate u = r*c1 - d/dt (r from the hardware rsqrt, about
// 2^-24, and one Newton step, about 2^-47; c1 = 1000/(cp dt)) differs
// from the reference's pre-rounding value by less than 2^-44 (|u| + |d/dt|
// + 1), against a certification margin of 2^-36 (same terms); rint(u) is
// taken when u lies farther than B = 2^-36 (nt + 2 + |d/dt|) from a
// rounding boundary (half-integer), which certifies it for every |u| <=
// nt + 2, and for larger |u| both values are out of range anyway.  The
// other contributions (about one in 10^7), and s = 0 or non-finite values,
// take the exact IEEE path: correctly rounded sqrt, two correctly rounded
// divisions and rint.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <type_traits>

#include "sir.hpp"
#include "tune.hpp"

namespace dream {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
// Sub-elements per block of independent index computations.
constexpr int kBlock = 4;
constexpr double kBoundScale = 0x1p-36;

// Sample index (as a double) of the sub-element at offset (rx, ry, rz), and
// q, proportional to 1/r, for its amplitude.  sir_fast takes s = v*v + w in
// one FMA (v the coordinate that varies along the walk, w the sum of the
// other two squares; within 2^-50 of the reference's (rx*rx + ry*ry) +
// rz*rz, far inside the certification margin), the hardware rsqrt seed y
// and one Newton step written as sqrt(s) = g (3 - g y) / 2 with g = s y --
// the Goldschmidt pair's value g (1 + e), e = (1 - g y) / 2, relative error
// 1.5 d^2 for a seed error d -- so that u = ((c1/2) g)(3 - g y) - c2 costs
// four FP64 operations after the seed; it reports whether rint(u) is
// certified (`lim` is 0.5 - B, see the file comment).  With REFINE, q = y (3
// - g y) = 2/r to about 2^-47 (FP64 output), else q = y = 1/r to the seed's
// 2^-24 (FP32 output, whose own rounding is of that size); the amplitude
// weight carries the matching 1/2 or 1.  sir_exact recomputes the index in
// the reference's order (ryy = ry*ry, rzz = rz*rz), with q = 2/re or 1/re;
// a NaN index (non-finite inputs) comes back as -1, out of range as the
// reference's integer conversion of NaN is.
template <bool REFINE>
__device__ __forceinline__ bool sir_fast(double v, double w, double c1h, double c2, double lim,
                                         double &ru, double &q)
{
  const double s = __builtin_fma(v, v, w);
  const double y = __builtin_amdgcn_rsq(s);
  const double g = s * y;
  const double t = __builtin_fma(-g, y, 3.0);
  const double u = __builtin_fma(c1h * g, t, -c2);
  ru = __builtin_rint(u);
  q = REFINE ? y * t : y;
  return __builtin_fabs(u - ru) <= lim;          // false also for s = 0, NaN, inf
}

// The reference's own sequence: correctly rounded sqrt and divisions.
template <bool REFINE>
__device__ __forceinline__ void sir_exact(double rx, double ryy, double rzz, double cp, double d,
                                          double dt, double &ru, double &q)
{
  const double re = __builtin_sqrt((rx * rx + ryy) + rzz);
  const double t = (re * 1000.0) / cp;
  ru = __builtin_rint((t - d) / dt);
  if (ru != ru) ru = -1.0;
  q = (REFINE ? 2.0 : 1.0) / re;
}

template <bool REFINE>
__device__ __forceinline__ double sir_index(double rx, double ryy, double rzz, double ryz,
                                            double c1h, double c2, double lim, double cp,
                                            double d, double dt, double &q)
{
  double ru;
  if (!sir_fast<REFINE>(rx, ryz, c1h, c2, lim, ru, q))
    sir_exact<REFINE>(rx, ryy, rzz, cp, d, dt, ru, q);
  return ru;
}

// v_cvt_i32_f64 saturates: values past the int32 range give INT_MIN or
// INT_MAX (where the C conversion is undefined), so for an integral x that
// is not NaN, x >= 0 && x < nt is one unsigned compare of the result.
__device__ __forceinline__ int sat_i32(double x)
{
  int r;
  asm("v_cvt_i32_f64 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// Address-space views of the plan's tables: uniform reads of the
// sub-element x table go through the scalar cache (constant space), and
// per-lane table reads are global loads, which -- unlike generic (flat)
// loads -- do not also hold up waits on LDS operations.
typedef const __attribute__((address_space(4))) double cdouble;
typedef const __attribute__((address_space(1))) double gdouble;

template <typename T>
__device__ __forceinline__ void lds_add(T *p, T v)
{
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <typename T>
__device__ __forceinline__ void global_add(T *p, T v)
{
  unsafeAtomicAdd(p, v);
}

__device__ __forceinline__ unsigned wave_sum(unsigned v)
{
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int find_problem(const SirDev *P, int n, int64_t item)
{
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (P[mid].item_begin <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Double sum over the wave, in a fixed order (the same on every work-group).
__device__ __forceinline__ double wave_sum_d(double v)
{
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Whether point (xo, yo, zo) of a rectangle problem walks columns (y inner)
// rather than rows (x inner).  A lane flushes its run into the histogram
// when the bin changes, which along x happens at a rate of about
// min(1, |rx|/r dx/l) per step (l = cp dt/1000, one sample of path) and
// along y at min(1, |ry|/r dy/l); with each element's mean offsets for rx
// and ry, the cheaper walk counts a flush as 4 contributions and charges
// each walk its partly filled lane rounds.  Every work-group of the point
// computes the same sums in the same order, so all its items agree.
__device__ __forceinline__ bool walk_columns(const SirDev &P, double xo, double yo, double zo, double *red)
{
  gdouble *el = (gdouble *) P.el;
  const int L = P.L;
  double fx = 0.0, fy = 0.0;
  for (int e = threadIdx.x; e < L; e += kThreads) {
    const double mx = __builtin_fmax(__builtin_fabs(xo - el[e]), P.wx);
    const double my = __builtin_fmax(__builtin_fabs(yo - el[L + e]), P.wy);
    const double mz = zo - el[2 * L + e];
    const double r = __builtin_sqrt(mx * mx + my * my + mz * mz);
    if (r > 0.0) {
      fx += __builtin_fmin(1.0, mx / r * P.sx);
      fy += __builtin_fmin(1.0, my / r * P.sy);
    }
  }
  fx = wave_sum_d(fx);
  fy = wave_sum_d(fy);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { red[w] = fx; red[kWaves + w] = fy; }
  __syncthreads();
  fx = red[0];
  fy = red[kWaves];
  for (int k = 1; k < kWaves; k++) {
    fx += red[k];
    fy += red[kWaves + k];
  }
  fx /= L;
  fy /= L;
  const double nx = P.nx, S = P.seg_len;
  const double rounds_r = __builtin_ceil((double) L * P.ny / P.chunks / kThreads);
  const double rounds_c = __builtin_ceil((double) L * P.nx * P.nseg / P.chunks / kThreads);
  const double cost_r = rounds_r * nx * (1.0 + 4.0 * (1.0 + (nx - 1.0) * fx) / nx);
  const double cost_c = rounds_c * S * (1.0 + 4.0 * (1.0 + (S - 1.0) * fy) / S);
  return __builtin_amdgcn_readfirstlane((int) (cost_c < cost_r)) != 0;
}

// One work item: observation point n of problem P, chunk c of its rows (or
// of its column segments).  The run-length sums and the LDS histogram are
// FP64 for both output precisions: gfx950's LDS FP64 atomic add is about 20
// times faster than its FP32 one, and the column is rounded to T once, when
// it is written.
template <typename T, bool ANY_GLOBAL>
__device__ __forceinline__ void sir_item(const SirDev &P, int64_t local, double *hist, double *red)
{
  const int64_t n = local / P.chunks;
  const int c = (int) (local - n * P.chunks);
  gdouble *Ro = (gdouble *) P.Ro;
  const double xo = Ro[n], yo = Ro[n + P.No], zo = Ro[n + 2 * P.No];
  const double dlay = P.delay_multi ? ((gdouble *) P.delay)[n] : ((gdouble *) P.delay)[0];
  const int ny = P.ny, nx = P.nx, L = P.L;
  const int64_t nt = P.nt;
  const double ntd = (double) nt;
  T *out = reinterpret_cast<T *>(P.h) + n * nt;
  // Plans without a global-mode problem compile the LDS path only.
  const bool use_lds = !ANY_GLOBAL || P.mode != SIR_GLOBAL;
  constexpr bool REFINE_H = !std::is_same<T, float>::value;
  const bool columns = P.nseg > 0 && (P.walk == 1 || walk_columns(P, xo, yo, zo, red));

  if (use_lds) {
    for (int i = threadIdx.x; i < nt; i += kThreads) hist[i] = 0.0;
    __syncthreads();
  }

  unsigned oob = 0;
  bool oob_last = false;
  // The run's bin.  With the LDS histogram it starts at the spare bin nt
  // past the column, so every flush is unconditional (the first adds 0
  // there); in global mode -1 marks "no run yet".
  int cur = use_lds ? (int) nt : -1;
  double acc = 0.0;
  // ai = amp / r = (amp / 2) q (or amp q without REFINE), added into the
  // run by one FMA; ampq is that weight.
  double ampq = 0.0;
  const double c1h = 0.5 * P.c1;
  gdouble *el = (gdouble *) P.el;
  gdouble *ys = (gdouble *) P.ys;
  // The run-length update of one in-range index.
  auto run = [&](int bin, double q) {
    if (bin != cur) {
      if (use_lds) lds_add(hist + cur, acc);
      else if (cur >= 0) global_add(out + cur, (T) acc);
      cur = bin;
      acc = 0.0;
    }
    acc = __builtin_fma(ampq, q, acc);
  };
  // One index: the run-length update, or the per-lane out-of-range count.
  auto accumulate = [&](int bin, double q) {
    if ((unsigned) bin < (unsigned) nt) run(bin, q);
    else oob++;
  };
  // kBlock indices in a loop whose trip count is the same on every active
  // lane: the out-of-range ones are counted for the whole wave from the
  // compare masks (scalar bit counts into cnt_w) and marked per lane in
  // out_mask, so counting costs no vector instruction per index.
  auto accumulate_block = [&](const double (&r)[kBlock], const double (&q)[kBlock],
                              unsigned &cnt_w, unsigned long long &out_mask) {
    int bin[kBlock];
    bool in[kBlock];
    const unsigned long long act = __ballot(1);
    unsigned long long all_in = act;
    int n_in = 0;
#pragma unroll
    for (int t = 0; t < kBlock; t++) {
      bin[t] = sat_i32(r[t]);
      in[t] = (unsigned) bin[t] < (unsigned) nt;
      const unsigned long long m = __ballot(in[t]);
      n_in += __popcll(m);
      all_in &= m;
    }
    cnt_w += (unsigned) (kBlock * __popcll(act) - n_in);
    out_mask |= act & ~all_in;
#pragma unroll
    for (int t = 0; t < kBlock; t++)
      if (in[t]) run(bin[t], q[t]);
  };
  // After such a loop: the wave's count goes to its lowest active lane (in
  // oob_s, apart from the per-lane counts in oob), and whether this lane saw
  // one comes back.
  unsigned oob_s = 0;
  auto settle = [&](unsigned cnt_w, unsigned long long out_mask) {
    const unsigned long long act = __ballot(1);
    const int lane = (int) (threadIdx.x & 63);
    if (lane == __ffsll((long long) act) - 1) oob_s += cnt_w;
    return ((out_mask >> lane) & 1ull) != 0;
  };
  if (!columns) {
    const int Q = L * ny;
    const int q0 = (int) ((int64_t) c * Q / P.chunks);
    const int q1 = (int) ((int64_t) (c + 1) * Q / P.chunks);
    cdouble *xs = (cdouble *) P.xs;
    // Rows q = e*ny + j are dealt to lanes with stride kThreads; (e, j, row)
    // advance incrementally (row = j*row_stride mod ny, unchanged by j -= ny).
    int q = q0 + (int) threadIdx.x;
    int e = q / ny;
    int j = q - e * ny;
    int row = (int) (((int64_t) j * P.row_stride) % ny);
    for (; q < q1; q += kThreads) {
      const double xe = xo - el[e];
      const double ye = yo - el[L + e];
      const double ze = zo - el[2 * L + e];
      const double d = (dlay - el[3 * L + e]) - el[4 * L + e];
      const double ry = ye - ys[row];
      const double ryy = ry * ry;
      // Cylinders: the row's depth on the curved surface (ze - 0.0 == ze
      // for flat elements).
      const double rz = ze - (P.shape == SIR_CYLIND ? ((gdouble *) P.zrow)[row] : 0.0);
      const double rzz = rz * rz;
      const double c2 = d * P.inv_dt;
      const double lim = 0.5 - kBoundScale * (ntd + 2.0 + __builtin_fabs(c2));
      const double ryz = ryy + rzz;
      ampq = (el[5 * L + e] * P.amp) * (REFINE_H ? 0.5 : 1.0);
      const unsigned oob_row = oob;
      bool row_out = false;
      unsigned cnt_w = 0;
      unsigned long long out_mask = 0;
      if (P.shape == SIR_CIRC) {
        // A disc row: the reference's accumulated walk from this row's start.
        double x = ((gdouble *) P.row_x0)[row];
        const int cnt = (int) ((gdouble *) P.row_n)[row];
        for (int i = 0; i < cnt; i++) {
          double h;
          const double ru = sir_index<REFINE_H>(xe - x, ryy, rzz, ryz, c1h, c2, lim, P.cp, d,
                                                P.dt, h);
          accumulate(sat_i32(ru), h);
          x += P.dx;
        }
      } else {
        int i = 0;
        for (; i + kBlock <= nx; i += kBlock) {
          double r[kBlock], w[kBlock];
          bool k[kBlock], all = true;
#pragma unroll
          for (int t = 0; t < kBlock; t++) {
            k[t] = sir_fast<REFINE_H>(xe - xs[i + t], ryz, c1h, c2, lim, r[t], w[t]);
            all &= k[t];
          }
          if (!all) {
#pragma unroll
            for (int t = 0; t < kBlock; t++)
              if (!k[t]) sir_exact<REFINE_H>(xe - xs[i + t], ryy, rzz, P.cp, d, P.dt, r[t], w[t]);
          }
          accumulate_block(r, w, cnt_w, out_mask);
        }
        row_out = settle(cnt_w, out_mask);
        for (; i < nx; i++) {
          double h;
          const double ru = sir_index<REFINE_H>(xe - xs[i], ryy, rzz, ryz, c1h, c2, lim, P.cp,
                                                d, P.dt, h);
          accumulate(sat_i32(ru), h);
        }
      }
      oob_last |= (e == L - 1) && (oob != oob_row || row_out);
      e += P.step_e;
      j += P.step_j;
      if (j >= ny) { j -= ny; e++; }
      row += P.step_row;
      if (row >= ny) row -= ny;
    }
  } else {
    // Column walk of a flat rectangle: unit u = sg*(L*nx) + i*L + e is
    // segment sg (rows [sg*seg_len, +seg_len)) of x column i of element e;
    // neighbouring lanes take neighbouring elements, so the lanes of a wave
    // share a segment (their row loads hit one address) and see different
    // bins.  The fast estimate takes s = ry*ry + (rx*rx + rz*rz), the exact
    // path the reference's (rx*rx + ry*ry) + rz*rz.
    const int QC = L * nx;
    const int U = QC * P.nseg;
    const int u0 = (int) ((int64_t) c * U / P.chunks);
    const int u1 = (int) ((int64_t) (c + 1) * U / P.chunks);
    gdouble *xs = (gdouble *) P.xs;
    int u = u0 + (int) threadIdx.x;
    int sg = u / QC;
    int i = (u - sg * QC) / L;
    int e = u - sg * QC - i * L;
    for (; u < u1; u += kThreads) {
      const double xe = xo - el[e];
      const double ye = yo - el[L + e];
      const double ze = zo - el[2 * L + e];
      const double d = (dlay - el[3 * L + e]) - el[4 * L + e];
      const double rx = xe - xs[i];
      const double rzz = ze * ze;
      const double rxz = rx * rx + rzz;
      const double c2 = d * P.inv_dt;
      const double lim = 0.5 - kBoundScale * (ntd + 2.0 + __builtin_fabs(c2));
      ampq = (el[5 * L + e] * P.amp) * (REFINE_H ? 0.5 : 1.0);
      const unsigned oob_unit = oob;
      unsigned cnt_w = 0;
      unsigned long long out_mask = 0;
      const int j0 = sg * P.seg_len;
      const int j1 = min(j0 + P.seg_len, ny);
      // Every segment has seg_len rows but the last, so the wave's whole
      // blocks of kBlock rows are a uniform count: the shorter length's when
      // any of its lanes holds a last segment.
      const int last_len = ny - (P.nseg - 1) * P.seg_len;
      const int nb = (__ballot(j1 - j0 < P.seg_len) ? last_len : P.seg_len) / kBlock;
      int j = j0;
      for (int blk = 0; blk < nb; blk++, j += kBlock) {
        double yv[kBlock], r[kBlock], w[kBlock];
        bool k[kBlock], all = true;
#pragma unroll
        for (int t = 0; t < kBlock; t++) yv[t] = ye - ys[j + t];
#pragma unroll
        for (int t = 0; t < kBlock; t++) {
          k[t] = sir_fast<REFINE_H>(yv[t], rxz, c1h, c2, lim, r[t], w[t]);
          all &= k[t];
        }
        if (!all) {
#pragma unroll
          for (int t = 0; t < kBlock; t++)
            if (!k[t]) sir_exact<REFINE_H>(rx, yv[t] * yv[t], rzz, P.cp, d, P.dt, r[t], w[t]);
        }
        accumulate_block(r, w, cnt_w, out_mask);
      }
      const bool unit_out = settle(cnt_w, out_mask);
      for (; j < j1; j++) {
        const double ry = ye - ys[j];
        double r, h;
        if (!sir_fast<REFINE_H>(ry, rxz, c1h, c2, lim, r, h))
          sir_exact<REFINE_H>(rx, ry * ry, rzz, P.cp, d, P.dt, r, h);
        accumulate(sat_i32(r), h);
      }
      oob_last |= (e == L - 1) && (oob != oob_unit || unit_out);
      e += P.cstep_e;
      i += P.cstep_i;
      if (e >= L) { e -= L; i++; }
      if (i >= nx) { i -= nx; sg++; }
      sg += P.cstep_sg;
    }
  }
  if (use_lds) lds_add(hist + cur, acc);
  else if (cur >= 0) global_add(out + cur, (T) acc);

  // Status: wave-reduced, one atomic per wave.
  const unsigned wsum = wave_sum(oob + oob_s);
  const bool any_last = __any(oob_last);
  if ((threadIdx.x & 63) == 0 && wsum) {
    atomicOr(&P.status->any_oob, 1u);
    if (any_last) atomicOr(&P.status->last_elem_oob, 1u);
    atomicAdd(&P.status->oob, (unsigned long long) wsum);
  }

  if (use_lds) {
    __syncthreads();
    if (P.mode == SIR_LDS_STORE) {
      for (int i = threadIdx.x; i < nt; i += kThreads) out[i] = (T) hist[i];
    } else {
      for (int i = threadIdx.x; i < nt; i += kThreads) {
        const double v = hist[i];
        if (v != 0.0) global_add(out + i, (T) v);
      }
    }
    __syncthreads();
  }
}

// Five waves per SIMD: the LDS histogram of a 4000-sample SIR already limits
// a CU to five 256-thread work-groups, and holding the registers to that
// occupancy (a few spilled) beat fewer, register-rich waves and larger
// work-groups on the uneven batch.
template <typename T, bool ANY_GLOBAL>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(5)))
sir_kernel(const SirDev *__restrict__ probs, int n_probs, int64_t item_lo, int64_t item_hi,
           unsigned long long *counter)
{
  extern __shared__ __align__(16) unsigned char smem[];
  double *hist = reinterpret_cast<double *>(smem);
  __shared__ double red[2 * kWaves];
  // This launch serves items [item_lo, item_hi) (a chunk of a long plan).
  if (counter == nullptr) {
    const int64_t item = item_lo + blockIdx.x;
    if (item >= item_hi) return;
    const int p = find_problem(probs, n_probs, item);
    sir_item<T, ANY_GLOBAL>(probs[p], item - probs[p].item_begin, hist, red);
    return;
  }
  // Persistent: pull items from counter[0]; the last work-group to leave
  // resets both counters for the next launch.
  __shared__ int64_t s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = item_lo + (int64_t) atomicAdd(&counter[0], 1ull);
    __syncthreads();
    const int64_t item = s_item;
    __syncthreads();
    if (item >= item_hi) break;
    const int p = find_problem(probs, n_probs, item);
    sir_item<T, ANY_GLOBAL>(probs[p], item - probs[p].item_begin, hist, red);
  }
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned long long done = atomicAdd(&counter[1], 1ull);
    if (done == gridDim.x - 1) {
      atomicExch(&counter[0], 0ull);
      atomicExch(&counter[1], 0ull);
    }
  }
}


__global__ void rect_indices_kernel(double xo, double yo, double zo, const double *xs, int nx,
                                    const double *ys, int ny, double cp, double dt,
                                    double delay, int64_t *it_out, int64_t cap)
{
  const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t) nx * ny || q >= cap) return;
  const int row = (int) (q / nx), col = (int) (q - (int64_t) row * nx);
  const double ry = yo - ys[row];
  const double rx = xo - xs[col];
  const double c1 = 1000.0 / (cp * dt);
  const double c2 = delay * (1.0 / dt);
  // Every index is wanted here, so the bound covers this contribution's |u|.
  double r, rinv;
  sqrt_rsqrt<1>((rx * rx + ry * ry) + zo * zo, r, rinv);
  const double umax = __builtin_fabs(__builtin_fma(r, c1, -c2));
  const double lim = 0.5 - kBoundScale * (umax + 2.0 + __builtin_fabs(c2));
  double w;
  it_out[q] = (int64_t) sir_index<true>(rx, ry * ry, zo * zo, ry * ry + zo * zo, 0.5 * c1, c2, lim,
                                        cp, delay, dt, w);
}

// The hardware FP64 rsqrt seed of each s (the first step of sir_fast and
// saft_pos), for tests of the accuracy the certification assumes.
__global__ void rsq_seed_kernel(const double *s, int64_t n, double *y)
{
  const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = __builtin_amdgcn_rsq(s[i]);
}

int gcd_int(int a, int b)
{
  while (b) { int t = a % b; a = b; b = t; }
  return a;
}

int device_cus()
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

// Up to kZeroRanges buffers zeroed by one launch (graph-safe; a plan's
// status block and the outputs it accumulates into): blockIdx.y picks the
// buffer, blockIdx.x strides over its 32-bit words.
constexpr int kZeroRanges = 16;
struct ZeroList {
  unsigned int *p[kZeroRanges];
  int64_t words[kZeroRanges];
};

__global__ void fill_zero_list_kernel(ZeroList z)
{
  unsigned int *p = z.p[blockIdx.y];
  const int64_t n = z.words[blockIdx.y];
  for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t) gridDim.x * blockDim.x)
    p[i] = 0u;
}

// Zero the buffers (pointer, bytes; bytes a multiple of 4) in as few
// launches as possible.
void launch_fill_zero_list(const std::vector<std::pair<void *, size_t>> &bufs, hipStream_t s)
{
  size_t k = 0;
  while (k < bufs.size()) {
    ZeroList z;
    int n = 0;
    int64_t most = 0;
    for (; k < bufs.size() && n < kZeroRanges; k++) {
      if (!bufs[k].second) continue;
      z.p[n] = reinterpret_cast<unsigned int *>(bufs[k].first);
      z.words[n] = (int64_t) (bufs[k].second / 4);
      most = std::max(most, z.words[n]);
      n++;
    }
    if (!n) break;
    const unsigned bx = (unsigned) std::min<int64_t>((most + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_zero_list_kernel, dim3(bx, (unsigned) n), dim3(256), 0, s, z);
    DREAM_HIP_CHECK(hipGetLastError());
  }
}


} // namespace

// ------------------------------------------------------------------ host helpers

void disc_rows(double R, double dx, double dy, std::vector<double> &ys, std::vector<double> &x0,
               std::vector<double> &count)
{
  ys.clear();
  x0.clear();
  count.clear();
  const double ysmin = -R, ysmax = R;
  double y = ysmin + dy / 2.0;
  while (y <= ysmax) {
    const double rxs = std::sqrt(R * R - y * y);
    const double xsmin = -rxs, xsmax = rxs;
    const double x = xsmin + dx / 2.0;
    double n = 0.0;
    for (double xs = x; xs <= xsmax; xs += dx) n += 1.0;
    ys.push_back(y);
    x0.push_back(x);
    count.push_back(n);
    y += dy;
  }
}

double cylinder_rows(double b, double Rc, double dy, std::vector<double> &ys,
                     std::vector<double> &zrow)
{
  ys.clear();
  zrow.clear();
  const double R = std::fabs(Rc);
  const double hb = b / 2.0;
  const double z_R = R - std::sqrt(Rc * Rc - hb * hb);
  const double span = 2.0 * std::atan(hb / (R - z_R));
  const double ph_min = -span / 2.0, ph_max = span / 2.0;
  const double dphi = std::asin(dy / R);
  double ph = ph_min + dphi / 2.0;
  double y = R * std::sin(ph);
  while (ph <= ph_max) {            // false throughout when b > 2|Rc| (NaN angles)
    const double zs = R - std::sqrt(R * R - y * y);
    ys.push_back(y);
    zrow.push_back(Rc >= 0.0 ? zs : -zs);
    ph += dphi;
    y = R * std::sin(ph);
  }
  return dphi;
}

void split_disc_rows(int64_t L, double dx, std::vector<double> &ys, std::vector<double> &x0,
                     std::vector<double> &count)
{
  // Rows are dealt one per lane, so a disc whose rows number fewer than the
  // lanes, or whose rows differ in length, leaves lanes idle.  Cut the rows
  // into segments of at most S sub-elements, S chosen to minimise (passes of
  // 256 lanes) x S over the point's L x segments; each segment starts at the
  // x the row's accumulated walk reaches there, so every position is the
  // reference's bit for bit.
  double mx = 0.0;
  for (double c : count) mx = std::max(mx, c);
  const int64_t maxc = (int64_t) mx;
  if (maxc <= 1 || L < 1) return;
  int64_t best_s = maxc;
  double best_t = 0.0;
  // Segments shorter than 48 cost more in per-row set-up than they gain in
  // balance (measured on MI355X: the annular benchmark's disc launch took
  // 2.15 ms unsplit, 1.83 ms with segments down to 9 sub-elements, and the
  // whole call was fastest with a floor of 48).
  const int64_t smin = std::max<int64_t>({(int64_t) 48, maxc / 16});
  for (int64_t S = maxc; S >= smin; S = std::min(S - 1, S * 31 / 32)) {
    int64_t q = 0;
    for (double c : count) q += ((int64_t) c + S - 1) / S;
    const double t = std::ceil((double) (q * L) / (double) kThreads) * (double) S;
    if (S == maxc || t < best_t) { best_t = t; best_s = S; }
  }
  if (best_s == maxc) return;
  std::vector<double> vy, vx, vn;
  for (size_t r = 0; r < ys.size(); r++) {
    double x = x0[r];
    int64_t left = (int64_t) count[r];
    if (left == 0) {
      vy.push_back(ys[r]); vx.push_back(x); vn.push_back(0.0);
      continue;
    }
    while (left > 0) {
      const int64_t c = std::min(best_s, left);
      vy.push_back(ys[r]);
      vx.push_back(x);
      vn.push_back((double) c);
      for (int64_t k = 0; k < c; k++) x += dx;
      left -= c;
    }
  }
  ys.swap(vy);
  x0.swap(vx);
  count.swap(vn);
}

std::vector<double> subgrid(double w, double d)
{
  std::vector<double> v;
  const double smax = w / 2.0;
  double s = -w / 2.0 + d / 2.0;
  while (s <= smax) {
    v.push_back(s);
    s += d;
  }
  return v;
}

// Array helpers (src/arr_functions.cc:60-262), evaluated per element in the
// reference's operation order so the delays -- and hence the sample
// indices -- are bit-identical.
namespace {

void aperture_extent(const double *gx, const double *gy, int64_t L, double &x_max,
                     double &y_max, double &r_max)
{
  x_max = std::fabs(gx[0]);
  for (int64_t i = 1; i < L; i++) if (std::fabs(gx[i]) > x_max) x_max = std::fabs(gx[i]);
  y_max = std::fabs(gy[0]);
  for (int64_t i = 1; i < L; i++) if (std::fabs(gy[i]) > y_max) y_max = std::fabs(gy[i]);
  r_max = std::sqrt(x_max * x_max + y_max * y_max);
}

double focus_delay(int met, double focal, double gx, double gy, double x_max, double y_max,
                   double r_max, double cp)
{
  switch (met) {
  case 1: { // x
    double outer = std::sqrt(x_max * x_max + focal * focal);
    return (outer - std::sqrt(gx * gx + focal * focal)) * 1.0e3 / cp;
  }
  case 3: { // y
    double outer = std::sqrt(y_max * y_max + focal * focal);
    return (outer - std::sqrt(gy * gy + focal * focal)) * 1.0e3 / cp;
  }
  case 5: { // xy
    double outer = std::sqrt(r_max * r_max + focal * focal);
    return (outer - std::sqrt(gx * gx + gy * gy + focal * focal)) * 1.0e3 / cp;
  }
  case 6: { // x+y
    double outer = std::sqrt(r_max * r_max + focal * focal);
    double px = std::sqrt(gx * gx + focal * focal);
    double py = std::sqrt(gy * gy + focal * focal);
    return (outer - (px + py)) * 1.0e3 / cp;
  }
  case 7:   // ud
    return focal;
  default:  // none, negx, negy
    return 0.0;
  }
}

double steer_delay(int met, double theta, double phi, double gx, double gy, double x_max,
                   double y_max, double cp)
{
  const double deg2rad = M_PI / (double) 180.0;
  auto leg = [&](double ang, double ext, double g) {
    double sn = std::sin(ang * deg2rad);
    double e = ext * sn;
    return (e + g * sn) * 1.0e3 / cp;
  };
  switch (met) {
  case 1: return leg(theta, x_max, gx);
  case 2: return leg(phi, y_max, gy);
  case 3: {
    double sx = leg(theta, x_max, gx);
    double sy = leg(phi, y_max, gy);
    return sx + sy;
  }
  default: return 0.0;
  }
}

double apod_weight(int met, int64_t n, const double *ud, double gx, double gy, double r_max,
                   double par)
{
  const double r = std::sqrt(gx * gx + gy * gy);
  switch (met) {
  case 7: return ud[n];
  case 0: return 1.0 - std::fabs(r) / r_max;
  case 1: return std::exp(-(par * r * r) / (r_max * r_max));
  case 2: return par + std::cos(r * M_PI / r_max);
  case 3: return 0.5 * (1.0 - std::cos((r + r_max) * M_PI / r_max));
  case 4: {
    const double a0 = 25.0 / 46.0;
    return a0 - (1.0 - a0) * std::cos((r + r_max) * M_PI / r_max);
  }
  case 5: return 1.0 - r * r / (r_max * r_max);
  case 6: return (1.0 - r * r / (r_max * r_max)) * (1.0 - r * r / (r_max * r_max));
  default: return 1.0;
  }
}

} // namespace

std::vector<double> element_table(const SirProblem &p)
{
  const int64_t L = (int64_t) (p.G.size() / 3);
  const double *gx = p.G.data(), *gy = gx + L, *gz = gy + L;
  double x_max, y_max, r_max;
  aperture_extent(gx, gy, L, x_max, y_max, r_max);
  double ds = p.dx * p.dy;
  if (p.shape == SIR_CYLIND) {
    std::vector<double> ys, zr;
    ds = std::fabs(p.R) * p.dx * cylinder_rows(p.b, p.R, p.dy, ys, zr);
  }
  std::vector<double> t((size_t) (6 * L));
  for (int64_t e = 0; e < L; e++) {
    const double foc = (p.foc_met == 7) ? p.focal[(size_t) e]
                     : focus_delay(p.foc_met, p.focal.empty() ? 0.0 : p.focal[0], gx[e], gy[e],
                                   x_max, y_max, r_max, p.cp);
    const double steer = steer_delay(p.steer_met, p.theta, p.phi, gx[e], gy[e], x_max, y_max, p.cp);
    const double w = p.do_apod ? apod_weight(p.apod_met, e, p.apod.data(), gx[e], gy[e], r_max,
                                             p.apod_par)
                               : 1.0;
    t[(size_t) e] = gx[e];
    t[(size_t) (L + e)] = gy[e];
    t[(size_t) (2 * L + e)] = gz[e];
    t[(size_t) (3 * L + e)] = foc;
    t[(size_t) (4 * L + e)] = steer;
    t[(size_t) (5 * L + e)] = w * p.v * ds;   // ((w v) ds), the numerator of ai
  }
  return t;
}

// ------------------------------------------------------------------ plan

// DREAM_HIP_SIR_WALK = rows | columns forces the sub-element walk of
// rectangle problems (for comparisons and tests): 1 rows, 2 columns, 0 the
// per-point choice.
int walk_override()
{
  const char *e = std::getenv("DREAM_HIP_SIR_WALK");
  if (!e) return 0;
  if (std::strcmp(e, "rows") == 0) return 1;
  if (std::strcmp(e, "columns") == 0) return 2;
  return 0;
}

SirPlan::SirPlan(const std::vector<SirProblem> &problems, int out_dtype, int target_items)
  : out_dtype_(out_dtype)
{
  require(out_dtype == DREAM_HIP_F64 || out_dtype == DREAM_HIP_F32,
          "SIR output precision must be DREAM_HIP_F64 or DREAM_HIP_F32");
  const size_t P = problems.size();
  const size_t esz = sizeof(double);          // LDS histogram entries are FP64
  const size_t lds_cap = 64 * 1024;
  const int cus = device_cus();
  if (target_items <= 0) target_items = 8 * cus;

  int64_t total_no = 0;
  for (auto &p : problems) total_no += p.No;

  // Host tables and per-problem shape.
  std::vector<std::vector<double>> xs(P), ys(P), el(P), rx0(P), rn(P), zr(P);
  std::vector<SirDev> dev(P);
  std::vector<double> cost(P, 0.0);
  size_t table_doubles = 0;
  int64_t max_nt_lds = 0;
  for (size_t i = 0; i < P; i++) {
    const SirProblem &p = problems[i];
    const int64_t L = (int64_t) (p.G.size() / 3);
    require(p.No >= 0 && p.nt >= 1 && p.nt < (1ll << 31), "SIR: need No >= 0 and 1 <= nt < 2^31");
    require(L >= 1, "SIR: need at least one element");
    require(p.dx > 0 && p.dy > 0 && p.dt > 0, "SIR: dx, dy and dt must be positive");
    require(p.No == 0 || (p.Ro && p.delay && p.h), "SIR: null Ro, delay or h");
    require(p.shape == SIR_RECT || p.shape == SIR_CIRC || p.shape == SIR_CYLIND,
            "SIR: unknown element shape");
    int64_t nx_cost;
    if (p.shape == SIR_CIRC) {
      disc_rows(p.R, p.dx, p.dy, ys[i], rx0[i], rn[i]);
      split_disc_rows((int64_t) (p.G.size() / 3), p.dx, ys[i], rx0[i], rn[i]);
      double sum = 0.0, mx = 0.0;
      for (double c : rn[i]) { sum += c; mx = std::max(mx, c); }
      xs[i].assign(1, 0.0);                     // unused by the disc walk
      nx_cost = ys[i].empty() ? 0 : (int64_t) (sum / (double) ys[i].size() + 1.0);
      if (mx == 0.0) nx_cost = 0;
    } else if (p.shape == SIR_CYLIND) {
      xs[i] = subgrid(p.a, p.dx);
      cylinder_rows(p.b, p.R, p.dy, ys[i], zr[i]);
      nx_cost = (int64_t) xs[i].size();
    } else {
      xs[i] = subgrid(p.a, p.dx);
      ys[i] = subgrid(p.b, p.dy);
      nx_cost = (int64_t) xs[i].size();
    }
    el[i] = element_table(p);
    const int64_t nx = nx_cost, ny = (int64_t) ys[i].size();
    require(L * ny < (1ll << 31) && nx < (1ll << 31), "SIR: element x row count too large");
    SirDev &d = dev[i];
    std::memset(&d, 0, sizeof d);
    d.Ro = p.Ro;
    d.No = p.No;
    d.delay = p.delay;
    d.delay_multi = p.delay_type != 0;
    d.L = (int32_t) L;
    d.nx = (int32_t) nx;
    d.ny = (int32_t) ny;
    d.shape = p.shape;
    d.dx = p.dx;
    int stride = 1;
    if (ny > 64) {
      stride = (int) (ny / 64);
      while (gcd_int(stride, (int) ny) != 1) stride++;
    }
    d.row_stride = stride;
    // An element with no rows (a side or radius below half a step) has no
    // work items; its walk steps are never used.
    const int64_t ny_w = std::max<int64_t>(ny, 1);
    d.step_e = (int32_t) (kThreads / ny_w);
    d.step_j = (int32_t) (kThreads % ny_w);
    d.step_row = (int32_t) (((int64_t) d.step_j * stride) % ny_w);
    d.nt = p.nt;
    d.cp = p.cp;
    d.dt = p.dt;
    d.inv_dt = 1.0 / p.dt;
    d.c1 = 1000.0 / (p.cp * p.dt);
    d.amp = 1000.0 / (2.0 * M_PI * p.dt);
    d.h = p.h;
    const int64_t Q = L * ny;
    int64_t chunks = 1;
    if (total_no < target_items) {
      chunks = (target_items + std::max<int64_t>(total_no, 1) - 1) / std::max<int64_t>(total_no, 1);
      chunks = std::min<int64_t>(chunks, std::max<int64_t>(Q / kThreads, 1));
    }
    d.chunks = (int32_t) std::max<int64_t>(chunks, 1);
    if (p.shape == SIR_RECT && nx > 0 && ny > 0 && walk_override() != 1) {
      d.walk = walk_override() == 2 ? 1 : 0;
      // Column segments: the count that minimises lane rounds per item x
      // rows per segment x the run-length cost of a segment (a flush per
      // segment plus one per eighth step or so), over at most ny/8
      // segments.
      const int64_t QC = L * nx;
      double best = 0.0;
      int64_t best_n = 1;
      for (int64_t ns = 1; ns <= std::max<int64_t>(ny / 8, 1); ns++) {
        const int64_t S = (ny + ns - 1) / ns;
        const int64_t nseg = (ny + S - 1) / S;
        const double rounds = std::ceil((double) (QC * nseg) / (double) d.chunks / kThreads);
        const double c = rounds * (double) S * (1.0 + 4.0 * (1.0 + (double) (S - 1) * 0.15) / (double) S);
        if (ns == 1 || c < best) { best = c; best_n = nseg; }
      }
      const int64_t S = (ny + best_n - 1) / best_n;
      require(QC * best_n < (1ll << 31) - kThreads, "SIR: element x column count too large");
      d.nseg = (int32_t) ((ny + S - 1) / S);
      d.seg_len = (int32_t) S;
      d.cstep_sg = (int32_t) (kThreads / QC);
      d.cstep_i = (int32_t) ((kThreads % QC) / L);
      d.cstep_e = (int32_t) ((kThreads % QC) % L);
      double sxm = 0.0, sym = 0.0;
      for (double v : xs[i]) sxm += std::fabs(v);
      for (double v : ys[i]) sym += std::fabs(v);
      const double lam = p.cp * p.dt / 1000.0;
      d.wx = sxm / (double) nx;
      d.wy = sym / (double) ny;
      d.sx = p.dx / lam;
      d.sy = p.dy / lam;
    }
    const bool fits = (size_t) (p.nt + 1) * esz <= lds_cap;     // + the spare bin
    d.mode = !fits ? SIR_GLOBAL : (d.chunks > 1 ? SIR_LDS_ATOMIC : SIR_LDS_STORE);
    if (fits) max_nt_lds = std::max(max_nt_lds, p.nt);
    cost[i] = (double) Q * (double) nx / d.chunks;
    table_doubles += xs[i].size() + ys[i].size() + el[i].size() + rx0[i].size() + rn[i].size()
                   + zr[i].size();
    levels_.push_back(p.err_level);
  }

  // Largest per-item cost first: one work-group per item is then served in
  // longest-processing-time order by the dispatcher (or by the counter).
  std::vector<size_t> order(P);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });

  DREAM_HIP_CHECK(hipMalloc(&d_tables_, std::max<size_t>(table_doubles, 1) * sizeof(double)));
  DREAM_HIP_CHECK(hipMalloc(&d_status_, std::max<size_t>(P, 1) * sizeof(Status)));
  DREAM_HIP_CHECK(hipMalloc(&d_probs_, std::max<size_t>(P, 1) * sizeof(SirDev)));
  DREAM_HIP_CHECK(hipMalloc(&d_counter_, 2 * sizeof(unsigned long long)));
  DREAM_HIP_CHECK(hipMemset(d_counter_, 0, 2 * sizeof(unsigned long long)));

  std::vector<double> tables;
  tables.reserve(table_doubles);
  host_.resize(P);
  item_count_.assign(P, 0);
  item_cost_.assign(P, 0.0);
  zero_len_.assign(P, 0);
  zero_ptr_.assign(P, nullptr);
  int64_t item = 0;
  for (size_t k = 0; k < P; k++) {
    const size_t i = order[k];
    SirDev d = dev[i];
    d.xs = d_tables_ + tables.size();
    tables.insert(tables.end(), xs[i].begin(), xs[i].end());
    d.ys = d_tables_ + tables.size();
    tables.insert(tables.end(), ys[i].begin(), ys[i].end());
    d.el = d_tables_ + tables.size();
    tables.insert(tables.end(), el[i].begin(), el[i].end());
    d.row_x0 = d_tables_ + tables.size();
    tables.insert(tables.end(), rx0[i].begin(), rx0[i].end());
    d.row_n = d_tables_ + tables.size();
    tables.insert(tables.end(), rn[i].begin(), rn[i].end());
    d.zrow = d_tables_ + tables.size();
    tables.insert(tables.end(), zr[i].begin(), zr[i].end());
    d.status = d_status_ + i;
    d.item_begin = item;
    // A problem with no sub-elements (or no points) contributes no items.
    const int64_t items = (d.nx > 0 && d.ny > 0) ? d.No * d.chunks : 0;
    item += items;
    host_[k] = d;
    item_count_[k] = items;
    item_cost_[k] = cost[i];
    if (d.mode != SIR_LDS_STORE || items == 0) {
      zero_len_[i] = d.No * d.nt;
      zero_ptr_[i] = d.h;
    }
  }
  n_items_ = item;
  lds_bytes_ = max_nt_lds ? (size_t) (max_nt_lds + 1) * esz : 0;
  if (!tables.empty())
    DREAM_HIP_CHECK(hipMemcpy(d_tables_, tables.data(), tables.size() * sizeof(double),
                              hipMemcpyHostToDevice));
  DREAM_HIP_CHECK(hipMemcpy(d_probs_, host_.data(), P * sizeof(SirDev), hipMemcpyHostToDevice));

  // Persistent grid when the list is long and uneven; otherwise one
  // work-group per item.
  const int wg_per_cu = lds_bytes_ ? (int) std::max<size_t>(1, std::min<size_t>(8, (160 * 1024) / std::max<size_t>(lds_bytes_, 1))) : 8;
  persistent_ = P > 1 && n_items_ > (int64_t) cus * wg_per_cu * 4;
  grid_ = persistent_ ? cus * wg_per_cu : (int) std::min<int64_t>(n_items_, 1ll << 31);
  // Launch shapes for the autotuner (the heuristic first): for batches, one
  // work-group per item and persistent grids of 1, 2 or wg_per_cu
  // work-groups per CU.
  shapes_.push_back(SirShape{persistent_, grid_});
  if (P > 1 && n_items_ < (1ll << 31)) {
    auto add = [&](bool pers, int grid) {
      for (const SirShape &x : shapes_) if (x.persistent == pers && x.grid == grid) return;
      shapes_.push_back(SirShape{pers, grid});
    };
    add(false, (int) n_items_);
    for (int k : {wg_per_cu, 2, 1}) add(true, cus * k);
  }
}

SirPlan::~SirPlan()
{
  if (graph_) (void) hipGraphExecDestroy(graph_);
  (void) hipFree(d_tables_);
  (void) hipFree(d_status_);
  (void) hipFree(d_probs_);
  (void) hipFree(d_counter_);
}

void SirPlan::choose(hipStream_t s)
{
  if (pick_ >= 0) return;
  int64_t total_no = 0;
  for (const SirDev &d : host_) total_no += d.No;
  char key[160];
  std::snprintf(key, sizeof key, "sir|%d|%zu|%lld|%lld|%zu", out_dtype_, host_.size(),
                (long long) n_items_, (long long) total_no, lds_bytes_);
  const int t = n_items_ ? autotune(key, (int) shapes_.size(),
                                    [&](int i, hipStream_t st) { launch_shape(shapes_[(size_t) i], st, nullptr); }, s)
                         : -1;
  pick_ = t >= 0 ? t : 0;
}

void SirPlan::launch(hipStream_t s, ChunkGate *gate)
{
  launch_shape(shapes_[(size_t) std::max(pick_, 0)], s, gate);
}

void SirPlan::launch_shape(const SirShape &sh, hipStream_t s, ChunkGate *gate)
{
  // The status blocks and the outputs that items accumulate into, zeroed
  // by one launch (usually).
  std::vector<std::pair<void *, size_t>> zero{{d_status_, host_.size() * sizeof(Status)}};
  const size_t esz = out_dtype_ == DREAM_HIP_F64 ? 8 : 4;
  for (size_t i = 0; i < zero_len_.size(); i++)
    if (zero_len_[i]) zero.emplace_back(zero_ptr_[i], (size_t) zero_len_[i] * esz);
  launch_fill_zero_list(zero, s);
  if (n_items_ == 0) return;
  unsigned long long *ctr = sh.persistent ? d_counter_ : nullptr;
  bool any_global = false;
  for (const SirDev &d : host_) any_global |= d.mode == SIR_GLOBAL;
  auto range = [&](int64_t lo, int64_t hi) {
    const int64_t n = hi - lo;
    const unsigned grid = (unsigned) (sh.persistent ? std::min<int64_t>(sh.grid, n) : n);
    const int np = (int) host_.size();
    if (out_dtype_ == DREAM_HIP_F64) {
      if (any_global)
        hipLaunchKernelGGL((sir_kernel<double, true>), dim3(grid), dim3(kThreads), lds_bytes_, s,
                           d_probs_, np, lo, hi, ctr);
      else
        hipLaunchKernelGGL((sir_kernel<double, false>), dim3(grid), dim3(kThreads), lds_bytes_,
                           s, d_probs_, np, lo, hi, ctr);
    } else {
      if (any_global)
        hipLaunchKernelGGL((sir_kernel<float, true>), dim3(grid), dim3(kThreads), lds_bytes_, s,
                           d_probs_, np, lo, hi, ctr);
      else
        hipLaunchKernelGGL((sir_kernel<float, false>), dim3(grid), dim3(kThreads), lds_bytes_, s,
                           d_probs_, np, lo, hi, ctr);
    }
    DREAM_HIP_CHECK(hipGetLastError());
  };
  if (!gate) {
    range(0, n_items_);
    return;
  }
  // Chunks of about chunk_work() contributions: items are in the plan's
  // order (largest first) and all items of one problem cost the same.
  const double W = (double) chunk_work();
  int64_t lo = 0, item = 0;
  double acc = 0.0;
  for (size_t k = 0; k < host_.size(); k++) {
    const int64_t items = item_count_[k];
    const double c = std::max(item_cost_[k], 1.0);
    int64_t left = items;
    while (left > 0) {
      const int64_t take = std::min<int64_t>(left, (int64_t) std::ceil((W - acc) / c));
      acc += (double) take * c;
      item += take;
      left -= take;
      if (acc >= W && item < n_items_) {
        gate->before();
        range(lo, item);
        gate->after();
        lo = item;
        acc = 0.0;
      }
    }
  }
  gate->before();
  range(lo, n_items_);
  gate->after();
}

void SirPlan::run(hipStream_t s, bool use_graph)
{
  if (cancelled()) throw Error(DREAM_HIP_ECANCELLED, "SIR cancelled");
  last_stream_ = s;
  ran_ = true;
  choose(s);
  if (!use_graph) {
    ChunkGate gate(s);
    launch(s, &gate);
    return;
  }
  if (!graph_) {
    // Capture on the library's own stream (the default stream cannot be
    // captured); the graph is then launched on the caller's stream.
    hipStream_t cs = lib_stream();
    hipGraph_t g = nullptr;
    DREAM_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    launch(cs, nullptr);
    DREAM_HIP_CHECK(hipStreamEndCapture(cs, &g));
    DREAM_HIP_CHECK(hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0));
    DREAM_HIP_CHECK(hipGraphDestroy(g));
  }
  DREAM_HIP_CHECK(hipGraphLaunch(graph_, s));
}

void SirPlan::status(int32_t *codes, int64_t *oob)
{
  const size_t P = host_.size();
  std::vector<Status> st(P);
  if (ran_) DREAM_HIP_CHECK(hipStreamSynchronize(last_stream_));
  if (P) DREAM_HIP_CHECK(hipMemcpy(st.data(), d_status_, P * sizeof(Status), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < P; i++) {
    const int lv = levels_[i];
    int code = E_NONE;
    if (lv == L_STOP || lv == L_PSTOP) {
      if (st[i].any_oob) code = E_OOB;
    } else if (st[i].last_elem_oob) {
      code = level_code(lv);
    }
    if (codes) codes[i] = code;
    if (oob) oob[i] = (int64_t) st[i].oob;
  }
}

void rsq_seed(const double *s, int64_t n, double *y)
{
  if (n <= 0) return;
  double *d = nullptr;
  DREAM_HIP_CHECK(hipMalloc(&d, 2 * (size_t) n * sizeof(double)));
  DREAM_HIP_CHECK(hipMemcpy(d, s, (size_t) n * sizeof(double), hipMemcpyHostToDevice));
  const int blocks = (int) ((n + 255) / 256);
  hipLaunchKernelGGL(rsq_seed_kernel, dim3(blocks), dim3(256), 0, lib_stream(), d, n, d + n);
  DREAM_HIP_CHECK(hipGetLastError());
  DREAM_HIP_CHECK(hipStreamSynchronize(lib_stream()));
  DREAM_HIP_CHECK(hipMemcpy(y, d + n, (size_t) n * sizeof(double), hipMemcpyDeviceToHost));
  (void) hipFree(d);
}

void rect_indices(double xo, double yo, double zo, double a, double b, double dx, double dy,
                  double dt, double delay, double cp, int64_t *it_out, int64_t cap,
                  int64_t *count)
{
  std::vector<double> xs = subgrid(a, dx), ys = subgrid(b, dy);
  const int64_t n = (int64_t) xs.size() * (int64_t) ys.size();
  *count = n;
  const int64_t m = std::min(n, cap);
  if (m <= 0) return;
  double *d_xy = nullptr;
  int64_t *d_it = nullptr;
  DREAM_HIP_CHECK(hipMalloc(&d_xy, (xs.size() + ys.size()) * sizeof(double)));
  DREAM_HIP_CHECK(hipMalloc(&d_it, (size_t) m * sizeof(int64_t)));
  DREAM_HIP_CHECK(hipMemcpy(d_xy, xs.data(), xs.size() * sizeof(double), hipMemcpyHostToDevice));
  DREAM_HIP_CHECK(hipMemcpy(d_xy + xs.size(), ys.data(), ys.size() * sizeof(double),
                            hipMemcpyHostToDevice));
  const int blocks = (int) ((m + 255) / 256);
  hipLaunchKernelGGL(rect_indices_kernel, dim3(blocks), dim3(256), 0, lib_stream(), xo, yo, zo,
                     d_xy, (int) xs.size(), d_xy + xs.size(), (int) ys.size(), cp, dt, delay,
                     d_it, m);
  DREAM_HIP_CHECK(hipGetLastError());
  DREAM_HIP_CHECK(hipStreamSynchronize(lib_stream()));
  DREAM_HIP_CHECK(hipMemcpy(it_out, d_it, (size_t) m * sizeof(int64_t), hipMemcpyDeviceToHost));
  (void) hipFree(d_xy);
  (void) hipFree(d_it);
}

} // namespace dream
