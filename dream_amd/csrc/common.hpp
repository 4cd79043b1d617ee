// common.hpp -- shared host/device helpers of the DREAM HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "dream_hip.h"

namespace dream {

// SIRError / ErrorLevel numbering of the reference (include/dream.h:42-55).
enum { E_NONE = 0, E_IGNORE = 1, E_WARN = 2, E_OOB = 3 };
enum { L_NONE = 0, L_IGNORE = 1, L_WARN = 2, L_STOP = 3, L_PSTOP = 4 };
enum { DAS_SAFT = 0, DAS_TFM = 1, DAS_RCA_COLTX = 2, DAS_RCA_ROWTX = 3 };

// Error level -> SIRError code (src/dream_error.cc:40-76).
inline int level_code(int level)
{
  if (level == L_IGNORE) return E_IGNORE;
  if (level == L_WARN) return E_WARN;
  if (level == L_STOP || level == L_PSTOP) return E_OOB;
  return E_NONE;
}

// Exception carrying one of the DREAM_HIP_E* codes; the C ABI converts it.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hip_fail(hipError_t e, const char *what, const char *file, int line);

#define DREAM_HIP_CHECK(expr)                                              \
  do {                                                                     \
    hipError_t e__ = (expr);                                               \
    if (e__ != hipSuccess) ::dream::hip_fail(e__, #expr, __FILE__, __LINE__); \
  } while (0)

inline void require(bool ok, const std::string &msg)
{
  if (!ok) throw Error(DREAM_HIP_EINVAL, msg);
}
// Literal messages: no string is built unless the check fails.
inline void require(bool ok, const char *msg)
{
  if (!ok) throw Error(DREAM_HIP_EINVAL, msg);
}

// Device-side status block of one SIR problem or DAS call.
struct Status {
  unsigned int any_oob;        // some contribution / pair was out of bounds
  unsigned int last_elem_oob;  // SIR: some point's last element had one
  unsigned long long oob;      // number of out-of-bounds contributions
};

bool cancelled();

// Work units (pixel-channel pairs, sub-element contributions) per chunk of a
// long call: about 20-30 ms of device time (default 5 x 10^10; each chunk
// launch costs a partly filled last wave of work-groups), or the value a
// test set through dream_hip_test_chunking.
int64_t chunk_work();

// Runs a long call as consecutive chunk launches on stream `s`, testing the
// cancellation flag before each one.  At most two chunks are queued ahead of
// the device (before chunk i the host waits for chunk i - 2), so a call
// cancelled from another thread stops within about two chunks and throws
// DREAM_HIP_ECANCELLED; the chunks already launched have written their
// outputs.  While `s` captures a graph nothing is waited for or tested.
class ChunkGate {
public:
  explicit ChunkGate(hipStream_t s);
  ~ChunkGate();
  ChunkGate(const ChunkGate &) = delete;
  ChunkGate &operator=(const ChunkGate &) = delete;
  void before();   // call before launching a chunk
  void after();    // call after launching it
private:
  hipStream_t s_;
  hipEvent_t ev_[2] = {nullptr, nullptr};
  int64_t n_ = 0;
  bool capturing_ = false;
};

// Thread-local stream of the library (created on first use, default
// priority).
hipStream_t lib_stream();

// ------------------------------------------------------------------ device

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)

// sqrt(s) and 1/sqrt(s) to a few ulp from the hardware reciprocal square
// root (relative error about 2^-23) and two Goldschmidt steps.  Not
// correctly rounded: callers use it only where a certified error bound
// decides the result, and fall back to the IEEE sqrt otherwise.
// STEPS = 2 (the default) leaves a relative error of a few ulp; STEPS = 1
// leaves about 1.5 e0^2, e0 being the seed's relative error.
template <int STEPS = 2>
__device__ __forceinline__ void sqrt_rsqrt(double s, double &r, double &rinv)
{
  double y = __builtin_amdgcn_rsq(s);
  double g = s * y;
  double h = 0.5 * y;
#pragma unroll
  for (int k = 0; k < STEPS; k++) {
    const double e = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
  }
  r = g;
  rinv = h + h;
}

#endif

} // namespace dream
