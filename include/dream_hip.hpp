// dream_hip.hpp -- the reference's C++ hot-path classes over the C ABI of
// libdream_hip.so, for the reference's language gateways (Octave .oct,
// MATLAB .mex, pybind11, Julia CxxWrap) to link unchanged.
//
// Include it after the reference's "dream.h" (which supplies SIRError,
// ErrorLevel, DelayType, FocusMet, SteerMet, ApodMet, DASType and
// dream_idx_type) in place of "dream_arr_rect.h", "dreamrect.h" and "das.h".
//
//   class ArrRect   include/dream_arr_rect.h:33-54 (dream_arr_rect, abort, is_running)
//   class Rect      include/dreamrect.h:32-66     (dreamrect, cl_dreamrect, abort, is_running)
//   class DAS<T>    include/das.h:52-94, 386-393  (ctor, das, cl_das, abort, is_running,
//                                                  set_running, das_setup_has_changed)
//
// Behaviour kept from the reference: the SIRError result; with
// ErrorLevel::stop an out-of-bounds sample clears the running flag, so the
// gateways' `if (!obj.is_running())` error branches fire as before; the
// flag is set again at the start of every call.  Differences: h is
// overwritten rather than accumulated into (every reference caller zeroes
// it first, python/src/py_dream_arr_rect.cc:188-189), and HIP errors
// throw std::runtime_error as the reference's OpenCL setup errors do
// (include/das.h:159-162).  Attenuation (alpha > 0) runs on the GPU.
#pragma once

#include <stdexcept>
#include <string>

#include "dream_hip.h"

namespace dream_hip_detail {
// false when dream_hip_abort() stopped the call: the running flag stays
// cleared, so the gateway's own `if (!obj.is_running())` branch reports the
// interruption as it does for the reference's threads.
inline bool check(int rc)
{
  if (rc == DREAM_HIP_ECANCELLED) return false;
  if (rc != DREAM_HIP_OK) throw std::runtime_error(std::string("dream_hip: ") + dream_hip_last_error());
  return true;
}
inline SIRError finish(int code)
{
  if (code == 3) dream_hip_abort(0);    // the reference clears `running` on out_of_bounds
  return static_cast<SIRError>(code);
}
} // namespace dream_hip_detail

class ArrRect
{
public:
  ArrRect() = default;
  ~ArrRect() = default;

  SIRError dream_arr_rect(double alpha, double *Ro, dream_idx_type No, double a, double b,
                          double dx, double dy, double dt, dream_idx_type nt,
                          DelayType delay_type, double *delay, double v, double cp,
                          dream_idx_type num_elements, double *G, FocusMet foc_met, double *focal,
                          SteerMet steer_met, double theta, double phi, double *apod,
                          bool do_apod, ApodMet apod_met, double apod_par, double *h,
                          ErrorLevel err_level)
  {
    dream_hip_set_running();
    int code = 0;
    if (!dream_hip_detail::check(dream_hip_arr_rect(
            alpha, Ro, (int64_t) No, a, b, dx, dy, dt, (int64_t) nt, (int) delay_type, delay, v,
            cp, (int64_t) num_elements, G, (int) foc_met, focal, (int) steer_met, theta, phi,
            apod, do_apod ? 1 : 0, (int) apod_met, apod_par, h, (int) err_level, &code)))
      return SIRError::none;
    return dream_hip_detail::finish(code);
  }

  static void abort(int signum) { dream_hip_abort(signum); }
  bool is_running() { return dream_hip_is_running() != 0; }
};

class Rect
{
public:
  Rect() = default;
  ~Rect() = default;

  SIRError dreamrect(double alpha, double *Ro, dream_idx_type No, double a, double b, double dx,
                     double dy, double dt, dream_idx_type nt, DelayType delay_type,
                     double *delay, double v, double cp, double *h, ErrorLevel err_level)
  {
    dream_hip_set_running();
    int code = 0;
    if (!dream_hip_detail::check(dream_hip_rect(alpha, Ro, (int64_t) No, a, b, dx, dy, dt,
                                                (int64_t) nt, (int) delay_type, delay, v, cp, h,
                                                (int) err_level, &code)))
      return SIRError::none;
    return dream_hip_detail::finish(code);
  }

  // The reference's OpenCL entry point (opencl/cl_dreamrect.cc:42): scalar
  // delay, no error reporting, returns 0.  No No % 64 restriction here.
  int cl_dreamrect(const double *Ro, int No, double a, double b, double dx, double dy,
                   double dt, int nt, double delay, double v, double cp, double *h)
  {
    int code = 0;
    dream_hip_detail::check(dream_hip_rect(0.0, Ro, No, a, b, dx, dy, dt, nt, 0, &delay, v, cp,
                                           h, 1, &code));
    return 0;
  }

  static void abort(int signum) { dream_hip_abort(signum); }
  bool is_running() { return dream_hip_is_running() != 0; }
};

template <class T>
class DAS
{
public:
  // float uses the reference's float arithmetic (DREAM_HIP_F32_REF) so the
  // sample indices equal DAS<float>'s; double is DREAM_HIP_F64.
  static constexpr int kPrecision = sizeof(T) == 8 ? DREAM_HIP_F64 : DREAM_HIP_F32_REF;

  DAS(DASType das_type, dream_idx_type a_scan_len, dream_idx_type No,
      dream_idx_type num_t_elements, dream_idx_type num_r_elements, bool use_gpu = false,
      bool verbose = false)
    : m_das_type(das_type), m_a_scan_len(a_scan_len), m_No(No),
      m_num_t_elements(num_t_elements), m_num_r_elements(num_r_elements), m_use_gpu(use_gpu)
  {
    (void) verbose;
    dream_hip_detail::check(dream_hip_das_create((int) das_type, (int64_t) a_scan_len,
                                                 (int64_t) No, (int64_t) num_t_elements,
                                                 (int64_t) num_r_elements, kPrecision, &m_h));
  }
  ~DAS() { dream_hip_das_destroy(m_h); }
  DAS(const DAS &) = delete;
  DAS &operator=(const DAS &) = delete;

  SIRError das(const T *Y, const T *Ro, const T *gt, const T *gr, T dt, DelayType delay_type,
               const T *delay, T cp, T *Im, ErrorLevel err_level)
  {
    int code = 0;
    if (!dream_hip_detail::check(dream_hip_das(m_h, Y, Ro, gt, gr, (double) dt,
                                               (int) delay_type, delay, (double) cp, Im,
                                               (int) err_level, &code)))
      return SIRError::none;
    return dream_hip_detail::finish(code);
  }

  // DAS<T>::cl_das (opencl/cl_das.cc:62-159): scalar delay, returns 0 and
  // reports no out-of-bounds samples.
  int cl_das(const T *Y, const T *Ro, const T *gt, const T *gr, T dt, T delay, T cp, T *Im)
  {
    int code = 0;
    dream_hip_detail::check(dream_hip_das(m_h, Y, Ro, gt, gr, (double) dt, 0, &delay,
                                          (double) cp, Im, 1, &code));
    return 0;
  }

  static void abort(int signum) { dream_hip_abort(signum); }
  bool is_running() { return dream_hip_is_running() != 0; }
  void set_running() { dream_hip_set_running(); }

  bool das_setup_has_changed(DASType das_type, dream_idx_type a_scan_len, dream_idx_type No,
                             dream_idx_type num_t_elements, dream_idx_type num_r_elements,
                             bool use_gpu)
  {
    return das_type != m_das_type || a_scan_len != m_a_scan_len || No != m_No ||
           num_t_elements != m_num_t_elements || num_r_elements != m_num_r_elements ||
           use_gpu != m_use_gpu;
  }

private:
  void *m_h = nullptr;
  DASType m_das_type;
  dream_idx_type m_a_scan_len, m_No, m_num_t_elements, m_num_r_elements;
  bool m_use_gpu;
};
