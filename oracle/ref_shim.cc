// ref_shim.cc -- C entry points over the reference's own DAS and array-helper
// sources, compiled from the DREAM source tree by oracle/Makefile (target
// `ref`) into oracle/_ref/.  TEST INFRASTRUCTURE ONLY: it lets the oracle
// restatement and the golden-vector script compare against the reference
// itself.  The reference sources are compiled where they lie; nothing of
// them is copied here.
//
// Interfaces wrapped:
//   DAS<T>::DAS / das        include/das.h:57-72, extra_src/das.cc:36-298
//   max_dim_arr, focusing, beamsteering, apodization
//                            include/arr_functions.h, src/arr_functions.cc:60-262

#include <cstdint>

#include "das.h"
#include "arr_functions.h"

extern "C" {

int ref_das_f64(int type, const double *Y, int64_t K, const double *Ro, int64_t No,
                const double *Gt, int64_t Lt, const double *Gr, int64_t Lr, double dt,
                int delay_type, const double *delay, double cp, double *Im, int level)
{
  DAS<double> das((DASType) type, K, No, Lt, Lr, false, false);
  das.set_running();
  return (int) das.das(Y, Ro, Gt, Gr, dt, (DelayType) delay_type, delay, cp, Im,
                       (ErrorLevel) level);
}

int ref_das_f32(int type, const float *Y, int64_t K, const float *Ro, int64_t No,
                const float *Gt, int64_t Lt, const float *Gr, int64_t Lr, float dt,
                int delay_type, const float *delay, float cp, float *Im, int level)
{
  DAS<float> das((DASType) type, K, No, Lt, Lr, false, false);
  das.set_running();
  return (int) das.das(Y, Ro, Gt, Gr, dt, (DelayType) delay_type, delay, cp, Im,
                       (ErrorLevel) level);
}

void ref_max_dim_arr(double *gx, double *gy, double *gz, int64_t L,
                     double *x_max, double *y_max, double *r_max)
{
  max_dim_arr(x_max, y_max, r_max, gx, gy, gz, L);
}

double ref_focusing(int foc_met, double focal, double gx, double gy, double x_max,
                    double y_max, double r_max, double cp)
{
  double d = 0.0;
  focusing((FocusMet) foc_met, focal, gx, gy, x_max, y_max, r_max, cp, &d);
  return d;
}

double ref_beamsteering(int steer_met, double theta, double phi, double gx, double gy,
                        double x_max, double y_max, double r_max, double cp)
{
  double d = 0.0;
  beamsteering((SteerMet) steer_met, theta, phi, gx, gy, x_max, y_max, r_max, cp, &d);
  return d;
}

double ref_apodization(int apod_met, int64_t n, double *apod_vec, double prev, double gx,
                       double gy, double r_max, double apod_par)
{
  double w = prev;
  apodization((ApodMet) apod_met, n, apod_vec, &w, gx, gy, r_max, apod_par);
  return w;
}

} // extern "C"
