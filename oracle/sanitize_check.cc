// Sanitizer driver of the threaded CPU oracle (test infrastructure): runs
// the oracle's SIR and DAS entry points with several threads on small
// inputs that hit the shared state -- the `running` flag an out-of-bounds
// `stop` clears for every thread, the per-thread status aggregation, the
// out-of-bounds counters, the adjoint's per-thread partial sums -- so that
// ThreadSanitizer and AddressSanitizer/UBSan builds (oracle/sanitize.mk)
// report any race or invalid access.  Prints "sanitize ok" at the end.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

typedef int64_t idx_t;

extern "C" {
void dro_set_alpha(double alpha);
int dro_dream_arr_rect(const double *Ro, idx_t No, double a, double b, double dx, double dy,
                       double dt, idx_t nt, int delay_type, const double *delay, double v,
                       double cp, idx_t L, const double *G, int foc_met, const double *focal,
                       int steer_met, double theta, double phi, const double *apod,
                       int do_apod, int apod_met, double apod_par, double *h, int level,
                       int nthreads, long long *oob_total);
int dro_dream_arr_circ(const double *Ro, idx_t No, double R, double dx, double dy, double dt,
                       idx_t nt, int delay_type, const double *delay, double v, double cp,
                       idx_t L, const double *G, int foc_met, const double *focal,
                       int steer_met, double theta, double phi, const double *apod, int do_apod,
                       int apod_met, double apod_par, double *h, int level, int nthreads,
                       long long *oob_total);
int dro_dream_arr_annu(const double *Ro, idx_t No, double dx, double dy, double dt, idx_t nt,
                       int delay_type, const double *delay, double v, double cp,
                       idx_t num_radii, const double *Gr, int foc_met, const double *focal,
                       const double *apod, int do_apod, int apod_met, double apod_par,
                       double *h, int level, int nthreads, long long *oob_total);
int dro_das_f64(int type, const double *Y, idx_t K, const double *Ro, idx_t No,
                const double *Gt, idx_t Lt, const double *Gr, idx_t Lr, double dt,
                int delay_type, const double *delay, double cp, double *Im, int level,
                int nthreads);
int dro_das_f32(int type, const float *Y, idx_t K, const float *Ro, idx_t No,
                const float *Gt, idx_t Lt, const float *Gr, idx_t Lr, float dt,
                int delay_type, const float *delay, float cp, float *Im, int level,
                int nthreads);
int dro_das_ext_f64(int type, int interp, const double *Y, idx_t K, const double *Ro, idx_t No,
                    const double *Gt, idx_t Lt, const double *Gr, idx_t Lr, double dt,
                    double delay, double cp, const double *apod_t, const double *apod_r,
                    double *Im, int nthreads);
void dro_das_ext_adj_f64(int type, int interp, const double *Y, idx_t K, const double *Ro,
                         idx_t No, const double *Gt, idx_t Lt, const double *Gr, idx_t Lr,
                         double dt, double delay, double cp, const double *apod_t,
                         const double *apod_r, const double *g, double *gY, double *g_apod_t,
                         double *g_apod_r, double *g_cp, int nthreads);
}

int main()
{
  const int T = 4;
  // SIR: 3 elements, 40 points, some of them late enough to fall past nt.
  const idx_t No = 40, nt = 120, L = 3;
  std::vector<double> Ro((size_t) 3 * No), G = {-1.0, 0.0, 1.0, 0, 0, 0, 0, 0, 0};
  for (idx_t n = 0; n < No; n++) {
    Ro[(size_t) n] = 0.1 * (double) (n % 7) - 0.3;
    Ro[(size_t) (n + No)] = 0.0;
    Ro[(size_t) (n + 2 * No)] = 2.0 + 0.2 * (double) n;
  }
  std::vector<double> delay = {1.0}, focal = {10.0}, apod = {1.0, 1.0, 1.0};
  std::vector<double> h((size_t) (nt * No));
  long long oob = 0;
  for (int level : {1, 2, 3}) {          // ignore, warn, stop (clears `running`)
    int e = dro_dream_arr_rect(Ro.data(), No, 0.5, 2.0, 0.05, 0.05, 0.05, nt, 0, delay.data(),
                               1.0, 1500.0, L, G.data(), 1, focal.data(), 0, 0.0, 0.0,
                               apod.data(), 1, 3, 0.0, h.data(), level, T, &oob);
    std::printf("arr_rect level %d: %d oob %lld\n", level, e, oob);
    e = dro_dream_arr_circ(Ro.data(), No, 0.6, 0.05, 0.05, 0.05, nt, 0, delay.data(), 1.0,
                           1500.0, L, G.data(), 0, focal.data(), 0, 0.0, 0.0, apod.data(), 0, 0,
                           0.0, h.data(), level, T, &oob);
    std::printf("arr_circ level %d: %d oob %lld\n", level, e, oob);
  }
  std::vector<double> Gr = {0.5, 0.6, 1.0, 1.1, 1.5};
  std::vector<double> rf = {0, 0, 0, 0, 0}, ra = {1, 1, 1, 1, 1};
  dro_set_alpha(0.5);                     // attenuated discs
  int e = dro_dream_arr_annu(Ro.data(), No, 0.05, 0.05, 0.05, nt, 0, delay.data(), 1.0, 1500.0,
                             (idx_t) Gr.size(), Gr.data(), 0, rf.data(), ra.data(), 0, 0, 0.0,
                             h.data(), 2, T, &oob);
  dro_set_alpha(0.0);
  std::printf("arr_annu (alpha 0.5): %d oob %lld\n", e, oob);

  // DAS: 6-element TFM, pixels partly out of the A-scans' range.
  const idx_t K = 200, Le = 6, Np = 64;
  std::vector<double> Y((size_t) (K * Le * Le)), Gt((size_t) 3 * Le, 0.0), P((size_t) 3 * Np);
  for (size_t i = 0; i < Y.size(); i++) Y[i] = std::sin(0.1 * (double) i);
  for (idx_t i = 0; i < Le; i++) Gt[(size_t) i] = -1.25 + 0.5 * (double) i;
  for (idx_t n = 0; n < Np; n++) {
    P[(size_t) n] = -2.0 + 0.0625 * (double) n;
    P[(size_t) (n + Np)] = 0.0;
    P[(size_t) (n + 2 * Np)] = 0.5 + 0.1 * (double) n;
  }
  std::vector<double> Im((size_t) Np), dl = {0.1};
  for (int type = 0; type < 4; type++)
    for (int level : {1, 2, 3}) {
      e = dro_das_f64(type, Y.data(), K, P.data(), Np, Gt.data(), Le, Gt.data(), Le, 0.02, 0,
                      dl.data(), 1500.0, Im.data(), level, T);
      std::printf("das f64 type %d level %d: %d\n", type, level, e);
    }
  std::vector<float> Yf(Y.begin(), Y.end()), Gtf(Gt.begin(), Gt.end()), Pf(P.begin(), P.end());
  std::vector<float> Imf((size_t) Np), dlf = {0.1f};
  e = dro_das_f32(1, Yf.data(), K, Pf.data(), Np, Gtf.data(), Le, Gtf.data(), Le, 0.02f, 0,
                  dlf.data(), 1500.0f, Imf.data(), 2, T);
  std::printf("das f32: %d\n", e);
  std::vector<double> g((size_t) Np, 1.0), gY(Y.size()), gat((size_t) Le), gar((size_t) Le);
  std::vector<double> at((size_t) Le, 1.0);
  double gcp = 0.0;
  for (int type = 0; type < 4; type++) {
    e = dro_das_ext_f64(type, 1, Y.data(), K, P.data(), Np, Gt.data(), Le, Gt.data(), Le, 0.02,
                        0.1, 1500.0, at.data(), at.data(), Im.data(), T);
    dro_das_ext_adj_f64(type, 1, Y.data(), K, P.data(), Np, Gt.data(), Le, Gt.data(), Le, 0.02,
                        0.1, 1500.0, at.data(), at.data(), g.data(), gY.data(), gat.data(),
                        gar.data(), &gcp, T);
    std::printf("das ext/adjoint type %d: oob %d gcp %.6g\n", type, e, gcp);
  }
  std::printf("sanitize ok\n");
  return 0;
}
