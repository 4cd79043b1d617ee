// Consumer of include/dream_hip.hpp as a reference gateway would be: the
// reference's dream.h first, then the drop-in classes, no other sources.
// With argument "run" it beamforms a tiny TFM case on the GPU through
// DAS<double> / DAS<float> and computes a one-element SIR through ArrRect,
// printing the results for the test to compare with the oracle.
#include <cstdio>
#include <cstring>
#include <vector>

#include "dream.h"
#include "dream_hip.hpp"

int main(int argc, char **argv)
{
  const bool run = argc > 1 && std::strcmp(argv[1], "run") == 0;
  if (!run) {
    // No device needed: the API surface compiles, links and reports errors.
    try {
      DAS<double> d(DASType::tfm, 64, 4, 2, 2);
      std::printf("created\n");
    } catch (const std::runtime_error &e) {
      std::printf("no-device: %s\n", e.what());
    }
    return 0;
  }
  const int K = 64, L = 2, No = 3;
  std::vector<double> Y(K * L * L), Gt = {-1.0, 1.0, 0, 0, 0, 0}, Ro = {0, 0.5, -0.5, 0, 0, 0,
                                                                        1.0, 1.2, 1.4};
  for (int i = 0; i < K * L * L; i++) Y[i] = (double) ((i * 37) % 101) - 50.0;
  std::vector<double> Im(No);
  double delay = 0.0;
  DAS<double> d(DASType::tfm, K, No, L, L);
  SIRError e = d.das(Y.data(), Ro.data(), Gt.data(), Gt.data(), 0.02, DelayType::single, &delay,
                     1500.0, Im.data(), ErrorLevel::ignore);
  std::printf("das64 %d %.17g %.17g %.17g\n", (int) e, Im[0], Im[1], Im[2]);
  std::vector<float> Yf(Y.begin(), Y.end()), Gf(Gt.begin(), Gt.end()), Rf(Ro.begin(), Ro.end());
  std::vector<float> Imf(No);
  float delayf = 0.0f;
  DAS<float> df(DASType::tfm, K, No, L, L);
  e = df.das(Yf.data(), Rf.data(), Gf.data(), Gf.data(), 0.02f, DelayType::single, &delayf,
             1500.0f, Imf.data(), ErrorLevel::ignore);
  std::printf("das32 %d %.9g %.9g %.9g\n", (int) e, Imf[0], Imf[1], Imf[2]);
  ArrRect ar;
  std::vector<double> h(200 * No, 0.0), G = {0, 0, 0}, focal = {0}, apod = {1};
  std::vector<double> dl = {0.0};
  e = ar.dream_arr_rect(0.0, Ro.data(), No, 0.5, 2.0, 0.05, 0.05, 0.025, 200, DelayType::single,
                        dl.data(), 1.0, 1500.0, 1, G.data(), FocusMet::none, focal.data(),
                        SteerMet::none, 0.0, 0.0, apod.data(), false, ApodMet::gauss, 0.0,
                        h.data(), ErrorLevel::stop);
  double s = 0;
  for (double v : h) s += v;
  std::printf("arr_rect %d %d %.17g\n", (int) e, (int) ar.is_running(), s);
  return 0;
}
