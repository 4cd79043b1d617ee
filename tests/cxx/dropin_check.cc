// Consumer of include/dream_hip.hpp as a reference gateway would be: the
// reference's dream.h first, then the drop-in classes, no other sources.
//
// Without arguments it only constructs a DAS<double> (no device needed).
// With argument "run" it computes on the GPU through every drop-in entry
// point -- DAS<double>/DAS<float>::das for the four methods, cl_das,
// ArrRect::dream_arr_rect, Rect::dreamrect, Rect::cl_dreamrect, a `stop`
// case that clears is_running(), and a call cancelled between two work
// chunks -- and prints the results, one record per line, for the test to
// compare with the oracle.  Inputs come from exact rational formulas the
// test restates.
#include <cstdio>
#include <cstring>
#include <vector>

#include "dream.h"
#include "dream_hip.hpp"

namespace {

// Y[k + K a] = ((k + 3 a) * 37 mod 101 - 50) / 8: exact in float and double.
template <class T>
std::vector<T> data(int K, int n_ascans)
{
  std::vector<T> Y((size_t) K * n_ascans);
  for (int a = 0; a < n_ascans; a++)
    for (int k = 0; k < K; k++) Y[(size_t) a * K + k] = (T) ((((k + 3 * a) * 37) % 101) - 50) / (T) 8;
  return Y;
}

// L elements along x (dim 0) or y (dim 1) at 0.5 mm pitch, centred.
template <class T>
std::vector<T> array(int L, int dim)
{
  std::vector<T> G((size_t) 3 * L, (T) 0);
  for (int i = 0; i < L; i++) G[(size_t) dim * L + i] = ((T) i - (T) (L - 1) / (T) 2) * (T) 0.5;
  return G;
}

// No = 64 pixels on an 8 x 8 x-z grid with a y offset, column-major [x | y | z].
template <class T>
std::vector<T> pixels()
{
  const int No = 64;
  std::vector<T> R((size_t) 3 * No);
  for (int n = 0; n < No; n++) {
    R[n] = ((T) (n % 8) - (T) 3.5) * (T) 0.375;
    R[n + No] = (T) ((n / 8) % 3) * (T) 0.25 - (T) 0.25;
    R[n + 2 * No] = (T) 2 + (T) (n / 8) * (T) 0.5;
  }
  return R;
}

const char *kMethods[4] = {"saft", "tfm", "rca_coltx", "rca_rowtx"};
const DASType kTypes[4] = {DASType::saft, DASType::tfm, DASType::rca_coltx, DASType::rca_rowtx};

template <class T>
void das_all(const char *tag, const char *fmt)
{
  const int K = 256, L = 4, No = 64;
  const std::vector<T> Ro = pixels<T>();
  for (int m = 0; m < 4; m++) {
    const bool saft = kTypes[m] == DASType::saft;
    const std::vector<T> Gt = array<T>(L, 0);
    const std::vector<T> Gr = array<T>(L, kTypes[m] == DASType::tfm ? 0 : 1);
    const std::vector<T> Y = data<T>(K, saft ? L : L * L);
    std::vector<T> Im((size_t) No);
    const T delay = (T) 0.125;
    DAS<T> d(kTypes[m], K, No, L, saft ? 0 : L);
    const SIRError e = d.das(Y.data(), Ro.data(), Gt.data(), saft ? nullptr : Gr.data(), (T) 0.02,
                             DelayType::single, &delay, (T) 1500, Im.data(), ErrorLevel::ignore);
    std::printf("%s %s %d", tag, kMethods[m], (int) e);
    for (T v : Im) std::printf(fmt, (double) v);
    std::printf("\n");
    if (kTypes[m] == DASType::tfm) {
      std::vector<T> Im2((size_t) No);
      const int rc = d.cl_das(Y.data(), Ro.data(), Gt.data(), Gr.data(), (T) 0.02, delay, (T) 1500,
                              Im2.data());
      std::printf("cl_%s tfm %d", tag, rc);
      for (T v : Im2) std::printf(fmt, (double) v);
      std::printf("\n");
    }
  }
}

void print_h(const char *tag, int code, const std::vector<double> &h)
{
  std::printf("%s %d", tag, code);
  for (double v : h) std::printf(" %.17g", v);
  std::printf("\n");
}

} // namespace

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
  das_all<double>("das64", " %.17g");
  das_all<float>("das32", " %.9g");

  // SIR: three 0.5 x 2 mm elements at 1 mm pitch, x focusing at 20 mm, Hann
  // apodization; 3 on-axis-ish points; nt = 200, dt = 0.025 us.
  const int No = 3, nt = 200, L = 3;
  std::vector<double> Ro = {0.0, 0.5, -0.75, 0.0, 0.25, 0.0, 3.0, 3.5, 4.0};
  std::vector<double> G = {-1.0, 0.0, 1.0, 0, 0, 0, 0, 0, 0};
  std::vector<double> focal = {20.0}, apod = {1.0, 1.0, 1.0};
  std::vector<double> dl = {1.5};
  std::vector<double> h((size_t) nt * No, 0.0);
  ArrRect ar;
  SIRError e = ar.dream_arr_rect(0.0, Ro.data(), No, 0.5, 2.0, 0.05, 0.05, 0.025, nt,
                                 DelayType::single, dl.data(), 1.0, 1500.0, L, G.data(),
                                 FocusMet::x, focal.data(), SteerMet::none, 0.0, 0.0, apod.data(),
                                 true, ApodMet::hann, 0.0, h.data(), ErrorLevel::stop);
  print_h("arr_rect", (int) e, h);
  std::printf("arr_rect_running %d\n", (int) ar.is_running());

  Rect r;
  std::vector<double> h1((size_t) nt * No, 0.0);
  e = r.dreamrect(0.0, Ro.data(), No, 0.5, 2.0, 0.05, 0.05, 0.025, nt, DelayType::single,
                  dl.data(), 1.0, 1500.0, h1.data(), ErrorLevel::stop);
  print_h("rect", (int) e, h1);
  std::vector<double> h2((size_t) nt * No, 0.0);
  const int rc = r.cl_dreamrect(Ro.data(), No, 0.5, 2.0, 0.05, 0.05, 0.025, nt, 1.5, 1.0, 1500.0,
                                h2.data());
  print_h("cl_rect", rc, h2);

  // `stop` with samples past nt: SIRError::out_of_bounds and is_running()
  // false, as the reference's threads leave it; the next call sets it again.
  std::vector<double> h3((size_t) nt * No, 0.0);
  std::vector<double> late = {-10.0};
  e = ar.dream_arr_rect(0.0, Ro.data(), No, 0.5, 2.0, 0.05, 0.05, 0.025, nt, DelayType::single,
                        late.data(), 1.0, 1500.0, L, G.data(), FocusMet::none, focal.data(),
                        SteerMet::none, 0.0, 0.0, apod.data(), false, ApodMet::hann, 0.0,
                        h3.data(), ErrorLevel::stop);
  std::printf("stop %d %d\n", (int) e, (int) ar.is_running());
  e = ar.dream_arr_rect(0.0, Ro.data(), No, 0.5, 2.0, 0.05, 0.05, 0.025, nt, DelayType::single,
                        dl.data(), 1.0, 1500.0, L, G.data(), FocusMet::none, focal.data(),
                        SteerMet::none, 0.0, 0.0, apod.data(), false, ApodMet::hann, 0.0,
                        h3.data(), ErrorLevel::stop);
  std::printf("after_stop %d %d\n", (int) e, (int) ar.is_running());

  // A call cancelled between two work chunks (the library's test hook cuts
  // it into one-tile chunks and aborts after the first): das returns, the
  // running flag is cleared, and the gateway's is_running() test sees it.
  {
    const int K = 256, Lt = 4;
    const std::vector<double> Gt = array<double>(Lt, 0);
    const std::vector<double> Y = data<double>(K, Lt * Lt);
    const std::vector<double> R = pixels<double>();
    std::vector<double> R4((size_t) 3 * 256);
    for (int c = 0; c < 3; c++)
      for (int n = 0; n < 256; n++) R4[(size_t) c * 256 + n] = R[(size_t) c * 64 + n % 64];
    std::vector<double> Im(256, -1.0);
    const double delay = 0.0;
    DAS<double> d(DASType::tfm, K, 256, Lt, Lt);
    d.set_running();
    dream_hip_test_chunking(64 * Lt * Lt, 1);
    const SIRError ec = d.das(Y.data(), R4.data(), Gt.data(), Gt.data(), 0.02, DelayType::single,
                              &delay, 1500.0, Im.data(), ErrorLevel::ignore);
    dream_hip_test_chunking(0, 0);
    std::printf("cancel %d %d\n", (int) ec, (int) d.is_running());
    d.set_running();
    std::printf("set_running %d\n", (int) d.is_running());
  }
  return 0;
}
