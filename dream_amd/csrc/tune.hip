// tune.hip -- launch-shape autotuner (see tune.hpp).

#include "tune.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <vector>

namespace dream {

namespace {

std::mutex g_mu;
std::map<std::string, int> g_cache;
bool g_loaded = false;
int g_enabled = -1;          // -1: not yet read from the environment

std::string arch_name()
{
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return "unknown";
  std::string name = prop.gcnArchName;
  const size_t colon = name.find(':');
  return colon == std::string::npos ? name : name.substr(0, colon);
}

const char *cache_file() { return std::getenv("DREAM_HIP_TUNE_CACHE"); }

void load_file_locked()
{
  if (g_loaded) return;
  g_loaded = true;
  const char *path = cache_file();
  if (!path) return;
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    const size_t sp = line.rfind(' ');
    if (sp == std::string::npos) continue;
    g_cache[line.substr(0, sp)] = std::atoi(line.c_str() + sp + 1);
  }
}

void append_file_locked(const std::string &key, int best)
{
  const char *path = cache_file();
  if (!path) return;
  std::ofstream out(path, std::ios::app);
  out << key << ' ' << best << '\n';
}

} // namespace

bool autotune_enabled()
{
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_enabled < 0) {
    const char *e = std::getenv("DREAM_HIP_AUTOTUNE");
    g_enabled = (e && e[0] == '0') ? 0 : 1;
  }
  return g_enabled != 0;
}

void set_autotune(bool on)
{
  std::lock_guard<std::mutex> lock(g_mu);
  g_enabled = on ? 1 : 0;
}

int autotune_cache_size()
{
  std::lock_guard<std::mutex> lock(g_mu);
  return (int) g_cache.size();
}

void autotune_clear()
{
  std::lock_guard<std::mutex> lock(g_mu);
  g_cache.clear();
}

namespace {
thread_local bool t_suspended = false;
} // namespace

SuspendAutotune::SuspendAutotune() : prev_(t_suspended) { t_suspended = true; }
SuspendAutotune::~SuspendAutotune() { t_suspended = prev_; }
bool autotune_suspended() { return t_suspended; }

int autotune(const std::string &key, int n, const std::function<void(int, hipStream_t)> &run,
             hipStream_t s)
{
  if (n <= 1) return n == 1 ? 0 : -1;
  if (!autotune_enabled() || t_suspended) return -1;
  const std::string full = arch_name() + "|" + key;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    load_file_locked();
    auto it = g_cache.find(full);
    if (it != g_cache.end() && it->second >= 0 && it->second < n) return it->second;
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return -1;

  // Short launches are timed in kRounds interleaved rounds of enough
  // back-to-back launches to last about kRoundMs, and each candidate keeps its fastest round, so clock
  // ramps and other transients early in the sweep do not decide the pick.  Long launches (tens of ms, where a first
  // launch's set-up is noise) are timed once, and not at all past 1.5 times
  // the best long time so far.
  constexpr int kRounds = 3, kMinReps = 5, kMaxReps = 64;
  constexpr float kRoundMs = 1.0f, kLongMs = 20.0f;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  DREAM_HIP_CHECK(hipEventCreate(&e0));
  DREAM_HIP_CHECK(hipEventCreate(&e1));
  auto timed = [&](int i, int reps) {
    DREAM_HIP_CHECK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++) run(i, s);
    DREAM_HIP_CHECK(hipEventRecord(e1, s));
    DREAM_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    DREAM_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    return ms / (float) reps;
  };
  const float kNone = -1.0f;
  std::vector<float> ms(n, kNone);              // per launch; kNone: not measured
  std::vector<int> reps(n, 0);                  // > 0: short candidate
  int best = 0;
  try {
    float best_long = kNone;
    for (int i = 0; i < n; i++) {
      const float warm = timed(i, 1);           // warm-up
      if (warm > kLongMs) {
        if (best_long >= 0.0f && warm > 1.5f * best_long) continue;
        ms[i] = timed(i, 1);
        if (best_long < 0.0f || ms[i] < best_long) best_long = ms[i];
      } else {
        reps[i] = std::min(kMaxReps, std::max(kMinReps, (int) (kRoundMs / std::max(warm, 1e-3f))));
      }
    }
    for (int round = 0; round < kRounds; round++)
      for (int i = 0; i < n; i++)
        if (reps[i] > 0) {
          const float t = timed(i, reps[i]);
          if (ms[i] < 0.0f || t < ms[i]) ms[i] = t;
        }
    for (int i = 0; i < n; i++)
      if (ms[i] >= 0.0f && (ms[best] < 0.0f || ms[i] < ms[best])) best = i;
  } catch (...) {
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    throw;
  }
  (void) hipEventDestroy(e0);
  (void) hipEventDestroy(e1);
  std::lock_guard<std::mutex> lock(g_mu);
  g_cache[full] = best;
  append_file_locked(full, best);
  return best;
}

} // namespace dream
