// tune.hpp -- launch-shape autotuner of the DREAM HIP library.
//
// A kernel with several valid launch shapes registers them as numbered
// candidates.  The first launch of a shape class times every candidate on
// the caller's stream (a warm-up, then interleaved rounds of timed
// launches, HIP events) and keeps the fastest; later launches of the class
// reuse it.
// Results are cached per process, keyed by the device architecture and the
// shape class, and -- when DREAM_HIP_TUNE_CACHE names a file -- read from and
// appended to that file.  Tuning is skipped (the caller's default shape is
// used) while the stream is being captured into a graph, or when autotuning
// is off (DREAM_HIP_AUTOTUNE=0 or dream_hip_set_autotune(0)).
#pragma once

#include <functional>
#include <string>

#include "common.hpp"

namespace dream {

bool autotune_enabled();
void set_autotune(bool on);

// Index of the best of n candidates for shape class `key` (a string the
// caller builds from everything that decides the best shape), or -1 when
// tuning is unavailable.  run(i, s) launches candidate i on stream s; every
// candidate must compute the same result.
int autotune(const std::string &key, int n, const std::function<void(int, hipStream_t)> &run,
             hipStream_t s);

// While one of these lives on a thread, launches on that thread use their
// default shapes: autotune() is not consulted (callers test
// autotune_suspended() before building its arguments), so nothing is timed,
// no events are created and no host memory is taken.  For entry points that
// must not allocate.
struct SuspendAutotune {
  SuspendAutotune();
  ~SuspendAutotune();
  SuspendAutotune(const SuspendAutotune &) = delete;
  SuspendAutotune &operator=(const SuspendAutotune &) = delete;
private:
  bool prev_;
};
bool autotune_suspended();

// Number of cached results (for tests) and removal of all of them.
int autotune_cache_size();
void autotune_clear();

} // namespace dream
