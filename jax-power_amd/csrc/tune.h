// The autotuner's table of choices.  Each tunable launch (knob) has a few candidate
// configurations; on the first launch of a case (a key naming the knob and the shapes that
// matter) every candidate is timed on the actual data with HIP events and the fastest is kept
// for the process.  Entries can be exported as JSON and imported again (a persistent cache),
// and a knob can be forced to one candidate (tests of every configuration).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "common.h"

namespace mp {

enum TuneKnob : int {
  TK_PAINT_ROWS = 0,  // row-block shape of the fused paint
  TK_XPASS_COLS = 1,  // columns per tile of the binning x pass
  TK_YPASS_COLS = 2,  // columns per tile of the y pass
  TK_SORT_CHUNK = 3,  // particles per workgroup of the row-sort scatter
  TK_COUNT = 4,
};
constexpr int TUNE_MAX_CAND = 4;

// Candidate chosen for `key`, or -1 when the case has not been tuned or imported.
int tune_get(const std::string& key);
// Records the choice for `key` (ms: the measured time of each candidate, < 0 where not run).
void tune_put(const std::string& key, int knob, int choice, const float* ms, int ncand);
// Candidate forced for a knob (tests), or -1.
int tune_forced(int knob);
// Whether a measurement may run now on `s`: tuning enabled, the stream not being captured into
// a graph, per-kernel event timing off.
bool tune_measure_ok(hipStream_t s);

// Times run(q) for every candidate q with ok[q] (two launches each, the faster kept) and returns
// the fastest candidate, or -1 on a HIP error (the error is set).  ms[q] < 0 for skipped ones.
template <typename F>
int tune_time(hipStream_t s, int ncand, const bool* ok, F run, float* ms) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) {
    set_error(MP_ERR_HIP, "tuner: hipEventCreate failed");
    return -1;
  }
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    set_error(MP_ERR_HIP, "tuner: hipEventCreate failed");
    return -1;
  }
  int best = -1;
  float tbest = 1e30f;
  for (int q = 0; q < ncand; ++q) {
    ms[q] = -1.f;
    if (!ok[q]) continue;
    float tmin = 1e30f;
    for (int rep = 0; rep < 2; ++rep) {
      if (hipEventRecord(e0, s) != hipSuccess || run(q) != MP_OK || hipEventRecord(e1, s) != hipSuccess ||
          hipEventSynchronize(e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return -1;
      }
      float t = 0.f;
      (void)hipEventElapsedTime(&t, e0, e1);
      if (t < tmin) tmin = t;
    }
    ms[q] = tmin;
    if (tmin < tbest) {
      tbest = tmin;
      best = q;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

}  // namespace mp
