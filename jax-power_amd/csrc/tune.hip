// The autotuner's table (host side) and its C entry points: export as JSON, import entry by
// entry, force a knob, enable or disable measurement.
#include "tune.h"

#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

namespace mp {

struct TuneEntry {
  int knob = 0;
  int choice = 0;
  int ncand = 0;
  float ms[TUNE_MAX_CAND] = {-1.f, -1.f, -1.f, -1.f};
  bool measured = false;  // false: imported
};

static std::mutex g_tune_mu;
static std::map<std::string, TuneEntry> g_tune;
static int g_tune_forced[TK_COUNT] = {-1, -1, -1, -1};
static int g_tune_enabled = 1;

static const char* knob_name(int knob) {
  switch (knob) {
    case TK_PAINT_ROWS: return "paint_rows";
    case TK_XPASS_COLS: return "xpass_cols";
    case TK_YPASS_COLS: return "ypass_cols";
    case TK_SORT_CHUNK: return "sort_chunk";
    default: return "unknown";
  }
}

int tune_get(const std::string& key) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto it = g_tune.find(key);
  return it == g_tune.end() ? -1 : it->second.choice;
}

void tune_put(const std::string& key, int knob, int choice, const float* ms, int ncand) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  TuneEntry e;
  e.knob = knob;
  e.choice = choice;
  e.ncand = ncand < TUNE_MAX_CAND ? ncand : TUNE_MAX_CAND;
  for (int q = 0; q < e.ncand; ++q) e.ms[q] = ms ? ms[q] : -1.f;
  e.measured = ms != nullptr;
  g_tune[key] = e;
}

int tune_forced(int knob) { return knob >= 0 && knob < TK_COUNT ? g_tune_forced[knob] : -1; }

bool tune_measure_ok(hipStream_t s) {
  if (!g_tune_enabled || timing_enabled()) return false;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cap) != hipSuccess) return false;
  return cap == hipStreamCaptureStatusNone;
}

// Legacy switch of the paint tuner: 0 off, 1 on, 100 + q forces row-block candidate q.
int paint_tuning_option(int enable) {
  const int old = g_tune_forced[TK_PAINT_ROWS] >= 0 ? 100 + g_tune_forced[TK_PAINT_ROWS] : g_tune_enabled;
  if (enable >= 100) {
    g_tune_forced[TK_PAINT_ROWS] = enable - 100;
  } else if (enable >= 0) {
    g_tune_forced[TK_PAINT_ROWS] = -1;
    g_tune_enabled = enable ? 1 : 0;
  }
  return old;
}

static std::string json_escape(const std::string& s) {
  std::string out;
  for (char c : s) {
    if (c == '"' || c == '\\') out += '\\';
    out += c;
  }
  return out;
}

static int copy_out(const std::string& text, char* buf, int64_t len) {
  if (buf && len > 0) {
    const size_t n = text.size() < (size_t)(len - 1) ? text.size() : (size_t)(len - 1);
    memcpy(buf, text.data(), n);
    buf[n] = 0;
  }
  return (int)text.size();
}

}  // namespace mp

using namespace mp;

// JSON object {"<key>": {"knob": "...", "choice": q, "ms": [...], "measured": true}, ...};
// returns the length of the full text (the buffer holds at most buflen - 1 characters of it).
MP_EXPORT int mp_tune_export(char* buf, int64_t buflen) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  std::string text = "{";
  char num[64];
  bool first = true;
  for (const auto& kv : g_tune) {
    const TuneEntry& e = kv.second;
    if (!first) text += ", ";
    first = false;
    text += "\"" + json_escape(kv.first) + "\": {\"knob\": \"" + knob_name(e.knob) + "\", \"choice\": ";
    snprintf(num, sizeof(num), "%d", e.choice);
    text += num;
    text += ", \"ms\": [";
    for (int q = 0; q < e.ncand; ++q) {
      snprintf(num, sizeof(num), q ? ", %.5f" : "%.5f", (double)e.ms[q]);
      text += num;
    }
    text += e.measured ? "], \"measured\": true}" : "], \"measured\": false}";
  }
  text += "}";
  return copy_out(text, buf, buflen);
}

// One entry of an exported table (its key and choice); knob as in mp_tune_force.
MP_EXPORT int mp_tune_set(const char* key, int knob, int choice) {
  if (!key || knob < 0 || knob >= TK_COUNT || choice < 0 || choice >= TUNE_MAX_CAND)
    return set_error(MP_ERR_INVALID, "mp_tune_set: invalid key, knob %d or choice %d", knob, choice);
  tune_put(key, knob, choice, nullptr, 0);
  return MP_OK;
}

MP_EXPORT int mp_tune_clear(void) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune.clear();
  return MP_OK;
}

// Forces candidate `choice` of a knob (-1 frees it); returns the previous forced candidate.
MP_EXPORT int mp_tune_force(int knob, int choice) {
  if (knob < 0 || knob >= TK_COUNT) return set_error(MP_ERR_INVALID, "mp_tune_force: knob %d not in 0..%d", knob, TK_COUNT - 1);
  const int old = g_tune_forced[knob];
  g_tune_forced[knob] = choice < 0 ? -1 : choice;
  return old;
}

// Measurement on (1) or off (0: cases not in the table take candidate 0); returns the previous
// setting, a negative argument only queries.
MP_EXPORT int mp_tune_enable(int enable) {
  const int old = g_tune_enabled;
  if (enable >= 0) g_tune_enabled = enable ? 1 : 0;
  return old;
}

// One line per entry: "<key> choice=q ms: t0 t1 ..." (the old paint-tuning report).
MP_EXPORT int mp_paint_tuning_report(char* buf, int64_t buflen) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  std::string text;
  char num[64];
  for (const auto& kv : g_tune) {
    const TuneEntry& e = kv.second;
    snprintf(num, sizeof(num), " choice=%d ms:", e.choice);
    text += kv.first + num;
    for (int q = 0; q < e.ncand; ++q) {
      snprintf(num, sizeof(num), " %.4f", (double)e.ms[q]);
      text += num;
    }
    text += "\n";
  }
  copy_out(text, buf, buflen);
  return (int)g_tune.size();
}

MP_EXPORT int mp_set_paint_tuning(int enable) { return paint_tuning_option(enable); }
