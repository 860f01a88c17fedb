// Error reporting and kernel timing for the C ABI (host side).
#include "common.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace mp {

static thread_local char g_last_error[1024] = "";

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

struct TimingRecord {
  std::string name;
  hipEvent_t start, stop;
};

struct TimingStats {
  double total_ms = 0.;
  int64_t count = 0;
};

static std::mutex g_timing_mutex;
static bool g_timing = false;
static std::vector<TimingRecord*> g_pending;
static std::vector<hipEvent_t> g_event_pool;
static std::map<std::string, TimingStats> g_stats;

bool timing_enabled() { return g_timing; }

static hipEvent_t pool_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void timing_begin(const char* name, hipStream_t stream, void** token) {
  std::lock_guard<std::mutex> lock(g_timing_mutex);
  TimingRecord* rec = new TimingRecord{name, pool_event(), pool_event()};
  if (!rec->start || !rec->stop) {
    delete rec;
    *token = nullptr;
    return;
  }
  (void)hipEventRecord(rec->start, stream);
  *token = rec;
}

void timing_end(void* token, hipStream_t stream) {
  TimingRecord* rec = static_cast<TimingRecord*>(token);
  (void)hipEventRecord(rec->stop, stream);
  std::lock_guard<std::mutex> lock(g_timing_mutex);
  g_pending.push_back(rec);
}

}  // namespace mp

using namespace mp;

MP_EXPORT const char* mp_last_error(void) { return g_last_error; }

MP_EXPORT int mp_version(void) { return 1; }

MP_EXPORT int mp_timing_enable(int enable) {
  std::lock_guard<std::mutex> lock(g_timing_mutex);
  g_timing = enable != 0;
  return MP_OK;
}

MP_EXPORT int mp_timing_collect(void) {
  std::lock_guard<std::mutex> lock(g_timing_mutex);
  for (TimingRecord* rec : g_pending) {
    hipError_t e = hipEventSynchronize(rec->stop);
    if (e != hipSuccess) return set_error(MP_ERR_HIP, "hipEventSynchronize: %s", hipGetErrorString(e));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, rec->start, rec->stop);
    TimingStats& s = g_stats[rec->name];
    s.total_ms += ms;
    s.count += 1;
    g_event_pool.push_back(rec->start);
    g_event_pool.push_back(rec->stop);
    delete rec;
  }
  g_pending.clear();
  return MP_OK;
}

MP_EXPORT int mp_timing_reset(void) {
  int rc = mp_timing_collect();
  std::lock_guard<std::mutex> lock(g_timing_mutex);
  g_stats.clear();
  return rc;
}

// Writes "name total_ms count\n" lines into buf (truncated to buflen); returns the number of
// kernels recorded.
MP_EXPORT int mp_timing_report(char* buf, int64_t buflen) {
  int rc = mp_timing_collect();
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(g_timing_mutex);
  std::string out;
  for (auto& kv : g_stats) {
    char line[256];
    snprintf(line, sizeof(line), "%s %.6f %lld\n", kv.first.c_str(), kv.second.total_ms, (long long)kv.second.count);
    out += line;
  }
  if (buf && buflen > 0) {
    size_t n = out.size() < (size_t)(buflen - 1) ? out.size() : (size_t)(buflen - 1);
    memcpy(buf, out.data(), n);
    buf[n] = 0;
  }
  return (int)g_stats.size();
}
