// Live kernel timing: HIP events recorded on the launch stream around every
// launch of one watched kernel (identified by its launch-site id), summed on
// demand.  Used by bench.py for the achieved-bandwidth figure of the dominant
// kernel; off (no events) unless a kernel is watched.
#include <vector>

#include "../../include/becgpp_hip.h"
#include "bec_internal.h"

namespace {
struct EvPair {
  hipEvent_t a, b;
};
int g_watch = 0;
std::vector<EvPair> g_pool;
size_t g_used = 0;
double g_ms = 0.0;
long long g_cnt = 0;

void collect() {
  for (size_t i = 0; i < g_used; ++i) {
    float ms = 0.f;
    if (hipEventSynchronize(g_pool[i].b) == hipSuccess &&
        hipEventElapsedTime(&ms, g_pool[i].a, g_pool[i].b) == hipSuccess) {
      g_ms += ms;
      ++g_cnt;
    }
  }
  g_used = 0;
}
}  // namespace

void prof_pre(int id, hipStream_t s) {
  if (g_watch == 0 || id != g_watch) return;
  if (g_used == g_pool.size()) {
    if (g_pool.size() < 4096) {
      EvPair e;
      if (hipEventCreate(&e.a) != hipSuccess || hipEventCreate(&e.b) != hipSuccess) return;
      g_pool.push_back(e);
    } else {
      collect();
    }
  }
  (void)hipEventRecord(g_pool[g_used].a, s);
}

bool prof_events(int id, hipEvent_t* start, hipEvent_t* stop) {
  if (g_watch == 0 || id != g_watch) return false;
  if (g_used == g_pool.size()) {
    if (g_pool.size() < 4096) {
      EvPair e;
      if (hipEventCreate(&e.a) != hipSuccess || hipEventCreate(&e.b) != hipSuccess) return false;
      g_pool.push_back(e);
    } else {
      collect();
    }
  }
  *start = g_pool[g_used].a;
  *stop = g_pool[g_used].b;
  ++g_used;
  return true;
}

void prof_post(int id, hipStream_t s) {
  if (g_watch == 0 || id != g_watch) return;
  (void)hipEventRecord(g_pool[g_used].b, s);
  ++g_used;
}

extern "C" int bec_prof_watch(int id) {
  collect();
  g_watch = id;
  g_ms = 0.0;
  g_cnt = 0;
  return 0;
}

extern "C" int bec_prof_read(double* total_ms, long long* count) {
  collect();
  if (total_ms) *total_ms = g_ms;
  if (count) *count = g_cnt;
  return 0;
}
