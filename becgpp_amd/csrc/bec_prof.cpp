// Live kernel timing: HIP events recorded on the launch stream around every
// launch of one watched kernel (identified by its launch-site id), summed on
// demand together with the algorithmic bytes each launch moved.  Used by
// bench.py for the achieved-bandwidth figure of the dominant kernel; off (no
// events) unless a kernel is watched.
#include <vector>

#include "../../include/becgpp_hip.h"
#include "bec_internal.h"

namespace {
struct EvPair {
  hipEvent_t a, b;
  double bytes;
};
int g_watch = 0;
unsigned long long* g_dcount = nullptr;    // device: instances covered by watched device-driven launches
std::vector<EvPair> g_pool;
size_t g_used = 0;
double g_ms = 0.0;
long long g_cnt = 0;
double g_bytes = 0.0;

void collect() {
  for (size_t i = 0; i < g_used; ++i) {
    float ms = 0.f;
    if (hipEventSynchronize(g_pool[i].b) == hipSuccess &&
        hipEventElapsedTime(&ms, g_pool[i].a, g_pool[i].b) == hipSuccess) {
      g_ms += ms;
      ++g_cnt;
      g_bytes += g_pool[i].bytes;
    }
  }
  g_used = 0;
}

EvPair* next_pair() {
  if (g_used == g_pool.size()) {
    if (g_pool.size() < 4096) {
      EvPair e{};
      if (hipEventCreate(&e.a) != hipSuccess || hipEventCreate(&e.b) != hipSuccess) return nullptr;
      g_pool.push_back(e);
    } else {
      collect();
    }
  }
  return &g_pool[g_used];
}
}  // namespace

void prof_pre(int id, hipStream_t s, double bytes) {
  if (g_watch == 0 || id != g_watch) return;
  EvPair* e = next_pair();
  if (!e) return;
  e->bytes = bytes;
  (void)hipEventRecord(e->a, s);
}

bool prof_events(int id, hipEvent_t* start, hipEvent_t* stop, double bytes) {
  if (g_watch == 0 || id != g_watch) return false;
  EvPair* e = next_pair();
  if (!e) return false;
  e->bytes = bytes;
  *start = e->a;
  *stop = e->b;
  ++g_used;
  return true;
}

void prof_post(int id, hipStream_t s) {
  if (g_watch == 0 || id != g_watch || g_used == g_pool.size()) return;
  (void)hipEventRecord(g_pool[g_used].b, s);
  ++g_used;
}

bool prof_active() { return g_watch != 0; }

unsigned long long* prof_counter(int id) {
  if (g_watch == 0 || id != g_watch) return nullptr;
  return g_dcount;
}

extern "C" int bec_prof_watch(int id) {
  collect();
  g_watch = id;
  g_ms = 0.0;
  g_cnt = 0;
  g_bytes = 0.0;
  if (id != 0) {
    if (!g_dcount && hipMalloc(&g_dcount, sizeof(unsigned long long)) != hipSuccess) {
      g_dcount = nullptr;
      return 2;
    }
    if (hipMemset(g_dcount, 0, sizeof(unsigned long long)) != hipSuccess) return 2;
  }
  return 0;
}

extern "C" int bec_prof_read(double* total_ms, long long* count, double* bytes) {
  collect();
  unsigned long long dev = 0;
  if (g_dcount && hipMemcpy(&dev, g_dcount, sizeof(dev), hipMemcpyDeviceToHost) != hipSuccess) return 2;
  if (total_ms) *total_ms = g_ms;
  if (count) *count = g_cnt;
  if (bytes) *bytes = g_bytes + (double)dev;
  return 0;
}
