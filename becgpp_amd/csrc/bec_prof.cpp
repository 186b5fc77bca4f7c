// Live kernel timing: HIP events recorded on the launch stream around every
// launch of the watched kernels (identified by their launch-site ids), summed
// per site on demand together with the algorithmic bytes each launch moved.
// Watching one site times that kernel alone; watching every site (id -1)
// gives the breakdown of a whole solver step, from which bench.py takes the
// kernel with the largest share of the step time.  Off (no events, and the
// solver loop replays hipGraphs) unless something is watched.
#include <map>
#include <vector>

#include "../../include/becgpp_hip.h"
#include "bec_internal.h"

namespace {
struct EvPair {
  hipEvent_t a, b;
  double bytes;
  int site;
};
struct SiteSum {
  double ms = 0.0;
  long long cnt = 0;
  double bytes = 0.0;
  int slot = -1;                     // index of its device byte counter
};
constexpr int kMaxSlots = 256;
int g_watch = 0;                       // 0 off, -1 every site, else one site id
unsigned long long* g_dcount = nullptr;    // device: algorithmic bytes of device-driven launches, per slot
std::vector<EvPair> g_pool;
size_t g_used = 0;
std::map<int, SiteSum> g_sites;
int g_nslots = 0;

bool watched(int id) { return g_watch == -1 || (g_watch != 0 && id == g_watch); }

void collect() {
  for (size_t i = 0; i < g_used; ++i) {
    float ms = 0.f;
    if (hipEventSynchronize(g_pool[i].b) == hipSuccess &&
        hipEventElapsedTime(&ms, g_pool[i].a, g_pool[i].b) == hipSuccess) {
      SiteSum& s = g_sites[g_pool[i].site];
      s.ms += ms;
      ++s.cnt;
      s.bytes += g_pool[i].bytes;
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
  if (!watched(id)) return;
  EvPair* e = next_pair();
  if (!e) return;
  e->bytes = bytes;
  e->site = id;
  (void)hipEventRecord(e->a, s);
}

bool prof_events(int id, hipEvent_t* start, hipEvent_t* stop, double bytes) {
  if (!watched(id)) return false;
  EvPair* e = next_pair();
  if (!e) return false;
  e->bytes = bytes;
  e->site = id;
  *start = e->a;
  *stop = e->b;
  ++g_used;
  return true;
}

void prof_post(int id, hipStream_t s) {
  if (!watched(id) || g_used == g_pool.size()) return;
  (void)hipEventRecord(g_pool[g_used].b, s);
  ++g_used;
}

bool prof_active() { return g_watch != 0; }

unsigned long long* prof_counter(int id) {
  if (!watched(id) || !g_dcount) return nullptr;
  SiteSum& s = g_sites[id];
  if (s.slot < 0) {
    if (g_nslots == kMaxSlots) return nullptr;
    s.slot = g_nslots++;
  }
  return g_dcount + s.slot;
}

extern "C" int bec_prof_watch(int id) {
  collect();
  g_watch = id;
  g_sites.clear();
  g_nslots = 0;
  if (id != 0) {
    if (!g_dcount && hipMalloc(&g_dcount, sizeof(unsigned long long) * kMaxSlots) != hipSuccess) {
      g_dcount = nullptr;
      return 2;
    }
    if (hipMemset(g_dcount, 0, sizeof(unsigned long long) * kMaxSlots) != hipSuccess) return 2;
  }
  return 0;
}

extern "C" int bec_prof_read_site(int id, double* total_ms, long long* count, double* bytes) {
  collect();
  double ms = 0.0, b = 0.0;
  long long c = 0;
  std::vector<unsigned long long> dev(kMaxSlots, 0ull);
  if (g_dcount && g_nslots > 0 &&
      hipMemcpy(dev.data(), g_dcount, sizeof(unsigned long long) * g_nslots, hipMemcpyDeviceToHost) != hipSuccess)
    return 2;
  for (const auto& kv : g_sites) {
    if (id != -1 && kv.first != id) continue;
    ms += kv.second.ms;
    c += kv.second.cnt;
    b += kv.second.bytes + (kv.second.slot >= 0 ? (double)dev[kv.second.slot] : 0.0);
  }
  if (total_ms) *total_ms = ms;
  if (count) *count = c;
  if (bytes) *bytes = b;
  return 0;
}

extern "C" int bec_prof_read(double* total_ms, long long* count, double* bytes) {
  return bec_prof_read_site(-1, total_ms, count, bytes);
}

extern "C" int bec_prof_sites(int* ids, int cap, int* n) {
  collect();
  int k = 0;
  for (const auto& kv : g_sites)
    if (kv.second.cnt > 0) {
      if (ids && k < cap) ids[k] = kv.first;
      ++k;
    }
  if (n) *n = k;
  return 0;
}
