// bpx_runtime.hip -- status strings and the kernel timer.
#include <mutex>
#include <vector>

#include "bpx_common.h"

namespace bpx {
namespace {

struct Pending {
  int kid;
  hipEvent_t start, stop;
};

std::mutex g_mu;
bool g_enabled = false;
std::vector<Pending> g_pending;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_free;
double g_total_ms[KID_COUNT];
int64_t g_count[KID_COUNT];

// Fold finished event pairs into the per-kernel totals.  Caller holds g_mu.
void drain(bool wait) {
  std::vector<Pending> keep;
  for (auto& p : g_pending) {
    if (wait) hipEventSynchronize(p.stop);
    if (hipEventQuery(p.stop) == hipSuccess) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
        g_total_ms[p.kid] += ms;
        g_count[p.kid] += 1;
      }
      g_free.emplace_back(p.start, p.stop);
    } else {
      keep.push_back(p);
    }
  }
  g_pending.swap(keep);
}

}  // namespace

void timer_acquire(int kid, hipEvent_t* start, hipEvent_t* stop) {
  *start = nullptr;
  *stop = nullptr;
  if (!g_enabled) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_enabled) return;
  if (g_pending.size() > 4096) drain(true);
  hipEvent_t s, e;
  if (!g_free.empty()) {
    s = g_free.back().first;
    e = g_free.back().second;
    g_free.pop_back();
  } else {
    if (hipEventCreate(&s) != hipSuccess) return;
    if (hipEventCreate(&e) != hipSuccess) { hipEventDestroy(s); return; }
  }
  g_pending.push_back({kid, s, e});
  *start = s;
  *stop = e;
}

}  // namespace bpx

extern "C" {

const char* bpx_status_string(int status) {
  switch (status) {
    case BPX_OK: return "ok";
    case BPX_ERR_DTYPE: return "unsupported dtype";
    case BPX_ERR_SHAPE: return "inconsistent shapes";
    case BPX_ERR_WORKSPACE: return "workspace missing or too small";
    case BPX_ERR_LAUNCH: return "kernel launch failed";
    case BPX_ERR_ARG: return "invalid argument";
  }
  return "unknown status";
}

const char* bpx_version(void) { return "bpx 0.1 gfx950"; }

int bpx_timer_enable(int on) {
  std::lock_guard<std::mutex> lk(bpx::g_mu);
  bpx::g_enabled = on != 0;
  return BPX_OK;
}

int bpx_timer_read(int kid, double* total_ms, int64_t* launches) {
  if (kid < 0 || kid >= bpx::KID_COUNT) return BPX_ERR_ARG;
  std::lock_guard<std::mutex> lk(bpx::g_mu);
  bpx::drain(true);
  if (total_ms) *total_ms = bpx::g_total_ms[kid];
  if (launches) *launches = bpx::g_count[kid];
  return BPX_OK;
}

int bpx_timer_reset(void) {
  std::lock_guard<std::mutex> lk(bpx::g_mu);
  bpx::drain(true);
  for (int i = 0; i < bpx::KID_COUNT; ++i) {
    bpx::g_total_ms[i] = 0.0;
    bpx::g_count[i] = 0;
  }
  return BPX_OK;
}

}  // extern "C"
