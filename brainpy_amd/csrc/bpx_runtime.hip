// bpx_runtime.hip -- status strings, the kernel timer and the copy kernel.
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

// A plain streaming copy: 16-byte vector loads and stores, grid-stride.
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n16; i += stride) dst[i] = src[i];
}

__global__ void k_copy1(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
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

int bpx_copy(void* stream, const void* src, void* dst, int64_t bytes) {
  if (bytes < 0) return BPX_ERR_ARG;
  if (bytes == 0) return BPX_OK;
  hipStream_t s = (hipStream_t)stream;
  bool aligned = (((uintptr_t)src | (uintptr_t)dst | (uintptr_t)bytes) & 15) == 0;
  if (aligned) {
    int64_t n16 = bytes / 16;
    int64_t blocks = (n16 + 255) / 256;
    if (blocks > 256 * 32) blocks = 256 * 32;
    BPX_LAUNCH(bpx::KID_COPY, bpx::k_copy16, dim3((unsigned)blocks), dim3(256), 0, s,
               (const uint4*)src, (uint4*)dst, n16);
  } else {
    int64_t blocks = (bytes + 255) / 256;
    hipLaunchKernelGGL(bpx::k_copy1, dim3((unsigned)blocks), dim3(256), 0, s, (const uint8_t*)src, (uint8_t*)dst, bytes);
  }
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

}  // extern "C"
