// bpx_common.h -- shared definitions of the gfx950 event-driven synaptic library.
//
// Everything in this library is written for CDNA4 (gfx950): 64-lane
// wavefronts, ballot/mbcnt-based compaction, hardware FP32/FP64 atomics in
// L2 (built with -munsafe-fp-atomics), no MFMA (the path is sparse and
// bandwidth/integer bound).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdint.h>

#include "../../include/bpx.h"

namespace bpx {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// element types
// ---------------------------------------------------------------------------
struct bf16_t { uint16_t bits; };   // raw bfloat16 storage (weights only)

__device__ __forceinline__ float to_f32(float v) { return v; }
__device__ __forceinline__ double to_f32(double v) { return v; }
__device__ __forceinline__ float to_f32(bf16_t v) { return __uint_as_float(((uint32_t)v.bits) << 16); }
__device__ __forceinline__ float to_f32(_Float16 v) { return (float)v; }

// accumulation type of a value type: f64 stays f64, the rest accumulate in f32
template <typename V> struct acc_of { using type = float; };
template <> struct acc_of<double> { using type = double; };

// ---------------------------------------------------------------------------
// events.  Bool events are bytes (truthy if non-zero).  Float events follow
// one of two reference rules: `event > 0.` (Numba kernels,
// event/_csr_matvec.py:349) or `event != 0.` (Taichi kernels,
// event/_csr_matvec_taichi.py:70).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool ev_true(uint8_t v, int) { return v != 0; }
__device__ __forceinline__ bool ev_true(float v, int rule) { return rule == BPX_RULE_POSITIVE ? v > 0.f : v != 0.f; }
__device__ __forceinline__ bool ev_true(double v, int rule) { return rule == BPX_RULE_POSITIVE ? v > 0.0 : v != 0.0; }
__device__ __forceinline__ bool ev_true(int32_t v, int) { return v != 0; }

__device__ __forceinline__ float ev_value(uint8_t v) { return v ? 1.f : 0.f; }
__device__ __forceinline__ float ev_value(float v) { return v; }
__device__ __forceinline__ double ev_value(double v) { return v; }

// ---------------------------------------------------------------------------
// atomics: hardware global_atomic_add_f32/f64 (no return value used)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void atomic_add(float* p, float v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_add(double* p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// number of set bits of `mask` below this lane (order-preserving slot)
__device__ __forceinline__ int lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

template <int WIDTH, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int o = WIDTH / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WIDTH);
  return v;
}

// ---------------------------------------------------------------------------
// kernel timer: when enabled, the dominant kernel of an entry point is
// launched with hipExtLaunchKernelGGL start/stop events, so its device time
// is measured on the stream it runs on (bpx_timer_*).
// ---------------------------------------------------------------------------
void timer_acquire(int kernel_id, hipEvent_t* start, hipEvent_t* stop);

enum KernelId {
  KID_EVENT_SCATTER = 0,
  KID_EVENT_GATHER = 1,
  KID_CSR_SCATTER = 2,
  KID_CSR_GATHER = 3,
  KID_JIT_OUTDIM = 4,
  KID_JIT_SCATTER = 5,
  KID_INFO = 6,
  KID_LIF = 7,
  KID_FIXED_ORDER = 8,
  KID_TILED_LIST = 9,
  KID_NET = 10,
  KID_TILED = 11,
  KID_COO = 12,
  KID_BSR = 13,
  KID_NET_RUN = 14,
  KID_TILED_GATHER = 15,
  KID_JIT_EMIT = 16,
  KID_COUNT = 17
};

}  // namespace bpx

#define BPX_LAUNCH(KID, KERNEL, GRID, BLOCK, SHMEM, STREAM, ...)                              \
  do {                                                                                       \
    hipEvent_t _bpx_s = nullptr, _bpx_e = nullptr;                                           \
    ::bpx::timer_acquire((KID), &_bpx_s, &_bpx_e);                                           \
    if (_bpx_s)                                                                              \
      hipExtLaunchKernelGGL(KERNEL, GRID, BLOCK, SHMEM, STREAM, _bpx_s, _bpx_e, 0, __VA_ARGS__); \
    else                                                                                     \
      hipLaunchKernelGGL(KERNEL, GRID, BLOCK, SHMEM, STREAM, __VA_ARGS__);                   \
  } while (0)

#define BPX_CHECK_LAUNCH()                                  \
  do {                                                      \
    if (hipGetLastError() != hipSuccess) return BPX_ERR_LAUNCH; \
  } while (0)
