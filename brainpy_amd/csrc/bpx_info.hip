// bpx_info.hip -- order-preserving spike compaction (bm.event.info).
//
// Three passes over tiles of 4096 events (256 threads x 16 rounds):
//   count : per-tile number of truthy events (block reduction)
//   scan  : exclusive scan of tile counts per batch row -> tile offsets, num
//   write : per round, each wave ballots its 64 events; mbcnt gives the slot
//           inside the wave, an LDS exchange of the 4 wave counts gives the
//           slot inside the block, the tile offset the slot in the row.
//           Positions at or past num are filled with -1 by their owners.
// The result equals the serial loop of event/_info_collection.py:143-154.
#include "bpx_common.h"

namespace bpx {

constexpr int kInfoRounds = 16;
constexpr int kInfoTile = 256 * kInfoRounds;

template <typename E>
__device__ __forceinline__ bool info_true(E v) { return ev_true(v, BPX_RULE_NONZERO); }

template <typename E>
__global__ __launch_bounds__(256) void k_info_count(const E* __restrict__ ev, int64_t n, int64_t ntiles,
                                                    int32_t* __restrict__ counts) {
  const int64_t b = blockIdx.y;
  const int64_t tile = blockIdx.x;
  const E* row = ev + b * n;
  int c = 0;
#pragma unroll
  for (int k = 0; k < kInfoRounds; ++k) {
    const int64_t i = tile * kInfoTile + (int64_t)k * 256 + threadIdx.x;
    if (i < n && info_true(row[i])) ++c;
  }
  __shared__ int s_w[4];
  const uint32_t w = threadIdx.x >> 6;
  int v = c;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (lane_id() == 0) s_w[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) counts[b * ntiles + tile] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

__global__ __launch_bounds__(1024) void k_info_scan(const int32_t* __restrict__ counts, int64_t ntiles,
                                                    int32_t* __restrict__ offsets, int32_t* __restrict__ num) {
  const int64_t b = blockIdx.x;
  __shared__ int s_w[16];
  __shared__ int s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  const uint32_t w = threadIdx.x >> 6;
  for (int64_t base = 0; base < ntiles; base += 1024) {
    const int64_t i = base + threadIdx.x;
    const int x = i < ntiles ? counts[b * ntiles + i] : 0;
    // inclusive wave scan
    int v = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      int y = __shfl_up(v, o, 64);
      if (lane_id() >= o) v += y;
    }
    if (lane_id() == 63) s_w[w] = v;
    __syncthreads();
    int wo = 0;
    for (uint32_t k = 0; k < w; ++k) wo += s_w[k];
    const int carry = s_carry;
    if (i < ntiles) offsets[b * ntiles + i] = carry + wo + v - x;
    __syncthreads();
    if (threadIdx.x == 1023) s_carry = carry + wo + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) num[b] = s_carry;
}

template <typename E>
__global__ __launch_bounds__(256) void k_info_write(const E* __restrict__ ev, int64_t n, int64_t ntiles,
                                                    const int32_t* __restrict__ offsets,
                                                    const int32_t* __restrict__ num, int32_t* __restrict__ ids) {
  const int64_t b = blockIdx.y;
  const int64_t tile = blockIdx.x;
  const E* row = ev + b * n;
  int32_t* out = ids + b * n;
  const int64_t total = num[b];
  int64_t off = offsets[b * ntiles + tile];
  __shared__ int s_w[kInfoRounds][4];
  const uint32_t w = threadIdx.x >> 6;
  bool act[kInfoRounds];
  int slot[kInfoRounds];
#pragma unroll
  for (int k = 0; k < kInfoRounds; ++k) {
    const int64_t i = tile * kInfoTile + (int64_t)k * 256 + threadIdx.x;
    act[k] = i < n && info_true(row[i]);
    const uint64_t mask = __ballot(act[k]);
    slot[k] = lanes_below(mask);
    if (lane_id() == 0) s_w[k][w] = __popcll(mask);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kInfoRounds; ++k) {
    const int64_t i = tile * kInfoTile + (int64_t)k * 256 + threadIdx.x;
    int wo = 0;
    for (uint32_t q = 0; q < w; ++q) wo += s_w[k][q];
    if (act[k]) out[off + wo + slot[k]] = (int32_t)i;
    if (i < n && i >= total) out[i] = -1;
    off += s_w[k][0] + s_w[k][1] + s_w[k][2] + s_w[k][3];
  }
}

template <typename E>
static int run_info(hipStream_t s, const void* events, int64_t batch, int64_t n, int32_t* ids, int32_t* num,
                    int32_t* ws) {
  const int64_t ntiles = (n + kInfoTile - 1) / kInfoTile;
  int32_t* counts = ws;
  int32_t* offsets = ws + batch * ntiles;
  dim3 grid((unsigned)ntiles, (unsigned)batch);
  hipLaunchKernelGGL(k_info_count<E>, grid, dim3(256), 0, s, (const E*)events, n, ntiles, counts);
  hipLaunchKernelGGL(k_info_scan, dim3((unsigned)batch), dim3(1024), 0, s, counts, ntiles, offsets, num);
  BPX_LAUNCH(KID_INFO, k_info_write<E>, grid, dim3(256), 0, s, (const E*)events, n, ntiles, offsets, num, ids);
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

__global__ void k_zero_i32(int32_t* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace bpx

using namespace bpx;

extern "C" {

size_t bpx_event_info_workspace(int64_t batch, int64_t n) {
  if (batch <= 0 || n <= 0) return 0;
  const int64_t ntiles = (n + kInfoTile - 1) / kInfoTile;
  return (size_t)(2 * batch * ntiles) * sizeof(int32_t);
}

int bpx_event_info(void* stream, const void* events, int ev_dtype, int64_t batch, int64_t n, int32_t* ids,
                   int32_t* num, void* workspace, size_t workspace_bytes) {
  if (batch < 0 || n < 0) return BPX_ERR_SHAPE;
  if (batch > 65535) return BPX_ERR_SHAPE;
  hipStream_t s = (hipStream_t)stream;
  if (batch == 0) return BPX_OK;
  if (n == 0) {
    hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, s, num, batch);
    BPX_CHECK_LAUNCH();
    return BPX_OK;
  }
  if (n > INT32_MAX) return BPX_ERR_SHAPE;
  if (!workspace || workspace_bytes < bpx_event_info_workspace(batch, n)) return BPX_ERR_WORKSPACE;
  switch (ev_dtype) {
    case BPX_BOOL: return run_info<uint8_t>(s, events, batch, n, ids, num, (int32_t*)workspace);
    case BPX_F32: return run_info<float>(s, events, batch, n, ids, num, (int32_t*)workspace);
    case BPX_F64: return run_info<double>(s, events, batch, n, ids, num, (int32_t*)workspace);
    case BPX_I32: return run_info<int32_t>(s, events, batch, n, ids, num, (int32_t*)workspace);
    default: return BPX_ERR_DTYPE;
  }
}

}  // extern "C"
