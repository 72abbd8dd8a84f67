// bpx_info.hip -- order-preserving spike compaction (bm.event.info).
//
// Three passes over tiles of 4096 events (256 threads x 16 rounds):
//   count : per-tile number of truthy events (block reduction)
//   scan  : exclusive scan of tile counts per batch row -> tile offsets, num
//   write : per round, each wave ballots its 64 events; mbcnt gives the slot
//           inside the wave, an LDS exchange of the 4 wave counts gives the
//           slot inside the block, the tile offset the slot in the row.
//           Positions at or past num are filled with -1 by their owners.
// Bool events whose rows start 16-byte aligned take k_info_count_b8 /
// k_info_write_b8: each lane loads its 16 consecutive events with one
// 16-byte load and tests the bytes four at a time; the write pass compacts a
// tile's ids into LDS (block scan of the lane counts) and stores them as one
// coalesced run.  With the full workspace, the count pass also stores each
// lane's 16 event bits, and the write pass reads those 2 bytes instead of the
// 16 event bytes (1e8 events: write pass 89 -> ~72 µs).
// The result equals the serial loop of event/_info_collection.py:143-154.
#include <mutex>

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

// exclusive scan of one batch row's tile counts: thread t owns the
// contiguous chunk [t*per, (t+1)*per) -- a serial sum, one block scan of the
// 1024 chunk sums, then the chunk's offsets
__global__ __launch_bounds__(1024) void k_info_scan(const int32_t* __restrict__ counts, int64_t ntiles,
                                                    int32_t* __restrict__ offsets, int32_t* __restrict__ num) {
  const int64_t b = blockIdx.x;
  const int32_t* cnt = counts + b * ntiles;
  int32_t* off = offsets + b * ntiles;
  const int64_t per = (ntiles + 1023) / 1024;
  const int64_t lo = (int64_t)threadIdx.x * per;
  const int64_t hi = lo + per < ntiles ? lo + per : ntiles;
  int sum = 0;
  for (int64_t i = lo; i < hi; ++i) sum += cnt[i];
  int v = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(v, o, 64);
    if (lane_id() >= o) v += y;
  }
  __shared__ int s_w[16];
  const uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 63) s_w[w] = v;
  __syncthreads();
  int run = v - sum;
  for (uint32_t k = 0; k < w; ++k) run += s_w[k];
  for (int64_t i = lo; i < hi; ++i) {
    off[i] = run;
    run += cnt[i];
  }
  if (threadIdx.x == 0) {
    int total = 0;
    for (int k = 0; k < 16; ++k) total += s_w[k];
    num[b] = total;
  }
}

// the same scan with the row's counts staged in LDS (coalesced load and
// store; ntiles <= kScanLdsTiles)
constexpr int64_t kScanLdsTiles = 32768;

__global__ __launch_bounds__(1024) void k_info_scan_lds(const int32_t* __restrict__ counts, int64_t ntiles,
                                                        int32_t* __restrict__ offsets, int32_t* __restrict__ num) {
  extern __shared__ int32_t s_c[];
  const int64_t b = blockIdx.x;
  const int32_t* cnt = counts + b * ntiles;
  int32_t* off = offsets + b * ntiles;
  // all of a thread's loads (ntiles <= 32 * 1024) are issued before any LDS
  // store, so the copy costs one memory round trip
  {
    int32_t v[kScanLdsTiles / 1024];
#pragma unroll
    for (int u = 0; u < kScanLdsTiles / 1024; ++u) {
      const int64_t i = threadIdx.x + (int64_t)u * 1024;
      v[u] = i < ntiles ? cnt[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < kScanLdsTiles / 1024; ++u) {
      const int64_t i = threadIdx.x + (int64_t)u * 1024;
      if (i < ntiles) s_c[i] = v[u];
    }
  }
  __syncthreads();
  const int64_t per = (ntiles + 1023) / 1024;
  const int64_t lo = (int64_t)threadIdx.x * per;
  const int64_t hi = lo + per < ntiles ? lo + per : ntiles;
  int sum = 0;
  for (int64_t i = lo; i < hi; ++i) sum += s_c[i];
  int v = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(v, o, 64);
    if (lane_id() >= o) v += y;
  }
  __shared__ int s_w[16];
  const uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 63) s_w[w] = v;
  __syncthreads();
  int run = v - sum;
  for (uint32_t k = 0; k < w; ++k) run += s_w[k];
  for (int64_t i = lo; i < hi; ++i) {
    const int c = s_c[i];
    s_c[i] = run;
    run += c;
  }
  __syncthreads();
  for (int64_t i = threadIdx.x; i < ntiles; i += 1024) off[i] = s_c[i];
  if (threadIdx.x == 0) {
    int total = 0;
    for (int k = 0; k < 16; ++k) total += s_w[k];
    num[b] = total;
  }
}

static void launch_info_scan(hipStream_t s, const int32_t* counts, int64_t ntiles, int64_t batch, int32_t* offsets,
                             int32_t* num) {
  if (ntiles <= kScanLdsTiles)
    hipLaunchKernelGGL(k_info_scan_lds, dim3((unsigned)batch), dim3(1024), (size_t)ntiles * 4, s, counts, ntiles,
                       offsets, num);
  else
    hipLaunchKernelGGL(k_info_scan, dim3((unsigned)batch), dim3(1024), 0, s, counts, ntiles, offsets, num);
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

// bool (0/1 byte) events: bit 7 of each byte of the result is set iff that
// byte of w is non-zero
__device__ __forceinline__ uint32_t nz_bytes(uint32_t w) {
  return (((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | w) & 0x80808080u;
}

// the 16 events of lane t of a tile: [tile*4096 + 16t, +16), zero past n
__device__ __forceinline__ void load16(const uint8_t* __restrict__ row, int64_t i0, int64_t n, uint32_t (&m)[4]) {
  if (i0 + 16 <= n) {
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    const u4 v = *(const u4*)(row + i0);
    m[0] = v[0]; m[1] = v[1]; m[2] = v[2]; m[3] = v[3];
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint32_t w = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t i = i0 + 4 * q + j;
        if (i < n) w |= (uint32_t)row[i] << (8 * j);
      }
      m[q] = w;
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) m[q] = nz_bytes(m[q]);
}

__global__ __launch_bounds__(256) void k_info_count_b8(const uint8_t* __restrict__ ev, int64_t n, int64_t ntiles,
                                                       int32_t* __restrict__ counts) {
  const int64_t b = blockIdx.y;
  const int64_t tile = blockIdx.x;
  uint32_t m[4];
  load16(ev + b * n, tile * kInfoTile + (int64_t)threadIdx.x * 16, n, m);
  int v = __popc(m[0]) + __popc(m[1]) + __popc(m[2]) + __popc(m[3]);
  __shared__ int s_w[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (lane_id() == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) counts[b * ntiles + tile] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// bits 7, 15, 23, 31 of m (the nz_bytes flags) -> bits 0..3
__device__ __forceinline__ uint32_t flags4(uint32_t m) { return (((m >> 7) & 0x01010101u) * 0x10204080u) >> 28; }

// count pass of TPB tiles per block (one measured fastest: more tiles per
// block, with all their loads in flight, took longer); lane t's 16 event
// flags of each tile go to bits[tile * 256 + t] (bit e <-> event
// tile*4096 + 16t + e)
template <int TPB>
__global__ __launch_bounds__(256) void k_info_count_bits(const uint8_t* __restrict__ ev, int64_t n, int64_t ntiles,
                                                         int32_t* __restrict__ counts, uint16_t* __restrict__ bits) {
  const int64_t b = blockIdx.y;
  const int64_t tile0 = (int64_t)blockIdx.x * TPB;
  uint32_t m[TPB][4];
  const uint8_t* row = ev + b * n;
  if ((tile0 + TPB) * kInfoTile <= n) {
    // the tiles inside the row: their 16-byte loads in flight together, no branch
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    u4 v[TPB];
#pragma unroll
    for (int u = 0; u < TPB; ++u) v[u] = *(const u4*)(row + (tile0 + u) * kInfoTile + (int64_t)threadIdx.x * 16);
#pragma unroll
    for (int u = 0; u < TPB; ++u)
#pragma unroll
      for (int q = 0; q < 4; ++q) m[u][q] = nz_bytes(v[u][q]);
  } else {
#pragma unroll
    for (int u = 0; u < TPB; ++u) {
      const int64_t i0 = (tile0 + u) * kInfoTile + (int64_t)threadIdx.x * 16;
      if (tile0 + u < ntiles) load16(row, i0, n, m[u]);
      else m[u][0] = m[u][1] = m[u][2] = m[u][3] = 0u;
    }
  }
  __shared__ int s_w[TPB][4];
#pragma unroll
  for (int u = 0; u < TPB; ++u) {
    const uint32_t mask = flags4(m[u][0]) | (flags4(m[u][1]) << 4) | (flags4(m[u][2]) << 8) |
                          (flags4(m[u][3]) << 12);
    // lanes 2k and 2k+1 store their two masks as one 32-bit word
    const uint32_t hi = (uint32_t)__shfl_down((int)mask, 1, 64);
    if (tile0 + u < ntiles && (threadIdx.x & 1) == 0)
      *(uint32_t*)(bits + (b * ntiles + tile0 + u) * 256 + threadIdx.x) = mask | (hi << 16);
    int v = __popc(mask);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane_id() == 0) s_w[u][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < TPB && tile0 + threadIdx.x < ntiles) {
    const int u = threadIdx.x;
    counts[b * ntiles + tile0 + u] = s_w[u][0] + s_w[u][1] + s_w[u][2] + s_w[u][3];
  }
}

// the write pass from the stored event bits
__global__ __launch_bounds__(256) void k_info_write_bits(const uint16_t* __restrict__ bits, int64_t n, int64_t ntiles,
                                                         const int32_t* __restrict__ offsets,
                                                         const int32_t* __restrict__ num, int32_t* __restrict__ ids) {
  const int64_t b = blockIdx.y;
  const int64_t tile = blockIdx.x;
  int32_t* out = ids + b * n;
  const int64_t t0 = tile * kInfoTile;
  const int64_t i0 = t0 + (int64_t)threadIdx.x * 16;
  uint32_t mask = bits[(b * ntiles + tile) * 256 + threadIdx.x];
  const int c = __popc(mask);
  int v = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(v, o, 64);
    if (lane_id() >= o) v += y;
  }
  __shared__ int s_w[4];
  __shared__ int32_t s_ids[kInfoTile];
  const uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 63) s_w[w] = v;
  __syncthreads();
  int pos = v - c;
  for (uint32_t q = 0; q < w; ++q) pos += s_w[q];
  const int cnt = s_w[0] + s_w[1] + s_w[2] + s_w[3];
  while (mask != 0u) {
    const int e = __builtin_ctz(mask);
    mask &= mask - 1u;
    s_ids[pos++] = (int32_t)(i0 + e);
  }
  __syncthreads();
  const int64_t off = offsets[b * ntiles + tile];
  for (int k = threadIdx.x; k < cnt; k += 256) out[off + k] = s_ids[k];
  // the -1 fill of this tile's positions at or past num, four at a time
  // where the row and the positions allow 16-byte stores
  const int64_t total = num[b];
  if (t0 + kInfoTile > total) {
    const bool vec = ((uintptr_t)out & 15u) == 0 && t0 + kInfoTile <= n;
#pragma unroll
    for (int k = 0; k < kInfoRounds / 4; ++k) {
      const int64_t i = t0 + (int64_t)k * 1024 + (int64_t)threadIdx.x * 4;
      if (vec && i >= total) {
        typedef int32_t i4 __attribute__((ext_vector_type(4)));
        *(i4*)(out + i) = (i4){-1, -1, -1, -1};
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (i + e < n && i + e >= total) out[i + e] = -1;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_info_write_b8(const uint8_t* __restrict__ ev, int64_t n, int64_t ntiles,
                                                       const int32_t* __restrict__ offsets,
                                                       const int32_t* __restrict__ num, int32_t* __restrict__ ids) {
  const int64_t b = blockIdx.y;
  const int64_t tile = blockIdx.x;
  int32_t* out = ids + b * n;
  const int64_t t0 = tile * kInfoTile;
  const int64_t i0 = t0 + (int64_t)threadIdx.x * 16;
  uint32_t m[4];
  load16(ev + b * n, i0, n, m);
  const int c = __popc(m[0]) + __popc(m[1]) + __popc(m[2]) + __popc(m[3]);
  // block exclusive scan of the lane counts
  int v = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(v, o, 64);
    if (lane_id() >= o) v += y;
  }
  __shared__ int s_w[4];
  __shared__ int32_t s_ids[kInfoTile];
  const uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 63) s_w[w] = v;
  __syncthreads();
  int pos = v - c;
  for (uint32_t q = 0; q < w; ++q) pos += s_w[q];
  const int cnt = s_w[0] + s_w[1] + s_w[2] + s_w[3];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (m[q] & (0x80u << (8 * j))) s_ids[pos++] = (int32_t)(i0 + 4 * q + j);
    }
  }
  __syncthreads();
  const int64_t off = offsets[b * ntiles + tile];
  for (int k = threadIdx.x; k < cnt; k += 256) out[off + k] = s_ids[k];
  // positions at or past num belong to the -1 fill; each tile fills its own range
  const int64_t total = num[b];
#pragma unroll
  for (int k = 0; k < kInfoRounds; ++k) {
    const int64_t i = t0 + (int64_t)k * 256 + threadIdx.x;
    if (i < n && i >= total) out[i] = -1;
  }
}

// workspace: tile counts and offsets (the minimum), then the event bits of
// the bool fast path
static size_t info_ws_base(int64_t batch, int64_t n) {
  const int64_t ntiles = (n + kInfoTile - 1) / kInfoTile;
  return (size_t)(2 * batch * ntiles) * sizeof(int32_t);
}
static size_t info_ws_bits(int64_t batch, int64_t n) {
  const int64_t ntiles = (n + kInfoTile - 1) / kInfoTile;
  return (size_t)(batch * ntiles) * 256 * sizeof(uint16_t);
}

template <typename E>
static int run_info(hipStream_t s, const void* events, int64_t batch, int64_t n, int32_t* ids, int32_t* num,
                    int32_t* ws, size_t ws_bytes) {
  const int64_t ntiles = (n + kInfoTile - 1) / kInfoTile;
  int32_t* counts = ws;
  int32_t* offsets = ws + batch * ntiles;
  dim3 grid((unsigned)ntiles, (unsigned)batch);
  // bool rows that start 16-byte aligned: 16 events per lane per load
  const bool b8 = sizeof(E) == 1 && ((uintptr_t)events & 15u) == 0 && (n & 15) == 0;
  if (b8 && ws_bytes >= info_ws_base(batch, n) + info_ws_bits(batch, n)) {
    uint16_t* bits = (uint16_t*)((char*)ws + info_ws_base(batch, n));
    hipLaunchKernelGGL(k_info_count_bits<1>, dim3((unsigned)ntiles, (unsigned)batch), dim3(256), 0, s,
                       (const uint8_t*)events, n, ntiles, counts, bits);
    launch_info_scan(s, counts, ntiles, batch, offsets, num);
    BPX_LAUNCH(KID_INFO, k_info_write_bits, grid, dim3(256), 0, s, (const uint16_t*)bits, n, ntiles, offsets, num,
               ids);
    BPX_CHECK_LAUNCH();
    return BPX_OK;
  }
  if (b8) {
    hipLaunchKernelGGL(k_info_count_b8, grid, dim3(256), 0, s, (const uint8_t*)events, n, ntiles, counts);
    launch_info_scan(s, counts, ntiles, batch, offsets, num);
    BPX_LAUNCH(KID_INFO, k_info_write_b8, grid, dim3(256), 0, s, (const uint8_t*)events, n, ntiles, offsets, num,
               ids);
    BPX_CHECK_LAUNCH();
    return BPX_OK;
  }
  hipLaunchKernelGGL(k_info_count<E>, grid, dim3(256), 0, s, (const E*)events, n, ntiles, counts);
  launch_info_scan(s, counts, ntiles, batch, offsets, num);
  BPX_LAUNCH(KID_INFO, k_info_write<E>, grid, dim3(256), 0, s, (const E*)events, n, ntiles, offsets, num, ids);
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

__global__ void k_zero_i32(int32_t* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Scratch of bpx_event_info_noalloc: the tile counts and tile offsets that
// bpx_event_info keeps in its workspace, in a static device array (one copy
// per GPU, held until the process ends).  Host code never passes `scratch`
// itself (on the host it only names the symbol): its device address is
// looked up once per device with hipGetSymbolAddress and kept in
// g_scratch_ptr.  Every call writes all the counts and offsets it reads, so
// nothing has to start at zero; calls that overlap on one GPU race.
constexpr size_t kInfoScratchBytes = BPX_EVENT_INFO_SCRATCH_BYTES;
static __device__ __align__(256) int32_t scratch[kInfoScratchBytes / sizeof(int32_t)];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static int32_t* g_scratch_ptr[kMaxDevices];

// device address of `scratch` on the current device (looked up on the
// first call on that device, with the device current)
static int info_scratch(int32_t** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return BPX_ERR_LAUNCH;
  if (dev < 0 || dev >= kMaxDevices) return BPX_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return BPX_ERR_LAUNCH;
    g_scratch_ptr[dev] = (int32_t*)p;
  }
  *out = g_scratch_ptr[dev];
  return BPX_OK;
}

}  // namespace bpx

using namespace bpx;

extern "C" {

size_t bpx_event_info_workspace(int64_t batch, int64_t n) {
  if (batch <= 0 || n <= 0) return 0;
  return info_ws_base(batch, n) + info_ws_bits(batch, n);
}

int bpx_event_info(void* stream, const void* events, int ev_dtype, int64_t batch, int64_t n, int32_t* ids,
                   int32_t* num, void* workspace, size_t workspace_bytes) {
  if (batch < 0 || n < 0) return BPX_ERR_SHAPE;
  if (batch > 65535) return BPX_ERR_SHAPE;
  hipStream_t s = (hipStream_t)stream;
  if (batch == 0) return BPX_OK;
  if (!num || (n > 0 && (!events || !ids))) return BPX_ERR_ARG;
  if (n == 0) {
    hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, s, num, batch);
    BPX_CHECK_LAUNCH();
    return BPX_OK;
  }
  if (n > INT32_MAX) return BPX_ERR_SHAPE;
  if (!workspace || workspace_bytes < info_ws_base(batch, n)) return BPX_ERR_WORKSPACE;
  int32_t* ws = (int32_t*)workspace;
  switch (ev_dtype) {
    case BPX_BOOL: return run_info<uint8_t>(s, events, batch, n, ids, num, ws, workspace_bytes);
    case BPX_F32: return run_info<float>(s, events, batch, n, ids, num, ws, workspace_bytes);
    case BPX_F64: return run_info<double>(s, events, batch, n, ids, num, ws, workspace_bytes);
    case BPX_I32: return run_info<int32_t>(s, events, batch, n, ids, num, ws, workspace_bytes);
    default: return BPX_ERR_DTYPE;
  }
}

int bpx_event_info_noalloc(void* stream, const void* events, int ev_dtype, int64_t batch, int64_t n, int32_t* ids,
                           int32_t* num) {
  if (batch < 0 || n < 0) return BPX_ERR_SHAPE;
  if (batch > 65535) return BPX_ERR_SHAPE;
  if (ev_dtype != BPX_BOOL && ev_dtype != BPX_F32 && ev_dtype != BPX_F64 && ev_dtype != BPX_I32) return BPX_ERR_DTYPE;
  if (batch == 0) return BPX_OK;
  if (!num || (n > 0 && (!events || !ids))) return BPX_ERR_ARG;
  if (n > INT32_MAX) return BPX_ERR_SHAPE;
  // the bits of the bool fast path when they fit the scratch, else the minimum
  size_t need = bpx_event_info_workspace(batch, n);
  if (need > kInfoScratchBytes) need = info_ws_base(batch, n);
  if (need > kInfoScratchBytes) return BPX_ERR_WORKSPACE;
  int32_t* ws = nullptr;
  if (n > 0) {
    const int rc = info_scratch(&ws);
    if (rc != BPX_OK) return rc;
  }
  return bpx_event_info(stream, events, ev_dtype, batch, n, ids, num, ws, need);
}

}  // extern "C"
