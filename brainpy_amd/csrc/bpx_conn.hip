// bpx_conn.hip -- FixedProb connectivity built on the device
// (FixedProb.build_csr, connect/random_conn.py:85-128 with pre_ratio = 1).
//
// Every pre-synaptic row draws k = int(n_post * prob) distinct targets
// uniformly (the reference's `rng.choice(n_post, k, replace=False)`), and
// with include_self = False drops its own index if drawn (row length k - 1).
// With allow_multi_conn the k targets are drawn with replacement (the
// reference's `randint(0, n_post, (n_pre, k))`).
//
// One wave per row.  The row's targets live in an LDS bitmap of n_post
// bits.  Lanes draw candidates from a counter-based stream (SplitMix64 of
// seed, row and draw index; Lemire's multiply-shift with rejection for an
// exactly uniform value in [0, n)), set their bit with an LDS atomic OR, and
// the wave keeps, in lane order, the first `remaining` candidates whose bit
// was clear: exactly the sequential rejection sampler (draw, keep if new,
// until k), which gives a uniform k-subset.  For k > n/2 the n - k excluded
// targets are drawn instead.  A scan of the bitmap (popcount prefix over the
// lanes) writes the row in ascending order and clears the words for the
// next row.  Rows of different length (include_self = False) take a counting
// pass, a scan for indptr, and the writing pass, which draws the same
// stream again.
#include "bpx_common.h"

namespace bpx {
namespace {

constexpr int kConnLdsWords = 40960;   // 160 KB of bitmap words per block

// SplitMix64 output function (Steele, Lea and Flood, OOPSLA 2014)
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// uniform in [0, n) for draw d of the row stream h: Lemire, "Fast random
// integer generation in an interval" (TOMS 2019) -- the low word of x * n
// below (2^32 - n) mod n is rejected and the draw repeated with the next
// sub-index; `thresh` = (2^32 - n) mod n
__device__ __forceinline__ uint32_t draw_below(uint64_t h, uint64_t d, uint32_t n, uint32_t thresh) {
  uint32_t v = 0;
  for (uint32_t a = 0; a < 8; ++a) {   // 8 rejections in a row: probability (n / 2^32)^8
    const uint32_t x = (uint32_t)(mix64(h ^ ((d << 3) | a)) >> 32);
    const uint64_t p = (uint64_t)x * n;
    v = (uint32_t)(p >> 32);
    if ((uint32_t)p >= thresh) break;
  }
  return v;
}

__device__ __forceinline__ uint64_t row_stream(uint64_t seed, int64_t row) {
  return mix64(mix64(seed) ^ (uint64_t)row);
}

// sample the row's set into the wave's bitmap (clear on entry); complement:
// the bitmap then holds the n - k excluded targets
__device__ void sample_row(uint32_t* bm, uint64_t h, uint32_t n, uint32_t want, uint32_t thresh) {
  const int lane = lane_id();
  uint32_t remaining = want;
  // bounded: with want <= n/2 a batch of 64 draws finds >= 32 new targets on
  // average, so the cap (4096 times the expected number of batches) is never
  // reached in practice but keeps every wave finite
  const uint64_t cap = ((uint64_t)want + kWave) * kWave;
  for (uint64_t batch = 0; remaining > 0 && batch < cap; ++batch) {
    const uint32_t c = draw_below(h, batch * kWave + (uint64_t)lane, n, thresh);
    const uint32_t bit = 1u << (c & 31u);
    const uint32_t old = atomicOr(bm + (c >> 5), bit);
    const bool fresh = (old & bit) == 0u;
    const uint64_t mask = __ballot(fresh);
    const uint32_t pos = (uint32_t)lanes_below(mask);
    if (fresh && pos >= remaining) atomicAnd(bm + (c >> 5), ~bit);   // past the k-th: not drawn
    const uint32_t got = (uint32_t)__popcll(mask);
    remaining -= got < remaining ? got : remaining;
  }
}

// the row's targets in ascending order: word w of the bitmap (inverted for
// the complement, bits >= n masked, the row's own bit dropped when
// no_self); writes them at out (when non-null), clears the bitmap, returns
// the count
__device__ uint32_t emit_row(uint32_t* bm, uint32_t n, bool complement, int64_t self, int32_t* out) {
  const int lane = lane_id();
  const uint32_t nw = (n + 31u) >> 5;
  uint32_t total = 0;
  for (uint32_t w0 = 0; w0 < nw; w0 += kWave) {
    const uint32_t w = w0 + (uint32_t)lane;
    uint32_t word = 0;
    if (w < nw) {
      word = bm[w];
      bm[w] = 0u;
      if (complement) {
        word = ~word;
        if (w == nw - 1 && (n & 31u)) word &= (1u << (n & 31u)) - 1u;
      }
      if (self >= 0 && (uint64_t)self >> 5 == w) word &= ~(1u << (self & 31));
    }
    const uint32_t cnt = (uint32_t)__popc(word);
    uint32_t incl = cnt;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t y = __shfl_up(incl, o, kWave);
      if (lane >= o) incl += y;
    }
    if (out) {
      uint32_t p = total + incl - cnt;
      while (word) {
        const uint32_t b = (uint32_t)__ffs(word) - 1u;
        out[p++] = (int32_t)(w * 32u + b);
        word &= word - 1u;
      }
    }
    total += __shfl(incl, kWave - 1, kWave);
  }
  return total;
}

// one wave per row (grid-stride over rows); mode 0: count into counts[row],
// mode 1: write the row at indices + indptr[row] (indptr from the count pass
// or row * k)
__global__ __launch_bounds__(256) void k_fixed_prob(int64_t m, uint32_t n, uint32_t k, uint64_t seed, int no_self,
                                                   int mode, const int32_t* __restrict__ indptr,
                                                   int32_t* __restrict__ counts, int32_t* __restrict__ indices,
                                                   int waves_per_block) {
  extern __shared__ uint32_t s_bm[];
  const int wave = threadIdx.x / kWave;
  if (wave >= waves_per_block) return;
  const uint32_t nw = (n + 31u) >> 5;
  uint32_t* bm = s_bm + (size_t)wave * nw;
  for (uint32_t w = (uint32_t)lane_id(); w < nw; w += kWave) bm[w] = 0u;
  const bool complement = 2ull * k > n;
  const uint32_t want = complement ? n - k : k;
  const uint32_t thresh = (0u - n) % n;
  const int64_t nwaves = (int64_t)gridDim.x * waves_per_block;
  for (int64_t row = (int64_t)blockIdx.x * waves_per_block + wave; row < m; row += nwaves) {
    const uint64_t h = row_stream(seed, row);
    sample_row(bm, h, n, want, thresh);
    const int64_t self = no_self ? row : -1;
    if (mode == 0) {
      const uint32_t c = emit_row(bm, n, complement, self, nullptr);
      if (lane_id() == 0) counts[row] = (int32_t)c;
    } else {
      const int64_t base = indptr ? (int64_t)indptr[row] : row * (int64_t)k;
      emit_row(bm, n, complement, self, indices + base);
    }
  }
}

// with replacement: lane l of the row's wave draws targets l, l + 64, ...
__global__ __launch_bounds__(256) void k_fixed_prob_multi(int64_t m, uint32_t n, uint32_t k, uint64_t seed,
                                                         int32_t* __restrict__ indices) {
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / kWave);
  const uint32_t thresh = (0u - n) % n;
  for (int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; row < m; row += nwaves) {
    const uint64_t h = row_stream(seed, row);
    for (uint32_t j = (uint32_t)lane_id(); j < k; j += kWave)
      indices[row * (int64_t)k + j] = (int32_t)draw_below(h, j, n, thresh);
  }
}

// indptr[0] = 0, indptr[i + 1] = indptr[i] + counts[i]: one block, thread t
// scans the contiguous chunk [t * per, (t + 1) * per)
__global__ __launch_bounds__(1024) void k_counts_to_indptr(const int32_t* __restrict__ counts, int64_t m,
                                                          int32_t* __restrict__ indptr) {
  const int64_t per = (m + 1023) / 1024;
  const int64_t lo = (int64_t)threadIdx.x * per;
  const int64_t hi = lo + per < m ? lo + per : m;
  int64_t sum = 0;
  for (int64_t i = lo; i < hi; ++i) sum += counts[i];
  int64_t v = sum;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int64_t y = __shfl_up(v, o, kWave);
    if (lane_id() >= o) v += y;
  }
  __shared__ int64_t s_w[16];
  const int w = threadIdx.x / kWave;
  if (lane_id() == kWave - 1) s_w[w] = v;
  __syncthreads();
  int64_t run = v - sum;
  for (int q = 0; q < w; ++q) run += s_w[q];
  if (threadIdx.x == 0) indptr[0] = 0;
  for (int64_t i = lo; i < hi; ++i) {
    run += counts[i];
    indptr[i + 1] = (int32_t)run;
  }
}

__global__ __launch_bounds__(256) void k_regular_indptr(int64_t m, int32_t k, int32_t* __restrict__ indptr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= m) indptr[i] = (int32_t)(i * k);
}

}  // namespace
}  // namespace bpx

using namespace bpx;

extern "C" {

size_t bpx_fixed_prob_workspace(int64_t m) { return m > 0 ? (size_t)m * sizeof(int32_t) : 0; }

int bpx_fixed_prob_csr(void* stream, int64_t m, int64_t n, int64_t k, uint64_t seed, int include_self,
                       int allow_multi_conn, int32_t* indices, int32_t* indptr, void* workspace,
                       size_t workspace_bytes) {
  if (m < 0 || n < 0 || k < 0 || k > n) return BPX_ERR_SHAPE;
  if (!include_self && m != n) return BPX_ERR_SHAPE;          // the reference raises for pre != post
  if (m * k > (int64_t)INT32_MAX || n > (int64_t)INT32_MAX) return BPX_ERR_SHAPE;
  if (!indptr || (m * k > 0 && !indices)) return BPX_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const bool ragged = !include_self && !allow_multi_conn && k > 0;
  if (!ragged) {
    hipLaunchKernelGGL(k_regular_indptr, dim3((unsigned)((m + 1 + 255) / 256)), dim3(256), 0, s, m, (int32_t)k,
                       indptr);
    BPX_CHECK_LAUNCH();
  }
  if (m == 0 || k == 0) return BPX_OK;
  if (allow_multi_conn) {
    int64_t blocks = (m + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fixed_prob_multi, dim3((unsigned)blocks), dim3(256), 0, s, m, (uint32_t)n, (uint32_t)k,
                       seed, indices);
    BPX_CHECK_LAUNCH();
    return BPX_OK;
  }
  const int64_t nw = (n + 31) / 32;
  if (nw > kConnLdsWords) return BPX_ERR_SHAPE;                // n_post above 1,310,720
  int wpb = (int)(kConnLdsWords / nw);
  if (wpb > 4) wpb = 4;
  const size_t shm = (size_t)wpb * (size_t)nw * sizeof(uint32_t);
  int64_t blocks = (m + wpb - 1) / wpb;
  if (blocks > 8192) blocks = 8192;
  if (ragged) {
    if (!workspace || workspace_bytes < bpx_fixed_prob_workspace(m)) return BPX_ERR_WORKSPACE;
    int32_t* counts = (int32_t*)workspace;
    hipLaunchKernelGGL(k_fixed_prob, dim3((unsigned)blocks), dim3(256), shm, s, m, (uint32_t)n, (uint32_t)k, seed,
                       1, 0, (const int32_t*)nullptr, counts, (int32_t*)nullptr, wpb);
    hipLaunchKernelGGL(k_counts_to_indptr, dim3(1), dim3(1024), 0, s, counts, m, indptr);
    hipLaunchKernelGGL(k_fixed_prob, dim3((unsigned)blocks), dim3(256), shm, s, m, (uint32_t)n, (uint32_t)k, seed,
                       1, 1, (const int32_t*)indptr, counts, indices, wpb);
  } else {
    hipLaunchKernelGGL(k_fixed_prob, dim3((unsigned)blocks), dim3(256), shm, s, m, (uint32_t)n, (uint32_t)k, seed,
                       0, 1, (const int32_t*)nullptr, (int32_t*)nullptr, indices, wpb);
  }
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

}  // extern "C"
