// bpx_tiled.hip -- XCD/LDS-tiled event scatter ("tiled CSR plan").
//
// The transpose event csrmv is bound by global FP32 atomics when every
// synaptic event becomes one atomic on a random output address: MI355X
// executes ~20 G 64-byte atomic segments per second, so 1e7 events cost
// ~0.5 ms whatever their bytes.  The tiled layout removes them:
//
//   * the output [n] is cut into P tiles of T entries, P = CUs per XCD, and
//     each XCD holds C full copies of the output (C = 2) spread over the LDS
//     of its CUs (block b runs on XCD b % 8; its tile is (b / 8) % P);
//   * the rows are dealt to the 8C copies in 64-row chunks (chunk q -> copy
//     q % 8C); inside a block each wave walks chunks of its copy, lists the
//     active rows with their tile-p segment (row-tile offsets precomputed),
//     and groups of 8 lanes stream the segments with vector loads (8 bytes
//     of indices, 16 of f32 weights; segments are padded to 4 entries),
//     accumulating at 16-bit tile-local
//     column indices with LDS integer atomics in 64-bit fixed point (gfx950's
//     LDS float atomics run ~20x slower than integer ones);
//   * at the end each block adds its non-zero LDS entries into the output
//     with coalesced global atomics: 8C copies x n/16 segments per call
//     instead of one segment per synaptic event.
//
// Every synapse is read exactly once per call (2 B index + weight bytes);
// events are read once from HBM and re-read from L2 by the P blocks of a
// copy.  The layout is built once per matrix (bpx_tiled_csr_build), like a
// sparse-library preprocess; the result equals bpx_event_csrmv transpose=1.
#include "bpx_common.h"
#include "bpx_rng.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

namespace bpx {

constexpr int kTiledThreads = 1024;
constexpr int kBuildThreads = 128;
constexpr int kMaxTiles = 64;        // tiles whose build counts fit the LDS of a build thread
constexpr int kMaxTilesBig = 1024;    // tiles built through global cursors (n <= 1024 * 4096)

static int device_cus() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) return 256;
  return cus;
}

// ---------------------------------------------------------------------------
// build: per row, count elements per tile -> row-tile offsets; then place
// tile-local indices (u16) and weights in tile-grouped order (stable)
// ---------------------------------------------------------------------------
// Segments are padded to a multiple of kVec elements so the exec kernel can
// fetch kVec tile-local indices and kVec weights with one vector load each
// (kVec = 4: 8 B of indices, 16 B of f32 weights; 8 measured slower at most
// spike rates, its padding costs ~11 % more segment bytes);
// padding entries carry the index 0xFFFF, which no tile reaches.
constexpr int kVec = 4;

template <typename I>
__global__ __launch_bounds__(kBuildThreads) void k_tiled_count(const I* __restrict__ indices,
                                                               const I* __restrict__ indptr, int64_t nnz, int64_t m,
                                                               int64_t n, int P, DivU32 dT, int32_t* __restrict__ rowtile,
                                                               int32_t* __restrict__ rowsize) {
  __shared__ int32_t s_cnt[kBuildThreads * kMaxTiles];
  int32_t* cnt = s_cnt + threadIdx.x * kMaxTiles;
  const int64_t r = (int64_t)blockIdx.x * kBuildThreads + threadIdx.x;
  if (r >= m) return;
  for (int p = 0; p < P; ++p) cnt[p] = 0;
  int64_t s = (int64_t)indptr[r], t = (int64_t)indptr[r + 1];
  s = s < 0 ? 0 : s;
  t = t > nnz ? nnz : t;
  for (int64_t j = s; j < t; ++j) {
    const int64_t c = (int64_t)indices[j];
    if ((uint64_t)c < (uint64_t)n) cnt[dT.div((uint32_t)c)] += 1;
  }
  int32_t* rt = rowtile + r * (P + 1);
  int32_t tot = 0;
  for (int p = 0; p < P; ++p) {
    const int32_t padded = (cnt[p] + kVec - 1) & ~(kVec - 1);
    rt[p] = padded;           // segment size for now; offsets after the row scan
    tot += padded;
  }
  rowsize[r] = tot;
}

// exclusive scan of the padded row sizes (one block), total in rowsize[m]
__global__ __launch_bounds__(1024) void k_scan_rows(int32_t* __restrict__ rowsize, int64_t m) {
  __shared__ int s_w[16];
  __shared__ int s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  const uint32_t w = threadIdx.x >> 6;
  for (int64_t base = 0; base < m; base += 1024) {
    const int64_t i = base + threadIdx.x;
    const int x = i < m ? rowsize[i] : 0;
    int v = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(v, o, 64);
      if (lane_id() >= o) v += y;
    }
    if (lane_id() == 63) s_w[w] = v;
    __syncthreads();
    int wo = 0;
    for (uint32_t k = 0; k < w; ++k) wo += s_w[k];
    const int carry = s_carry;
    if (i < m) rowsize[i] = carry + wo + v - x;
    __syncthreads();
    if (threadIdx.x == 1023) s_carry = carry + wo + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) rowsize[m] = s_carry;
}

__global__ __launch_bounds__(256) void k_tiled_offsets(const int32_t* __restrict__ rowbase, int64_t m, int P,
                                                       int32_t* __restrict__ rowtile) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= m) return;
  int32_t* rt = rowtile + r * (P + 1);
  int32_t acc = rowbase[r];
  for (int p = 0; p < P; ++p) {
    const int32_t sz = rt[p];
    rt[p] = acc;
    acc += sz;
  }
  rt[P] = acc;
}

template <typename I, typename V, typename W>
__global__ __launch_bounds__(kBuildThreads) void k_tiled_fill(const V* __restrict__ data, int heter,
                                                              const I* __restrict__ indices,
                                                              const I* __restrict__ indptr, int64_t nnz, int64_t m,
                                                              int64_t n, int P, DivU32 dT,
                                                              const int32_t* __restrict__ rowtile,
                                                              uint16_t* __restrict__ lidx, W* __restrict__ w) {
  __shared__ int32_t s_cur[kBuildThreads * kMaxTiles];
  int32_t* cur = s_cur + threadIdx.x * kMaxTiles;
  const int64_t r = (int64_t)blockIdx.x * kBuildThreads + threadIdx.x;
  if (r >= m) return;
  const int32_t* rt = rowtile + r * (P + 1);
  for (int p = 0; p < P; ++p) cur[p] = rt[p];
  int64_t s = (int64_t)indptr[r], t = (int64_t)indptr[r + 1];
  s = s < 0 ? 0 : s;
  t = t > nnz ? nnz : t;
  for (int64_t j = s; j < t; ++j) {
    const int64_t c = (int64_t)indices[j];
    if ((uint64_t)c >= (uint64_t)n) continue;
    const uint32_t p = dT.div((uint32_t)c);
    const int32_t pos = cur[p]++;
    lidx[pos] = (uint16_t)((uint32_t)c - p * dT.d);
    if (heter) {
      const float v = to_f32(data[j]);
      if constexpr (sizeof(W) == 2) {
        // round-to-nearest-even f32 -> bf16
        uint32_t u = __float_as_uint(v);
        u += 0x7FFFu + ((u >> 16) & 1u);
        w[pos].bits = (uint16_t)(u >> 16);
      } else {
        w[pos] = v;
      }
    }
  }
  // pad every segment to its vector boundary
  for (int p = 0; p < P; ++p) {
    for (int32_t pos = cur[p]; pos < rt[p + 1]; ++pos) {
      lidx[pos] = 0xFFFF;
      if (heter) {
        if constexpr (sizeof(W) == 2) w[pos].bits = 0;
        else w[pos] = 0.f;
      }
    }
  }
}

// More than kMaxTiles tiles: a block of tb threads (one row each) keeps
// every thread's P counters in dynamic LDS (tb * P * 4 <= 64 KiB), writes
// the counts into the rowtile row, k_tiled_offsets_big turns a row's counts
// into segment starts (one wave per row, coalesced), and k_tiled_fill_big
// places the elements in order from LDS cursors.  Padding slots were filled
// with 0xFFFF beforehand (k_tiled_prefill).
constexpr int kBigBuildLds = 65536;

static int big_build_threads(int P) {
  int tb = kBigBuildLds / (4 * P);
  return tb > 64 ? 64 : (tb < 1 ? 1 : tb);
}

template <typename I>
__global__ __launch_bounds__(64) void k_tiled_count_big(const I* __restrict__ indices, const I* __restrict__ indptr,
                                                        int64_t nnz, int64_t m, int64_t n, int P, DivU32 dT,
                                                        int32_t* __restrict__ rowtile,
                                                        int32_t* __restrict__ rowsize) {
  extern __shared__ int32_t s_big[];
  int32_t* cnt = s_big + threadIdx.x * P;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  for (int p = 0; p < P; ++p) cnt[p] = 0;
  int64_t s = (int64_t)indptr[r], t = (int64_t)indptr[r + 1];
  s = s < 0 ? 0 : s;
  t = t > nnz ? nnz : t;
  for (int64_t j = s; j < t; ++j) {
    const int64_t c = (int64_t)indices[j];
    if ((uint64_t)c < (uint64_t)n) cnt[dT.div((uint32_t)c)] += 1;
  }
  int32_t* rt = rowtile + r * (P + 1);
  int32_t tot = 0;
  for (int p = 0; p < P; ++p) {
    rt[p] = cnt[p];
    tot += (cnt[p] + kVec - 1) & ~(kVec - 1);
  }
  rowsize[r] = tot;
}

// one wave per row: rt[p] = the start of segment p (counts padded to kVec),
// rt[P] = the row's end
__global__ __launch_bounds__(256) void k_tiled_offsets_big(const int32_t* __restrict__ rowbase, int64_t m, int P,
                                                           int32_t* __restrict__ rowtile) {
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= m) return;
  const int lane = threadIdx.x & 63;
  int32_t* rt = rowtile + r * (P + 1);
  int32_t carry = rowbase[r];
  for (int p0 = 0; p0 < P; p0 += 64) {
    const int p = p0 + lane;
    const int32_t c = p < P ? rt[p] : 0;
    const int32_t sz = (c + kVec - 1) & ~(kVec - 1);
    int32_t incl = sz;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    if (p < P) rt[p] = carry + incl - sz;
    carry += __shfl(incl, 63, 64);
  }
  if (lane == 0) rt[P] = carry;
}

// every slot below the plan's total (rowsize[m], read on the device) holds
// the padding index and a zero weight
template <typename W>
__global__ __launch_bounds__(256) void k_tiled_prefill(const int32_t* __restrict__ total, int heter,
                                                       uint16_t* __restrict__ lidx, W* __restrict__ w) {
  const int64_t len = *total;
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (; i < len; i += stride) {
    lidx[i] = 0xFFFF;
    if (heter) {
      if constexpr (sizeof(W) == 2) w[i].bits = 0;
      else w[i] = 0.f;
    }
  }
}

template <typename I, typename V, typename W>
__global__ __launch_bounds__(64) void k_tiled_fill_big(const V* __restrict__ data, int heter,
                                                       const I* __restrict__ indices, const I* __restrict__ indptr,
                                                       int64_t nnz, int64_t m, int64_t n, int P, DivU32 dT,
                                                       const int32_t* __restrict__ rowtile,
                                                       uint16_t* __restrict__ lidx, W* __restrict__ w) {
  extern __shared__ int32_t s_big[];
  int32_t* cur = s_big + threadIdx.x * P;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const int32_t* rt = rowtile + r * (P + 1);
  for (int p = 0; p < P; ++p) cur[p] = rt[p];
  int64_t s = (int64_t)indptr[r], t = (int64_t)indptr[r + 1];
  s = s < 0 ? 0 : s;
  t = t > nnz ? nnz : t;
  for (int64_t j = s; j < t; ++j) {
    const int64_t c = (int64_t)indices[j];
    if ((uint64_t)c >= (uint64_t)n) continue;
    const uint32_t p = dT.div((uint32_t)c);
    const int32_t pos = cur[p]++;
    lidx[pos] = (uint16_t)((uint32_t)c - p * dT.d);
    if (heter) {
      const float v = to_f32(data[j]);
      if constexpr (sizeof(W) == 2) {
        uint32_t u = __float_as_uint(v);
        u += 0x7FFFu + ((u >> 16) & 1u);
        w[pos].bits = (uint16_t)(u >> 16);
      } else {
        w[pos] = v;
      }
    }
  }
}

// out[0] = max |v| (NaN propagates, or with FINITE_ONLY the max over finite
// values), out[1] = min |v| over the finite non-zero values (+inf if none);
// out must hold {0, +inf} before (k_init_minmax)
template <typename V, bool FINITE_ONLY = false>
__global__ __launch_bounds__(256) void k_absmax(const V* __restrict__ data, int64_t len, uint32_t* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * 256;
  float mx = 0.f, mn = __uint_as_float(0x7F800000u);
  for (; i < len; i += stride) {
    float v = fabsf((float)to_f32(data[i]));
    const bool fin = v <= 3.402823466e38f;
    if (fin && v > 0.f) mn = fminf(mn, v);
    if constexpr (FINITE_ONLY) v = fin ? v : 0.f;
    mx = (v > mx || v != v) ? v : mx;    // NaN propagates
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const float y = __shfl_xor(mx, o, 64);
    mx = (y > mx || y != y) ? y : mx;
    mn = fminf(mn, __shfl_xor(mn, o, 64));
  }
  // non-negative floats (and NaN/inf) order like their bit patterns
  if ((threadIdx.x & 63) == 0) {
    atomicMax(out, __float_as_uint(mx));
    atomicMin(out + 1, __float_as_uint(mn));
  }
}

__global__ void k_init_minmax(uint32_t* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    p[0] = 0u;
    p[1] = 0x7F800000u;
  }
}

// The fixed-point grid (fx_inv, one unit in weight or product units)
// rounds a term to within 1e-6 of itself when the term is at least 5e5
// units.  Smaller terms (and, for a dense x, products that are inf or NaN)
// are not put on the grid: they are added to their output directly with an
// FP32 atomic.  For events the plan's flag says whether any weight is that
// small (the COARSE kernels test each weight); for a dense x every product
// is tested.
__device__ __forceinline__ float grid_floor(double fx_inv) { return (float)(5e5 * fx_inv); }

__global__ void k_store_scalar(const void* src, int src_dtype, float* dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float v;
    if (src_dtype == BPX_F64) v = (float)*(const double*)src;
    else if (src_dtype == BPX_BF16) v = to_f32(*(const bf16_t*)src);
    else v = *(const float*)src;
    *dst = v;
  }
}

// ---------------------------------------------------------------------------
// exec
//
// Each wave walks its chunks of 64 rows four at a time: the next four
// chunks' events are loaded before the current ones are processed.  Active
// rows of a batch with a non-empty tile-p segment (ballot + mbcnt) are
// listed in LDS with the segment [start, end); then 8 groups of 8 lanes
// take one listed row each per round, a lane loading 4 indices (8 bytes)
// and 4 weights per pass before its LDS atomics.  Every load of a phase is
// issued without a branch around it, so the four chunks' loads are in
// flight together.
// ---------------------------------------------------------------------------
constexpr int kChunksPerBatch = 4;
constexpr int kListCap = 64 * kChunksPerBatch;
constexpr int kWavesPerBlock = kTiledThreads / 64;

// weight -> 64-bit fixed point: round(w * 2^shift) to nearest even, exact
// for |w| >= 2^(23-shift).  The plan's shift keeps every term below 2^50, so
// one fma adding 1.5 * 2^52 (whose ulp is 1) rounds the scaled weight to an
// integer, and the difference of the two bit patterns is that integer.
__device__ __forceinline__ long long fixed_of(float w, double scale) {
  const double y = fma((double)w, scale, 6755399441055744.0);
  return (long long)__double_as_longlong(y) - 0x4338000000000000ll;
}


// DENSE: `ev` is a dense f32 vector x (bm.sparse.csrmv with transpose): the
// rows with x != 0 are listed with x[row] beside their segment, each weight
// is multiplied by it (one f32 rounding, as the plain product), and the
// fixed-point shift comes from the plan's less ceil(log2 max|x|) over the
// finite x (device scalar xmax: the f32 bits of that max), so the sums still
// stay below 2^61.  Products below the grid floor, and products that are inf
// or NaN (non-finite x, or an f32 overflow), go to their outputs with FP32
// atomics, so an output follows IEEE addition (inf + finite = inf, +inf +
// -inf = NaN).  `homo_w` points at the plan's scalars {homogeneous weight,
// max |w|, min non-zero |w|} and `xmax` at x's {max finite |x|, min non-zero
// |x|}.
// COARSE (events; the plan's flag): weights below the grid floor go to out
// with FP32 atomics, the others to the fixed-point sums.
// LISTED: the active rows were listed once for the call (k_tiled_list:
// lcount[v] rows of vector v in lrows[v][...], any order), and the waves walk
// that list instead of scanning every event of their copy's rows: with many
// tiles per copy each row's event would otherwise be read by every tile
// block.  Sums are the same (fixed point is exact in any order).
template <int WK, typename E, int R, bool DENSE, bool COARSE, bool LISTED = false>
__global__ __launch_bounds__(kTiledThreads) void k_tiled_scatter(const int32_t* __restrict__ rowtile,
                                                                 const uint16_t* __restrict__ lidx,
                                                                 const void* __restrict__ wptr,
                                                                 const float* __restrict__ homo_w,
                                                                 const E* __restrict__ ev, int rule, int64_t m,
                                                                 int64_t n, int P, int T, int ncopy, double fx_scale,
                                                                 double fx_inv, const uint32_t* __restrict__ xmax,
                                                                 float* __restrict__ out,
                                                                 const int32_t* __restrict__ lcount = nullptr,
                                                                 const int32_t* __restrict__ lrows = nullptr) {
  extern __shared__ unsigned long long s_acc[];
  __shared__ int2 s_seg[kWavesPerBlock][kListCap];
  __shared__ float s_xv[DENSE ? kWavesPerBlock : 1][DENSE ? kListCap : 1];
  if constexpr (DENSE) {
    const float mx = __uint_as_float(xmax[0]);
    int ex = 0;
    if (mx > 0.f && mx <= 3.402823466e38f) frexpf(mx, &ex);   // mx < 2^ex
    // products |w * x| < max|w| * 2^ex: scaling by 2^(shift - ex) keeps every
    // term below 2^50 and every sum below 2^61 for any ex, and for |x| < 1
    // (ex <= 0) it raises the resolution with the products' magnitude
    if (ex != 0) {
      fx_scale = ldexp(fx_scale, -ex);
      fx_inv = ldexp(fx_inv, ex);
    }
  }
  const float floor_term = grid_floor(fx_inv);
  // a batch of event vectors: vector blockIdx.y of [B][m] into [B][n]
  ev += (int64_t)blockIdx.y * m;
  out += (int64_t)blockIdx.y * n;
  // rows to walk: the listed active rows, or every row
  int64_t nrows = m;
  if constexpr (LISTED) {
    nrows = lcount[blockIdx.y];
    lrows += (int64_t)blockIdx.y * m;
  }
  // XCD x (blocks are dispatched to the XCDs round-robin) takes the
  // (copy, tile) pairs [x * ncopy * P / 8, (x + 1) * ncopy * P / 8): whole
  // copies while ncopy >= 8, else a range of one copy's tiles, so the blocks
  // of an XCD read the same rows' events and the same part of their
  // row-tile offsets through its L2
  const int b = blockIdx.x;
  const int x = b & 7;
  const int jb = b >> 3;
  const int pair = x * (ncopy * P / 8) + jb;
  const int p = pair % P;              // output tile
  const int copy = pair / P;           // output copy
  const int64_t t0 = (int64_t)p * T;
  const int tlen = (int)((n - t0) < (int64_t)T ? (n - t0) : (int64_t)T);
  for (int k = threadIdx.x; k < T; k += kTiledThreads) s_acc[k] = 0ull;
  __syncthreads();
  float hw = 0.f;
  if constexpr (WK == 0) hw = *homo_w;
  const long long hfx = fixed_of(hw, fx_scale);
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  constexpr int GV = 32 / kVec;         // lanes per row segment, kVec elements each
  constexpr int NGRP = 64 / GV;         // lane groups per wave
  const int g = lane / GV, lv = lane % GV;
  int2* seg = s_seg[wave];
  const int64_t nchunks = (nrows + 63) >> 6;
  const int64_t cstep = (int64_t)ncopy * kWavesPerBlock;    // chunk stride of one wave
  int64_t q0 = copy + (int64_t)ncopy * wave;                // first chunk of this wave

  // loads are issued for every lane (rows past the end read row m-1 and are
  // masked), so the four chunks' loads are in flight together.  LISTED:
  // position k of the list holds the row; rid keeps it for the offsets
  auto load_batch = [&](int64_t qb, bool* act, float* xv, int32_t* rid) {
    if constexpr (LISTED) {
#pragma unroll
      for (int u = 0; u < kChunksPerBatch; ++u) {
        const int64_t k = ((qb + cstep * u) << 6) + lane;
        act[u] = k < nrows;
        rid[u] = act[u] ? lrows[k] : 0;
      }
      if constexpr (DENSE) {
#pragma unroll
        for (int u = 0; u < kChunksPerBatch; ++u) xv[u] = (float)ev[rid[u]];
      }
    } else {
      E v[kChunksPerBatch];
#pragma unroll
      for (int u = 0; u < kChunksPerBatch; ++u) {
        const int64_t r = ((qb + cstep * u) << 6) + lane;
        v[u] = ev[r < m ? r : m - 1];
      }
#pragma unroll
      for (int u = 0; u < kChunksPerBatch; ++u) {
        const int64_t r = ((qb + cstep * u) << 6) + lane;
        act[u] = r < m && ev_true(v[u], rule);
        if constexpr (DENSE) xv[u] = (float)v[u];
      }
    }
  };
  bool act[kChunksPerBatch];
  float xv[kChunksPerBatch];
  int32_t rid[kChunksPerBatch];
  load_batch(q0, act, xv, rid);
  while (q0 < nchunks) {
    // list the batch's active rows with their tile-p segments
    // (inactive lanes read row 0's pair: one shared line, no branch, so the
    // four chunks' loads are in flight together)
    int cnt = 0;
    bool any = false;
#pragma unroll
    for (int u = 0; u < kChunksPerBatch; ++u) any = any || act[u];
    if (__ballot(any) != 0ull) {   // wave-uniform: a batch without active rows skips the offset loads
      int2 sgl[kChunksPerBatch];
#pragma unroll
      for (int u = 0; u < kChunksPerBatch; ++u) {
        const int64_t r = !act[u] ? 0 : LISTED ? (int64_t)rid[u] : ((q0 + cstep * u) << 6) + lane;
        const int32_t* rt = rowtile + r * (P + 1) + p;
        sgl[u] = make_int2(rt[0], rt[1]);
      }
#pragma unroll
      for (int u = 0; u < kChunksPerBatch; ++u) {
        const bool nz = act[u] && sgl[u].y > sgl[u].x;
        const uint64_t mask = __ballot(nz);
        if (nz) {
          seg[cnt + lanes_below(mask)] = sgl[u];
          if constexpr (DENSE) s_xv[wave][cnt + lanes_below(mask)] = xv[u];
        }
        cnt += __popcll(mask);
      }
    }
    // prefetch the next batch's events
    const int64_t qn = q0 + cstep * kChunksPerBatch;
    bool nact[kChunksPerBatch];
    float nxv[kChunksPerBatch];
    int32_t nrid[kChunksPerBatch];
    load_batch(qn, nact, nxv, nrid);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // each group of GV lanes takes R listed rows per round; a lane fetches
    // kVec indices and kVec weights (one vector load each) per pass
    for (int e = g; e < cnt; e += R * NGRP) {
      int2 sg[R];
      float xr[R];
      long long hq[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        sg[q] = (e + q * NGRP < cnt) ? seg[e + q * NGRP] : make_int2(0, 0);
        xr[q] = 1.f;
        if constexpr (DENSE) xr[q] = (e + q * NGRP < cnt) ? s_xv[wave][e + q * NGRP] : 0.f;
        hq[q] = hfx;
        if constexpr (DENSE && WK == 0) hq[q] = fixed_of(hw * xr[q], fx_scale);
      }
      int32_t len = 0;
#pragma unroll
      for (int q = 0; q < R; ++q) len = max(len, sg[q].y - sg[q].x);
      for (int32_t k = kVec * lv; k < len; k += kVec * GV) {
        constexpr int NW = kVec / 2;    // 32-bit words of kVec u16 indices
        uint32_t li[R][NW];
        float wv[R][kVec];
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int32_t j = sg[q].x + k;
#pragma unroll
          for (int u = 0; u < NW; ++u) li[q][u] = 0xFFFFFFFFu;
          if (j < sg[q].y) {
            if constexpr (kVec == 8) {
              const uint4 a = *reinterpret_cast<const uint4*>(lidx + j);
              li[q][0] = a.x; li[q][1] = a.y; li[q][2] = a.z; li[q][3] = a.w;
            } else {
              const uint2 a = *reinterpret_cast<const uint2*>(lidx + j);
              li[q][0] = a.x; li[q][1] = a.y;
            }
            if constexpr (WK == 1) {
#pragma unroll
              for (int h = 0; h < kVec / 4; ++h) {
                const float4 a = *reinterpret_cast<const float4*>((const float*)wptr + j + 4 * h);
                wv[q][4 * h] = a.x; wv[q][4 * h + 1] = a.y; wv[q][4 * h + 2] = a.z; wv[q][4 * h + 3] = a.w;
              }
            } else if constexpr (WK == 2) {
              uint32_t hw4[NW];
              if constexpr (kVec == 8) {
                const uint4 h = *reinterpret_cast<const uint4*>((const bf16_t*)wptr + j);
                hw4[0] = h.x; hw4[1] = h.y; hw4[2] = h.z; hw4[3] = h.w;
              } else {
                const uint2 h = *reinterpret_cast<const uint2*>((const bf16_t*)wptr + j);
                hw4[0] = h.x; hw4[1] = h.y;
              }
#pragma unroll
              for (int u = 0; u < NW; ++u) {
                wv[q][2 * u] = __uint_as_float(hw4[u] << 16);
                wv[q][2 * u + 1] = __uint_as_float(hw4[u] & 0xFFFF0000u);
              }
            }
          }
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const uint32_t* l4 = li[q];
#pragma unroll
          for (int u = 0; u < kVec; ++u) {
            const uint32_t c = (l4[u >> 1] >> (16 * (u & 1))) & 0xFFFFu;
            if (c < (uint32_t)tlen) {
              if constexpr (DENSE || COARSE) {
                // on the grid when at least its floor; else (small, inf or
                // NaN) an FP32 atomic to the output
                float t = (WK == 0) ? hw : wv[q][u];
                if constexpr (DENSE) t *= xr[q];
                const float at = fabsf(t);
                if (at >= floor_term && at <= 3.402823466e38f) {
                  const long long fx = (WK == 0 && DENSE) ? hq[q] : (WK == 0) ? hfx : fixed_of(t, fx_scale);
                  __hip_atomic_fetch_add(s_acc + c, (unsigned long long)fx, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
                } else if (t != 0.f) {
                  atomic_add(out + t0 + c, t);
                }
              } else {
                const long long fx = (WK == 0) ? hfx : fixed_of(wv[q][u], fx_scale);
                __hip_atomic_fetch_add(s_acc + c, (unsigned long long)fx, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
              }
            }
          }
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < kChunksPerBatch; ++u) {
      act[u] = nact[u];
      xv[u] = nxv[u];
      rid[u] = nrid[u];
    }
    q0 = qn;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < tlen; k += kTiledThreads) {
    const long long a = (long long)s_acc[k];
    if (a != 0) atomic_add(out + t0 + k, (float)((double)a * fx_inv));
  }
}

// ---------------------------------------------------------------------------
// gather on the plan: y[m] = A @ v (bm.sparse.csrmv / bm.event.csrmv with
// transpose=False).  Block (row chunk, tile p) stages v's tile p (T values,
// or the events as 0/1 by the rule) in LDS, and groups of 8 lanes walk the
// tile-p segment of each row of the chunk (4 tile-local indices and weights
// per lane and pass), each lane summing w * v[c] in element order, the 8
// lanes reduced by a fixed tree; the row's partial of tile p goes to
// partial[p][row], and k_tiled_gather_sum adds the P partials of a row in
// tile order.  The matrix is read once at 2 + w bytes per synapse and every
// v read is an LDS read; the result is deterministic.  Blocks of one row
// chunk take consecutive indices (p fastest), so the tiles read the
// segments of the same rows at about the same time.
// XM: 0 dense f32 v, 1 bool events, 2 f32 events (ev_rule)
// ---------------------------------------------------------------------------
constexpr int kGatherRows = 256;    // rows per block

template <int WK, int XM>
__global__ __launch_bounds__(256) void k_tiled_gather(const int32_t* __restrict__ rowtile,
                                                      const uint16_t* __restrict__ lidx,
                                                      const void* __restrict__ wptr, const float* __restrict__ homo_w,
                                                      const void* __restrict__ v, int rule, int64_t m, int64_t n,
                                                      int P, int T, float* __restrict__ partial) {
  extern __shared__ float s_v[];
  const int p = (int)(blockIdx.x % (unsigned)P);
  const int64_t chunk = blockIdx.x / (unsigned)P;
  const int64_t t0 = (int64_t)p * T;
  const int tlen = (int)((n - t0) < (int64_t)T ? (n - t0) : (int64_t)T);
  for (int k = threadIdx.x; k < tlen; k += 256) {
    float x;
    if constexpr (XM == 0) x = ((const float*)v)[t0 + k];
    else if constexpr (XM == 1) x = ((const uint8_t*)v)[t0 + k] ? 1.f : 0.f;
    else x = ev_true(((const float*)v)[t0 + k], rule) ? 1.f : 0.f;
    s_v[k] = x;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int lv = lane & 7;                       // lane within its group of 8
  const int gbase = lane & ~7;                   // the group's first lane
  const int grp = threadIdx.x >> 3;              // 32 groups per block
  const int64_t r0 = chunk * kGatherRows;
  const int64_t r_end = min(r0 + kGatherRows, m);
  float hw = 0.f;
  if constexpr (WK == 0) hw = *homo_w;
  // the group's 8 rows r0 + grp + 32k (k = 0..7): lane k loads row k's
  // segment bounds, so all 8 are one memory round trip
  int32_t sb = 0, se = 0;
  {
    const int64_t row = r0 + grp + 32 * lv;
    if (row < r_end) {
      const int32_t* rt = rowtile + row * (P + 1) + p;
      sb = rt[0];
      se = rt[1];
    }
  }
  auto term = [&](uint32_t c, float wv, float& acc) {
    // padding entries carry 0xFFFF, past every tile
    if (c < (uint32_t)tlen) {
      if constexpr (WK == 0) acc += s_v[c];
      else acc += wv * s_v[c];
    }
  };
  auto load_w = [&](int32_t j, float (&w4)[4]) {
    if constexpr (WK == 1) {
      const float4 f = *reinterpret_cast<const float4*>((const float*)wptr + j);
      w4[0] = f.x; w4[1] = f.y; w4[2] = f.z; w4[3] = f.w;
    } else if constexpr (WK == 2) {
      const uint2 h = *reinterpret_cast<const uint2*>((const bf16_t*)wptr + j);
      w4[0] = __uint_as_float(h.x << 16); w4[1] = __uint_as_float(h.x & 0xFFFF0000u);
      w4[2] = __uint_as_float(h.y << 16); w4[3] = __uint_as_float(h.y & 0xFFFF0000u);
    } else {
      w4[0] = w4[1] = w4[2] = w4[3] = 1.f;
    }
  };
  constexpr int RB = 4;                          // rows whose first pass is in flight together
#pragma unroll
  for (int k0 = 0; k0 < 8; k0 += RB) {
    int32_t s[RB], e[RB];
    uint2 a[RB];
    float w4[RB][4];
#pragma unroll
    for (int q = 0; q < RB; ++q) {
      s[q] = __shfl(sb, gbase + k0 + q, 64);
      e[q] = __shfl(se, gbase + k0 + q, 64);
      const int32_t j = s[q] + 4 * lv;
      const bool ok = j < e[q];
      const int32_t js = ok ? j : 0;
      a[q] = *reinterpret_cast<const uint2*>(lidx + js);
      load_w(js, w4[q]);
      if (!ok) a[q] = make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
    }
#pragma unroll
    for (int q = 0; q < RB; ++q) {
      const int64_t row = r0 + grp + 32 * (k0 + q);
      float acc = 0.f;
      term(a[q].x & 0xFFFFu, w4[q][0], acc);
      term(a[q].x >> 16, w4[q][1], acc);
      term(a[q].y & 0xFFFFu, w4[q][2], acc);
      term(a[q].y >> 16, w4[q][3], acc);
      // the rest of a segment longer than 32 entries
      for (int32_t j = s[q] + 32 + 4 * lv; j < e[q]; j += 32) {
        const uint2 a2 = *reinterpret_cast<const uint2*>(lidx + j);
        float v4[4];
        load_w(j, v4);
        term(a2.x & 0xFFFFu, v4[0], acc);
        term(a2.x >> 16, v4[1], acc);
        term(a2.y & 0xFFFFu, v4[2], acc);
        term(a2.y >> 16, v4[3], acc);
      }
      // fixed reduction tree over the group's 8 lanes
      acc += __shfl_xor(acc, 1, 64);
      acc += __shfl_xor(acc, 2, 64);
      acc += __shfl_xor(acc, 4, 64);
      if constexpr (WK == 0) acc = hw * acc;
      if (lv == 0 && row < r_end) partial[(int64_t)p * m + row] = acc;
    }
  }
}

// The same partials from blocks that each stage a range of NT consecutive
// column tiles of v in LDS (up to ~150 KB) and walk, for their slice of the
// rows, every (row, tile) segment of the range: a wave takes the items
// (row, tile) in row-major order, 64 at a time (lane l loads item l's
// segment bounds), and group g of 8 lanes takes items g, g + 8, ..., so the
// 8 groups of a round read consecutive tiles' segments of one row, one
// contiguous span of the plan, RB rounds' loads in flight together.  Each
// (tile, row) partial is summed as in k_tiled_gather (entries in lane
// order, the fixed 8-lane tree), so the results are bit-identical to it.
// Block b takes range b % NR and row slice b / NR.
constexpr int kGatherRangeThreads = 1024;
#ifndef BPX_GATHER_RANGE
#define BPX_GATHER_RANGE 1
#endif

template <int WK, int XM>
__global__ __launch_bounds__(kGatherRangeThreads) void k_tiled_gather_rng(
    const int32_t* __restrict__ rowtile, const uint16_t* __restrict__ lidx, const void* __restrict__ wptr,
    const float* __restrict__ homo_w, const void* __restrict__ v, int rule, int64_t m, int64_t n, int P, int T,
    int NT, int NR, int64_t rows_per_slice, float* __restrict__ partial) {
  extern __shared__ float s_v[];
  const int k = (int)(blockIdx.x % (unsigned)NR);
  const int64_t slice = blockIdx.x / (unsigned)NR;
  const int ta = k * NT;                                   // the range's first tile
  const int nt = min(NT, P - ta);                          // its tiles
  const int64_t c0 = (int64_t)ta * T;
  const int64_t clen = min((int64_t)nt * T, n - c0);       // its columns
  for (int64_t c = threadIdx.x; c < clen; c += kGatherRangeThreads) {
    float x;
    if constexpr (XM == 0) x = ((const float*)v)[c0 + c];
    else if constexpr (XM == 1) x = ((const uint8_t*)v)[c0 + c] ? 1.f : 0.f;
    else x = ev_true(((const float*)v)[c0 + c], rule) ? 1.f : 0.f;
    s_v[c] = x;
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int lv = lane & 7;
  const int g = lane >> 3;
  const int64_t r0 = slice * rows_per_slice;
  const int64_t r1 = min(r0 + rows_per_slice, m);
  if (r0 >= r1) return;
  float hw = 0.f;
  if constexpr (WK == 0) hw = *homo_w;
  auto load_w = [&](int32_t j, float (&w4)[4]) {
    if constexpr (WK == 1) {
      const float4 f = *reinterpret_cast<const float4*>((const float*)wptr + j);
      w4[0] = f.x; w4[1] = f.y; w4[2] = f.z; w4[3] = f.w;
    } else if constexpr (WK == 2) {
      const uint2 h = *reinterpret_cast<const uint2*>((const bf16_t*)wptr + j);
      w4[0] = __uint_as_float(h.x << 16); w4[1] = __uint_as_float(h.x & 0xFFFF0000u);
      w4[2] = __uint_as_float(h.y << 16); w4[3] = __uint_as_float(h.y & 0xFFFF0000u);
    } else {
      w4[0] = w4[1] = w4[2] = w4[3] = 1.f;
    }
  };
  // tile-local column c of tile (ta + tl) sits at s_v[tl * T + c]; padding
  // entries (0xFFFF) and columns past the tile's end are skipped
  auto term = [&](uint32_t c, int tl, int tlen, float wv, float& acc) {
    if (c < (uint32_t)tlen) {
      if constexpr (WK == 0) acc += s_v[tl * T + c];
      else acc += wv * s_v[tl * T + c];
    }
  };
  const int64_t nitems = (r1 - r0) * nt;
  constexpr int RB = 4;
  for (int64_t i0 = (int64_t)wave * 64; i0 < nitems; i0 += (int64_t)(kGatherRangeThreads / 64) * 64) {
    // lane l: bounds of item i0 + l
    int32_t sb = 0, se = 0;
    {
      const int64_t it = i0 + lane;
      if (it < nitems) {
        const int64_t row = r0 + it / nt;
        const int32_t* rt = rowtile + row * (P + 1) + ta + (int)(it % nt);
        sb = rt[0];
        se = rt[1];
      }
    }
#pragma unroll
    for (int q0 = 0; q0 < 8; q0 += RB) {
      // the first 64 entries of each segment (segments average n / P per
      // row; 1000 / 32 here) in flight together, the rest after
      int32_t s_[RB], e_[RB];
      uint2 a[RB][2];
      float w4[RB][2][4];
#pragma unroll
      for (int q = 0; q < RB; ++q) {
        const int src = 8 * (q0 + q) + g;                  // the group's item of round q0 + q
        s_[q] = __shfl(sb, src, 64);
        e_[q] = __shfl(se, src, 64);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int32_t j = s_[q] + 32 * h + 4 * lv;
          const bool ok = j < e_[q];
          const int32_t js = ok ? j : 0;
          a[q][h] = *reinterpret_cast<const uint2*>(lidx + js);
          load_w(js, w4[q][h]);
          if (!ok) a[q][h] = make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
        }
      }
#pragma unroll
      for (int q = 0; q < RB; ++q) {
        const int64_t it = i0 + 8 * (q0 + q) + g;
        const int tl = (int)(it % nt);
        const int64_t row = r0 + it / nt;
        const int tlen = (int)min((int64_t)T, n - (int64_t)(ta + tl) * T);
        float acc = 0.f;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          term(a[q][h].x & 0xFFFFu, tl, tlen, w4[q][h][0], acc);
          term(a[q][h].x >> 16, tl, tlen, w4[q][h][1], acc);
          term(a[q][h].y & 0xFFFFu, tl, tlen, w4[q][h][2], acc);
          term(a[q][h].y >> 16, tl, tlen, w4[q][h][3], acc);
        }
        for (int32_t j = s_[q] + 64 + 4 * lv; j < e_[q]; j += 32) {
          const uint2 a2 = *reinterpret_cast<const uint2*>(lidx + j);
          float v4[4];
          load_w(j, v4);
          term(a2.x & 0xFFFFu, tl, tlen, v4[0], acc);
          term(a2.x >> 16, tl, tlen, v4[1], acc);
          term(a2.y & 0xFFFFu, tl, tlen, v4[2], acc);
          term(a2.y >> 16, tl, tlen, v4[3], acc);
        }
        acc += __shfl_xor(acc, 1, 64);
        acc += __shfl_xor(acc, 2, 64);
        acc += __shfl_xor(acc, 4, 64);
        if constexpr (WK == 0) acc = hw * acc;
        if (lv == 0 && it < nitems) partial[(int64_t)(ta + tl) * m + row] = acc;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_tiled_gather_sum(const float* __restrict__ partial, int64_t m, int P,
                                                          float* __restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= m) return;
  float acc = 0.f;
  // the loads of 8 tiles in flight together, the adds in tile order
  int p = 0;
  for (; p + 8 <= P; p += 8) {
    float t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = partial[(int64_t)(p + u) * m + r];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += t[u];
  }
  for (; p < P; ++p) acc += partial[(int64_t)p * m + r];
  out[r] = acc;
}

template <typename O>
__global__ __launch_bounds__(256) void k_fill_zero_t(O* __restrict__ out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = O(0);
}

// zero-fill of the output (n floats, none when accumulating) and of the
// list counters (nc ints) in one launch
__global__ __launch_bounds__(256) void k_fill_zero_out_counts(float* __restrict__ out, int64_t n,
                                                              int32_t* __restrict__ cnt, int64_t nc) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = i; k < nc; k += stride) cnt[k] = 0;
  for (; i < n; i += stride) out[i] = 0.f;
}

// the active rows of event vector blockIdx.y (ev_true by rule; for a dense
// vector the rule != 0) appended to lrows[y][...]: a block takes 4096 rows
// (4 per thread), scans its counts and takes its place in the list with one
// counter atomic (a counter shared by every wave serialises its atomics:
// ~11k of them cost ~130 us on a million rows); blocks land in arrival order
constexpr int kListThreads = 1024;
constexpr int kListRows = 4 * kListThreads;

template <typename E>
__global__ __launch_bounds__(kListThreads) void k_tiled_list(const E* __restrict__ ev, int rule, int64_t m,
                                                             int32_t* __restrict__ lcount,
                                                             int32_t* __restrict__ lrows) {
  __shared__ int32_t s_wave[kListThreads / 64];
  __shared__ int32_t s_base;
  ev += (int64_t)blockIdx.y * m;
  lrows += (int64_t)blockIdx.y * m;
  const int64_t r0 = (int64_t)blockIdx.x * kListRows + 4 * (int64_t)threadIdx.x;
  uint32_t on = 0;
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (r0 + u < m && ev_true(ev[r0 + u], rule)) on |= 1u << u;
  const int32_t cnt = __popc(on);
  int32_t incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(incl, o, 64);
    if (lane_id() >= (uint32_t)o) incl += y;
  }
  const int wv = threadIdx.x >> 6;
  if (lane_id() == 63) s_wave[wv] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t tot = 0;
    for (int q = 0; q < kListThreads / 64; ++q) {
      const int32_t t = s_wave[q];
      s_wave[q] = tot;
      tot += t;
    }
    s_base = tot > 0 ? atomicAdd(lcount + blockIdx.y, tot) : 0;
  }
  __syncthreads();
  int32_t pos = s_base + s_wave[wv] + incl - cnt;
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (on & (1u << u)) lrows[pos++] = (int32_t)(r0 + u);
}

// output copies of a call: two per XCD up to 64 tiles (two 1024-thread
// blocks per CU hide the memory latency of the row walk), then fewer as the
// tiles grow, so that a call keeps ~1024 blocks and its flush (one atomic
// per output and copy) stays a small part of its traffic; ncopy * P is a
// multiple of 8 (P > 64 is rounded up to a multiple of 8)
static int tiled_copies(int P) {
  if (P <= kMaxTiles) return 16;
  if (P <= 128) return 8;
  if (P <= 256) return 4;
  if (P <= 512) return 2;
  return 1;
}


struct TiledShape {
  int P, T;
};


static int max_tile_len();

// one tile per CU of an XCD, or more tiles when a tile would not fit the
// block's LDS (beyond kMaxTiles a multiple of 8, at most kMaxTilesBig)
static TiledShape tiled_shape(int64_t n) {
  int cus = device_cus();
  int P = cus / 8;
  if (P < 1) P = 1;
  const int64_t tmax = max_tile_len();
  if (tmax > 0 && (n + P - 1) / P > tmax) {
    const int64_t pn = (n + tmax - 1) / tmax;
    P = (int)(pn > kMaxTilesBig + 1 ? kMaxTilesBig + 1 : pn);
  }
  if (P > kMaxTiles) P = (P + 7) & ~7;
  int64_t T = (n + P - 1) / P;
  if (T < 1) T = 1;
  return {P, (int)T};
}

// the device's LDS per block (bytes)
static int shm_bytes() {
  static int shm = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
      return 65536;
    return v;
  }();
  return shm;
}

// largest tile the device's per-block LDS holds (floats), capped by the
// 16-bit tile-local index
static int max_tile_len() {
  int dev = 0, shm = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 16384;
  if (hipDeviceGetAttribute(&shm, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || shm <= 0)
    return 16384;
  // the segment lists of the exec kernel take kWavesPerBlock * kListCap int2
  // of static LDS; keep the whole block within 64 KiB
  if (shm > 65536) shm = 65536;
  int t = (shm - (int)(kWavesPerBlock * kListCap * sizeof(int2))) / (int)sizeof(unsigned long long);
  return t < BPX_TILED_MAX_TILE ? t : BPX_TILED_MAX_TILE;
}

}  // namespace bpx

using namespace bpx;

extern "C" {

int bpx_tiled_csr_layout(int64_t m, int64_t n, int64_t nnz, int weight_dtype, bpx_tiled_csr* plan,
                         size_t* bytes) {
  if (!plan || !bytes) return BPX_ERR_ARG;
  if (m < 0 || n < 1 || nnz < 0 || nnz > INT32_MAX) return BPX_ERR_SHAPE;
  if (weight_dtype != BPX_F32 && weight_dtype != BPX_BF16 && weight_dtype != BPX_TILED_HOMO) return BPX_ERR_DTYPE;
  TiledShape ts = tiled_shape(n);
  if (ts.P > kMaxTilesBig || ts.T > max_tile_len()) return BPX_ERR_SHAPE;
  // beyond kMaxTiles tiles a row's offsets (4 (P + 1) bytes) and read of
  // them per active row and tile grow with P: refuse rows shorter than P / 4
  // synapses on average, for which the stateless kernels do as well
  if (ts.P > kMaxTiles && m > 0 && nnz * 4 < m * (int64_t)ts.P) return BPX_ERR_SHAPE;
  plan->m = m;
  plan->n = n;
  plan->nnz = nnz;
  plan->tiles = ts.P;
  plan->tile_len = ts.T;
  plan->weight_dtype = weight_dtype;
  plan->coarse = 0;
  plan->reserved = 0;
  size_t off = 0;
  auto take = [&](size_t b) {
    size_t o = off;
    off += (b + 255) & ~(size_t)255;
    return o;
  };
  // every (row, tile) segment may gain up to kVec-1 padding entries
  const int64_t cap = nnz + (int64_t)(kVec - 1) * m * ts.P;
  if (cap > INT32_MAX) return BPX_ERR_SHAPE;
  plan->capacity = cap;
  size_t o_rt = take((size_t)m * (ts.P + 1) * sizeof(int32_t));
  size_t o_aux = take((size_t)(m + 1) * sizeof(int32_t));
  size_t o_li = take((size_t)cap * sizeof(uint16_t));
  size_t wb = weight_dtype == BPX_F32 ? 4 : weight_dtype == BPX_BF16 ? 2 : 0;
  size_t o_w = take((size_t)cap * wb);
  size_t o_h = take(4 * sizeof(float));     // homogeneous weight, max |w|, min non-zero |w|
  plan->aux_offset = (int64_t)o_aux;
  plan->rowtile_offset = (int64_t)o_rt;
  plan->lidx_offset = (int64_t)o_li;
  plan->weights_offset = (int64_t)o_w;
  plan->homo_offset = (int64_t)o_h;
  *bytes = off;
  return BPX_OK;
}

int bpx_tiled_csr_build(void* stream, const void* data, int data_dtype, int64_t data_len, const void* indices,
                        const void* indptr, int index_dtype, bpx_tiled_csr* plan, void* mem, size_t mem_bytes) {
  if (!plan || !mem) return BPX_ERR_ARG;
  size_t need = 0;
  bpx_tiled_csr chk = *plan;
  int rc = bpx_tiled_csr_layout(plan->m, plan->n, plan->nnz, plan->weight_dtype, &chk, &need);
  if (rc != BPX_OK) return rc;
  if (mem_bytes < need) return BPX_ERR_WORKSPACE;
  *plan = chk;
  plan->mem = mem;
  const int64_t m = plan->m, n = plan->n, nnz = plan->nnz;
  const bool homo = plan->weight_dtype == BPX_TILED_HOMO;
  if (homo ? data_len != 1 : data_len != nnz) return BPX_ERR_SHAPE;
  if (data_dtype != BPX_F32 && data_dtype != BPX_F64 && data_dtype != BPX_BF16) return BPX_ERR_DTYPE;
  hipStream_t s = (hipStream_t)stream;
  char* base = (char*)mem;
  int32_t* rowtile = (int32_t*)(base + plan->rowtile_offset);
  int32_t* aux = (int32_t*)(base + plan->aux_offset);
  uint16_t* lidx = (uint16_t*)(base + plan->lidx_offset);
  void* w = base + plan->weights_offset;
  float* hw = (float*)(base + plan->homo_offset);
  const DivU32 dT = DivU32::make((uint32_t)plan->tile_len);
  const int P = plan->tiles;
  const unsigned blocks = (unsigned)((m + kBuildThreads - 1) / kBuildThreads);
  if (homo) hipLaunchKernelGGL(k_store_scalar, dim3(1), dim3(64), 0, s, data, data_dtype, hw);
  // fixed-point scale of the LDS accumulators: |sum| <= max|w| * m < 2^62
  {
    uint32_t* amax = (uint32_t*)(base + plan->homo_offset + sizeof(float));
    hipLaunchKernelGGL(k_init_minmax, dim3(1), dim3(64), 0, s, amax);
    const int64_t len = homo ? 1 : nnz;
    int64_t ab = (len + 255) / 256;
    if (ab > 1024) ab = 1024;
    if (ab < 1) ab = 1;
    if (data_dtype == BPX_F64) hipLaunchKernelGGL(k_absmax<double>, dim3((unsigned)ab), dim3(256), 0, s, (const double*)data, len, amax);
    else if (data_dtype == BPX_BF16) hipLaunchKernelGGL(k_absmax<bf16_t>, dim3((unsigned)ab), dim3(256), 0, s, (const bf16_t*)data, len, amax);
    else hipLaunchKernelGGL(k_absmax<float>, dim3((unsigned)ab), dim3(256), 0, s, (const float*)data, len, amax);
    uint32_t bits[2] = {0u, 0u};
    if (hipMemcpyAsync(bits, amax, sizeof(bits), hipMemcpyDeviceToHost, s) != hipSuccess) return BPX_ERR_LAUNCH;
    if (hipStreamSynchronize(s) != hipSuccess) return BPX_ERR_LAUNCH;
    float mx, mn;
    memcpy(&mx, &bits[0], sizeof(float));
    memcpy(&mn, &bits[1], sizeof(float));
    if (!(mx == mx) || mx == INFINITY) return BPX_ERR_ARG;      // non-finite weights
    double bound = (double)mx * (double)(m > 1 ? m : 1);
    int e = bound > 0.0 ? (int)ceil(log2(bound)) : -100;
    int shift = 61 - e;
    // every term |w| * 2^shift below 2^50 (the exec kernel's conversion is
    // exact there); binds only for m < 2^11
    const int cap = 50 - (mx > 0.f ? (int)ceil(log2((double)mx)) : -100);
    if (shift > cap) shift = cap;
    if (shift > 200) shift = 200;
    plan->fx_shift = shift;
    // the grid's unit against the smallest non-zero weight
    plan->coarse = (mn <= 3.402823466e38f && ldexp(1.0, -shift) > 2e-6 * (double)mn) ? 1 : 0;
  }
  if (m == 0) {
    BPX_CHECK_LAUNCH();
    return BPX_OK;
  }
#define FILL_BIG(I, V, W)                                                                                     \
  do {                                                                                                        \
    hipLaunchKernelGGL((k_tiled_prefill<W>), dim3(2048), dim3(256), 0, s, (const int32_t*)(aux + m), homo ? 0 : 1, \
                       lidx, (W*)w);                                                                          \
    hipLaunchKernelGGL((k_tiled_fill_big<I, V, W>), dim3(bblocks), dim3(tb), bshm, s, (const V*)data,         \
                       homo ? 0 : 1, (const I*)indices, (const I*)indptr, nnz, m, n, P, dT, (const int32_t*)rowtile, \
                       lidx, (W*)w);                                                                          \
  } while (0)
#define BUILD_BIG(I)                                                                                          \
  do {                                                                                                        \
    const int tb = big_build_threads(P);                                                                      \
    const unsigned bblocks = (unsigned)((m + tb - 1) / tb);                                                   \
    const size_t bshm = (size_t)tb * P * sizeof(int32_t);                                                     \
    hipLaunchKernelGGL(k_tiled_count_big<I>, dim3(bblocks), dim3(tb), bshm, s, (const I*)indices,             \
                       (const I*)indptr, nnz, m, n, P, dT, rowtile, aux);                                     \
    hipLaunchKernelGGL(k_scan_rows, dim3(1), dim3(1024), 0, s, aux, m);                                       \
    hipLaunchKernelGGL(k_tiled_offsets_big, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, s,                   \
                       (const int32_t*)aux, m, P, rowtile);                                                   \
    const bool wf = homo || plan->weight_dtype == BPX_F32;                                                    \
    if (data_dtype == BPX_F64) {                                                                              \
      if (wf) FILL_BIG(I, double, float); else FILL_BIG(I, double, bf16_t);                                   \
    } else if (data_dtype == BPX_BF16) {                                                                      \
      if (wf) FILL_BIG(I, bf16_t, float); else FILL_BIG(I, bf16_t, bf16_t);                                   \
    } else {                                                                                                  \
      if (wf) FILL_BIG(I, float, float); else FILL_BIG(I, float, bf16_t);                                     \
    }                                                                                                         \
  } while (0)
#define BUILD_IDX(I)                                                                                           \
  do {                                                                                                         \
    hipLaunchKernelGGL(k_tiled_count<I>, dim3(blocks), dim3(kBuildThreads), 0, s, (const I*)indices,          \
                       (const I*)indptr, nnz, m, n, P, dT, rowtile, aux);                                      \
    hipLaunchKernelGGL(k_scan_rows, dim3(1), dim3(1024), 0, s, aux, m);                                        \
    hipLaunchKernelGGL(k_tiled_offsets, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, (const int32_t*)aux, \
                       m, P, rowtile);                                                                         \
    if (homo || plan->weight_dtype == BPX_F32) {                                                               \
      if (data_dtype == BPX_F64)                                                                               \
        hipLaunchKernelGGL((k_tiled_fill<I, double, float>), dim3(blocks), dim3(kBuildThreads), 0, s,          \
                           (const double*)data, homo ? 0 : 1, (const I*)indices, (const I*)indptr, nnz, m, n, P, \
                           dT, rowtile, lidx, (float*)w);                                                      \
      else if (data_dtype == BPX_BF16)                                                                         \
        hipLaunchKernelGGL((k_tiled_fill<I, bf16_t, float>), dim3(blocks), dim3(kBuildThreads), 0, s,          \
                           (const bf16_t*)data, homo ? 0 : 1, (const I*)indices, (const I*)indptr, nnz, m, n, P, \
                           dT, rowtile, lidx, (float*)w);                                                      \
      else                                                                                                     \
        hipLaunchKernelGGL((k_tiled_fill<I, float, float>), dim3(blocks), dim3(kBuildThreads), 0, s,           \
                           (const float*)data, homo ? 0 : 1, (const I*)indices, (const I*)indptr, nnz, m, n, P, \
                           dT, rowtile, lidx, (float*)w);                                                      \
    } else {                                                                                                   \
      if (data_dtype == BPX_F64)                                                                               \
        hipLaunchKernelGGL((k_tiled_fill<I, double, bf16_t>), dim3(blocks), dim3(kBuildThreads), 0, s,         \
                           (const double*)data, 1, (const I*)indices, (const I*)indptr, nnz, m, n, P, dT,       \
                           rowtile, lidx, (bf16_t*)w);                                                          \
      else if (data_dtype == BPX_BF16)                                                                         \
        hipLaunchKernelGGL((k_tiled_fill<I, bf16_t, bf16_t>), dim3(blocks), dim3(kBuildThreads), 0, s,         \
                           (const bf16_t*)data, 1, (const I*)indices, (const I*)indptr, nnz, m, n, P, dT,       \
                           rowtile, lidx, (bf16_t*)w);                                                          \
      else                                                                                                     \
        hipLaunchKernelGGL((k_tiled_fill<I, float, bf16_t>), dim3(blocks), dim3(kBuildThreads), 0, s,          \
                           (const float*)data, 1, (const I*)indices, (const I*)indptr, nnz, m, n, P, dT,        \
                           rowtile, lidx, (bf16_t*)w);                                                          \
    }                                                                                                          \
  } while (0)
  if (index_dtype != BPX_I32 && index_dtype != BPX_I64) return BPX_ERR_DTYPE;
  if (P > kMaxTiles) {
    if (index_dtype == BPX_I32) BUILD_BIG(int32_t);
    else BUILD_BIG(int64_t);
  } else {
    if (index_dtype == BPX_I32) BUILD_IDX(int32_t);
    else BUILD_IDX(int64_t);
  }
#undef BUILD_IDX
#undef BUILD_BIG
#undef FILL_BIG
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

int bpx_event_csrmv_tiled(void* stream, const bpx_tiled_csr* plan, const void* events, int ev_dtype, int ev_rule,
                          int accumulate, float* out) {
  return bpx_event_csrmv_tiled_batch(stream, plan, events, ev_dtype, ev_rule, 1, accumulate, out);
}

int bpx_event_csrmv_tiled_batch(void* stream, const bpx_tiled_csr* plan, const void* events, int ev_dtype,
                                int ev_rule, int64_t batch, int accumulate, float* out) {
  return bpx_event_csrmv_tiled_ws(stream, plan, events, ev_dtype, ev_rule, batch, accumulate, out, nullptr, 0);
}

// list counters at the start (a 256-byte multiple), then [batch][m] rows
static size_t list_count_bytes(int64_t batch) { return ((size_t)batch * sizeof(int32_t) + 255) & ~(size_t)255; }

size_t bpx_tiled_csr_call_workspace(const bpx_tiled_csr* plan, int64_t batch) {
  if (!plan || plan->tiles <= kMaxTiles || plan->m <= 0 || batch < 1) return 0;
  return list_count_bytes(batch) + (size_t)batch * (size_t)plan->m * sizeof(int32_t);
}

int bpx_event_csrmv_tiled_ws(void* stream, const bpx_tiled_csr* plan, const void* events, int ev_dtype, int ev_rule,
                             int64_t batch, int accumulate, float* out, void* workspace, size_t workspace_bytes) {
  if (!plan || !plan->mem) return BPX_ERR_ARG;
  if (ev_rule != BPX_RULE_POSITIVE && ev_rule != BPX_RULE_NONZERO) return BPX_ERR_ARG;
  if (ev_dtype != BPX_BOOL && ev_dtype != BPX_F32) return BPX_ERR_DTYPE;
  if (batch < 0 || batch > 65535) return BPX_ERR_SHAPE;
  if (batch == 0) return BPX_OK;
  hipStream_t s = (hipStream_t)stream;
  const int64_t m = plan->m, n = plan->n;
  if (!out || (m > 0 && !events)) return BPX_ERR_ARG;
  // plans of more than kMaxTiles tiles list each vector's active rows once
  // when the caller passes the workspace for it
  const size_t lws = bpx_tiled_csr_call_workspace(plan, batch);
  const bool listed = lws > 0 && workspace && workspace_bytes >= lws;
  int32_t* lcount = listed ? (int32_t*)workspace : nullptr;
  int32_t* lrows = listed ? (int32_t*)((char*)workspace + list_count_bytes(batch)) : nullptr;
  if (!accumulate || listed) {
    const int64_t zn = accumulate ? 0 : batch * n;
    int64_t zb = ((zn > batch ? zn : batch) + 255) / 256;
    if (zb > 4096) zb = 4096;
    hipLaunchKernelGGL(k_fill_zero_out_counts, dim3((unsigned)zb), dim3(256), 0, s, out, zn, lcount,
                       listed ? batch : 0);
  }
  if (m == 0) {
    BPX_CHECK_LAUNCH();
    return BPX_OK;
  }
  if (listed) {
    const int64_t lb = (m + kListRows - 1) / kListRows;
    if (lb > 0x7FFFFFFFll) return BPX_ERR_SHAPE;
    if (ev_dtype == BPX_BOOL)
      BPX_LAUNCH(KID_TILED_LIST, k_tiled_list<uint8_t>, dim3((unsigned)lb, (unsigned)batch), dim3(kListThreads), 0,
                 s, (const uint8_t*)events, ev_rule, m, lcount, lrows);
    else if (ev_dtype == BPX_F32)
      BPX_LAUNCH(KID_TILED_LIST, k_tiled_list<float>, dim3((unsigned)lb, (unsigned)batch), dim3(kListThreads), 0,
                 s, (const float*)events, ev_rule, m, lcount, lrows);
    else
      return BPX_ERR_DTYPE;
  }
  const char* base = (const char*)plan->mem;
  const int32_t* rowtile = (const int32_t*)(base + plan->rowtile_offset);
  const uint16_t* lidx = (const uint16_t*)(base + plan->lidx_offset);
  const void* w = base + plan->weights_offset;
  const float* hw = (const float*)(base + plan->homo_offset);
  const int P = plan->tiles, T = plan->tile_len;
  const int ncopy = tiled_copies(P);
  const unsigned blocks = (unsigned)(ncopy * P);
  const size_t shm = (size_t)T * sizeof(unsigned long long);
  const double fx_scale = ldexp(1.0, plan->fx_shift), fx_inv = ldexp(1.0, -plan->fx_shift);
#define TLR(WK, E, CO, LI)                                                                                \
  BPX_LAUNCH(KID_TILED, (k_tiled_scatter<WK, E, 1, false, CO, LI>), dim3(blocks, (unsigned)batch), dim3(kTiledThreads), shm, s, rowtile, \
             lidx, w, hw, (const E*)events, ev_rule, m, n, P, T, ncopy, fx_scale, fx_inv,                         \
             (const uint32_t*)nullptr, out, (const int32_t*)lcount, (const int32_t*)lrows)
#define TL(WK, E)                                \
  do {                                           \
    if (listed) {                                \
      if (plan->coarse) TLR(WK, E, true, true);  \
      else TLR(WK, E, false, true);              \
    } else {                                     \
      if (plan->coarse) TLR(WK, E, true, false); \
      else TLR(WK, E, false, false);             \
    }                                            \
  } while (0)
#define TL_E(WK)                                     \
  do {                                               \
    if (ev_dtype == BPX_BOOL) TL(WK, uint8_t);       \
    else if (ev_dtype == BPX_F32) TL(WK, float);     \
    else return BPX_ERR_DTYPE;                       \
  } while (0)
  if (plan->weight_dtype == BPX_TILED_HOMO) TL_E(0);
  else if (plan->weight_dtype == BPX_F32) TL_E(1);
  else TL_E(2);
#undef TL_E
#undef TL
#undef TLR
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

int bpx_csrmv_tiled(void* stream, const bpx_tiled_csr* plan, const void* x, int x_dtype, int accumulate, float* out,
                    void* workspace, size_t workspace_bytes) {
  if (!plan || !plan->mem) return BPX_ERR_ARG;
  if (x_dtype != BPX_F32) return BPX_ERR_DTYPE;
  if (!workspace || workspace_bytes < 2 * sizeof(uint32_t)) return BPX_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  const int64_t m = plan->m, n = plan->n;
  // {max finite |x|, min non-zero finite |x|} first; with 256 more bytes
  // than bpx_tiled_csr_call_workspace(plan, 1), the rows with x != 0 are
  // listed once (plans of more than kMaxTiles tiles)
  const size_t lws = bpx_tiled_csr_call_workspace(plan, 1);
  const bool listed = lws > 0 && workspace_bytes >= 256 + lws;
  int32_t* lcount = listed ? (int32_t*)((char*)workspace + 256) : nullptr;
  int32_t* lrows = listed ? (int32_t*)((char*)workspace + 256 + list_count_bytes(1)) : nullptr;
  if (!accumulate || listed) {
    const int64_t zn = accumulate ? 0 : n;
    int64_t zb = ((zn > 1 ? zn : 1) + 255) / 256;
    if (zb > 4096) zb = 4096;
    hipLaunchKernelGGL(k_fill_zero_out_counts, dim3((unsigned)zb), dim3(256), 0, s, out, zn, lcount, listed ? 1 : 0);
  }
  if (m == 0) {
    BPX_CHECK_LAUNCH();
    return BPX_OK;
  }
  uint32_t* xmax = (uint32_t*)workspace;
  hipLaunchKernelGGL(k_init_minmax, dim3(1), dim3(64), 0, s, xmax);
  int64_t ab = (m + 255) / 256;
  if (ab > 1024) ab = 1024;
  hipLaunchKernelGGL((k_absmax<float, true>), dim3((unsigned)ab), dim3(256), 0, s, (const float*)x, m, xmax);
  const char* base = (const char*)plan->mem;
  const int32_t* rowtile = (const int32_t*)(base + plan->rowtile_offset);
  const uint16_t* lidx = (const uint16_t*)(base + plan->lidx_offset);
  const void* w = base + plan->weights_offset;
  const float* hw = (const float*)(base + plan->homo_offset);
  const int P = plan->tiles, T = plan->tile_len;
  const int ncopy = tiled_copies(P);
  const unsigned blocks = (unsigned)(ncopy * P);
  const size_t shm = (size_t)T * sizeof(unsigned long long);
  const double fx_scale = ldexp(1.0, plan->fx_shift), fx_inv = ldexp(1.0, -plan->fx_shift);
  if (listed) {
    const int64_t lb = (m + kListRows - 1) / kListRows;
    if (lb > 0x7FFFFFFFll) return BPX_ERR_SHAPE;
    BPX_LAUNCH(KID_TILED_LIST, k_tiled_list<float>, dim3((unsigned)lb), dim3(kListThreads), 0, s, (const float*)x,
               BPX_RULE_NONZERO, m, lcount, lrows);
  }
#define TLD(WK, LI)                                                                                           \
  BPX_LAUNCH(KID_TILED, (k_tiled_scatter<WK, float, 1, true, false, LI>), dim3(blocks), dim3(kTiledThreads), shm, s, \
             rowtile, lidx, w, hw, (const float*)x, BPX_RULE_NONZERO, m, n, P, T, ncopy, fx_scale, fx_inv, xmax, out, \
             (const int32_t*)lcount, (const int32_t*)lrows)
#define TLDR(WK)                 \
  do {                           \
    if (listed) TLD(WK, true);   \
    else TLD(WK, false);         \
  } while (0)
  if (plan->weight_dtype == BPX_TILED_HOMO) TLDR(0);
  else if (plan->weight_dtype == BPX_F32) TLDR(1);
  else TLDR(2);
#undef TLDR
#undef TLD
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

size_t bpx_tiled_csr_gather_workspace(const bpx_tiled_csr* plan) {
  if (!plan || plan->m <= 0) return 0;
  return (size_t)plan->tiles * (size_t)plan->m * sizeof(float);
}

int bpx_tiled_csr_gather(void* stream, const bpx_tiled_csr* plan, const void* vec, int vec_dtype, int event,
                         int ev_rule, float* out, void* workspace, size_t workspace_bytes) {
  if (!plan || !plan->mem) return BPX_ERR_ARG;
  if (event ? (vec_dtype != BPX_BOOL && vec_dtype != BPX_F32) : vec_dtype != BPX_F32) return BPX_ERR_DTYPE;
  if (event && ev_rule != BPX_RULE_POSITIVE && ev_rule != BPX_RULE_NONZERO) return BPX_ERR_ARG;
  const int64_t m = plan->m, n = plan->n;
  if (m == 0) return BPX_OK;
  if (!out || (n > 0 && !vec)) return BPX_ERR_ARG;
  if (!workspace || workspace_bytes < bpx_tiled_csr_gather_workspace(plan)) return BPX_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  const char* base = (const char*)plan->mem;
  const int32_t* rowtile = (const int32_t*)(base + plan->rowtile_offset);
  const uint16_t* lidx = (const uint16_t*)(base + plan->lidx_offset);
  const void* w = base + plan->weights_offset;
  const float* hw = (const float*)(base + plan->homo_offset);
  const int P = plan->tiles, T = plan->tile_len;
  float* partial = (float*)workspace;
  const int64_t chunks = (m + kGatherRows - 1) / kGatherRows;
  const int64_t blocks = chunks * P;
  if (blocks > 0x7FFFFFFFll) return BPX_ERR_SHAPE;
  const int xm = event ? (vec_dtype == BPX_BOOL ? 1 : 2) : 0;
  // tiles per LDS range: as many as the block's LDS holds (at most P), the
  // ranges then evened out; the row slices keep ~one block per CU
  const int64_t ntmax = (int64_t)shm_bytes() / ((int64_t)T * (int64_t)sizeof(float));
  if (BPX_GATHER_RANGE && ntmax >= 2) {
    const int NR = (int)((P + std::min<int64_t>(ntmax, P) - 1) / std::min<int64_t>(ntmax, P));
    const int NT = (P + NR - 1) / NR;
    int64_t slices = device_cus() / NR;
    if (slices < 1) slices = 1;
    const int64_t rps = (m + slices - 1) / slices;
    slices = (m + rps - 1) / rps;
    const int64_t rblocks = slices * NR;
    if (rblocks > 0x7FFFFFFFll) return BPX_ERR_SHAPE;
    const size_t rshm = (size_t)NT * T * sizeof(float);
#define TGR(WK, XM)                                                                                              \
  BPX_LAUNCH(KID_TILED_GATHER, (k_tiled_gather_rng<WK, XM>), dim3((unsigned)rblocks), dim3(kGatherRangeThreads), rshm, \
             s, rowtile, lidx, w, hw, vec, ev_rule, m, n, P, T, NT, NR, rps, partial)
#define TGR_X(WK)                  \
  do {                             \
    if (xm == 0) TGR(WK, 0);       \
    else if (xm == 1) TGR(WK, 1);  \
    else TGR(WK, 2);               \
  } while (0)
    if (plan->weight_dtype == BPX_TILED_HOMO) TGR_X(0);
    else if (plan->weight_dtype == BPX_F32) TGR_X(1);
    else TGR_X(2);
#undef TGR_X
#undef TGR
    hipLaunchKernelGGL(k_tiled_gather_sum, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, (const float*)partial,
                       m, P, out);
    BPX_CHECK_LAUNCH();
    return BPX_OK;
  }
  const size_t shm = (size_t)T * sizeof(float);
#define TG(WK, XM)                                                                                          \
  BPX_LAUNCH(KID_TILED_GATHER, (k_tiled_gather<WK, XM>), dim3((unsigned)blocks), dim3(256), shm, s, rowtile, lidx, w, \
             hw, vec, ev_rule, m, n, P, T, partial)
#define TG_X(WK)                   \
  do {                             \
    if (xm == 0) TG(WK, 0);        \
    else if (xm == 1) TG(WK, 1);   \
    else TG(WK, 2);                \
  } while (0)
  if (plan->weight_dtype == BPX_TILED_HOMO) TG_X(0);
  else if (plan->weight_dtype == BPX_F32) TG_X(1);
  else TG_X(2);
#undef TG_X
#undef TG
  hipLaunchKernelGGL(k_tiled_gather_sum, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, (const float*)partial, m,
                     P, out);
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

}  // extern "C"
