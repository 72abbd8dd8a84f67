// bpx_batch.hip -- the event operators over a batch of B event vectors in
// one launch (the batching rule of the reference's primitives is a
// sequential scan over the batch, op_register/utils.py:16-37).
//
// The B vectors are first packed into per-index bit masks (bit b of
// masks[i] = vector b is truthy at i; 32 vectors per pass).  Then:
//   * jitconn, row-parallel: each row's connections are regenerated once
//     (the LFSR88 draws, the cost of this operator, are shared by the whole
//     batch); per connection one 4-byte mask read and a predicated add per
//     vector.  Row sums are formed in the single-vector kernel's order
//     (stream partials in ascending stream order), so every vector's result
//     is bit-identical to the single call for bool events and finite weights.
//   * jitconn, event-driven scatter: a column active in any vector
//     regenerates its targets once and adds into every vector it fires in.
//   * event csrmv gather: each row's indices and weights are read once for
//     all vectors (shuffle reduction per vector).
//   * event csrmv scatter: each row active in any vector is read once, one
//     FP32 atomic per (vector, synapse).
#include "bpx_common.h"
#include "bpx_rng.h"

#include <math.h>

namespace bpx {

constexpr int kMaskBits = 32;

template <typename E>
__global__ __launch_bounds__(256) void k_pack_masks(const E* __restrict__ ev, int64_t len, int b0, int nb, int rule,
                                                    uint32_t* __restrict__ masks) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= len) return;
  uint32_t m = 0;
  for (int b = 0; b < nb; ++b) m |= (ev_true(ev[(int64_t)(b0 + b) * len + i], rule) ? 1u : 0u) << b;
  masks[i] = m;
}

template <typename T>
__global__ __launch_bounds__(256) void k_zero_n(T* __restrict__ out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (; i < n; i += stride) out[i] = T(0);
}

template <int DIST>
__device__ __forceinline__ float draw_w(Lfsr88& k, float w0, float w1) {
  if constexpr (DIST == BPX_DIST_UNIFORM) {
    const float r = (float)k.next() * 2.3283064365386963e-10f;
    return r * (w1 - w0) + w0;
  } else if constexpr (DIST == BPX_DIST_NORMAL) {
    float u1 = (float)k.next() * 2.3283064365386963e-10f;
    const float u2 = (float)k.next() * 2.3283064365386963e-10f;
    u1 = u1 > 1e-10f ? u1 : 1e-10f;
    const float ln_u1 = __builtin_amdgcn_logf(u1) * 0.6931471805599453f;
    const float mag = __builtin_amdgcn_sqrtf(-2.0f * ln_u1);
    return w0 + w1 * (mag * __builtin_amdgcn_sinf(u2));
  } else {
    return w0;
  }
}

__device__ __forceinline__ float repeat_add(uint32_t c, float w) {
  float r = 0.f;
  for (uint32_t k = 0; k < c; ++k) r += w;
  return r;
}

// ---------------------------------------------------------------------------
// jitconn, outdim_parallel, one thread per row, NB vectors of the batch.
// LANE32: 32 streams per row (key seed + row*32 + l, lane l walking
// (step*l - 1, step*(l+1) - 1)); otherwise one stream per row (key
// seed + row, first column bits % clen).  FIN: finite weights, so `w * e`
// with e in {0, 1} is `e ? w : +0` and a homogeneous partial is the
// repeated sum of its active terms.
// ---------------------------------------------------------------------------
template <int DIST, int NB, bool LANE32, bool FIN>
__global__ __launch_bounds__(256) void k_jit_outdim_batch(const uint32_t* __restrict__ masks, int nb, float w0,
                                                          float w1, DivU32 dv, uint32_t seed, int64_t num_row,
                                                          int64_t num_col, uint32_t step, int64_t ostride,
                                                          float* __restrict__ out) {
  const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (row >= num_row) return;
  // the single-vector kernels: lane32 sums w * e per connection (e ? w : +0
  // for finite weights), the row layout adds w when the event is on
  constexpr bool SELECT = FIN || !LANE32;
  constexpr bool HITS = DIST == BPX_DIST_HOMO && SELECT;
  const uint32_t ncol = (uint32_t)num_col;
  float racc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) racc[b] = 0.f;
  constexpr int NS = LANE32 ? 32 : 1;
  for (int l = 0; l < NS; ++l) {
    float p[NB];
    uint32_t hits[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      p[b] = 0.f;
      hits[b] = 0u;
    }
    uint32_t c, end;
    Lfsr88 key(LANE32 ? seed + (uint32_t)(row * 32) + (uint32_t)l : seed + (uint32_t)row);
    if constexpr (LANE32) {
      c = step * (uint32_t)l - 1u;
      end = c + step < ncol ? c + step : ncol;
      c += (uint32_t)rand_inc(key, dv);
    } else {
      c = (uint32_t)rand_first(key, dv);
      end = ncol;
    }
    while (c < end) {
      const uint32_t mk = masks[c];
      float w = w0;
      if constexpr (DIST != BPX_DIST_HOMO) w = draw_w<DIST>(key, w0, w1);
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const bool on = (mk >> b) & 1u;
        if constexpr (HITS) hits[b] += on ? 1u : 0u;
        else if constexpr (SELECT) p[b] += on ? w : 0.f;
        else p[b] += w * (on ? 1.f : 0.f);
      }
      c += (uint32_t)rand_inc(key, dv);
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      if constexpr (HITS) racc[b] += repeat_add(hits[b], w0);
      else racc[b] += p[b];
    }
  }
#pragma unroll
  for (int b = 0; b < NB; ++b)
    if (b < nb) out[(int64_t)b * ostride + row] = racc[b];
}

// jitconn, event-driven scatter (outdim_parallel=False): thread per (column,
// lane) (LANE32) or per column; columns active in no vector return at once
template <int DIST, int NB, bool LANE32>
__global__ __launch_bounds__(256) void k_jit_scatter_batch(const uint32_t* __restrict__ masks, float w0, float w1,
                                                           DivU32 dv, uint32_t seed, int64_t num_row,
                                                           int64_t num_col, uint32_t step, int64_t ostride,
                                                           float* __restrict__ out) {
  const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t col = LANE32 ? gid >> 5 : gid;
  if (col >= num_col) return;
  const uint32_t mk = masks[col];
  if (mk == 0u) return;
  const uint32_t l = LANE32 ? (uint32_t)(gid & 31) : 0u;
  const uint32_t nrow = (uint32_t)num_row;
  uint32_t r, end;
  Lfsr88 key(LANE32 ? seed + (uint32_t)(col * 32) + l : seed + (uint32_t)col);
  if constexpr (LANE32) {
    r = step * l - 1u;
    end = r + step < nrow ? r + step : nrow;
    r += (uint32_t)rand_inc(key, dv);
  } else {
    r = (uint32_t)rand_first(key, dv);
    end = nrow;
  }
  while (r < end) {
    const float w = draw_w<DIST>(key, w0, w1);
#pragma unroll
    for (int b = 0; b < NB; ++b)
      if ((mk >> b) & 1u) atomic_add(out + (int64_t)b * ostride + r, w);
    r += (uint32_t)rand_inc(key, dv);
  }
}

// ---------------------------------------------------------------------------
// event csrmv gather (transpose=False): G lanes per row, each row's indices
// and weights read once, a shuffle reduction per vector; lane b of the group
// stores vector b's sum
// ---------------------------------------------------------------------------
template <int G, int NB, bool HOMO, typename I, typename V>
__global__ __launch_bounds__(256) void k_csr_gather_batch(const V* __restrict__ data, const I* __restrict__ indices,
                                                          const I* __restrict__ indptr, int64_t nnz,
                                                          const uint32_t* __restrict__ masks, int nb, int64_t m,
                                                          int64_t n, int64_t ostride, float* __restrict__ out) {
  constexpr int NG = 256 / G;
  const int64_t row = (int64_t)blockIdx.x * NG + (int)threadIdx.x / G;
  const int lane = (int)threadIdx.x % G;
  const float hw = HOMO ? to_f32(data[0]) : 0.f;
  float acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = 0.f;
  if (row < m) {
    int64_t s = (int64_t)indptr[row], t = (int64_t)indptr[row + 1];
    s = s < 0 ? 0 : s;
    t = t > nnz ? nnz : t;
    for (int64_t j = s + lane; j < t; j += G) {
      const int64_t c = (int64_t)indices[j];
      const float w = HOMO ? hw : to_f32(data[j]);
      const uint32_t mk = (uint64_t)c < (uint64_t)n ? masks[c] : 0u;
#pragma unroll
      for (int b = 0; b < NB; ++b) acc[b] += ((mk >> b) & 1u) ? w : 0.f;
    }
  }
  float mine = 0.f;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const float v = group_sum<G>(acc[b]);
    if (lane == (b % G)) mine = v;
    if (b % G == G - 1 || b == NB - 1) {
      const int b0 = b - b % G;
      if (row < m && lane <= b - b0 && b0 + lane < nb) out[(int64_t)(b0 + lane) * ostride + row] = mine;
    }
  }
}

// event csrmv scatter (transpose=True): blocks of 64 rows, rows active in
// any vector listed in LDS with their masks, G lanes per listed row
template <int G, int NB, bool HOMO, typename I, typename V>
__global__ __launch_bounds__(256) void k_csr_scatter_batch(const V* __restrict__ data, const I* __restrict__ indices,
                                                           const I* __restrict__ indptr, int64_t nnz,
                                                           const uint32_t* __restrict__ masks, int64_t m, int64_t n,
                                                           int64_t ostride, float* __restrict__ out) {
  __shared__ int s_rows[64];
  __shared__ uint32_t s_mask[64];
  __shared__ int s_cnt;
  const int64_t row0 = (int64_t)blockIdx.x * 64;
  if (threadIdx.x < 64) {
    const int64_t r = row0 + threadIdx.x;
    const uint32_t mk = r < m ? masks[r] : 0u;
    const uint64_t bal = __ballot(mk != 0u);
    if (mk) {
      s_rows[lanes_below(bal)] = (int)threadIdx.x;
      s_mask[lanes_below(bal)] = mk;
    }
    if (threadIdx.x == 0) s_cnt = __popcll(bal);
  }
  __syncthreads();
  const float hw = HOMO ? to_f32(data[0]) : 0.f;
  constexpr int NGR = 256 / G;
  const int grp = (int)threadIdx.x / G, lane = (int)threadIdx.x % G;
  for (int e = grp; e < s_cnt; e += NGR) {
    const int64_t row = row0 + s_rows[e];
    const uint32_t mk = s_mask[e];
    int64_t s = (int64_t)indptr[row], t = (int64_t)indptr[row + 1];
    s = s < 0 ? 0 : s;
    t = t > nnz ? nnz : t;
    for (int64_t j = s + lane; j < t; j += G) {
      const int64_t c = (int64_t)indices[j];
      if ((uint64_t)c >= (uint64_t)n) continue;
      const float w = HOMO ? hw : to_f32(data[j]);
#pragma unroll
      for (int b = 0; b < NB; ++b)
        if ((mk >> b) & 1u) atomic_add(out + (int64_t)b * ostride + c, w);
    }
  }
}

template <typename E>
static void pack(hipStream_t s, const void* ev, int64_t len, int b0, int nb, int rule, uint32_t* masks) {
  hipLaunchKernelGGL(k_pack_masks<E>, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, (const E*)ev, len, b0, nb,
                     rule, masks);
}

static int pack_masks(hipStream_t s, const void* ev, int ev_dtype, int64_t len, int b0, int nb, int rule,
                      uint32_t* masks) {
  if (len <= 0) return BPX_OK;
  switch (ev_dtype) {
    case BPX_BOOL: pack<uint8_t>(s, ev, len, b0, nb, rule, masks); break;
    case BPX_F32: pack<float>(s, ev, len, b0, nb, rule, masks); break;
    default: return BPX_ERR_DTYPE;
  }
  return BPX_OK;
}

static void zero_f32(hipStream_t s, float* out, int64_t n) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_zero_n<float>, dim3((unsigned)blocks), dim3(256), 0, s, out, n);
}

// NB in {8, 16, 32}: the smallest that holds nb vectors
#define BPX_NB_SWITCH(nb, ...)            \
  do {                                    \
    if ((nb) <= 8) {                      \
      constexpr int NB = 8;               \
      __VA_ARGS__;                        \
    } else if ((nb) <= 16) {              \
      constexpr int NB = 16;              \
      __VA_ARGS__;                        \
    } else {                              \
      constexpr int NB = 32;              \
      __VA_ARGS__;                        \
    }                                     \
  } while (0)

template <int G, bool HOMO, typename I, typename V>
static void csr_batch_chunk(hipStream_t s, int transpose, int nb, const void* data, const void* indices,
                            const void* indptr, int64_t nnz, const uint32_t* masks, int64_t m, int64_t n,
                            float* out) {
  const int64_t ostride = transpose ? n : m;
  if (transpose) {
    BPX_NB_SWITCH(nb, hipLaunchKernelGGL((k_csr_scatter_batch<G, NB, HOMO, I, V>),
                                         dim3((unsigned)((m + 63) / 64)), dim3(256), 0, s, (const V*)data,
                                         (const I*)indices, (const I*)indptr, nnz, masks, m, n, ostride, out));
  } else {
    constexpr int NG = 256 / G;
    BPX_NB_SWITCH(nb, hipLaunchKernelGGL((k_csr_gather_batch<G, NB, HOMO, I, V>),
                                         dim3((unsigned)((m + NG - 1) / NG)), dim3(256), 0, s, (const V*)data,
                                         (const I*)indices, (const I*)indptr, nnz, masks, nb, m, n, ostride, out));
  }
}

template <bool HOMO, typename I, typename V>
static void csr_batch_g(hipStream_t s, int transpose, int nb, const void* data, const void* indices,
                        const void* indptr, int64_t nnz, const uint32_t* masks, int64_t m, int64_t n, float* out) {
  const double avg = m > 0 ? (double)nnz / (double)m : 0.0;
  if (avg <= 48.0) csr_batch_chunk<16, HOMO, I, V>(s, transpose, nb, data, indices, indptr, nnz, masks, m, n, out);
  else csr_batch_chunk<64, HOMO, I, V>(s, transpose, nb, data, indices, indptr, nnz, masks, m, n, out);
}

template <int DIST>
static void jit_batch_chunk(hipStream_t s, int nb, const uint32_t* masks, float w0, float w1, int32_t clen,
                            uint32_t seed, int64_t num_row, int64_t num_col, int outdim, int layout, float* out) {
  const DivU32 dv = DivU32::make((uint32_t)clen);
  const uint32_t step = (uint32_t)(((num_row + 1) >> 5) > 1 ? ((num_row + 1) >> 5) : 1);
  const bool fin = isfinite(w0) && (DIST == BPX_DIST_HOMO || isfinite(w1));
  const int64_t ostride = num_row;
  const bool l32 = layout == BPX_LAYOUT_LANE32;
  if (outdim) {
    const dim3 grid((unsigned)((num_row + 255) / 256));
#define JOB(L32, F)                                                                                              \
  BPX_NB_SWITCH(nb, hipLaunchKernelGGL((k_jit_outdim_batch<DIST, NB, L32, F>), grid, dim3(256), 0, s, masks, nb, w0, \
                                       w1, dv, seed, num_row, num_col, step, ostride, out))
    if (l32) {
      if (fin) JOB(true, true);
      else JOB(true, false);
    } else {
      if (fin) JOB(false, true);
      else JOB(false, false);
    }
#undef JOB
  } else {
    const int64_t threads = l32 ? num_col * 32 : num_col;
    const dim3 grid((unsigned)((threads + 255) / 256));
    if (l32)
      BPX_NB_SWITCH(nb, hipLaunchKernelGGL((k_jit_scatter_batch<DIST, NB, true>), grid, dim3(256), 0, s, masks, w0,
                                           w1, dv, seed, num_row, num_col, step, ostride, out));
    else
      BPX_NB_SWITCH(nb, hipLaunchKernelGGL((k_jit_scatter_batch<DIST, NB, false>), grid, dim3(256), 0, s, masks, w0,
                                           w1, dv, seed, num_row, num_col, step, ostride, out));
  }
}

}  // namespace bpx

using namespace bpx;

extern "C" {

size_t bpx_batch_workspace(int64_t len) { return len > 0 ? (size_t)len * sizeof(uint32_t) : 0; }

int bpx_event_csrmv_batch(void* stream, const void* data, int data_dtype, int64_t data_len, const void* indices,
                          const void* indptr, int index_dtype, int64_t nnz, const void* events, int ev_dtype,
                          int ev_rule, int64_t batch, int64_t m, int64_t n, int transpose, float* out,
                          void* workspace, size_t workspace_bytes) {
  if (m < 0 || n < 0 || nnz < 0 || batch < 0) return BPX_ERR_SHAPE;
  if (data_len != 1 && data_len != nnz) return BPX_ERR_SHAPE;
  if (ev_rule != BPX_RULE_POSITIVE && ev_rule != BPX_RULE_NONZERO) return BPX_ERR_ARG;
  if (data_dtype != BPX_F32 && data_dtype != BPX_BF16) return BPX_ERR_DTYPE;
  if (index_dtype != BPX_I32 && index_dtype != BPX_I64) return BPX_ERR_DTYPE;
  if (ev_dtype != BPX_BOOL && ev_dtype != BPX_F32) return BPX_ERR_DTYPE;
  const int64_t ev_len = transpose ? m : n, out_len = transpose ? n : m;
  if (batch == 0 || out_len == 0) return BPX_OK;
  if (!out || (nnz > 0 && (!data || !indices || !indptr || (!events && ev_len > 0)))) return BPX_ERR_ARG;
  if (workspace_bytes < bpx_batch_workspace(ev_len) || (ev_len > 0 && !workspace)) return BPX_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  if (transpose || nnz == 0) zero_f32(s, out, batch * out_len);
  if (nnz == 0 || ev_len == 0) {
    if (!transpose) zero_f32(s, out, batch * out_len);
    BPX_CHECK_LAUNCH();
    return BPX_OK;
  }
  uint32_t* masks = (uint32_t*)workspace;
  const bool homo = data_len == 1;
  for (int64_t b0 = 0; b0 < batch; b0 += kMaskBits) {
    const int nb = (int)(batch - b0 < kMaskBits ? batch - b0 : kMaskBits);
    const size_t esz = ev_dtype == BPX_BOOL ? 1 : 4;
    int rc = pack_masks(s, (const char*)events + (size_t)b0 * ev_len * esz, ev_dtype, ev_len, 0, nb, ev_rule, masks);
    if (rc != BPX_OK) return rc;
    float* o = out + b0 * out_len;
#define CB(HOMO, I, V) csr_batch_g<HOMO, I, V>(s, transpose, nb, data, indices, indptr, nnz, masks, m, n, o)
    if (index_dtype == BPX_I32) {
      if (data_dtype == BPX_F32) { if (homo) CB(true, int32_t, float); else CB(false, int32_t, float); }
      else { if (homo) CB(true, int32_t, bf16_t); else CB(false, int32_t, bf16_t); }
    } else {
      if (data_dtype == BPX_F32) { if (homo) CB(true, int64_t, float); else CB(false, int64_t, float); }
      else { if (homo) CB(true, int64_t, bf16_t); else CB(false, int64_t, bf16_t); }
    }
#undef CB
  }
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

int bpx_jitconn_mv_batch(void* stream, const void* events, int ev_dtype, int dist, float w0, float w1,
                         int32_t clen, uint32_t seed, int64_t batch, int64_t num_row, int64_t num_col,
                         int outdim_parallel, int layout, float* out, void* workspace, size_t workspace_bytes) {
  if (num_row < 0 || num_col < 0 || batch < 0) return BPX_ERR_SHAPE;
  if (num_row > INT32_MAX || num_col > INT32_MAX || num_row * 32 > (int64_t)UINT32_MAX ||
      num_col * 32 > (int64_t)UINT32_MAX)
    return BPX_ERR_SHAPE;
  if (clen < 1) return BPX_ERR_ARG;
  if (layout != BPX_LAYOUT_LANE32 && layout != BPX_LAYOUT_ROW) return BPX_ERR_ARG;
  if (dist < BPX_DIST_HOMO || dist > BPX_DIST_NORMAL) return BPX_ERR_ARG;
  if (ev_dtype != BPX_BOOL) return BPX_ERR_DTYPE;
  if (batch == 0 || num_row == 0) return BPX_OK;
  if (!out || (num_col > 0 && !events)) return BPX_ERR_ARG;
  if (workspace_bytes < bpx_batch_workspace(num_col) || (num_col > 0 && !workspace)) return BPX_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  if (!outdim_parallel || num_col == 0) zero_f32(s, out, batch * num_row);
  if (num_col == 0) {
    BPX_CHECK_LAUNCH();
    return BPX_OK;
  }
  uint32_t* masks = (uint32_t*)workspace;
  for (int64_t b0 = 0; b0 < batch; b0 += kMaskBits) {
    const int nb = (int)(batch - b0 < kMaskBits ? batch - b0 : kMaskBits);
    int rc = pack_masks(s, (const uint8_t*)events + (size_t)b0 * num_col, BPX_BOOL, num_col, 0, nb,
                        BPX_RULE_NONZERO, masks);
    if (rc != BPX_OK) return rc;
    float* o = out + b0 * num_row;
    if (dist == BPX_DIST_HOMO)
      jit_batch_chunk<BPX_DIST_HOMO>(s, nb, masks, w0, w1, clen, seed, num_row, num_col, outdim_parallel, layout, o);
    else if (dist == BPX_DIST_UNIFORM)
      jit_batch_chunk<BPX_DIST_UNIFORM>(s, nb, masks, w0, w1, clen, seed, num_row, num_col, outdim_parallel, layout, o);
    else
      jit_batch_chunk<BPX_DIST_NORMAL>(s, nb, masks, w0, w1, clen, seed, num_row, num_col, outdim_parallel, layout, o);
  }
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

}  // extern "C"
