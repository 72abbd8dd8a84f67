// bpx_coo.hip -- products with a COO sparse matrix: bm.sparse.coomv
// (sparse/_coo_mv.py:17-89, whose semantics are _coomv_impl at :99-105:
// out.at[row].add(data * v[col]), row and col swapped when transposed) and
// bm.sparse.seg_matmul (sparse/_jax_prim.py:97-166: a gather of the dense
// operand's rows or columns, a product with the values, a segment sum).
//
// Both are one pass over the nse (row, col, value) triples: HBM-bound
// streams of 12 bytes per element (f32 values, int32 indices) plus a gather
// of the dense operand, which stays in L2 / the MALL for the sizes the
// operators are used at.
//
// k_coo_seg: a wave takes a tile of 256 consecutive elements, 4 per lane
// (dwordx4 loads when the arrays are 16-byte aligned), and sums runs of equal
// output keys with a segmented scan: serial over a lane's 4 elements, then
// over the 64 lanes' trailing-run partials (shuffles).  Every run end issues
// one atomic, so a row-sorted matrix costs one atomic per (row, tile), an
// unsorted one at most one per element.  A dense operand with several
// columns is handled by looping over the columns with the key structure
// (run heads, ends, lane segments) computed once per tile.
//
// k_coo_rows: for a left sparse operand times a dense matrix of >= 16
// columns, groups of 16/32/64 lanes lie across the dense columns and walk a
// chunk of elements serially, accumulating in registers while the output row
// stays the same (coalesced row gathers, one coalesced atomic row per run).
//
// Index contract: row/col inside the matrix shape.  Outside it no fault
// occurs: output keys out of range are dropped and gather indices are clamped
// (JAX's scatter and gather defaults).
#include "bpx_common.h"

namespace bpx {
namespace {

constexpr int kPer = 4;               // elements per lane in k_coo_seg
constexpr int kTile = kWave * kPer;   // elements per wave tile

// arithmetic of the value type: int32 wraps like the reference's int32 math
__device__ __forceinline__ float vmul(float a, float b) { return a * b; }
__device__ __forceinline__ double vmul(double a, double b) { return a * b; }
__device__ __forceinline__ int32_t vmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
__device__ __forceinline__ float vadd(float a, float b) { return a + b; }
__device__ __forceinline__ double vadd(double a, double b) { return a + b; }
__device__ __forceinline__ int32_t vadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

__device__ __forceinline__ void vatomic(float* p, float v) { atomic_add(p, v); }
__device__ __forceinline__ void vatomic(double* p, double v) { atomic_add(p, v); }
__device__ __forceinline__ void vatomic(int32_t* p, int32_t v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ __forceinline__ void load4(const T* __restrict__ p, T (&v)[4]) {
  typedef T vec4 __attribute__((ext_vector_type(4)));
  const vec4 x = *(const vec4*)p;
  v[0] = x[0]; v[1] = x[1]; v[2] = x[2]; v[3] = x[3];
}

template <typename O>
__global__ __launch_bounds__(256) void k_coo_zero(O* __restrict__ out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = O(0);
}

template <typename O>
int zero_out(hipStream_t s, O* out, int64_t n) {
  if (n <= 0) return BPX_OK;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_coo_zero<O>, dim3((unsigned)blocks), dim3(256), 0, s, out, n);
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

// out[okey[e]*os + c*oj] += w_e * src[gkey[e]*gs + c*gj] for c < ncols.
// Output keys outside [0, out_keys) are dropped, gather keys are clamped to
// [0, src_keys).
template <typename V, typename I, bool HOMO, bool VEC>
__global__ __launch_bounds__(256) void k_coo_seg(const V* __restrict__ data, const I* __restrict__ okey,
                                                 const I* __restrict__ gkey, int64_t nse,
                                                 const V* __restrict__ src, int64_t src_keys, int64_t gs,
                                                 int64_t gj, int64_t ncols, V* __restrict__ out,
                                                 int64_t out_keys, int64_t os, int64_t oj) {
  const int lane = lane_id();
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / kWave);
  const int64_t ntiles = (nse + kTile - 1) / kTile;
  const V w0 = HOMO ? data[0] : V(0);
  for (int64_t t = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; t < ntiles; t += nwaves) {
    const int64_t e0 = t * kTile + (int64_t)lane * kPer;
    I k[kPer], g[kPer];
    V w[kPer];
    bool loaded = false;
    if constexpr (VEC) {
      if (e0 + kPer <= nse) {
        load4(okey + e0, k);
        load4(gkey + e0, g);
        if (!HOMO) load4(data + e0, w);
        loaded = true;
      }
    }
    if (!loaded) {
#pragma unroll
      for (int j = 0; j < kPer; ++j) {
        const bool ok = e0 + j < nse;
        const int64_t e = ok ? e0 + j : 0;
        k[j] = ok ? okey[e] : I(-1);
        g[j] = gkey[e];
        if (!HOMO) w[j] = data[e];
      }
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      if (k[j] < 0 || (int64_t)k[j] >= out_keys) k[j] = I(-1);
      g[j] = g[j] < 0 ? I(0) : ((int64_t)g[j] >= src_keys ? (I)(src_keys - 1) : g[j]);
      if (HOMO) w[j] = w0;
    }
    // key structure of the tile
    const I kprev = __shfl_up(k[kPer - 1], 1);      // last key of lane - 1
    const I knext = __shfl_down(k[0], 1);           // first key of lane + 1
    bool same[kPer];                                // element j continues element j-1's run
    same[0] = false;
#pragma unroll
    for (int j = 1; j < kPer; ++j) same[j] = k[j] == k[j - 1];
    const bool whole = same[1] && same[2] && same[3];
    const bool joins = lane > 0 && kprev == k[0];   // first run continues lane - 1's last run
    int seg = (whole && joins) ? 0 : lane;          // lane-segment head: max-scan of head lanes
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const int o = __shfl_up(seg, d);
      if (lane >= d) seg = o > seg ? o : seg;
    }
    bool first[kPer], end[kPer];                    // first run of the lane; run ends
    first[0] = true;
#pragma unroll
    for (int j = 1; j < kPer; ++j) first[j] = first[j - 1] && same[j];
#pragma unroll
    for (int j = 0; j < kPer - 1; ++j) end[j] = !same[j + 1] && k[j] >= 0;
    end[kPer - 1] = (lane == kWave - 1 || knext != k[kPer - 1]) && k[kPer - 1] >= 0;

    for (int64_t c = 0; c < ncols; ++c) {
      V s[kPer];
#pragma unroll
      for (int j = 0; j < kPer; ++j) s[j] = vmul(w[j], src[(int64_t)g[j] * gs + c * gj]);
#pragma unroll
      for (int j = 1; j < kPer; ++j) s[j] = same[j] ? vadd(s[j - 1], s[j]) : s[j];
      V X = s[kPer - 1];
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const V o = __shfl_up(X, d);
        if (lane - d >= seg) X = vadd(X, o);
      }
      const V Xp = __shfl_up(X, 1);
      const V cin = joins ? Xp : V(0);
#pragma unroll
      for (int j = 0; j < kPer; ++j) {
        const V v = first[j] ? vadd(s[j], cin) : s[j];
        if (end[j]) vatomic(out + (int64_t)k[j] * os + c * oj, v);
      }
    }
  }
}

// out[okey[e], :] += w_e * src[gkey[e], :] for row-major src (src_rows, K)
// and out (out_rows, K).  G lanes cover G*CT columns; a group walks kChunk
// elements, kU at a time (all gathers of a batch in flight before the serial
// accumulation).
constexpr int kChunk = 64;
constexpr int kU = 4;

template <typename V, typename I, bool HOMO, int G, int CT>
__global__ __launch_bounds__(256) void k_coo_rows(const V* __restrict__ data, const I* __restrict__ okey,
                                                  const I* __restrict__ gkey, int64_t nse,
                                                  const V* __restrict__ src, int64_t src_rows, int64_t K,
                                                  V* __restrict__ out, int64_t out_rows) {
  const int gl = threadIdx.x % G;
  const int64_t group = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int64_t ngroups = (int64_t)gridDim.x * blockDim.x / G;
  const int64_t nchunks = (nse + kChunk - 1) / kChunk;
  const int64_t ncb = (K + G * CT - 1) / (G * CT);
  const V w0 = HOMO ? data[0] : V(0);
  for (int64_t wk = group; wk < nchunks * ncb; wk += ngroups) {
    const int64_t ch = wk / ncb, cb = wk - ch * ncb;
    const int64_t eb = ch * kChunk;
    const int64_t ee = eb + kChunk < nse ? eb + kChunk : nse;
    const int64_t c0 = cb * G * CT + gl;
    V acc[CT];
#pragma unroll
    for (int t = 0; t < CT; ++t) acc[t] = V(0);
    I cur = I(-1);
    for (int64_t e = eb; e < ee; e += kU) {
      I r[kU], cidx[kU];
      V w[kU], x[kU][CT];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const bool ok = e + u < ee;
        const int64_t eu = ok ? e + u : eb;
        I rr = okey[eu], cc = gkey[eu];
        r[u] = (!ok || rr < 0 || (int64_t)rr >= out_rows) ? I(-2) : rr;   // -2: skipped, never flushed
        if (!ok) r[u] = I(-3);                                           // past the chunk end
        cidx[u] = cc < 0 ? I(0) : ((int64_t)cc >= src_rows ? (I)(src_rows - 1) : cc);
        w[u] = HOMO ? w0 : data[eu];
      }
#pragma unroll
      for (int u = 0; u < kU; ++u)
#pragma unroll
        for (int t = 0; t < CT; ++t) {
          const int64_t col = c0 + (int64_t)t * G;
          x[u][t] = col < K ? src[(int64_t)cidx[u] * K + col] : V(0);
        }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        if (r[u] == I(-3)) break;
        if (r[u] != cur) {
          if (cur >= 0) {
#pragma unroll
            for (int t = 0; t < CT; ++t) {
              const int64_t col = c0 + (int64_t)t * G;
              if (col < K) vatomic(out + (int64_t)cur * K + col, acc[t]);
              acc[t] = V(0);
            }
          }
#pragma unroll
          for (int t = 0; t < CT; ++t) acc[t] = V(0);
          cur = r[u];
        }
#pragma unroll
        for (int t = 0; t < CT; ++t) acc[t] = vadd(acc[t], vmul(w[u], x[u][t]));
      }
    }
    if (cur >= 0) {
#pragma unroll
      for (int t = 0; t < CT; ++t) {
        const int64_t col = c0 + (int64_t)t * G;
        if (col < K) vatomic(out + (int64_t)cur * K + col, acc[t]);
      }
    }
  }
}

// out[e] = sum_c a[row[e]*as + c*aj] * b[col[e]*bs + c*bj]: the value
// cotangent of the COO products (sampled dense-dense product).
template <typename V, typename I>
__global__ __launch_bounds__(256) void k_coo_sddmm(const I* __restrict__ row, const I* __restrict__ col, int64_t nse,
                                                   const V* __restrict__ a, int64_t a_keys, int64_t as,
                                                   int64_t aj, const V* __restrict__ b, int64_t b_keys,
                                                   int64_t bs, int64_t bj, int64_t ncols, V* __restrict__ out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; e < nse; e += stride) {
    const I r = row[e], c = col[e];
    V acc = V(0);
    if (r >= 0 && (int64_t)r < a_keys && c >= 0 && (int64_t)c < b_keys)
      for (int64_t j = 0; j < ncols; ++j)
        acc = vadd(acc, vmul(a[(int64_t)r * as + j * aj], b[(int64_t)c * bs + j * bj]));
    out[e] = acc;
  }
}

unsigned grid_for(int64_t work_items, int per_block) {
  int64_t blocks = (work_items + per_block - 1) / per_block;
  if (blocks < 1) blocks = 1;
  if (blocks > 65536) blocks = 65536;
  return (unsigned)blocks;
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// the segmented-scan kernel on one dtype/index combination
template <typename V, typename I>
int launch_seg(hipStream_t s, const V* data, int64_t data_len, const I* okey, const I* gkey, int64_t nse,
               const V* src, int64_t src_keys, int64_t gs, int64_t gj, int64_t ncols, V* out, int64_t out_keys,
               int64_t os, int64_t oj) {
  const bool homo = data_len == 1;
  constexpr bool kCanVec = sizeof(I) == 4 && sizeof(V) == 4;
  const bool vec = kCanVec && aligned16(okey) && aligned16(gkey) && (homo || aligned16(data));
  const int64_t ntiles = (nse + kTile - 1) / kTile;
  const dim3 grid(grid_for(ntiles, 256 / kWave)), block(256);
#define SEG_GO(H_, VEC_)                                                                                  \
  BPX_LAUNCH(KID_COO, (k_coo_seg<V, I, H_, VEC_>), grid, block, 0, s, data, okey, gkey, nse, src, src_keys, \
             gs, gj, ncols, out, out_keys, os, oj)
  if (homo) {
    if (vec) SEG_GO(true, kCanVec); else SEG_GO(true, false);
  } else {
    if (vec) SEG_GO(false, kCanVec); else SEG_GO(false, false);
  }
#undef SEG_GO
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

template <typename V, typename I, bool HOMO>
int launch_rows(hipStream_t s, const V* data, const I* okey, const I* gkey, int64_t nse, const V* src,
                int64_t src_rows, int64_t K, V* out, int64_t out_rows) {
  constexpr int CT = 2;
  const int64_t nchunks = (nse + kChunk - 1) / kChunk;
#define ROWS_GO(G_)                                                                                          \
  do {                                                                                                       \
    const int64_t ncb = (K + (G_) * CT - 1) / ((G_) * CT);                                                   \
    BPX_LAUNCH(KID_COO, (k_coo_rows<V, I, HOMO, G_, CT>), dim3(grid_for(nchunks * ncb, 256 / (G_))),          \
               dim3(256), 0, s, data, okey, gkey, nse, src, src_rows, K, out, out_rows);                     \
  } while (0)
  if (K > 32 * CT) ROWS_GO(64);
  else if (K > 16 * CT) ROWS_GO(32);
  else ROWS_GO(16);
#undef ROWS_GO
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

}  // namespace
}  // namespace bpx

using namespace bpx;

#define BPX_COO_IDX_SWITCH(IDT, ...)                     \
  switch (IDT) {                                         \
    case BPX_I32: { using I = int32_t; __VA_ARGS__; }    \
    case BPX_I64: { using I = int64_t; __VA_ARGS__; }    \
    default: return BPX_ERR_DTYPE;                       \
  }

#define BPX_COO_VAL_SWITCH(VDT, ...)                     \
  switch (VDT) {                                         \
    case BPX_F32: { using V = float; __VA_ARGS__; }      \
    case BPX_F64: { using V = double; __VA_ARGS__; }     \
    case BPX_I32: { using V = int32_t; __VA_ARGS__; }    \
    default: return BPX_ERR_DTYPE;                       \
  }

namespace {

template <typename V, typename I>
int coomv_typed(hipStream_t s, const void* data, int64_t data_len, const void* row, const void* col, int64_t nse,
                const void* vector, int64_t m, int64_t n, int transpose, void* out) {
  const int64_t out_len = transpose ? n : m, in_len = transpose ? m : n;
  int rc = zero_out<V>(s, (V*)out, out_len);
  if (rc != BPX_OK || nse == 0 || in_len == 0) return rc;
  const I* okey = (const I*)(transpose ? col : row);
  const I* gkey = (const I*)(transpose ? row : col);
  return launch_seg<V, I>(s, (const V*)data, data_len, okey, gkey, nse, (const V*)vector, in_len, 1, 0, 1,
                          (V*)out, out_len, 1, 0);
}

template <typename V, typename I>
int matmat_typed(hipStream_t s, const void* data, int64_t data_len, const void* row, const void* col, int64_t nse,
                 const void* dense, int64_t m, int64_t n, int64_t k, int sparse_left, void* out) {
  // sparse_left: out (m, k) = S (m, n) @ dense (n, k); else out (k, n) = dense (k, m) @ S (m, n)
  int rc = zero_out<V>(s, (V*)out, sparse_left ? m * k : k * n);
  // an empty gather dimension leaves nothing to gather (and no row to clamp to)
  if (rc != BPX_OK || nse == 0 || k == 0 || (sparse_left ? n : m) == 0) return rc;
  const V* d = (const V*)data;
  if (sparse_left) {
    if (k >= 16) {
      if (data_len == 1)
        return launch_rows<V, I, true>(s, d, (const I*)row, (const I*)col, nse, (const V*)dense, n, k, (V*)out, m);
      return launch_rows<V, I, false>(s, d, (const I*)row, (const I*)col, nse, (const V*)dense, n, k, (V*)out, m);
    }
    return launch_seg<V, I>(s, d, data_len, (const I*)row, (const I*)col, nse, (const V*)dense, n, k, 1, k,
                            (V*)out, m, k, 1);
  }
  // out[i, col[e]] += w_e * dense[i, row[e]]
  return launch_seg<V, I>(s, d, data_len, (const I*)col, (const I*)row, nse, (const V*)dense, m, 1, m, k,
                          (V*)out, n, 1, n);
}

}  // namespace

extern "C" {

int bpx_coomv(void* stream, const void* data, int data_dtype, int64_t data_len, const void* row, const void* col,
              int index_dtype, int64_t nse, const void* vector, int vec_dtype, int64_t m, int64_t n, int transpose,
              void* out) {
  if (m < 0 || n < 0 || nse < 0) return BPX_ERR_SHAPE;
  if (data_len != 1 && data_len != nse) return BPX_ERR_SHAPE;
  if (vec_dtype != data_dtype) return BPX_ERR_DTYPE;
  if (data_dtype != BPX_F32 && data_dtype != BPX_F64 && data_dtype != BPX_I32) return BPX_ERR_DTYPE;
  // (a vector of length 0 may come as a null pointer)
  if ((!out && (transpose ? n : m) > 0) ||
      (nse > 0 && (!data || !row || !col || (!vector && (transpose ? m : n) > 0))))
    return BPX_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  BPX_COO_VAL_SWITCH(data_dtype, BPX_COO_IDX_SWITCH(index_dtype,
      return (coomv_typed<V, I>(s, data, data_len, row, col, nse, vector, m, n, transpose, out))));
}

int bpx_coo_matmat(void* stream, const void* data, int data_dtype, int64_t data_len, const void* row,
                   const void* col, int index_dtype, int64_t nse, const void* dense, int dense_dtype, int64_t m,
                   int64_t n, int64_t k, int sparse_left, void* out) {
  if (m < 0 || n < 0 || k < 0 || nse < 0) return BPX_ERR_SHAPE;
  if (data_len != 1 && data_len != nse) return BPX_ERR_SHAPE;
  if (dense_dtype != data_dtype) return BPX_ERR_DTYPE;
  if (data_dtype != BPX_F32 && data_dtype != BPX_F64 && data_dtype != BPX_I32) return BPX_ERR_DTYPE;
  if ((!out && (sparse_left ? m : n) * k > 0) ||
      (nse > 0 && k > 0 && (!data || !row || !col || (!dense && (sparse_left ? n : m) > 0))))
    return BPX_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  BPX_COO_VAL_SWITCH(data_dtype, BPX_COO_IDX_SWITCH(index_dtype,
      return (matmat_typed<V, I>(s, data, data_len, row, col, nse, dense, m, n, k, sparse_left, out))));
}

int bpx_coo_sddmm(void* stream, const void* row, const void* col, int index_dtype, int64_t nse, const void* a,
                  int64_t a_keys, int64_t a_stride, int64_t a_col_stride, const void* b, int64_t b_keys,
                  int64_t b_stride, int64_t b_col_stride, int64_t ncols, int dtype, void* out) {
  if (nse < 0 || ncols < 0 || a_keys < 0 || b_keys < 0) return BPX_ERR_SHAPE;
  if (nse == 0) return BPX_OK;
  hipStream_t s = (hipStream_t)stream;
  BPX_COO_VAL_SWITCH(dtype, BPX_COO_IDX_SWITCH(index_dtype, {
    hipLaunchKernelGGL((k_coo_sddmm<V, I>), dim3(grid_for(nse, 256)), dim3(256), 0, s, (const I*)row,
                       (const I*)col, nse, (const V*)a, a_keys, a_stride, a_col_stride, (const V*)b, b_keys,
                       b_stride, b_col_stride, ncols, (V*)out);
    BPX_CHECK_LAUNCH();
    return BPX_OK;
  }));
}

}  // extern "C"
