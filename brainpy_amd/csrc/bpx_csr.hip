// bpx_csr.hip -- event-driven and dense CSR matvec, spike compaction and the
// fixed-order (CSC gather) scatter for gfx950.
//
// Scatter (transpose=True, the event-driven direction):
//   one 256-thread block per tile of 64 pre-synaptic rows.  Wave 0 reads the
//   tile's 64 events with one coalesced load and ballots them; the active row
//   offsets land in LDS in ascending order (mbcnt slot).  The block's threads
//   form groups of G lanes; each group takes active rows round-robin and
//   streams the row's column indices and weights with U-way unrolled,
//   coalesced loads (G*U elements in flight per group), then issues hardware
//   FP32 atomics into the post-synaptic output (resident in L2).
//   Inactive rows cost one byte of event traffic and nothing else.
// Gather (transpose=False): groups of G lanes per row, unrolled loads, a
//   shuffle reduction inside the group, one store per row.
#include "bpx_common.h"

#include <stdlib.h>

namespace bpx {

// ---------------------------------------------------------------------------
template <typename O>
__global__ __launch_bounds__(256) void k_fill_zero(O* __restrict__ out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = O(0);
}

template <typename O>
int launch_fill_zero(hipStream_t s, O* out, int64_t n) {
  if (n <= 0) return BPX_OK;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_zero<O>, dim3((unsigned)blocks), dim3(256), 0, s, out, n);
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

constexpr int kTileRows = 64;

// ---------------------------------------------------------------------------
// scatter: out[indices[j]] += w_j (EVENT: rows with a truthy event;
// dense: every row, w_j scaled by vec[row])
// ---------------------------------------------------------------------------
template <int G, int U, bool EVENT, bool HOMO, typename I, typename V, typename E>
__global__ __launch_bounds__(256) void k_csr_scatter(const V* __restrict__ data,
                                                     const I* __restrict__ indices,
                                                     const I* __restrict__ indptr, int64_t nnz,
                                                     const E* __restrict__ vec, int rule, int64_t m, int64_t n,
                                                     typename acc_of<V>::type* __restrict__ out) {
  using A = typename acc_of<V>::type;
  const A homo_w = HOMO ? (A)to_f32(data[0]) : A(0);
  __shared__ int s_rows[kTileRows];
  __shared__ int s_cnt;
  const int64_t row0 = (int64_t)blockIdx.x * kTileRows;
  if (threadIdx.x < kTileRows) {
    const int64_t r = row0 + threadIdx.x;
    bool act = false;
    if (r < m) {
      if constexpr (EVENT) act = ev_true(vec[r], rule);
      else act = true;
    }
    const uint64_t mask = __ballot(act);
    if (act) s_rows[lanes_below(mask)] = (int)threadIdx.x;
    if (threadIdx.x == 0) s_cnt = __popcll(mask);
  }
  __syncthreads();
  const int cnt = s_cnt;
  constexpr int NG = 256 / G;
  const int grp = (int)threadIdx.x / G;
  const int lane = (int)threadIdx.x % G;
  for (int e = grp; e < cnt; e += NG) {
    const int64_t row = row0 + s_rows[e];
    int64_t s = (int64_t)indptr[row];
    int64_t t = (int64_t)indptr[row + 1];
    s = s < 0 ? 0 : s;
    t = t > nnz ? nnz : t;
    A scale = A(1);
    if constexpr (!EVENT) scale = (A)vec[row];
    for (int64_t base = s; base < t; base += (int64_t)G * U) {
      int64_t idx[U];
      A val[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t j = base + lane + (int64_t)u * G;
        idx[u] = -1;
        if (j < t) {
          idx[u] = (int64_t)indices[j];
          if constexpr (HOMO) val[u] = homo_w;
          else val[u] = (A)to_f32(data[j]);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if ((uint64_t)idx[u] < (uint64_t)n) {
          A v = val[u];
          if constexpr (!EVENT) v = HOMO ? v * scale : scale * v;
          atomic_add(out + idx[u], v);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// gather: out[row] = sum_j w_j * [event(vec[indices[j]])]  (EVENT)
//                  = sum_j w_j * vec[indices[j]]           (dense)
// ---------------------------------------------------------------------------
template <int G, int U, bool EVENT, bool HOMO, typename I, typename V, typename E>
__global__ __launch_bounds__(256) void k_csr_gather(const V* __restrict__ data,
                                                    const I* __restrict__ indices,
                                                    const I* __restrict__ indptr, int64_t nnz,
                                                    const E* __restrict__ vec, int rule, int64_t m, int64_t n,
                                                    typename acc_of<V>::type* __restrict__ out) {
  using A = typename acc_of<V>::type;
  const A homo_w = HOMO ? (A)to_f32(data[0]) : A(0);
  constexpr int NG = 256 / G;
  const int64_t row = (int64_t)blockIdx.x * NG + (int)threadIdx.x / G;
  const int lane = (int)threadIdx.x % G;
  A acc = A(0);
  if (row < m) {
    int64_t s = (int64_t)indptr[row];
    int64_t t = (int64_t)indptr[row + 1];
    s = s < 0 ? 0 : s;
    t = t > nnz ? nnz : t;
    for (int64_t base = s; base < t; base += (int64_t)G * U) {
      // indices and weights are loaded together (the weights do not wait for
      // the gathered value), then the gathers
      int64_t idx[U];
      A wv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t j = base + lane + (int64_t)u * G;
        idx[u] = j < t ? (int64_t)indices[j] : -1;
        wv[u] = HOMO ? homo_w : (j < t ? (A)to_f32(data[j]) : A(0));
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if ((uint64_t)idx[u] < (uint64_t)n) {
          const E x = vec[idx[u]];
          if constexpr (EVENT) {
            if (ev_true(x, rule)) acc += wv[u];
          } else {
            acc += wv[u] * (A)x;
          }
        }
      }
    }
  }
  acc = group_sum<G>(acc);
  if (row < m && lane == 0) out[row] = acc;
}

// ---------------------------------------------------------------------------
// event gather with the events as a bitmask in LDS (bool events, n bits that
// fit the block's LDS): each block packs the whole event vector once (two
// 16-byte loads per 32 events, bytes tested four at a time) and walks rows
// grid-stride, so a synapse costs its index and weight loads plus an LDS bit
// test instead of a gathered global read.  Rows sum in the same lane order
// as k_csr_gather.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t nz_flags4(uint32_t w) {
  // bit 7 of each byte set iff the byte is non-zero, then gathered into 4 bits
  const uint32_t t = ((((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | w) & 0x80808080u) >> 7;
  return ((t * 0x00204081u) >> 21) & 0xFu;
}

// 1024-thread blocks, 2 per CU: a CU packs the event bitmask twice per call
// (each block reads the whole event vector) instead of once per 256-thread
// block, with the same 32 waves of rows in flight
constexpr int kGatherBitsThreads = 1024;

template <int G, int U, bool HOMO, typename I, typename V, typename E>
__global__ __launch_bounds__(kGatherBitsThreads, 2) void k_csr_gather_bits(const V* __restrict__ data,
                                                         const I* __restrict__ indices,
                                                         const I* __restrict__ indptr, int64_t nnz,
                                                         const E* __restrict__ ev, int rule, int64_t m, int64_t n,
                                                         typename acc_of<V>::type* __restrict__ out) {
  using A = typename acc_of<V>::type;
  extern __shared__ uint32_t s_ev[];
  const int64_t nw = (n + 31) >> 5;
  if constexpr (sizeof(E) == 1) {
    // bool events (16-byte aligned): a thread packs 32 events from two
    // 16-byte loads
    // every block reads the whole event vector; each starts at its own
    // offset so the blocks' first reads of a line do not all miss at once
    const int64_t rot = ((int64_t)blockIdx.x * 97) % nw;
    for (int64_t wr = threadIdx.x; wr < nw; wr += kGatherBitsThreads) {
      const int64_t w = wr + rot < nw ? wr + rot : wr + rot - nw;
      const int64_t c0 = w << 5;
      uint32_t bits = 0;
      if (c0 + 32 <= n) {
        const uint4 a = *reinterpret_cast<const uint4*>(ev + c0);
        const uint4 b = *reinterpret_cast<const uint4*>(ev + c0 + 16);
        const uint32_t q[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int k = 0; k < 8; ++k) bits |= nz_flags4(q[k]) << (4 * k);
      } else {
        for (int64_t c = c0; c < n; ++c) bits |= (ev[c] != 0 ? 1u : 0u) << (c - c0);
      }
      s_ev[w] = bits;
    }
  } else {
    // float events (tested with the operator's rule, never multiplied): a
    // wave tests 64 consecutive events per load, several loads in flight,
    // and a ballot gives two bitmask words
    constexpr int kIn = 8;
    const int lane = (int)(threadIdx.x & 63);
    const int64_t wave = threadIdx.x >> 6;
    constexpr int64_t kWaves = kGatherBitsThreads / 64;
    const int64_t nchunk = (n + 63) >> 6;
    const int64_t rot = ((int64_t)blockIdx.x * 97) % nchunk;   // as above
    for (int64_t c = wave; c < nchunk; c += kWaves * kIn) {
      // all loads first (unconditional, in bounds), then the tests: a
      // short-circuit test around each load would serialise them
      E v[kIn];
      int64_t cc[kIn];
#pragma unroll
      for (int k = 0; k < kIn; ++k) {
        const int64_t cr = c + kWaves * k;
        cc[k] = cr < nchunk ? (cr + rot < nchunk ? cr + rot : cr + rot - nchunk) : nchunk;
        const int64_t i = (cc[k] << 6) + lane;
        v[k] = ev[i < n ? i : 0];
      }
#pragma unroll
      for (int k = 0; k < kIn; ++k) {
        const int64_t i = (cc[k] << 6) + lane;
        const bool on = (i < n) & ev_true(v[k], rule);
        const uint64_t mask = __ballot(on);
        const int64_t cw = cc[k] << 1;                 // first of the chunk's two words
        if (lane == 0 && cw < nw) s_ev[cw] = (uint32_t)mask;
        if (lane == 1 && cw + 1 < nw) s_ev[cw + 1] = (uint32_t)(mask >> 32);
      }
    }
  }
  __syncthreads();
  const A homo_w = HOMO ? (A)to_f32(data[0]) : A(0);
  constexpr int NG = kGatherBitsThreads / G;
  const int lane = (int)threadIdx.x % G;
  const int64_t groups = (m + NG - 1) / NG;
  for (int64_t gi = blockIdx.x; gi < groups; gi += gridDim.x) {
    const int64_t row = gi * NG + (int)threadIdx.x / G;
    A acc = A(0);
    if (row < m) {
      int64_t s = (int64_t)indptr[row];
      int64_t t = (int64_t)indptr[row + 1];
      s = s < 0 ? 0 : s;
      t = t > nnz ? nnz : t;
      for (int64_t base = s; base < t; base += (int64_t)G * U) {
        // the weights are loaded with the indices, not after the event test
        // (with 10 % events most of their lines are touched anyway), so a
        // pass costs one memory round trip
        int64_t idx[U];
        A wv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int64_t j = base + lane + (int64_t)u * G;
          idx[u] = j < t ? (int64_t)indices[j] : -1;
          wv[u] = HOMO ? homo_w : (j < t ? (A)to_f32(data[j]) : A(0));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if ((uint64_t)idx[u] < (uint64_t)n && ((s_ev[idx[u] >> 5] >> (idx[u] & 31)) & 1u)) acc += wv[u];
        }
      }
    }
    acc = group_sum<G>(acc);
    if (row < m && lane == 0) out[row] = acc;
  }
}

// ---------------------------------------------------------------------------
// fixed order: one thread per post-synaptic column, ascending CSC rows
// ---------------------------------------------------------------------------
template <bool HOMO, typename V, typename E>
__global__ __launch_bounds__(256) void k_fixed_order(const V* __restrict__ data,
                                                     const int32_t* __restrict__ rows,
                                                     const int32_t* __restrict__ colptr, int64_t nnz,
                                                     const E* __restrict__ ev, int rule, int64_t m, int64_t n,
                                                     int accumulate, typename acc_of<V>::type* __restrict__ out) {
  using A = typename acc_of<V>::type;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const A homo_w = HOMO ? (A)to_f32(data[0]) : A(0);
  int64_t s = colptr[c], t = colptr[c + 1];
  s = s < 0 ? 0 : s;
  t = t > nnz ? nnz : t;
  A acc = A(0);
  for (int64_t k = s; k < t; ++k) {
    const int64_t r = rows[k];
    if ((uint64_t)r < (uint64_t)m && ev_true(ev[r], rule)) acc += HOMO ? homo_w : (A)to_f32(data[k]);
  }
  out[c] = accumulate ? out[c] + acc : acc;
}

// ---------------------------------------------------------------------------
// host dispatch
// ---------------------------------------------------------------------------
// blocks of the bitmask gather: 2 per CU (each packs the event vector once)
static int64_t gather_bits_grid() {
  static int64_t g = [] {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
    return (int64_t)cus * 2;
  }();
  return g;
}

static int pick_group(int64_t nnz, int64_t m) {
  const double avg = m > 0 ? (double)nnz / (double)m : 0.0;
  if (avg <= 48.0) return 16;
  if (avg <= 384.0) return 64;
  return 256;
}

template <int G, bool EVENT, bool HOMO, typename I, typename V, typename E>
static int launch_csr(hipStream_t s, int transpose, const void* data, const void* indices,
                      const void* indptr, int64_t nnz, const void* vec, int rule, int64_t m, int64_t n, void* out) {
  using A = typename acc_of<V>::type;
  constexpr int U = G >= 256 ? 4 : 8;
  if (transpose) {
    const int64_t blocks = (m + kTileRows - 1) / kTileRows;
    if (blocks > 0) {
      BPX_LAUNCH(EVENT ? KID_EVENT_SCATTER : KID_CSR_SCATTER, (k_csr_scatter<G, U, EVENT, HOMO, I, V, E>),
                 dim3((unsigned)blocks), dim3(256), 0, s, (const V*)data, (const I*)indices,
                 (const I*)indptr, nnz, (const E*)vec, rule, m, n, (A*)out);
    }
  } else {
    constexpr int NG = 256 / G;
    const int64_t blocks = (m + NG - 1) / NG;
    const int64_t nw = (n + 31) / 32;
    // events whose bitmask fits 64 KB of LDS (n <= 524,288); bool events
    // 16-byte aligned
    const bool bits = EVENT && nw * 4 <= 65536 && n > 0 && (sizeof(E) != 1 || ((uintptr_t)vec & 15u) == 0);
    if (blocks > 0 && bits) {
      int64_t grid = gather_bits_grid();
      if (grid > blocks) grid = blocks;
      BPX_LAUNCH(KID_EVENT_GATHER, (k_csr_gather_bits<G, U, HOMO, I, V, E>), dim3((unsigned)grid),
                 dim3(kGatherBitsThreads),
                 (size_t)nw * 4, s, (const V*)data, (const I*)indices, (const I*)indptr, nnz,
                 (const E*)vec, rule, m, n, (A*)out);
    } else if (blocks > 0) {
      BPX_LAUNCH(EVENT ? KID_EVENT_GATHER : KID_CSR_GATHER, (k_csr_gather<G, U, EVENT, HOMO, I, V, E>),
                 dim3((unsigned)blocks), dim3(256), 0, s, (const V*)data, (const I*)indices,
                 (const I*)indptr, nnz, (const E*)vec, rule, m, n, (A*)out);
    }
  }
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

template <bool EVENT, bool HOMO, typename I, typename V, typename E>
static int dispatch_g(hipStream_t s, int transpose, const void* data, const void* indices,
                      const void* indptr, int64_t nnz, const void* vec, int rule, int64_t m, int64_t n, void* out) {
  // the gather's groups read one row each; the scatter's groups read active rows
  int g = pick_group(nnz, m);
  if (!transpose && g > kWave) g = kWave;  // gather groups reduce with wave shuffles
  if (g == 16) return launch_csr<16, EVENT, HOMO, I, V, E>(s, transpose, data, indices, indptr, nnz, vec, rule, m, n, out);
  if (g == 64) return launch_csr<64, EVENT, HOMO, I, V, E>(s, transpose, data, indices, indptr, nnz, vec, rule, m, n, out);
  return launch_csr<256, EVENT, HOMO, I, V, E>(s, transpose, data, indices, indptr, nnz, vec, rule, m, n, out);
}

template <bool EVENT, typename I, typename V, typename E>
static int dispatch_homo(hipStream_t s, int transpose, const void* data, int64_t data_len, const void* indices,
                         const void* indptr, int64_t nnz, const void* vec, int rule, int64_t m, int64_t n, void* out,
                         bool accumulate = false) {
  using A = typename acc_of<V>::type;
  int64_t nout = transpose ? n : m;
  if (!accumulate) {
    int rc = launch_fill_zero<A>(s, (A*)out, nout);
    if (rc != BPX_OK) return rc;
  }
  if (data_len == 1)
    return dispatch_g<EVENT, true, I, V, E>(s, transpose, data, indices, indptr, nnz, vec, rule, m, n, out);
  return dispatch_g<EVENT, false, I, V, E>(s, transpose, data, indices, indptr, nnz, vec, rule, m, n, out);
}

}  // namespace bpx

using namespace bpx;

#define BPX_IDX_SWITCH(IDT, ...)                         \
  switch (IDT) {                                         \
    case BPX_I32: { using I = int32_t; __VA_ARGS__; }    \
    case BPX_I64: { using I = int64_t; __VA_ARGS__; }    \
    default: return BPX_ERR_DTYPE;                       \
  }

extern "C" {

int bpx_event_csrmv(void* stream, const void* data, int data_dtype, int64_t data_len, const void* indices,
                    const void* indptr, int index_dtype, int64_t nnz, const void* events, int ev_dtype, int ev_rule,
                    int64_t m, int64_t n, int transpose, void* out) {
  if (m < 0 || n < 0 || nnz < 0) return BPX_ERR_SHAPE;
  if (data_len != 1 && data_len != nnz) return BPX_ERR_SHAPE;
  if (ev_rule != BPX_RULE_POSITIVE && ev_rule != BPX_RULE_NONZERO) return BPX_ERR_ARG;
  // (a vector of length 0 may come as a null pointer)
  if ((!out && (transpose ? n : m) > 0) ||
      (nnz > 0 && (!data || !indices || !indptr || (!events && (transpose ? m : n) > 0))))
    return BPX_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (nnz == 0) {
    if (data_dtype == BPX_F64) return launch_fill_zero<double>(s, (double*)out, transpose ? n : m);
    return launch_fill_zero<float>(s, (float*)out, transpose ? n : m);
  }
#define EV_SWITCH(V_)                                                                                         \
  switch (ev_dtype) {                                                                                         \
    case BPX_BOOL: return dispatch_homo<true, I, V_, uint8_t>(s, transpose, data, data_len, indices, indptr, nnz, events, ev_rule, m, n, out); \
    case BPX_F32: return dispatch_homo<true, I, V_, float>(s, transpose, data, data_len, indices, indptr, nnz, events, ev_rule, m, n, out);   \
    case BPX_F64: return dispatch_homo<true, I, V_, double>(s, transpose, data, data_len, indices, indptr, nnz, events, ev_rule, m, n, out);  \
    default: return BPX_ERR_DTYPE;                                                                            \
  }
  switch (data_dtype) {
    case BPX_F32: BPX_IDX_SWITCH(index_dtype, EV_SWITCH(float));
    case BPX_F64: BPX_IDX_SWITCH(index_dtype, EV_SWITCH(double));
    case BPX_BF16: BPX_IDX_SWITCH(index_dtype, EV_SWITCH(bf16_t));
    default: return BPX_ERR_DTYPE;
  }
#undef EV_SWITCH
}

int bpx_event_csrmv_add(void* stream, const void* data, int data_dtype, int64_t data_len, const void* indices,
                        const void* indptr, int index_dtype, int64_t nnz, const void* events, int ev_dtype,
                        int ev_rule, int64_t m, int64_t n, void* out) {
  if (m < 0 || n < 0 || nnz < 0) return BPX_ERR_SHAPE;
  if (data_len != 1 && data_len != nnz) return BPX_ERR_SHAPE;
  if (ev_rule != BPX_RULE_POSITIVE && ev_rule != BPX_RULE_NONZERO) return BPX_ERR_ARG;
  if (nnz > 0 && (!data || !indices || !indptr || (!out && n > 0) || (!events && m > 0))) return BPX_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (nnz == 0) return BPX_OK;
#define EV_SWITCH_ADD(V_)                                                                                      \
  switch (ev_dtype) {                                                                                          \
    case BPX_BOOL: return dispatch_homo<true, I, V_, uint8_t>(s, 1, data, data_len, indices, indptr, nnz, events, ev_rule, m, n, out, true); \
    case BPX_F32: return dispatch_homo<true, I, V_, float>(s, 1, data, data_len, indices, indptr, nnz, events, ev_rule, m, n, out, true);   \
    default: return BPX_ERR_DTYPE;                                                                             \
  }
  switch (data_dtype) {
    case BPX_F32: BPX_IDX_SWITCH(index_dtype, EV_SWITCH_ADD(float));
    case BPX_BF16: BPX_IDX_SWITCH(index_dtype, EV_SWITCH_ADD(bf16_t));
    default: return BPX_ERR_DTYPE;
  }
#undef EV_SWITCH_ADD
}

int bpx_csrmv(void* stream, const void* data, int data_dtype, int64_t data_len, const void* indices,
              const void* indptr, int index_dtype, int64_t nnz, const void* vector, int vec_dtype, int64_t m,
              int64_t n, int transpose, void* out) {
  if (m < 0 || n < 0 || nnz < 0) return BPX_ERR_SHAPE;
  if (data_len != 1 && data_len != nnz) return BPX_ERR_SHAPE;
  hipStream_t s = (hipStream_t)stream;
  const bool f64 = data_dtype == BPX_F64;
  if (vec_dtype != (f64 ? BPX_F64 : BPX_F32)) return BPX_ERR_DTYPE;
  if ((!out && (transpose ? n : m) > 0) ||
      (nnz > 0 && (!data || !indices || !indptr || (!vector && (transpose ? m : n) > 0))))
    return BPX_ERR_ARG;
  if (nnz == 0) {
    if (f64) return launch_fill_zero<double>(s, (double*)out, transpose ? n : m);
    return launch_fill_zero<float>(s, (float*)out, transpose ? n : m);
  }
  switch (data_dtype) {
    case BPX_F32: BPX_IDX_SWITCH(index_dtype, return (dispatch_homo<false, I, float, float>(s, transpose, data, data_len, indices, indptr, nnz, vector, 0, m, n, out)));
    case BPX_F64: BPX_IDX_SWITCH(index_dtype, return (dispatch_homo<false, I, double, double>(s, transpose, data, data_len, indices, indptr, nnz, vector, 0, m, n, out)));
    case BPX_BF16: BPX_IDX_SWITCH(index_dtype, return (dispatch_homo<false, I, bf16_t, float>(s, transpose, data, data_len, indices, indptr, nnz, vector, 0, m, n, out)));
    default: return BPX_ERR_DTYPE;
  }
}

int bpx_event_csrmv_fixed_order(void* stream, const void* csc_data, int data_dtype, int64_t data_len,
                                const int32_t* csc_rows, const int32_t* csc_colptr, int64_t nnz,
                                const void* events, int ev_dtype, int ev_rule, int64_t m, int64_t n, int accumulate,
                                void* out) {
  if (m < 0 || n < 0 || nnz < 0) return BPX_ERR_SHAPE;
  if (data_len != 1 && data_len != nnz) return BPX_ERR_SHAPE;
  hipStream_t s = (hipStream_t)stream;
  if (n == 0) return BPX_OK;
  if (!out || !csc_colptr || (nnz > 0 && (!csc_data || !csc_rows || (!events && m > 0)))) return BPX_ERR_ARG;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  const bool homo = data_len == 1;
  if (data_dtype != BPX_F32) return BPX_ERR_DTYPE;
#define FO_LAUNCH(E_)                                                                                            \
  do {                                                                                                           \
    if (homo)                                                                                                    \
      BPX_LAUNCH(KID_FIXED_ORDER, (k_fixed_order<true, float, E_>), dim3(blocks), dim3(256), 0, s,               \
                 (const float*)csc_data, csc_rows, csc_colptr, nnz, (const E_*)events, ev_rule, m, n, accumulate, (float*)out); \
    else                                                                                                         \
      BPX_LAUNCH(KID_FIXED_ORDER, (k_fixed_order<false, float, E_>), dim3(blocks), dim3(256), 0, s,              \
                 (const float*)csc_data, csc_rows, csc_colptr, nnz, (const E_*)events, ev_rule, m, n, accumulate, (float*)out); \
  } while (0)
  switch (ev_dtype) {
    case BPX_BOOL: FO_LAUNCH(uint8_t); break;
    case BPX_F32: FO_LAUNCH(float); break;
    default: return BPX_ERR_DTYPE;
  }
#undef FO_LAUNCH
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

}  // extern "C"
