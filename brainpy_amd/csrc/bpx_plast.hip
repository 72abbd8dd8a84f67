// bpx_plast.hip -- event-driven STDP weight updates on stored connectivity
// (dnn/linear.py:238-290, 650-711).
//
// On every spike of a pre-synaptic neuron i the weights of its synapses are
// moved by the post-synaptic trace, w[k] = min(max(w[k] + trace[post(k)],
// w_min), w_max) for k in row i (CSR, csr_on_pre_update); on every spike of
// a post-synaptic neuron j the weights of the synapses into j are moved by the
// pre-synaptic trace, through the CSC view (pre ids, column pointers and the
// synapse id of each CSC entry; csc_on_post_update).  Each synapse belongs to
// one row and to one column, so every weight is written at most once per
// call and no atomics are needed; the result does not depend on scheduling.
// The dense forms update rows (dense_on_pre) or columns (dense_on_post) of a
// [pre, post] matrix with np.clip.
//
// Launch: waves own 64-row chunks of the spike vector; the active rows of a
// chunk (ballot) are walked by the whole wave, 64 synapses per step, so the
// index, weight and trace accesses of a row are coalesced (trace reads are
// gathers).  The update runs in place on `out` after `w` is copied into it
// when the two differ.
#include "bpx_common.h"
#include <stdlib.h>

namespace bpx {

constexpr int kPlastThreads = 256;
constexpr int kPlastU = 8;   // entries per lane per step of a row walk

template <typename W>
__device__ __forceinline__ W clamp_mm(W x, W lo, W hi) {
  // np.minimum(np.maximum(x, lo), hi)
  x = x > lo ? x : (x == x ? lo : x);
  return x < hi ? x : (x == x ? hi : x);
}

// CSR (pre) or CSC (post) form: rows of `ptr` are the spiking side; `other`
// maps an entry to the trace index; `wid` maps an entry to its synapse
// (null: the entry index itself).
template <typename I, typename E, typename W, bool WID>
__global__ __launch_bounds__(kPlastThreads) void k_sparse_on_spike(const I* __restrict__ other,
                                                                   const I* __restrict__ ptr,
                                                                   const I* __restrict__ wid,
                                                                   const E* __restrict__ spike, int rule,
                                                                   int64_t nrows, const W* __restrict__ trace,
                                                                   int64_t ntrace, W w_min, W w_max,
                                                                   W* __restrict__ out, int64_t nnz, int rw) {
  // a wave owns chunks of rw (1..64) consecutive rows; lanes below rw read
  // their row's spike flag and the wave walks the active rows one by one
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * kPlastThreads + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kPlastThreads) >> 6;
  for (int64_t base = wave * rw; base < nrows; base += nwaves * rw) {
    const int64_t r = base + lane;
    const bool act = lane < rw && r < nrows && ev_true(spike[r], rule);
    uint64_t mask = __ballot(act);
    while (mask) {
      const int b = __builtin_ctzll(mask);
      mask &= mask - 1;
      const int64_t row = base + b;
      int64_t s = (int64_t)ptr[row], t = (int64_t)ptr[row + 1];
      s = s < 0 ? 0 : s;
      t = t > nnz ? nnz : t;
      // kPlastU entries per lane per step, all loads of a step issued before
      // any of their uses (a row of 1000 synapses takes two steps)
      for (int64_t k0 = s + lane; k0 < t; k0 += 64 * kPlastU) {
        int64_t j[kPlastU], l[kPlastU];
        bool ok[kPlastU];
        W wv[kPlastU], tv[kPlastU];
#pragma unroll
        for (int u = 0; u < kPlastU; ++u) {
          const int64_t k = k0 + 64 * u;
          ok[u] = k < t;
          const int64_t kc = ok[u] ? k : s;
          j[u] = (int64_t)other[kc];
          l[u] = WID ? (int64_t)wid[kc] : kc;
          // CSR form: the weight sits at the entry itself, so its load goes
          // out with the index load instead of after it
          if constexpr (!WID) wv[u] = out[kc];
        }
#pragma unroll
        for (int u = 0; u < kPlastU; ++u) {
          ok[u] = ok[u] && (uint64_t)j[u] < (uint64_t)ntrace && (uint64_t)l[u] < (uint64_t)nnz;
          if constexpr (WID) wv[u] = ok[u] ? out[l[u]] : W(0);
          tv[u] = ok[u] ? trace[j[u]] : W(0);
        }
#pragma unroll
        for (int u = 0; u < kPlastU; ++u)
          if (ok[u]) out[l[u]] = clamp_mm(wv[u] + tv[u], w_min, w_max);
      }
    }
  }
}

// dense [pre, post]: ROWS = true updates the rows of spiking pre neurons with
// trace[post]; false the columns of spiking post neurons with trace[pre]
template <bool ROWS, typename E, typename W>
__global__ __launch_bounds__(kPlastThreads) void k_dense_on_spike(const E* __restrict__ spike, int rule,
                                                                  int64_t npre, int64_t npost,
                                                                  const W* __restrict__ trace, W w_min, W w_max,
                                                                  W* __restrict__ out) {
  const int64_t total = npre * npost;
  for (int64_t i = (int64_t)blockIdx.x * kPlastThreads + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * kPlastThreads) {
    const int64_t r = i / npost, c = i - r * npost;
    const bool on = ROWS ? ev_true(spike[r], rule) : ev_true(spike[c], rule);
    if (on) {
      // np.clip(x, lo, hi)
      W x = out[i] + trace[ROWS ? c : r];
      x = x < w_min ? w_min : x;
      x = x > w_max ? w_max : x;
      out[i] = x;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_copy_t(const T* __restrict__ src, T* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}

template <typename W>
static void copy_weights(hipStream_t s, const void* w, void* out, int64_t n) {
  if (w == out || n <= 0) return;
  int64_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  hipLaunchKernelGGL(k_copy_t<W>, dim3((unsigned)b), dim3(256), 0, s, (const W*)w, (W*)out, n);
}

// rows per wave chunk: the largest power of two (at most 64) that still gives
// about 16384 waves, so that enough rows' dependent index -> weight chains
// are in flight
static int plast_rows_per_wave(int64_t nrows) {
  int rw = 64;
  while (rw > 1 && nrows / rw < 16384) rw >>= 1;
  return rw;
}

static unsigned plast_blocks(int64_t nrows, int rw) {
  int64_t waves = (nrows + rw - 1) / rw;
  int64_t b = (waves + kPlastThreads / 64 - 1) / (kPlastThreads / 64);
  if (b > 4096) b = 4096;
  return (unsigned)(b < 1 ? 1 : b);
}

template <typename I, typename W>
static int sparse_on_spike(hipStream_t s, const void* w, const void* other, const void* ptr, const void* wid,
                           int64_t nnz, const void* spike, int sp_dtype, int64_t nrows, const void* trace,
                           int64_t ntrace, double w_min, double w_max, void* out) {
  copy_weights<W>(s, w, out, nnz);
  if (nrows == 0 || nnz == 0) return BPX_OK;
  const int rw = plast_rows_per_wave(nrows);
  const unsigned blocks = plast_blocks(nrows, rw);
#define LAUNCH_EW(E, HW)                                                                                        \
  hipLaunchKernelGGL((k_sparse_on_spike<I, E, W, HW>), dim3(blocks), dim3(kPlastThreads), 0, s, (const I*)other, \
                     (const I*)ptr, (const I*)wid, (const E*)spike, BPX_RULE_NONZERO, nrows, (const W*)trace,    \
                     ntrace, (W)w_min, (W)w_max, (W*)out, nnz, rw)
#define LAUNCH_E(E)            \
  do {                         \
    if (wid) LAUNCH_EW(E, true); \
    else LAUNCH_EW(E, false);  \
  } while (0)
  if (sp_dtype == BPX_BOOL) LAUNCH_E(uint8_t);
  else if (sp_dtype == BPX_F32) LAUNCH_E(float);
  else if (sp_dtype == BPX_F64) LAUNCH_E(double);
  else return BPX_ERR_DTYPE;
#undef LAUNCH_E
#undef LAUNCH_EW
  return BPX_OK;
}

template <typename W>
static int sparse_on_spike_i(hipStream_t s, const void* w, const void* other, const void* ptr, const void* wid,
                             int index_dtype, int64_t nnz, const void* spike, int sp_dtype, int64_t nrows,
                             const void* trace, int64_t ntrace, double w_min, double w_max, void* out) {
  if (index_dtype == BPX_I32)
    return sparse_on_spike<int32_t, W>(s, w, other, ptr, wid, nnz, spike, sp_dtype, nrows, trace, ntrace, w_min,
                                       w_max, out);
  if (index_dtype == BPX_I64)
    return sparse_on_spike<int64_t, W>(s, w, other, ptr, wid, nnz, spike, sp_dtype, nrows, trace, ntrace, w_min,
                                       w_max, out);
  return BPX_ERR_DTYPE;
}

template <bool ROWS, typename W>
static int dense_on_spike(hipStream_t s, const void* w, const void* spike, int sp_dtype, int64_t npre,
                          int64_t npost, const void* trace, double w_min, double w_max, void* out) {
  copy_weights<W>(s, w, out, npre * npost);
  if (npre == 0 || npost == 0) return BPX_OK;
  int64_t b = (npre * npost + kPlastThreads - 1) / kPlastThreads;
  if (b > 8192) b = 8192;
#define LAUNCH_E(E)                                                                                          \
  hipLaunchKernelGGL((k_dense_on_spike<ROWS, E, W>), dim3((unsigned)b), dim3(kPlastThreads), 0, s,           \
                     (const E*)spike, BPX_RULE_NONZERO, npre, npost, (const W*)trace, (W)w_min, (W)w_max,    \
                     (W*)out)
  if (sp_dtype == BPX_BOOL) LAUNCH_E(uint8_t);
  else if (sp_dtype == BPX_F32) LAUNCH_E(float);
  else if (sp_dtype == BPX_F64) LAUNCH_E(double);
  else return BPX_ERR_DTYPE;
#undef LAUNCH_E
  return BPX_OK;
}

}  // namespace bpx

using namespace bpx;

extern "C" {

int bpx_csr_on_pre_update(void* stream, const void* w, int w_dtype, const void* indices, const void* indptr,
                          int index_dtype, int64_t nnz, const void* spike, int spike_dtype, int64_t num_pre,
                          const void* trace, int64_t num_post, double w_min, double w_max, void* out) {
  if (!w || !out || (nnz > 0 && (!indices || !indptr)) || (num_pre > 0 && (!spike || !indptr)) ||
      (num_post > 0 && !trace))
    return BPX_ERR_ARG;
  if (nnz < 0 || num_pre < 0 || num_post < 0) return BPX_ERR_SHAPE;
  hipStream_t s = (hipStream_t)stream;
  int rc;
  if (w_dtype == BPX_F32)
    rc = sparse_on_spike_i<float>(s, w, indices, indptr, nullptr, index_dtype, nnz, spike, spike_dtype, num_pre,
                                  trace, num_post, w_min, w_max, out);
  else if (w_dtype == BPX_F64)
    rc = sparse_on_spike_i<double>(s, w, indices, indptr, nullptr, index_dtype, nnz, spike, spike_dtype, num_pre,
                                   trace, num_post, w_min, w_max, out);
  else
    return BPX_ERR_DTYPE;
  if (rc != BPX_OK) return rc;
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

int bpx_csc_on_post_update(void* stream, const void* w, int w_dtype, const void* pre_ids, const void* colptr,
                           const void* w_ids, int index_dtype, int64_t nnz, const void* spike, int spike_dtype,
                           int64_t num_post, const void* trace, int64_t num_pre, double w_min, double w_max,
                           void* out) {
  if (!w || !out || (nnz > 0 && (!pre_ids || !colptr || !w_ids)) || (num_post > 0 && (!spike || !colptr)) ||
      (num_pre > 0 && !trace))
    return BPX_ERR_ARG;
  if (nnz < 0 || num_pre < 0 || num_post < 0) return BPX_ERR_SHAPE;
  hipStream_t s = (hipStream_t)stream;
  int rc;
  if (w_dtype == BPX_F32)
    rc = sparse_on_spike_i<float>(s, w, pre_ids, colptr, w_ids, index_dtype, nnz, spike, spike_dtype, num_post,
                                  trace, num_pre, w_min, w_max, out);
  else if (w_dtype == BPX_F64)
    rc = sparse_on_spike_i<double>(s, w, pre_ids, colptr, w_ids, index_dtype, nnz, spike, spike_dtype, num_post,
                                   trace, num_pre, w_min, w_max, out);
  else
    return BPX_ERR_DTYPE;
  if (rc != BPX_OK) return rc;
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

int bpx_dense_on_pre(void* stream, const void* w, int w_dtype, const void* spike, int spike_dtype, int64_t num_pre,
                     int64_t num_post, const void* trace, double w_min, double w_max, void* out) {
  if (!w || !out || (num_pre > 0 && !spike) || (num_post > 0 && !trace)) return BPX_ERR_ARG;
  if (num_pre < 0 || num_post < 0) return BPX_ERR_SHAPE;
  hipStream_t s = (hipStream_t)stream;
  int rc;
  if (w_dtype == BPX_F32)
    rc = dense_on_spike<true, float>(s, w, spike, spike_dtype, num_pre, num_post, trace, w_min, w_max, out);
  else if (w_dtype == BPX_F64)
    rc = dense_on_spike<true, double>(s, w, spike, spike_dtype, num_pre, num_post, trace, w_min, w_max, out);
  else
    return BPX_ERR_DTYPE;
  if (rc != BPX_OK) return rc;
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

int bpx_dense_on_post(void* stream, const void* w, int w_dtype, const void* spike, int spike_dtype, int64_t num_pre,
                      int64_t num_post, const void* trace, double w_min, double w_max, void* out) {
  if (!w || !out || (num_post > 0 && !spike) || (num_pre > 0 && !trace)) return BPX_ERR_ARG;
  if (num_pre < 0 || num_post < 0) return BPX_ERR_SHAPE;
  hipStream_t s = (hipStream_t)stream;
  int rc;
  if (w_dtype == BPX_F32)
    rc = dense_on_spike<false, float>(s, w, spike, spike_dtype, num_pre, num_post, trace, w_min, w_max, out);
  else if (w_dtype == BPX_F64)
    rc = dense_on_spike<false, double>(s, w, spike, spike_dtype, num_pre, num_post, trace, w_min, w_max, out);
  else
    return BPX_ERR_DTYPE;
  if (rc != BPX_OK) return rc;
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

}  // extern "C"
