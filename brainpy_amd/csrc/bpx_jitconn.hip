// bpx_jitconn.hip -- just-in-time connectivity matvec on gfx950.
//
// The matrix is regenerated on the fly from LFSR88 streams and never stored.
// Connectivity is the Taichi kernels' (jitconn/_event_matvec_taichi.py,
// jitconn/_matvec_taichi.py), reproduced bit-exactly:
//   lane32 (GPU kernels): 32 streams per row (outdim_parallel) or per column,
//     stream key seed + i*32 + l, lane l walking (step*l - 1, step*(l+1) - 1)
//     in u32 with step = max((num_row+1) >> 5, 1) taken from num_row.
//   row (CPU kernels): one stream per row / column, key seed + i, first index
//     bits % clen, then increments bits % clen + 1.
// The u32 modulo by clen is an exact multiply-high (DivU32).
//
// Launch shapes:
//   outdim_parallel, lane32: one thread per (row, lane): a wave covers two
//     rows; the 32 lane partials are combined with a 32-wide shuffle sum and
//     written once (no atomics, no zero-fill).
//   scatter (outdim_parallel=False), lane32: 256-thread blocks own 256
//     columns; events are balloted into an LDS list of active columns, then 8
//     groups of 32 lanes regenerate the rows of one active column each and
//     add into the zero-filled output with FP32 atomics.  Inactive columns
//     cost one event byte.
//   row layout: one thread per row (outdim_parallel) or per active column.
#include "bpx_common.h"
#include "bpx_rng.h"

namespace bpx {

enum { VM_DENSE = 0, VM_BOOL = 1, VM_FLOAT = 2 };

template <int DIST>
__device__ __forceinline__ float draw_weight(Lfsr88& k, float w0, float w1) {
  if constexpr (DIST == BPX_DIST_UNIFORM) {
    // lfsr88_uniform (tifunc.py:156-168): f32(bits)*2^-32, then r*(hi-lo)+lo
    const float r = (float)k.next() * 2.3283064365386963e-10f;
    return r * (w1 - w0) + w0;
  } else if constexpr (DIST == BPX_DIST_NORMAL) {
    // lfsr88_normal (tifunc.py:89-133): Box-Muller, sine branch
    float u1 = (float)k.next() * 2.3283064365386963e-10f;
    const float u2 = (float)k.next() * 2.3283064365386963e-10f;
    u1 = u1 > 1e-10f ? u1 : 1e-10f;
    const float mag = sqrtf(-2.0f * logf(u1));
    const float z = mag * sinf(6.283185307179586f * u2);
    return w0 + w1 * z;
  } else {
    return w0;
  }
}

template <int VM>
__device__ __forceinline__ float vec_val(const void* v, int64_t i) {
  if constexpr (VM == VM_BOOL) return ((const uint8_t*)v)[i] ? 1.f : 0.f;
  else return ((const float*)v)[i];
}

template <int VM>
__device__ __forceinline__ bool vec_on(const void* v, int64_t i) {
  if constexpr (VM == VM_BOOL) return ((const uint8_t*)v)[i] != 0;
  else return ((const float*)v)[i] != 0.f;
}

// ---------------------------------------------------------------------------
// lane32, outdim_parallel: thread (row, lane)
// ---------------------------------------------------------------------------
template <int DIST, int VM>
__global__ __launch_bounds__(256) void k_jit_lane32_outdim(const void* __restrict__ vec, float w0, float w1,
                                                           DivU32 dv, uint32_t seed, int64_t num_row,
                                                           int64_t num_col, uint32_t step, float* __restrict__ out) {
  const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t row = gid >> 5;
  const uint32_t l = (uint32_t)(gid & 31);
  float acc = 0.f;
  if (row < num_row) {
    uint32_t c = step * l - 1u;
    uint32_t end = c + step;
    if ((int64_t)end > num_col) end = (uint32_t)num_col;
    Lfsr88 k(seed + (uint32_t)gid);
    c += (uint32_t)rand_inc(k, dv);
    while (c < end) {
      if constexpr (VM == VM_DENSE) {
        if constexpr (DIST == BPX_DIST_HOMO) acc += vec_val<VM>(vec, c);
        else { const float w = draw_weight<DIST>(k, w0, w1); acc += vec_val<VM>(vec, c) * w; }
      } else {
        const float w = draw_weight<DIST>(k, w0, w1);
        acc += w * vec_val<VM>(vec, c);
      }
      c += (uint32_t)rand_inc(k, dv);
    }
    if constexpr (VM == VM_DENSE && DIST == BPX_DIST_HOMO) acc = w0 * acc;
  }
  acc = group_sum<32>(acc);
  if (row < num_row && l == 0) out[row] = acc;
}

// ---------------------------------------------------------------------------
// lane32, scatter: 256 columns per block, 8 groups of 32 lanes
// ---------------------------------------------------------------------------
template <int DIST, int VM>
__global__ __launch_bounds__(256) void k_jit_lane32_scatter(const void* __restrict__ vec, float w0, float w1,
                                                            DivU32 dv, uint32_t seed, int64_t num_row,
                                                            int64_t num_col, uint32_t step, float* __restrict__ out) {
  __shared__ int s_cols[256];
  __shared__ int s_wcnt[4];
  const int64_t col0 = (int64_t)blockIdx.x * 256;
  {
    const int64_t c = col0 + threadIdx.x;
    bool act = c < num_col;
    if constexpr (VM != VM_DENSE) act = act && vec_on<VM>(vec, c);
    const uint64_t mask = __ballot(act);
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 0) s_wcnt[w] = __popcll(mask);
    __syncthreads();
    int off = 0;
    for (uint32_t q = 0; q < w; ++q) off += s_wcnt[q];
    if (act) s_cols[off + lanes_below(mask)] = (int)threadIdx.x;
    __syncthreads();
  }
  const int cnt = s_wcnt[0] + s_wcnt[1] + s_wcnt[2] + s_wcnt[3];
  const int grp = threadIdx.x >> 5;
  const uint32_t l = threadIdx.x & 31;
  for (int e = grp; e < cnt; e += 8) {
    const int64_t col = col0 + s_cols[e];
    float cv = 0.f;
    if constexpr (VM == VM_DENSE) cv = ((const float*)vec)[col];
    uint32_t r = step * l - 1u;
    uint32_t end = r + step;
    if ((int64_t)end > num_row) end = (uint32_t)num_row;
    Lfsr88 k(seed + (uint32_t)(col * 32 + l));
    r += (uint32_t)rand_inc(k, dv);
    while (r < end) {
      float v;
      if constexpr (VM == VM_DENSE) {
        if constexpr (DIST == BPX_DIST_HOMO) v = w0 * cv;
        else v = draw_weight<DIST>(k, w0, w1) * cv;
      } else {
        v = draw_weight<DIST>(k, w0, w1);
      }
      atomic_add(out + r, v);
      r += (uint32_t)rand_inc(k, dv);
    }
  }
}

// ---------------------------------------------------------------------------
// row layout
// ---------------------------------------------------------------------------
template <int DIST, int VM>
__global__ __launch_bounds__(256) void k_jit_row_outdim(const void* __restrict__ vec, float w0, float w1, DivU32 dv,
                                                        uint32_t seed, int64_t num_row, int64_t num_col,
                                                        float* __restrict__ out) {
  const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (row >= num_row) return;
  float acc = 0.f;
  Lfsr88 k(seed + (uint32_t)row);
  int32_t c = rand_first(k, dv);
  while (c < num_col) {
    if constexpr (VM == VM_DENSE) {
      if constexpr (DIST == BPX_DIST_HOMO) acc += vec_val<VM>(vec, c);
      else { const float w = draw_weight<DIST>(k, w0, w1); acc += vec_val<VM>(vec, c) * w; }
    } else {
      const float w = draw_weight<DIST>(k, w0, w1);
      if (vec_on<VM>(vec, c)) acc += w;
    }
    c += rand_inc(k, dv);
  }
  if constexpr (VM == VM_DENSE && DIST == BPX_DIST_HOMO) acc = acc * w0;
  out[row] = acc;
}

template <int DIST, int VM>
__global__ __launch_bounds__(256) void k_jit_row_scatter(const void* __restrict__ vec, float w0, float w1, DivU32 dv,
                                                         uint32_t seed, int64_t num_row, int64_t num_col,
                                                         float* __restrict__ out) {
  const int64_t col = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (col >= num_col) return;
  if constexpr (VM != VM_DENSE) {
    if (!vec_on<VM>(vec, col)) return;
  }
  float cv = 0.f;
  if constexpr (VM == VM_DENSE) cv = ((const float*)vec)[col];
  const float hv = cv * w0;
  Lfsr88 k(seed + (uint32_t)col);
  int32_t r = rand_first(k, dv);
  while (r < num_row) {
    float v;
    if constexpr (VM == VM_DENSE) {
      if constexpr (DIST == BPX_DIST_HOMO) v = hv;
      else v = cv * draw_weight<DIST>(k, w0, w1);
    } else {
      v = draw_weight<DIST>(k, w0, w1);
    }
    atomic_add(out + r, v);
    r += rand_inc(k, dv);
  }
}

__global__ __launch_bounds__(256) void k_zero_f32(float* __restrict__ out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = 0.f;
}

template <int DIST, int VM>
static int run_jit(hipStream_t s, const void* vec, float w0, float w1, int32_t clen, uint32_t seed, int64_t num_row,
                   int64_t num_col, int outdim, int layout, float* out) {
  const DivU32 dv = DivU32::make((uint32_t)clen);
  const uint32_t step = (uint32_t)(((num_row + 1) >> 5) > 1 ? ((num_row + 1) >> 5) : 1);
  if (outdim) {
    if (layout == BPX_LAYOUT_LANE32) {
      const int64_t blocks = (num_row * 32 + 255) / 256;
      BPX_LAUNCH(KID_JIT_OUTDIM, (k_jit_lane32_outdim<DIST, VM>), dim3((unsigned)blocks), dim3(256), 0, s, vec, w0,
                 w1, dv, seed, num_row, num_col, step, out);
    } else {
      const int64_t blocks = (num_row + 255) / 256;
      BPX_LAUNCH(KID_JIT_OUTDIM, (k_jit_row_outdim<DIST, VM>), dim3((unsigned)blocks), dim3(256), 0, s, vec, w0, w1,
                 dv, seed, num_row, num_col, out);
    }
  } else {
    int64_t zb = (num_row + 255) / 256;
    if (zb > 4096) zb = 4096;
    hipLaunchKernelGGL(k_zero_f32, dim3((unsigned)zb), dim3(256), 0, s, out, num_row);
    if (num_col > 0) {
      const int64_t blocks = (num_col + 255) / 256;
      if (layout == BPX_LAYOUT_LANE32)
        BPX_LAUNCH(KID_JIT_SCATTER, (k_jit_lane32_scatter<DIST, VM>), dim3((unsigned)blocks), dim3(256), 0, s, vec,
                   w0, w1, dv, seed, num_row, num_col, step, out);
      else
        BPX_LAUNCH(KID_JIT_SCATTER, (k_jit_row_scatter<DIST, VM>), dim3((unsigned)blocks), dim3(256), 0, s, vec, w0,
                   w1, dv, seed, num_row, num_col, out);
    }
  }
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

template <int DIST>
static int run_jit_vm(int vm, hipStream_t s, const void* vec, float w0, float w1, int32_t clen, uint32_t seed,
                      int64_t num_row, int64_t num_col, int outdim, int layout, float* out) {
  if (vm == VM_DENSE) return run_jit<DIST, VM_DENSE>(s, vec, w0, w1, clen, seed, num_row, num_col, outdim, layout, out);
  if (vm == VM_BOOL) return run_jit<DIST, VM_BOOL>(s, vec, w0, w1, clen, seed, num_row, num_col, outdim, layout, out);
  return run_jit<DIST, VM_FLOAT>(s, vec, w0, w1, clen, seed, num_row, num_col, outdim, layout, out);
}

}  // namespace bpx

using namespace bpx;

extern "C" int bpx_jitconn_mv(void* stream, const void* vec, int vec_dtype, int event, int dist, float w0, float w1,
                              int32_t clen, uint32_t seed, int64_t num_row, int64_t num_col, int outdim_parallel,
                              int layout, float* out) {
  if (num_row < 0 || num_col < 0) return BPX_ERR_SHAPE;
  if (num_row > INT32_MAX || num_col > INT32_MAX || num_row * 32 > (int64_t)UINT32_MAX ||
      num_col * 32 > (int64_t)UINT32_MAX)
    return BPX_ERR_SHAPE;
  if (clen < 1) return BPX_ERR_ARG;
  if (layout != BPX_LAYOUT_LANE32 && layout != BPX_LAYOUT_ROW) return BPX_ERR_ARG;
  if (dist < BPX_DIST_HOMO || dist > BPX_DIST_NORMAL) return BPX_ERR_ARG;
  int vm;
  if (!event) {
    if (vec_dtype != BPX_F32) return BPX_ERR_DTYPE;
    vm = VM_DENSE;
  } else if (vec_dtype == BPX_BOOL) {
    vm = VM_BOOL;
  } else if (vec_dtype == BPX_F32) {
    vm = VM_FLOAT;
  } else {
    return BPX_ERR_DTYPE;
  }
  hipStream_t s = (hipStream_t)stream;
  if (num_row == 0) return BPX_OK;
  const int o = outdim_parallel ? 1 : 0;
  switch (dist) {
    case BPX_DIST_HOMO: return run_jit_vm<BPX_DIST_HOMO>(vm, s, vec, w0, w1, clen, seed, num_row, num_col, o, layout, out);
    case BPX_DIST_UNIFORM: return run_jit_vm<BPX_DIST_UNIFORM>(vm, s, vec, w0, w1, clen, seed, num_row, num_col, o, layout, out);
    default: return run_jit_vm<BPX_DIST_NORMAL>(vm, s, vec, w0, w1, clen, seed, num_row, num_col, o, layout, out);
  }
}
