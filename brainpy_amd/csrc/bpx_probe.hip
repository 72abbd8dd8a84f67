// bpx_probe.hip -- measurement probes, built into libbpx_probe.so (not part
// of libbpx.so or of the product ABI in include/bpx.h; loaded only by the
// benchmark).
//
//   int bpx_probe_copy(void* stream, const void* src, void* dst, int64_t bytes);
//     A plain streaming copy (16-byte vector loads and stores, grid-stride):
//     the known-size kernel that calibrates the FETCH_SIZE / WRITE_SIZE
//     counters against bytes.
//
//   int bpx_probe_coo(void* stream, const int32_t* row, const int32_t* col, const float* data,
//                     const float* src, int64_t src_len, int64_t nse, int mode, float* out,
//                     int64_t out_len);
//     Ceilings of the COO / CSR-gather access pattern on a given (row, col,
//     data) set: mode 0 streams row, col and data (12 B per entry), mode 1
//     streams col and gathers src[col], mode 2 does both; one f32 sum per
//     thread is written to out (out_len a multiple of 256, one 256-thread
//     block per 256 entries of out, at most 65536 blocks).  nse must be a
//     multiple of 4, the arrays 16-byte aligned, and col values inside
//     [0, src_len).
//
//   int bpx_probe_valu(void* stream, int kind, int64_t iters, int blocks, int threads,
//                      uint64_t* out);
//     Integer issue-rate probe: every thread runs `iters` iterations of one
//     loop body and folds its results into out[thread] (so nothing is dead
//     code); thread 0 of each block stores its shader-clock and 100 MHz
//     wall-clock ticks at start and end into out[total_threads + 4 * block
//     ...] (4 words per block), from which the shader clock under load
//     follows.  kind 0: one LFSR88 draw per iteration (one generator,
//     three independent components); kind 1: two generators interleaved;
//     kind 2: the draw plus the multiply-high modulo and the column add of
//     the jitconn kernels; kind 3: one LFSR88 draw of a wave-uniform
//     generator (scalar ALU); kind 4: kinds 0 and 3 interleaved; kinds 5,
//     6, 7: kind 2 plus 8 independent FP32 fused multiply-adds, FP64
//     fused multiply-adds or 32-bit selects per iteration.
//
//   int bpx_probe_segstream(void* stream, const int32_t* desc_off, const int32_t* desc,
//                           const uint16_t* lidx, const float* w, int P, int ncopy, int R,
//                           int mode, uint32_t* out);
//     The segment reads of the tiled event scatter without its listing:
//     block b takes the (copy, tile) pair of the scatter kernel and streams
//     the segments desc[2 * e], desc[2 * e + 1] (start, end into lidx / w)
//     for e in [desc_off[copy * P + tile], desc_off[copy * P + tile + 1]),
//     8 lanes per segment, R segments in flight per lane group; mode 0
//     folds the loaded words into one checksum per thread, mode 1 adds
//     each weight into a 3200-entry LDS array as 64-bit integers (the
//     scatter's accumulation).  out has ncopy * P * 1024 words.
//
// All return 0 or a negative bpx status code.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bpx.h"
#include "bpx_rng.h"

namespace {

__global__ __launch_bounds__(256) void k_probe_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                      int64_t n16) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n16; i += stride) dst[i] = src[i];
}

__global__ void k_probe_copy1(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// per lane, 4 consecutive entries per dwordx4 load, grid-stride
__global__ __launch_bounds__(256) void k_probe_coo(const int32_t* __restrict__ row, const int32_t* __restrict__ col,
                                                   const float* __restrict__ w, const float* __restrict__ src,
                                                   int64_t n4, int mode, float* __restrict__ out) {
  typedef int32_t i4 __attribute__((ext_vector_type(4)));
  typedef float f4 __attribute__((ext_vector_type(4)));
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  float acc = 0.f;
  for (int64_t i = tid; i < n4; i += stride) {
    const i4 c = ((const i4*)col)[i];
    if (mode != 1) {
      const i4 r = ((const i4*)row)[i];
      const f4 x = ((const f4*)w)[i];
      if (mode == 0)
        acc += (float)(r.x ^ c.y) + x.x + x.y + x.z + x.w + (float)(r.w + c.w);
      else
        acc += x.x * src[c.x] + x.y * src[c.y] + x.z * src[c.z] + x.w * src[c.w] + (float)(r.x & 1);
    } else {
      acc += src[c.x] + src[c.y] + src[c.z] + src[c.w];
    }
  }
  out[tid] = acc;
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// the host form of the taus88 step (shifts, masks, xors): on a wave-uniform
// state the compiler keeps it on the scalar ALU
__device__ __forceinline__ uint32_t taus88_scalar(uint32_t& s0, uint32_t& s1, uint32_t& s2) {
  uint32_t b = ((s0 << 13) ^ s0) >> 19;
  s0 = ((s0 & 4294967294u) << 12) ^ b;
  b = ((s1 << 2) ^ s1) >> 25;
  s1 = ((s1 & 4294967288u) << 4) ^ b;
  b = ((s2 << 3) ^ s2) >> 11;
  s2 = ((s2 & 4294967280u) << 17) ^ b;
  return s0 ^ s1 ^ s2;
}

template <int R, int MODE>
__global__ __launch_bounds__(1024) void k_probe_segstream(const int32_t* __restrict__ desc_off,
                                                          const int32_t* __restrict__ desc,
                                                          const uint16_t* __restrict__ lidx,
                                                          const float* __restrict__ w, int P, int ncopy,
                                                          uint32_t* __restrict__ out) {
  __shared__ unsigned long long s_acc[3200];
  if (MODE == 1)
    for (int k = threadIdx.x; k < 3200; k += 1024) s_acc[k] = 0ull;
  __syncthreads();
  const int b = blockIdx.x;
  const int pair = (b & 7) * (ncopy * P / 8) + (b >> 3);
  const int blk = (pair / P) * P + pair % P;
  const int d0 = desc_off[blk], d1 = desc_off[blk + 1];
  const int g = threadIdx.x >> 3, lv = threadIdx.x & 7;
  uint32_t acc = 0;
  for (int e = d0 + g; e < d1; e += R * 128) {
    int32_t st[R], en[R];
    int32_t len = 0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const bool ok = e + q * 128 < d1;
      st[q] = ok ? desc[2 * (e + q * 128)] : 0;
      en[q] = ok ? desc[2 * (e + q * 128) + 1] : 0;
      len = max(len, en[q] - st[q]);
    }
    for (int k = 4 * lv; k < len; k += 32) {
      uint2 li[R];
      float4 wv[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        li[q] = make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
        wv[q] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (st[q] + k < en[q]) {
          li[q] = *reinterpret_cast<const uint2*>(lidx + st[q] + k);
          wv[q] = *reinterpret_cast<const float4*>(w + st[q] + k);
        }
      }
#pragma unroll
      for (int q = 0; q < R; ++q) {
        if constexpr (MODE == 0) {
          acc ^= li[q].x ^ li[q].y ^ __float_as_uint(wv[q].x) ^ __float_as_uint(wv[q].y) ^
                 __float_as_uint(wv[q].z) ^ __float_as_uint(wv[q].w);
        } else {
          const float wf[4] = {wv[q].x, wv[q].y, wv[q].z, wv[q].w};
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t c = ((u < 2 ? li[q].x : li[q].y) >> (16 * (u & 1))) & 0xFFFFu;
            if (c < 3200u)
              __hip_atomic_fetch_add(s_acc + c, (unsigned long long)(long long)(wf[u] * 1048576.0f),
                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        }
      }
    }
  }
  if constexpr (MODE == 1) {
    __syncthreads();
    for (int k = threadIdx.x; k < 3200; k += 1024) acc ^= (uint32_t)s_acc[k];
  }
  out[(size_t)b * 1024 + threadIdx.x] = acc;
}

template <int KIND>
__global__ void k_probe_valu(int64_t iters, bpx::DivShort ds, uint64_t* __restrict__ out) {
  const uint64_t t0 = clock64(), w0 = wall_clock64();
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  bpx::Lfsr88 a(tid * 2u), b(tid * 2u + 1u);
  uint32_t acc = 0, c = tid;
  const int32_t negd = -(int32_t)ds.d;
  uint32_t u0 = __builtin_amdgcn_readfirstlane(blockIdx.x + 1u), u1 = u0 + 6u, u2 = u0 + 14u, uacc = 0;
  float f[8];
  double g[8];
  uint32_t sel[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    f[k] = (float)(tid + k);
    g[k] = (double)(tid + k);
    sel[k] = tid * (k + 1);
  }
  const float fm = 1.0000001f + (float)(tid & 1) * 1e-7f;
  const double gm = 1.0000000001 + (double)(tid & 1) * 1e-12;
  for (int64_t i = 0; i < iters; ++i) {
    if constexpr (KIND >= 5) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if constexpr (KIND == 5) f[k] = __builtin_fmaf(f[k], fm, 0.5f);
        else if constexpr (KIND == 6) g[k] = __builtin_fma(g[k], gm, 0.5);
        else sel[k] = ((sel[k] ^ (uint32_t)i) & 1u) ? sel[k] + 3u : sel[k] ^ 5u;
      }
    }
    if constexpr (KIND == 0) {
      acc ^= a.next();
    } else if constexpr (KIND == 3) {
      uacc += taus88_scalar(u0, u1, u2);
    } else if constexpr (KIND == 4) {
      acc ^= a.next();
      uacc += taus88_scalar(u0, u1, u2);
    } else if constexpr (KIND == 1) {
      acc ^= a.next();
      acc += b.next();
    } else if constexpr (KIND == 2 || KIND >= 5) {
      const uint32_t x = a.next();
      const uint32_t q = ds.div(x);
      int32_t r;
      asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(q), "v"(negd), "v"(x));
      c = c + (uint32_t)r + 1u;
    }
  }
  uint32_t extra = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) extra += (uint32_t)f[k] + (uint32_t)g[k] + sel[k];
  out[tid] = (uint64_t)acc + c + uacc + extra;
  const uint64_t t1 = clock64(), w1 = wall_clock64();
  if (threadIdx.x == 0) {
    uint64_t* o = out + (uint64_t)gridDim.x * blockDim.x + 4u * blockIdx.x;
    o[0] = t0; o[1] = t1; o[2] = w0; o[3] = w1;
  }
}

}  // namespace

extern "C" {

int bpx_probe_copy(void* stream, const void* src, void* dst, int64_t bytes) {
  if (bytes < 0) return BPX_ERR_ARG;
  if (bytes == 0) return BPX_OK;
  if (!src || !dst) return BPX_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const bool aligned = (((uintptr_t)src | (uintptr_t)dst | (uintptr_t)bytes) & 15) == 0;
  if (aligned) {
    const int64_t n16 = bytes / 16;
    int64_t blocks = (n16 + 255) / 256;
    if (blocks > 256 * 32) blocks = 256 * 32;
    hipLaunchKernelGGL(k_probe_copy16, dim3((unsigned)blocks), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, n16);
  } else {
    const int64_t blocks = (bytes + 255) / 256;
    hipLaunchKernelGGL(k_probe_copy1, dim3((unsigned)blocks), dim3(256), 0, s, (const uint8_t*)src, (uint8_t*)dst,
                       bytes);
  }
  return hipGetLastError() == hipSuccess ? BPX_OK : BPX_ERR_LAUNCH;
}

int bpx_probe_coo(void* stream, const int32_t* row, const int32_t* col, const float* data, const float* src,
                  int64_t src_len, int64_t nse, int mode, float* out, int64_t out_len) {
  if (nse < 0 || (nse & 3) || out_len < 256 || (out_len & 255) || mode < 0 || mode > 2) return BPX_ERR_SHAPE;
  if (!col || !out || (mode != 1 && (!row || !data)) || (mode != 0 && (!src || src_len <= 0))) return BPX_ERR_ARG;
  for (const void* p : {(const void*)row, (const void*)col, (const void*)data})
    if (p && !aligned16(p)) return BPX_ERR_ARG;
  int64_t blocks = out_len / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_probe_coo, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, row, col, data, src,
                     nse / 4, mode, out);
  return hipGetLastError() == hipSuccess ? BPX_OK : BPX_ERR_LAUNCH;
}

int bpx_probe_valu(void* stream, int kind, int64_t iters, int blocks, int threads, uint64_t* out) {
  if (kind < 0 || kind > 7 || iters < 0 || blocks <= 0 || threads <= 0 || threads > 1024 || (threads & 63))
    return BPX_ERR_ARG;
  if (!out) return BPX_ERR_ARG;
  bpx::DivShort ds = {1u, 0u, 0u};
  if (!bpx::DivShort::make(19999u, &ds)) return BPX_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  switch (kind) {
    case 0: hipLaunchKernelGGL(k_probe_valu<0>, dim3(blocks), dim3(threads), 0, s, iters, ds, out); break;
    case 1: hipLaunchKernelGGL(k_probe_valu<1>, dim3(blocks), dim3(threads), 0, s, iters, ds, out); break;
    case 2: hipLaunchKernelGGL(k_probe_valu<2>, dim3(blocks), dim3(threads), 0, s, iters, ds, out); break;
    case 3: hipLaunchKernelGGL(k_probe_valu<3>, dim3(blocks), dim3(threads), 0, s, iters, ds, out); break;
    case 4: hipLaunchKernelGGL(k_probe_valu<4>, dim3(blocks), dim3(threads), 0, s, iters, ds, out); break;
    case 5: hipLaunchKernelGGL(k_probe_valu<5>, dim3(blocks), dim3(threads), 0, s, iters, ds, out); break;
    case 6: hipLaunchKernelGGL(k_probe_valu<6>, dim3(blocks), dim3(threads), 0, s, iters, ds, out); break;
    default: hipLaunchKernelGGL(k_probe_valu<7>, dim3(blocks), dim3(threads), 0, s, iters, ds, out); break;
  }
  return hipGetLastError() == hipSuccess ? BPX_OK : BPX_ERR_LAUNCH;
}

int bpx_probe_segstream(void* stream, const int32_t* desc_off, const int32_t* desc, const uint16_t* lidx,
                        const float* w, int P, int ncopy, int R, int mode, uint32_t* out) {
  if (!desc_off || !desc || !lidx || !w || !out || P <= 0 || ncopy <= 0 || ((ncopy * P) & 7)) return BPX_ERR_ARG;
  if (mode < 0 || mode > 1) return BPX_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((unsigned)(ncopy * P));
#define SEG(RV, MV) hipLaunchKernelGGL((k_probe_segstream<RV, MV>), grid, dim3(1024), 0, s, desc_off, desc, lidx, w, P, ncopy, out)
  if (R == 1) { if (mode) SEG(1, 1); else SEG(1, 0); }
  else if (R == 2) { if (mode) SEG(2, 1); else SEG(2, 0); }
  else if (R == 4) { if (mode) SEG(4, 1); else SEG(4, 0); }
  else return BPX_ERR_ARG;
#undef SEG
  return hipGetLastError() == hipSuccess ? BPX_OK : BPX_ERR_LAUNCH;
}

}  // extern "C"
