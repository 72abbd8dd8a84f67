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
// Both return 0 or a negative bpx status code.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bpx.h"

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

}  // extern "C"
