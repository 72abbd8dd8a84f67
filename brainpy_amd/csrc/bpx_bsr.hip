// bpx_bsr.hip -- products with a block-sparse (BSR) matrix:
//
//   * cusparse_bcsr_matvec (sparse/_bsr_mv.py:60-91): y = A x (transpose:
//     y = A^T x) for A of (mb x nb) blocks of bs x bs, stored as
//     data[nnzb][bs][bs] (row-major blocks; the block's row p and column q,
//     as the reference's transpose rule reads them, _bsr_mv.py:175-199),
//     block columns indices[nnzb] and block-row pointers indptr[mb + 1];
//   * bcsrmm (sparse/_bsr_mm.py:210-259): C = A @ B for a dense A (m x k)
//     and B (k x n) stored by 32-column blocks: B_ptr[n/32 + 1] over column
//     blocks, B_indices[nblk][2] = (k block, data block), B_data[nblk][32][32]
//     with B[ki*32 + b][ni*32 + a] = B_data[bi][a][b] (the loop of
//     _bsr_mm.py:292-307).
//
// k_bsr_mv: one wave per block row streams its blocks' values with 16-byte
// loads (4 elements per lane, 256 elements per wave-wide window: a window
// holds 256/bs^2 whole blocks when bs <= 8 and a 256/bs-row slice of one
// block otherwise), gathers the matching 4 entries of x, and keeps one
// register sum per window; lanes of the same output row are summed with
// shuffles at the end of the row.  The transposed form multiplies each
// block row by its x slice, sums over the block's rows with shuffles, moves
// column q of the block to lane q and adds the bs values per block with one
// atomic instruction.  Other block sizes use
// k_bsr_mv_any (a thread per output row, or per block-row column when
// transposed).
//
// k_bsr_mm: f32 MFMA (32x32x2): a wave owns a 32 x 32 output tile per
// m-tile it covers; for each B block of its column block, each lane loads
// half of its A row and half of its B_data row (16 floats each, 16-byte
// loads) and feeds 16 MFMAs; the 16 accumulators per lane are stored at
// the end.
//
// Index contract: blocks whose column index lies outside the shape are
// skipped (no fault); pointers are clamped to [0, nnzb].
#include "bpx_common.h"

namespace bpx {
namespace {

constexpr int kBsrThreads = 256;
constexpr int kBsrU = 4;   // blocks (or windows of blocks) in flight per lane

template <typename V>
__device__ __forceinline__ void load4(const V* p, V (&d)[4]) {
  if constexpr (sizeof(V) == 4) {
    const float4 v = *reinterpret_cast<const float4*>(p);
    d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
  } else {
    const double2 a = *reinterpret_cast<const double2*>(p);
    const double2 b = *reinterpret_cast<const double2*>(p + 2);
    d[0] = a.x; d[1] = a.y; d[2] = b.x; d[3] = b.y;
  }
}

template <typename I>
__device__ __forceinline__ void row_range(const I* indptr, int64_t i, int64_t nnzb, int64_t& s, int64_t& t) {
  s = (int64_t)indptr[i];
  t = (int64_t)indptr[i + 1];
  s = s < 0 ? 0 : (s > nnzb ? nnzb : s);
  t = t > nnzb ? nnzb : t;
}

// BS in {4, 8, 16, 32, 64}; data 16-byte aligned (checked on the host)
template <int BS, typename V, typename I, bool TRANS>
__global__ __launch_bounds__(kBsrThreads) void k_bsr_mv(const V* __restrict__ data, const I* __restrict__ indices,
                                                        const I* __restrict__ indptr, const V* __restrict__ x,
                                                        int64_t mb, int64_t nb, int64_t nnzb, V* __restrict__ y) {
  constexpr int BS2 = BS * BS;
  constexpr int WIN = 256;
  constexpr int NB = BS2 < WIN ? WIN / BS2 : 1;   // blocks per window
  constexpr int NT = BS2 > WIN ? BS2 / WIN : 1;   // windows per block
  constexpr int RL = BS / 4;                      // lanes per block row-row
  constexpr int U = NT >= 16 ? 1 : (NT >= 4 ? 2 : kBsrU);   // blocks in flight (register budget)
  const int lane = threadIdx.x & 63;
  const int e0 = 4 * lane;
  const int blk = NB > 1 ? e0 / BS2 : 0;          // block of this lane inside a window
  const int w = NB > 1 ? e0 % BS2 : e0;           // element offset inside the block
  const int c = w % BS;                           // first of the lane's 4 columns (same for every window)
  const int64_t wave = ((int64_t)blockIdx.x * kBsrThreads + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBsrThreads) >> 6;
  for (int64_t i = wave; i < mb; i += nwaves) {
    int64_t s, t;
    row_range(indptr, i, nnzb, s, t);
    if constexpr (!TRANS) {
      V acc[NT];
#pragma unroll
      for (int q = 0; q < NT; ++q) acc[q] = V(0);
      for (int64_t jj = s; jj < t; jj += NB * U) {
        V d[U][NT][4], xv[U][4];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int64_t j = jj + u * NB + blk;
          const int64_t col = j < t ? (int64_t)indices[j] : -1;
          ok[u] = (uint64_t)col < (uint64_t)nb;
          const int64_t jc = ok[u] ? j : s;
#pragma unroll
          for (int q = 0; q < NT; ++q) load4(data + jc * BS2 + w + WIN * q, d[u][q]);
          const int64_t xb = (ok[u] ? col : 0) * BS + c;
#pragma unroll
          for (int r = 0; r < 4; ++r) xv[u][r] = x[xb + r];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int q = 0; q < NT; ++q) {
            V p = d[u][q][0] * xv[u][0];
            p += d[u][q][1] * xv[u][1];
            p += d[u][q][2] * xv[u][2];
            p += d[u][q][3] * xv[u][3];
            acc[q] += ok[u] ? p : V(0);
          }
      }
#pragma unroll
      for (int q = 0; q < NT; ++q) {
        V v = acc[q];
#pragma unroll
        for (int o = 1; o < RL; o <<= 1) v += __shfl_xor(v, o, 64);
        if constexpr (NB > 1) {
#pragma unroll
          for (int o = BS2 / 4; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
        }
        acc[q] = v;
      }
      if ((lane & (RL - 1)) == 0 && blk == 0) {
#pragma unroll
        for (int q = 0; q < NT; ++q) y[i * BS + (w + WIN * q) / BS] = acc[q];
      }
    } else {
      // rows of the lane's elements, one x value per window
      V xr[NT];
#pragma unroll
      for (int q = 0; q < NT; ++q) xr[q] = x[i * BS + (w + WIN * q) / BS];
      for (int64_t jj = s; jj < t; jj += NB * U) {
        V d[U][NT][4];
        int64_t colv[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int64_t j = jj + u * NB + blk;
          const int64_t col = j < t ? (int64_t)indices[j] : -1;
          ok[u] = (uint64_t)col < (uint64_t)nb;
          colv[u] = col;
          const int64_t jc = ok[u] ? j : s;
#pragma unroll
          for (int q = 0; q < NT; ++q) load4(data + jc * BS2 + w + WIN * q, d[u][q]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          V p[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            V v = d[u][0][r] * xr[0];
#pragma unroll
            for (int q = 1; q < NT; ++q) v += d[u][q][r] * xr[q];
            // sum over the block rows held by other lanes of this window
#pragma unroll
            for (int o = RL; o < (BS2 < WIN ? BS2 : WIN) / 4; o <<= 1) v += __shfl_xor(v, o, 64);
            p[r] = v;
          }
          // the lanes holding row 0 of the window now hold the block's column
          // sums, 4 each; spread them so that lane q of the block holds
          // column q, and add the block's bs columns with one atomic
          // instruction (one 64-byte segment per 16 columns)
          constexpr int PER = NB > 1 ? BS2 / 4 : 64;   // lanes per block in a window
          const int tq = lane % PER;
          const int src = (lane - tq) + tq / 4;        // row-0 lane holding columns 4 (tq / 4) ..
          V val = V(0);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const V tv = __shfl(p[r], src, 64);
            val = (tq & 3) == r ? tv : val;
          }
          if (ok[u] && tq < BS) atomic_add(y + colv[u] * BS + tq, val);
        }
      }
    }
  }
}

// any block size: a thread per output row (or per block row and block
// column when transposed)
template <typename V, typename I, bool TRANS>
__global__ __launch_bounds__(kBsrThreads) void k_bsr_mv_any(const V* __restrict__ data, const I* __restrict__ indices,
                                                            const I* __restrict__ indptr, const V* __restrict__ x,
                                                            int64_t mb, int64_t nb, int64_t bs, int64_t nnzb,
                                                            V* __restrict__ y) {
  const int64_t total = mb * bs;
  for (int64_t g = (int64_t)blockIdx.x * kBsrThreads + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * kBsrThreads) {
    const int64_t i = g / bs, r = g - i * bs;
    int64_t s, t;
    row_range(indptr, i, nnzb, s, t);
    if constexpr (!TRANS) {
      V acc = V(0);
      for (int64_t j = s; j < t; ++j) {
        const int64_t col = (int64_t)indices[j];
        if ((uint64_t)col >= (uint64_t)nb) continue;
        const V* row = data + (j * bs + r) * bs;
        const V* xb = x + col * bs;
        V p = V(0);
        for (int64_t q = 0; q < bs; ++q) p += row[q] * xb[q];
        acc += p;
      }
      y[g] = acc;
    } else {
      // r is the block column here
      for (int64_t j = s; j < t; ++j) {
        const int64_t col = (int64_t)indices[j];
        if ((uint64_t)col >= (uint64_t)nb) continue;
        V p = V(0);
        for (int64_t q = 0; q < bs; ++q) p += data[(j * bs + q) * bs + r] * x[i * bs + q];
        atomic_add(y + col * bs + r, p);
      }
    }
  }
}

// gradient of the block values: d[j][p][q] = ct[i*bs + p] * x[col*bs + q]
// (transposed: x[i*bs + p] * ct[col*bs + q]); i = block row of block j
template <typename V, typename I>
__global__ __launch_bounds__(kBsrThreads) void k_bsr_sddmm(const I* __restrict__ indices, const I* __restrict__ indptr,
                                                           const V* __restrict__ rowv, const V* __restrict__ colv,
                                                           int64_t mb, int64_t nb, int64_t bs, int64_t nnzb,
                                                           V* __restrict__ out) {
  const int64_t bs2 = bs * bs;
  const int64_t total = nnzb * bs2;
  for (int64_t g = (int64_t)blockIdx.x * kBsrThreads + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * kBsrThreads) {
    const int64_t j = g / bs2, e = g - j * bs2;
    const int64_t p = e / bs, q = e - p * bs;
    // block row: the last i with indptr[i] <= j (upper bound - 1)
    int64_t lo = 0, hi = mb;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((int64_t)indptr[mid + 1] <= j) lo = mid + 1;
      else hi = mid;
    }
    const int64_t col = (int64_t)indices[j];
    V v = V(0);
    if (lo < mb && (uint64_t)col < (uint64_t)nb) v = rowv[lo * bs + p] * colv[col * bs + q];
    out[g] = v;
  }
}

// ---------------------------------------------------------------------------
// bcsrmm, f32: MFMA 32x32x2.  Operand layouts of v_mfma_f32_32x32x2f32:
// A[i][k] in lane i + 32k, B[k][j] in lane j + 32k, D[i][j] in lane
// j + 32 * ((i / 4) % 2), accumulator (i % 4) + 4 * (i / 8).
// ---------------------------------------------------------------------------
typedef float v16f __attribute__((ext_vector_type(16)));
constexpr int kMmWaves = 4;   // waves per block, each on its own m-tile(s)
constexpr int kMmMT = 2;      // 32-row m-tiles per wave (they share the B operand)

template <typename I>
__global__ __launch_bounds__(kMmWaves * 64) void k_bsr_mm_f32(const float* __restrict__ A,
                                                               const float* __restrict__ Bd,
                                                               const I* __restrict__ Bidx,
                                                               const I* __restrict__ Bptr, int64_t m, int64_t k,
                                                               int64_t n, int64_t nblk, float* __restrict__ C) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t ni = blockIdx.x;                  // column block
  const int64_t kb = k / 32;
  const int64_t m0 = ((int64_t)blockIdx.y * kMmWaves + wave) * (32 * kMmMT);
  if (m0 >= m) return;
  const int li = lane & 31, g = lane >> 5;
  int64_t s, t;
  row_range(Bptr, ni, nblk, s, t);
  v16f acc[kMmMT];
#pragma unroll
  for (int mt = 0; mt < kMmMT; ++mt)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[mt][v] = 0.f;
  // rows of A this lane feeds (clamped rows past m read row m-1; their
  // outputs are not stored)
  int64_t arow[kMmMT];
#pragma unroll
  for (int mt = 0; mt < kMmMT; ++mt) {
    const int64_t r = m0 + 32 * mt + li;
    arow[mt] = r < m ? r : m - 1;
  }
  for (int64_t e = s; e < t; ++e) {
    const int64_t ki = (int64_t)Bidx[2 * e];
    const int64_t bi = (int64_t)Bidx[2 * e + 1];
    if ((uint64_t)ki >= (uint64_t)kb || (uint64_t)bi >= (uint64_t)nblk) continue;
    // the block's k index 16 g + kk feeds MFMA step kk in lane half g, so a
    // lane reads 16 contiguous values of its A row and of its B_data row
    // (B operand: B[k][j = li] = B_data[bi][li][k])
    float b[16];
    const float* brow = Bd + (bi * 32 + li) * 32 + 16 * g;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 v = *reinterpret_cast<const float4*>(brow + 4 * q);
      b[4 * q] = v.x; b[4 * q + 1] = v.y; b[4 * q + 2] = v.z; b[4 * q + 3] = v.w;
    }
#pragma unroll
    for (int mt = 0; mt < kMmMT; ++mt) {
      float a[16];
      const float* ar = A + arow[mt] * k + ki * 32 + 16 * g;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 v = *reinterpret_cast<const float4*>(ar + 4 * q);
        a[4 * q] = v.x; a[4 * q + 1] = v.y; a[4 * q + 2] = v.z; a[4 * q + 3] = v.w;
      }
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) acc[mt] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[kk], b[kk], acc[mt], 0, 0, 0);
    }
  }
#pragma unroll
  for (int mt = 0; mt < kMmMT; ++mt)
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int64_t r = m0 + 32 * mt + (v % 4) + 4 * g + 8 * (v / 4);
      if (r < m) C[r * n + ni * 32 + li] = acc[mt][v];
    }
}

// f64 (and any layout the MFMA kernel does not take): a thread per output
// element of the column block
template <typename V, typename I>
__global__ __launch_bounds__(kBsrThreads) void k_bsr_mm_any(const V* __restrict__ A, const V* __restrict__ Bd,
                                                            const I* __restrict__ Bidx, const I* __restrict__ Bptr,
                                                            int64_t m, int64_t k, int64_t n, int64_t nblk,
                                                            V* __restrict__ C) {
  const int64_t total = m * n;
  const int64_t kb = k / 32;
  for (int64_t gi = (int64_t)blockIdx.x * kBsrThreads + threadIdx.x; gi < total;
       gi += (int64_t)gridDim.x * kBsrThreads) {
    const int64_t r = gi / n, col = gi - r * n;
    const int64_t ni = col / 32, a = col - ni * 32;
    int64_t s, t;
    row_range(Bptr, ni, nblk, s, t);
    V acc = V(0);
    for (int64_t e = s; e < t; ++e) {
      const int64_t ki = (int64_t)Bidx[2 * e];
      const int64_t bi = (int64_t)Bidx[2 * e + 1];
      if ((uint64_t)ki >= (uint64_t)kb || (uint64_t)bi >= (uint64_t)nblk) continue;
      const V* ar = A + r * k + ki * 32;
      const V* br = Bd + (bi * 32 + a) * 32;
      for (int q = 0; q < 32; ++q) acc += ar[q] * br[q];
    }
    C[gi] = acc;
  }
}

template <typename O>
__global__ __launch_bounds__(256) void k_bsr_zero(O* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) out[i] = O(0);
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

unsigned grid_of(int64_t work, int per_block, int64_t cap = 65536) {
  int64_t b = (work + per_block - 1) / per_block;
  if (b > cap) b = cap;
  return (unsigned)(b < 1 ? 1 : b);
}

template <typename V, typename I, bool TRANS>
int bsrmv_launch(hipStream_t s, const V* data, const I* indices, const I* indptr, const V* x, int64_t mb, int64_t nb,
                 int64_t bs, int64_t nnzb, V* y) {
  const int64_t ylen = (TRANS ? nb : mb) * bs;
  if (TRANS && ylen > 0) {
    hipLaunchKernelGGL(k_bsr_zero<V>, dim3(grid_of(ylen, 256, 8192)), dim3(256), 0, s, y, ylen);
  }
  if (mb == 0 || ylen == 0) return BPX_OK;
  // no block can contribute (and with nb == 0 the vector is empty): zeros
  if (!TRANS && (nnzb == 0 || nb == 0)) {
    hipLaunchKernelGGL(k_bsr_zero<V>, dim3(grid_of(ylen, 256, 8192)), dim3(256), 0, s, y, ylen);
    BPX_CHECK_LAUNCH();
    return BPX_OK;
  }
  const bool vec_ok = aligned16(data);
  const unsigned gw = grid_of(mb, kBsrThreads / 64);
#define BSR_VEC(B)                                                                                              \
  case B:                                                                                                      \
    BPX_LAUNCH(KID_BSR, (k_bsr_mv<B, V, I, TRANS>), dim3(gw), dim3(kBsrThreads), 0, s, data, indices, indptr, x, \
               mb, nb, nnzb, y);                                                                               \
    BPX_CHECK_LAUNCH();                                                                                        \
    return BPX_OK;
  if (vec_ok) {
    switch (bs) {
      BSR_VEC(4)
      BSR_VEC(8)
      BSR_VEC(16)
      BSR_VEC(32)
      BSR_VEC(64)
      default: break;
    }
  }
#undef BSR_VEC
  BPX_LAUNCH(KID_BSR, (k_bsr_mv_any<V, I, TRANS>), dim3(grid_of(mb * bs, kBsrThreads)), dim3(kBsrThreads), 0, s,
             data, indices, indptr, x, mb, nb, bs, nnzb, y);
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

}  // namespace
}  // namespace bpx

using namespace bpx;

#define BPX_BSR_IDX_SWITCH(IDT, ...)                     \
  switch (IDT) {                                         \
    case BPX_I32: { using I = int32_t; __VA_ARGS__; }    \
    case BPX_I64: { using I = int64_t; __VA_ARGS__; }    \
    default: return BPX_ERR_DTYPE;                       \
  }

#define BPX_BSR_VAL_SWITCH(VDT, ...)                     \
  switch (VDT) {                                         \
    case BPX_F32: { using V = float; __VA_ARGS__; }      \
    case BPX_F64: { using V = double; __VA_ARGS__; }     \
    default: return BPX_ERR_DTYPE;                       \
  }

extern "C" {

int bpx_bsrmv(void* stream, const void* data, int dtype, const void* indices, const void* indptr, int index_dtype,
              int64_t nnzb, int64_t blocksize, const void* vector, int64_t mb, int64_t nb, int transpose, void* out) {
  if (mb < 0 || nb < 0 || nnzb < 0 || blocksize < 1) return BPX_ERR_SHAPE;
  if ((mb > 0 && !indptr) || (nnzb > 0 && (!data || !indices)) || (!out && (transpose ? nb : mb) > 0))
    return BPX_ERR_ARG;
  if (!vector && mb > 0 && nb > 0 && nnzb > 0) return BPX_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  BPX_BSR_VAL_SWITCH(dtype, BPX_BSR_IDX_SWITCH(index_dtype, {
    if (transpose)
      return bsrmv_launch<V, I, true>(s, (const V*)data, (const I*)indices, (const I*)indptr, (const V*)vector, mb, nb,
                                      blocksize, nnzb, (V*)out);
    return bsrmv_launch<V, I, false>(s, (const V*)data, (const I*)indices, (const I*)indptr, (const V*)vector, mb, nb,
                                     blocksize, nnzb, (V*)out);
  }));
}

int bpx_bsr_sddmm(void* stream, const void* indices, const void* indptr, int index_dtype, int64_t nnzb,
                  int64_t blocksize, const void* rowv, const void* colv, int dtype, int64_t mb, int64_t nb,
                  void* out) {
  if (mb < 0 || nb < 0 || nnzb < 0 || blocksize < 1) return BPX_ERR_SHAPE;
  if (nnzb == 0) return BPX_OK;
  if (!indices || !indptr || !rowv || !colv || !out) return BPX_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  BPX_BSR_VAL_SWITCH(dtype, BPX_BSR_IDX_SWITCH(index_dtype, {
    hipLaunchKernelGGL((k_bsr_sddmm<V, I>), dim3(grid_of(nnzb * blocksize * blocksize, kBsrThreads)),
                       dim3(kBsrThreads), 0, s, (const I*)indices, (const I*)indptr, (const V*)rowv, (const V*)colv,
                       mb, nb, blocksize, nnzb, (V*)out);
    BPX_CHECK_LAUNCH();
    return BPX_OK;
  }));
}

int bpx_bcsrmm(void* stream, const void* a, int dtype, const void* b_data, const void* b_indices, const void* b_ptr,
               int index_dtype, int64_t nblk, int64_t m, int64_t k, int64_t n, void* out) {
  if (m < 0 || k < 0 || n < 0 || nblk < 0 || (k % 32) || (n % 32)) return BPX_ERR_SHAPE;
  if (m == 0 || n == 0) return BPX_OK;
  if (!out || !b_ptr || (nblk > 0 && (!b_data || !b_indices)) || (k > 0 && !a)) return BPX_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  BPX_BSR_IDX_SWITCH(index_dtype, {
    if (dtype == BPX_F32 && aligned16(a) && aligned16(b_data) && k > 0) {
      const int64_t rows_per_block = 32 * kMmMT * kMmWaves;
      const int64_t gy = (m + rows_per_block - 1) / rows_per_block;
      if (gy > 65535 || n / 32 > 0x7FFFFFFF) return BPX_ERR_SHAPE;
      BPX_LAUNCH(KID_BSR, (k_bsr_mm_f32<I>), dim3((unsigned)(n / 32), (unsigned)gy), dim3(kMmWaves * 64), 0, s,
                 (const float*)a, (const float*)b_data, (const I*)b_indices, (const I*)b_ptr, m, k, n, nblk,
                 (float*)out);
      BPX_CHECK_LAUNCH();
      return BPX_OK;
    }
    BPX_BSR_VAL_SWITCH(dtype, {
      BPX_LAUNCH(KID_BSR, (k_bsr_mm_any<V, I>), dim3(grid_of(m * n, kBsrThreads)), dim3(kBsrThreads), 0, s,
                 (const V*)a, (const V*)b_data, (const I*)b_indices, (const I*)b_ptr, m, k, n, nblk, (V*)out);
      BPX_CHECK_LAUNCH();
      return BPX_OK;
    });
  });
}

}  // extern "C"
