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
//   outdim_parallel, lane32, bool events, homo/uniform weights: one thread
//     per row walking the row's 32 streams in one flattened loop, event
//     bitmask in LDS (k_jit_rows_bits).
//   outdim_parallel, lane32, otherwise: one thread per (row, lane): a wave
//     covers two rows; the 32 lane partials are combined with a 32-wide
//     shuffle sum and written once (no atomics, no zero-fill).
//   scatter (outdim_parallel=False), lane32: 256-thread blocks own 256
//     columns; events are balloted into an LDS list of active columns, then 8
//     groups of 32 lanes regenerate the rows of one active column each and
//     add into the zero-filled output with FP32 atomics.  Inactive columns
//     cost one event byte.
//   row layout: one thread per row (outdim_parallel) or per active column.
#include "bpx_common.h"
#include "bpx_rng.h"

#include <math.h>
#include <stdlib.h>


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
    // hardware transcendentals: v_log_f32 is log2 (ln u1 = log2 u1 * ln 2),
    // v_sin_f32 takes revolutions (sin(2 pi u2) from u2 itself, no rounding
    // of the 2 pi product), v_sqrt_f32; each within a few ulp, the weights
    // are pinned to a tolerance
    const float ln_u1 = __builtin_amdgcn_logf(u1) * 0.6931471805599453f;
    const float mag = __builtin_amdgcn_sqrtf(-2.0f * ln_u1);
    const float z = mag * __builtin_amdgcn_sinf(u2);
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
// lane32, outdim_parallel, bool events, event bitmask in LDS.
//
// Events are packed once into a bitmask (k_pack_bits); every block of this
// persistent kernel copies the bitmask into LDS (1M columns = 125 KB of the
// 160 KB gfx950 gives a block) so the random event test of each generated
// column is an LDS read instead of a scattered global byte load.  Each
// thread runs the streams of KS rows for one lane, one after another, which
// evens out the per-lane draw counts a wave waits on; the 32 lanes of each
// row are then summed with a 32-wide shuffle reduction.
// ---------------------------------------------------------------------------
constexpr int kJitThreads = 1024;
constexpr int kJitKS = 4;                       // rows per thread (per lane)
constexpr int kRowsPerPass = kJitThreads / 32 * kJitKS;

__global__ __launch_bounds__(256) void k_pack_bits(const uint8_t* __restrict__ ev, int64_t n,
                                                   uint32_t* __restrict__ bits) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool on = i < n && ev[i] != 0;
  const uint64_t mask = __ballot(on);
  const int64_t w = i >> 5;
  if ((threadIdx.x & 63) == 0 && (w << 5) < n) bits[w] = (uint32_t)mask;
  if ((threadIdx.x & 63) == 32 && (w << 5) < n) bits[w] = (uint32_t)(mask >> 32);
}

__device__ __forceinline__ float repeat_sum_f(uint32_t c, float w) {
  float r = 0.f;
  for (uint32_t k = 0; k < c; ++k) r += w;
  return r;
}

// copy of the packed event bitmask into LDS by a whole block: each thread's
// loads are issued in batches before their LDS stores (a plain copy loop
// waits for every load), and each block starts at its own offset so the
// blocks' first reads of a line do not all miss L2 together
__device__ __forceinline__ void copy_bits_to_lds(const uint32_t* __restrict__ gbits, int64_t nwords,
                                                 uint32_t* s_bits) {
  constexpr int kB = 16;
  const int64_t rot = ((int64_t)blockIdx.x * 997) % (nwords > 0 ? nwords : 1);
  for (int64_t base = threadIdx.x; base < nwords; base += (int64_t)kJitThreads * kB) {
    uint32_t v[kB];
    int64_t kk[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int64_t k = base + (int64_t)u * kJitThreads;
      kk[u] = k < nwords ? (k + rot < nwords ? k + rot : k + rot - nwords) : -1;
      v[u] = gbits[kk[u] >= 0 ? kk[u] : 0];
    }
#pragma unroll
    for (int u = 0; u < kB; ++u)
      if (kk[u] >= 0) s_bits[kk[u]] = v[u];
  }
}

template <int DIST, bool Q24>
__global__ __launch_bounds__(kJitThreads) void k_jit_lane32_outdim_bits(const uint32_t* __restrict__ gbits,
                                                                        int64_t nwords, float w0, float w1, DivU32 dv,
                                                                        uint32_t seed, int64_t num_row,
                                                                        int64_t num_col, uint32_t step,
                                                                        float* __restrict__ out) {
  extern __shared__ uint32_t s_bits[];
  copy_bits_to_lds(gbits, nwords, s_bits);
  __syncthreads();
  const uint32_t l = threadIdx.x & 31;
  const int half = threadIdx.x >> 5;            // 32 halves of 32 lanes
  uint32_t c0 = step * l - 1u;
  uint32_t end = c0 + step;
  if ((int64_t)end > num_col) end = (uint32_t)num_col;
  for (int64_t base = (int64_t)blockIdx.x * kRowsPerPass; base < num_row; base += (int64_t)gridDim.x * kRowsPerPass) {
    float acc[kJitKS];
#pragma unroll
    for (int k = 0; k < kJitKS; ++k) {
      const int64_t row = base + (int64_t)half * kJitKS + k;
      acc[k] = 0.f;
      uint32_t hits = 0;
      if (row < num_row) {
        Lfsr88 key(seed + (uint32_t)(row * 32) + l);
        uint32_t c = c0 + rand_inc_u<Q24>(key, dv);
        while (c < end) {
          const uint32_t bit = (s_bits[c >> 5] >> (c & 31)) & 1u;
          if constexpr (DIST == BPX_DIST_HOMO) {
            hits += bit;
          } else {
            // r += w * e with e in {0, 1}: w * 1 = w and w * 0 adds +0
            const float w = draw_weight<DIST>(key, w0, w1);
            acc[k] += bit ? w : 0.f;
          }
          c += rand_inc_u<Q24>(key, dv);
        }
      }
      // the reference lane adds weight0 * event per connection: a run of
      // `hits` additions of weight0 (zero terms leave the sum unchanged)
      if constexpr (DIST == BPX_DIST_HOMO) acc[k] = repeat_sum_f(hits, w0);
    }
#pragma unroll
    for (int k = 0; k < kJitKS; ++k) {
      const float v = group_sum<32>(acc[k]);
      const int64_t row = base + (int64_t)half * kJitKS + k;
      if (l == 0 && row < num_row) out[row] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// lane32, outdim_parallel, bool events, one thread per row.
//
// A thread walks the 32 streams of its row one after another in a single
// flattened loop: each iteration either tests the current column (and, for
// uniform/normal weights, draws the connection's weight) or, when the stream
// has passed its end, folds the stream's partial into the row sum and starts
// the next stream; then every lane takes the one shared LFSR88 increment
// draw.  A row is ~32 x (1 + connections per stream) draws, so the lanes of a
// wave stay busy together (a wave waits for its longest row, not for its
// longest 3-draw stream), and the row sum is formed in ascending stream
// order, `out[row] = (((0 + p_0) + p_1) + ... + p_31)` with
// `p_l = sum of w * e over stream l's connections in order` (the Taichi
// kernel's `r += weight0 * events[i_col]` per stream, then `out[i_row] += r`;
// jitconn/_event_matvec_taichi.py:140-155).  Rows are dealt to the grid's
// threads in turn; the event bitmask sits in LDS as above.
// ---------------------------------------------------------------------------
// Modulo variants of the increment draw: 0 exact multiply-high (DivU32),
// 1 the same with a 24-bit product, 2 the short form (DivShort) with one
// signed 24-bit multiply-add for the remainder.
enum { MOD_GM = 0, MOD_GM24 = 1, MOD_SHORT24 = 2 };

template <int MK>
__device__ __forceinline__ uint32_t mod_clen(uint32_t x, const DivU32& dv, const DivShort& ds) {
  if constexpr (MK == MOD_SHORT24) return ds.mod24(x);
  else if constexpr (MK == MOD_GM24) return dv.mod24(x);
  else return dv.mod(x);
}

// FIN: the weights are finite, so `w * e` with e in {0, 1} equals
// `e ? w : +0`, adding +0 leaves a partial unchanged (it starts at +0 and
// never becomes -0), and a homogeneous stream's partial is the sum of its
// `hits` active terms, ((0 + w) + w) + ...; non-finite weights take the
// multiply.
// c + 1 + x % clen, the next connection's column: for MOD_SHORT24 the
// remainder x - q * clen is one 24-bit multiply-add (q < 2^23 and
// clen < 2^23 there), which the compiler otherwise splits into a multiply
// and an add
template <int MK>
__device__ __forceinline__ uint32_t next_col(uint32_t c, uint32_t x, const DivU32& dv, const DivShort& ds,
                                             int32_t negd) {
  if constexpr (MK == MOD_SHORT24) {
    const uint32_t q = ds.div(x);
    int32_t r;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(q), "s"(negd), "v"(x));
    return c + (uint32_t)r + 1u;
  } else {
    return c + mod_clen<MK>(x, dv, ds) + 1u;
  }
}

// CLIP: some stream of a row ends past num_col (32 * step > num_col), so a
// stream's end is min(next start, num_col); otherwise the next start itself
template <int DIST, int MK, bool FIN, bool CLIP>
__global__ __launch_bounds__(kJitThreads) void k_jit_rows_bits(const uint32_t* __restrict__ gbits, int64_t nwords,
                                                                float w0, float w1, DivU32 dv, DivShort ds,
                                                                uint32_t seed, int64_t num_row, int64_t num_col,
                                                                uint32_t step, float* __restrict__ out) {
  extern __shared__ uint32_t s_bits[];
  // word 0 is zero (the event word of lanes between streams), the bitmask
  // follows from word 1
  copy_bits_to_lds(gbits, nwords, s_bits + 1);
  if (threadIdx.x == 0) s_bits[0] = 0u;
  __syncthreads();
  // homogeneous weights add w0 per active connection like the others
  // (e ? w0 : +0 with finite weights): the partial's sequence of additions
  // is the reference lane's r += weight0 * e, and a stream switch is one add
  //
  // A thread takes a contiguous block of rows (num_row < 2^31, checked by
  // the launcher), so the stream key seed + row * 32 + l simply counts up
  // across its rows.  Columns are kept 32 above the reference's index
  // (c = step * l - 1 + 32 at a stream's start): c >> 5 is then the index of
  // c's event word in s_bits, and the bit is still c % 32.
  constexpr uint32_t OFF = 32u;
  // Draws per loop iteration: the stream switch, which some lane of a wave
  // takes in nearly every iteration, is tested after the last of them only.
  // With ~3 connections per stream, two slots cost a lane ~0.5 idle slot
  // per stream and save the switch block every other draw (homogeneous
  // 138 -> 128 us, uniform 193 -> 187 us on 1M x 1M at p = 1e-4); normal
  // weights keep one (the idle slot's weight draw costs more than the
  // switch).  Same draws, same sums, so the same bits.
  constexpr int kSlots = DIST == BPX_DIST_NORMAL ? 1 : 2;
  const uint32_t nthreads = gridDim.x * kJitThreads;
  const uint32_t tid = blockIdx.x * kJitThreads + threadIdx.x;
  const uint32_t rq = (uint32_t)num_row / nthreads, rr = (uint32_t)num_row % nthreads;
  uint32_t row = tid * rq + min(tid, rr);
  const uint32_t row_end = row + rq + (tid < rr ? 1u : 0u);
  bool live = row < row_end;
  const uint32_t ncol = (uint32_t)num_col + OFF;  // num_col < 2^32 - 64 (checked by the launcher)
  const uint32_t c_first = OFF - 1u;              // step * 0 - 1
  const uint32_t c0_last = step * 32u - 1u + OFF; // step * l - 1 at l = 32: the row is done
  const int32_t negd = -(int32_t)ds.d;
  uint32_t kseed = seed + row * 32u;              // stream key seed + row*32 + l
  Lfsr88 key(kseed);
  uint32_t c = c_first;
  uint32_t cn = c + step;                         // step * (l + 1) - 1: the next stream's start
  uint32_t end = CLIP ? min(cn, ncol) : cn;
  float acc = 0.f, racc = 0.f;
  c = next_col<MK>(c, key.next(), dv, ds, negd);
  while (live) {
#pragma unroll
    for (int sl = 1; sl < kSlots; ++sl) {
      // a draw with no switch after it: a lane whose stream ends on this
      // test fails the next test too (c only grows within a stream) and
      // switches there, the slot between drawing from the key it drops
      const bool in = c < end;
      const uint32_t word = s_bits[in ? c >> 5 : 0u];
      const uint32_t cb = c;
      float w = w0;
      if constexpr (DIST != BPX_DIST_HOMO) {
        if (FIN || in) w = draw_weight<DIST>(key, w0, w1);
      }
      c = next_col<MK>(c, key.next(), dv, ds, negd);
      const uint32_t bit = __builtin_amdgcn_ubfe(word, cb, 1);
      if constexpr (FIN) {
        acc = __builtin_fmaf(w, (float)bit, acc);
      } else if (in) {
        acc += w * (float)bit;
      }
    }
    const bool in = c < end;
    // in flight during the draws; lanes past their stream's end read the zero word
    const uint32_t word = s_bits[in ? c >> 5 : 0u];
    const uint32_t cb = c;                               // its bit: cb % 32 (c moves on below)
    float w = w0;
    if constexpr (DIST != BPX_DIST_HOMO) {
      // drawn by every lane: a lane past its stream's end advances a key it
      // discards at the switch below, and its zero event bit drops the weight
      if (FIN || in) w = draw_weight<DIST>(key, w0, w1);
    }
    if (!in) {
      racc += acc;
      acc = 0.f;
      ++kseed;
      // unclipped, a stream's end is the next stream's start (one register)
      if constexpr (CLIP) {
        c = cn;
        cn += step;
      } else {
        c = end;
        end += step;
      }
      if (c == c0_last) {
        out[row] = racc;
        racc = 0.f;
        ++row;
        live = row < row_end;
        c = c_first;
        if constexpr (CLIP) cn = c_first + step;
        else end = c_first + step;
      }
      key = Lfsr88(kseed);
      if constexpr (CLIP) end = min(cn, ncol);
    }
    c = next_col<MK>(c, key.next(), dv, ds, negd);
    // bit-field extract of one bit (the offset's low 5 bits select it)
    const uint32_t bit = __builtin_amdgcn_ubfe(word, cb, 1);
    if constexpr (FIN) {
      // acc + w * e with e in {0, 1}: w * e is exact, so one fused
      // multiply-add rounds exactly as the reference's product-then-add
      // does; e = 0 (also every lane past its stream) adds a zero
      acc = __builtin_fmaf(w, (float)bit, acc);
    } else if (in) {
      acc += w * (float)bit;
    }
  }
}

// ---------------------------------------------------------------------------
// lane32 scatter (outdim_parallel=False), bool events: exclusive row tiles.
//
// Stream (col, l) of the reference's event kernel covers the rows
// [step*l, step*l + step - 1) of the output (jitconn/_event_matvec_taichi.py
// event kernels: i_row = step*index - 1, end = min(i_row + step, num_row)).
// Block (l, k) owns the rows [lo, hi) = step*l + [k*T, (k+1)*T) (within the
// stream's range and num_row) and, for every active column, regenerates
// stream l from its start up to hi, adding the connections that fall in
// [lo, hi) in LDS: each output row belongs to one block, so the results are
// stored once (no global atomics, no zero-fill).  Homogeneous weights count
// the connections (u32) and store S(count) = ((0 + w) + w) + ..., what the
// reference's `out[i_row] += weight` gives in any order; uniform and normal
// weights are drawn per connection in the reference's order (a skipped
// connection still consumes its draws) and summed with LDS float atomics.
// The price is the regeneration of the rows before lo: the last tile of a
// stream walks the whole stream.
//
// The active columns come as a list (k_list_events, any order); each lane
// walks the list entries tid, tid + 1024, ... in one flattened loop (one
// draw per iteration: a connection of the current stream, or the first
// draw of the next entry's stream), with the next entry's column loaded
// while the current stream runs.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_list_events(const uint8_t* __restrict__ ev, int64_t n,
                                                      uint32_t* __restrict__ count, uint32_t* __restrict__ list) {
  // 4 columns per thread (one 32-bit load when aligned), one counter atomic per block
  __shared__ uint32_t s_wave[16];
  __shared__ uint32_t s_base;
  const int64_t c0 = ((int64_t)blockIdx.x * 1024 + threadIdx.x) * 4;
  uint32_t on = 0;
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (c0 + u < n && ev[c0 + u] != 0) on |= 1u << u;
  const uint32_t cnt = (uint32_t)__popc(on);
  // inclusive wave scan of the counts
  uint32_t incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o, 64);
    if (lane_id() >= (uint32_t)o) incl += y;
  }
  const uint32_t wv = threadIdx.x >> 6;
  if (lane_id() == 63) s_wave[wv] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int q = 0; q < 16; ++q) {
      const uint32_t t = s_wave[q];
      s_wave[q] = tot;
      tot += t;
    }
    s_base = tot > 0 ? atomicAdd(count, tot) : 0u;
  }
  __syncthreads();
  uint32_t pos = s_base + s_wave[wv] + incl - cnt;
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (on & (1u << u)) list[pos++] = (uint32_t)(c0 + u);
}

__global__ void k_zero_count(uint32_t* count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *count = 0u;
}

template <int DIST, int MK>
__global__ __launch_bounds__(kJitThreads) void k_jit_scatter_tiles(const uint32_t* __restrict__ count,
                                                                   const uint32_t* __restrict__ list, float w0,
                                                                   float w1, DivU32 dv, DivShort ds, uint32_t seed,
                                                                   int64_t num_row, uint32_t step, int K, int T,
                                                                   float* __restrict__ out) {
  extern __shared__ uint32_t s_acc[];            // homo: counts; else f32 sums
  const uint32_t l = blockIdx.x / (uint32_t)K, k = blockIdx.x % (uint32_t)K;
  const uint64_t s0 = (uint64_t)step * l;        // first row of stream l's range
  const uint64_t lo64 = s0 + (uint64_t)k * (uint64_t)T;
  uint64_t hi64 = lo64 + (uint64_t)T;
  if (hi64 > s0 + step) hi64 = s0 + step;
  if (hi64 > (uint64_t)num_row) hi64 = (uint64_t)num_row;
  if (lo64 < hi64) {
    const uint32_t lo = (uint32_t)lo64, hi = (uint32_t)hi64;
    // connections reach rows below min(step*l - 1 + step, num_row)
    const uint32_t send = (uint32_t)min((int64_t)(s0 + step - 1), num_row);
    const uint32_t hie = min(hi, send);
    const uint32_t r0 = (uint32_t)s0 - 1u;       // step*l - 1 (u32 wrap at l = 0)
    for (int i = threadIdx.x; i < T; i += kJitThreads) s_acc[i] = 0u;
    __syncthreads();
    const uint32_t cnt = *count;
    uint32_t e = threadIdx.x;
    bool live = e < cnt;
    uint32_t col = live ? list[e] : 0u;
    uint32_t ncol = e + kJitThreads < cnt ? list[e + kJitThreads] : 0u;
    Lfsr88 key(seed + col * 32u + l);
    uint32_t r = r0 + mod_clen<MK>(key.next(), dv, ds) + 1u;
    while (live) {
      if (r < hie) {
        float v = w0;
        if constexpr (DIST != BPX_DIST_HOMO) v = draw_weight<DIST>(key, w0, w1);
        if (r >= lo) {
          if constexpr (DIST == BPX_DIST_HOMO)
            __hip_atomic_fetch_add(s_acc + (r - lo), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          else
            __hip_atomic_fetch_add(reinterpret_cast<float*>(s_acc) + (r - lo), v, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      } else {
        // the next entry: its column was loaded one entry ahead
        e += kJitThreads;
        live = e < cnt;
        col = ncol;
        if (e + kJitThreads < cnt) ncol = list[e + kJitThreads];
        key = Lfsr88(seed + col * 32u + l);
        r = r0;
      }
      r += mod_clen<MK>(key.next(), dv, ds) + 1u;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; lo + i < hi; i += kJitThreads) {
      if constexpr (DIST == BPX_DIST_HOMO) out[lo + i] = repeat_sum_f(s_acc[i], w0);
      else out[lo + i] = __uint_as_float(s_acc[i]);
    }
  }
  // rows past every stream's range (32 * step <= row < num_row) get no
  // connection: the last block stores their zeros
  if (blockIdx.x == gridDim.x - 1)
    for (int64_t row = (int64_t)step * 32 + threadIdx.x; row < num_row; row += kJitThreads) out[row] = 0.f;
}

// ---------------------------------------------------------------------------
// lane32 scatter (outdim_parallel=False), bool events, in two passes.
//
// k_jit_scatter_tiles regenerates every active stream up to the end of its
// block's tile, so the last tile of each stream range walks whole streams
// and bounds the call.  Here each active stream is walked once:
//   pass 1, k_jit_emit, block (l, j): stream l of the active columns of
//     column slice j, each to its end; every connection appends its
//     tile-local row (and, for uniform and normal weights, its weight, drawn
//     in the reference's order) to list (l, j, k) of its row tile k;
//   pass 2, k_jit_sum_lists, block (l, k): adds the kEmitSlices lists of
//     tile k into LDS (counts for homogeneous weights, FP32 sums otherwise)
//     and stores the tile's rows once.
// The lists have a fixed capacity and their counts are stored whole: a tile
// with an overflowing list regenerates its streams instead (as
// k_jit_scatter_tiles does), so every input is served.
// ---------------------------------------------------------------------------
constexpr int kEmitSlices = 8;          // column slices per stream range
// w * scale rounded to the nearest integer (ties to even) as int64, for
// |w * scale| < 2^51: one fma adding 1.5 * 2^52 leaves the integer in the
// low mantissa bits
__device__ __forceinline__ long long jit_fixed(float w, double scale) {
  const double y = fma((double)w, scale, 6755399441055744.0);
  return (long long)__double_as_longlong(y) - 0x4338000000000000ll;
}
constexpr int kEmitList = 16384;        // active columns held in LDS per round

// f(c) for every column c of thread t's share of the slice [sb, se) with
// ev[c] != 0: the 16-byte pieces sb + 16 t + 16384 v (v < nv), so a wave's
// loads cover 1 KB of consecutive events; 8 pieces are loaded together.
template <typename F>
__device__ __forceinline__ void for_active(const uint8_t* __restrict__ ev, int64_t sb, int64_t se, int64_t nv,
                                           F&& f) {
  const bool al = (reinterpret_cast<uintptr_t>(ev + sb) & 15u) == 0;
  for (int64_t v0 = 0; v0 < nv; v0 += 8) {
    uint4 q[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t b = sb + (v0 + u) * (16 * kJitThreads) + 16 * (int64_t)threadIdx.x;
      q[u] = make_uint4(0u, 0u, 0u, 0u);
      if (al && v0 + u < nv && b + 16 <= se) q[u] = *reinterpret_cast<const uint4*>(ev + b);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t b = sb + (v0 + u) * (16 * kJitThreads) + 16 * (int64_t)threadIdx.x;
      if (v0 + u >= nv || b >= se) continue;
      if (al && b + 16 <= se) {
        const uint32_t w[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          // the nonzero bytes of the word, lowest first
          uint32_t x = w[h];
          while (x != 0u) {
            const uint32_t y = (uint32_t)__builtin_ctz(x) >> 3;
            f(b + 4 * h + y);
            x &= ~(0xFFu << (8 * y));
          }
        }
      } else {
        for (int64_t c = b; c < min(b + 16, se); ++c)
          if (ev[c] != 0) f(c);
      }
    }
  }
}

template <int DIST, int MK>
__global__ __launch_bounds__(kJitThreads) void k_jit_emit(const uint8_t* __restrict__ ev, int64_t num_col, float w0,
                                                          float w1, DivU32 dv, DivShort ds, uint32_t seed,
                                                          int64_t num_row, uint32_t step, int K, int T, DivU32 dT,
                                                          int64_t span, uint32_t cap, uint32_t* __restrict__ counts,
                                                          uint16_t* __restrict__ lrow, float* __restrict__ lw) {
  __shared__ uint32_t s_list[kEmitList];
  __shared__ uint32_t s_w[kJitThreads / 64];
  __shared__ uint32_t s_tot;
  extern __shared__ uint32_t s_cnt[];     // K list counters
  const uint32_t l = blockIdx.x / (uint32_t)kEmitSlices, j = blockIdx.x % (uint32_t)kEmitSlices;
  const uint64_t s0 = (uint64_t)step * l;
  const uint32_t send = (uint32_t)min((int64_t)(s0 + step - 1), num_row);   // rows below it
  const uint32_t r0 = (uint32_t)s0 - 1u;                                    // step*l - 1 (u32 wrap at l = 0)
  const int32_t negd = -(int32_t)ds.d;
  for (int k = threadIdx.x; k < K; k += kJitThreads) s_cnt[k] = 0u;
  // slice j: columns [sb, se); its active columns get positions by a block
  // prefix over the threads' counts.  A thread keeps its first kHold active
  // columns in s_list[t * kHold + i] while counting (the list is filled
  // from those; a thread with more scans its share again).
  constexpr int kHold = kEmitList / kJitThreads;
  const int64_t sb = min((int64_t)j * span * kJitThreads, num_col);
  const int64_t se = min(sb + span * kJitThreads, num_col);
  const int64_t nv = (span * kJitThreads + 16 * kJitThreads - 1) / (16 * kJitThreads);
  uint32_t mine = 0;
  for_active(ev, sb, se, nv, [&](int64_t c) {
    if (mine < (uint32_t)kHold) s_list[threadIdx.x * kHold + mine] = (uint32_t)c;
    ++mine;
  });
  uint32_t incl = mine;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o, 64);
    if (lane_id() >= (uint32_t)o) incl += y;
  }
  const uint32_t wv = threadIdx.x >> 6;
  if (lane_id() == 63) s_w[wv] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int q = 0; q < kJitThreads / 64; ++q) {
      const uint32_t x = s_w[q];
      s_w[q] = t;
      t += x;
    }
    s_tot = t;
  }
  __syncthreads();
  const uint32_t off = s_w[wv] + incl - mine;
  const uint32_t total = s_tot;
  uint32_t held[kHold];
#pragma unroll
  for (int i = 0; i < kHold; ++i) held[i] = (uint32_t)i < mine ? s_list[threadIdx.x * kHold + i] : 0u;
  __syncthreads();
  // this block's K lists (32-bit offsets within them: K * cap < 2^31)
  uint16_t* const brow = lrow + (size_t)blockIdx.x * (size_t)K * cap;
  float* const bw = DIST == BPX_DIST_HOMO ? nullptr : lw + (size_t)blockIdx.x * (size_t)K * cap;
  for (uint32_t rb = 0; rb < total; rb += kEmitList) {
    // the round's active columns (positions [rb, rb + kEmitList)) into LDS
#pragma unroll
    for (int i = 0; i < kHold; ++i) {
      const uint32_t p = off + i;
      if ((uint32_t)i < mine && p >= rb && p < rb + kEmitList) s_list[p - rb] = held[i];
    }
    if (mine > (uint32_t)kHold && off + mine > rb && off + kHold < rb + kEmitList) {
      uint32_t p = off;
      for_active(ev, sb, se, nv, [&](int64_t c) {
        if (p >= off + kHold && p >= rb && p < rb + kEmitList) s_list[p - rb] = (uint32_t)c;
        ++p;
      });
    }
    __syncthreads();
    const uint32_t n = min(total - rb, (uint32_t)kEmitList);
    // each connection's list slot is reserved by an LDS atomic whose result
    // is used one connection later, so its round trip overlaps the next draw
    uint32_t pat = 0xFFFFFFFFu, ploc = 0u, ppos = 0u;
    float pv = 0.f;
    auto put = [&]() {
      if (pat != 0xFFFFFFFFu && ppos < cap) {
        const uint32_t at = pat + ppos;
        brow[at] = (uint16_t)ploc;
        if constexpr (DIST != BPX_DIST_HOMO) bw[at] = pv;
      }
    };
    for (uint32_t e = threadIdx.x; e < n; e += kJitThreads) {
      Lfsr88 key(seed + s_list[e] * 32u + l);
      uint32_t r = next_col<MK>(r0, key.next(), dv, ds, negd);
      while (r < send) {
        float v = w0;
        if constexpr (DIST != BPX_DIST_HOMO) v = draw_weight<DIST>(key, w0, w1);
        const uint32_t orow = r - (uint32_t)s0;            // row within the stream range
        const uint32_t k = dT.div(orow);
        const uint32_t pos = atomicAdd(s_cnt + k, 1u);
        put();
        pat = k * cap;
        ploc = orow - k * (uint32_t)T;
        pv = v;
        ppos = pos;
        r = next_col<MK>(r, key.next(), dv, ds, negd);
      }
    }
    put();
    __syncthreads();
  }
  // every count stored, so the pass after needs no cleared memory
  for (int k = threadIdx.x; k < K; k += kJitThreads) counts[(size_t)blockIdx.x * K + k] = s_cnt[k];
}

// FX (uniform and normal weights whose draws are all finite and bounded by
// the host's bound): the weights are summed in 64-bit fixed point (scale
// fxs, unit fxi; exact in any order, one rounding at the store) instead of
// FP32 LDS atomics, which run ~20x slower than integer ones
template <int DIST, int MK, bool FX>
__global__ __launch_bounds__(kJitThreads) void k_jit_sum_lists(const uint8_t* __restrict__ ev, int64_t num_col,
                                                               float w0, float w1, DivU32 dv, DivShort ds,
                                                               uint32_t seed, int64_t num_row, uint32_t step, int K,
                                                               int T, uint32_t cap,
                                                               const uint32_t* __restrict__ counts,
                                                               const uint16_t* __restrict__ lrow,
                                                               const float* __restrict__ lw, double fxs, double fxi,
                                                               float* __restrict__ out) {
  // homo: counts; FX: int64 sums (the array is 16-byte aligned, so the
  // 64-bit LDS atomics stay aligned whatever static LDS precedes it); else
  // f32 sums
  extern __shared__ __align__(16) uint32_t s_acc[];
  unsigned long long* s_acc64 = reinterpret_cast<unsigned long long*>(s_acc);
  auto add = [&](uint32_t i, float v) {
    if constexpr (DIST == BPX_DIST_HOMO)
      __hip_atomic_fetch_add(s_acc + i, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else if constexpr (FX)
      __hip_atomic_fetch_add(s_acc64 + i, (unsigned long long)jit_fixed(v, fxs), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
    else
      __hip_atomic_fetch_add(reinterpret_cast<float*>(s_acc) + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  };
  const uint32_t l = blockIdx.x / (uint32_t)K, k = blockIdx.x % (uint32_t)K;
  const uint64_t s0 = (uint64_t)step * l;
  const uint64_t lo64 = s0 + (uint64_t)k * (uint64_t)T;
  uint64_t hi64 = lo64 + (uint64_t)T;
  if (hi64 > s0 + step) hi64 = s0 + step;
  if (hi64 > (uint64_t)num_row) hi64 = (uint64_t)num_row;
  if (lo64 < hi64) {
    const uint32_t lo = (uint32_t)lo64, hi = (uint32_t)hi64;
    for (int i = threadIdx.x; i < (FX ? 2 * T : T); i += kJitThreads) s_acc[i] = 0u;
    __syncthreads();
    // the tile's list counts, loaded together
    uint32_t nl[kEmitSlices];
#pragma unroll
    for (int jj = 0; jj < kEmitSlices; ++jj) nl[jj] = counts[((size_t)l * kEmitSlices + jj) * K + k];
    bool over = false;
#pragma unroll
    for (int jj = 0; jj < kEmitSlices; ++jj) over = over || nl[jj] > cap;
    if (!over) {
      // one entry of each list per thread and round: the 8 loads of a
      // round are in flight together
      uint32_t nmax = 0;
#pragma unroll
      for (int jj = 0; jj < kEmitSlices; ++jj) nmax = max(nmax, nl[jj]);
      for (uint32_t e0 = 0; e0 < nmax; e0 += kJitThreads) {
        const uint32_t e = e0 + threadIdx.x;
        uint32_t loc[kEmitSlices];
        float wv[kEmitSlices];
#pragma unroll
        for (int jj = 0; jj < kEmitSlices; ++jj) {
          const size_t at = (((size_t)l * kEmitSlices + jj) * K + k) * cap + e;
          loc[jj] = e < nl[jj] ? (uint32_t)lrow[at] : 0xFFFFFFFFu;
          wv[jj] = 0.f;
          if constexpr (DIST != BPX_DIST_HOMO) wv[jj] = e < nl[jj] ? lw[at] : 0.f;
        }
#pragma unroll
        for (int jj = 0; jj < kEmitSlices; ++jj)
          if (loc[jj] != 0xFFFFFFFFu) add(loc[jj], wv[jj]);
      }
    } else {
      // a list overflowed: regenerate every active stream l up to hi
      const uint32_t send = (uint32_t)min((int64_t)(s0 + step - 1), num_row);
      const uint32_t hie = min(hi, send);
      const uint32_t r0 = (uint32_t)s0 - 1u;
      const int32_t negd = -(int32_t)ds.d;
      for (int64_t c = threadIdx.x; c < num_col; c += kJitThreads) {
        if (ev[c] == 0) continue;
        Lfsr88 key(seed + (uint32_t)c * 32u + l);
        uint32_t r = next_col<MK>(r0, key.next(), dv, ds, negd);
        while (r < hie) {
          float v = w0;
          if constexpr (DIST != BPX_DIST_HOMO) v = draw_weight<DIST>(key, w0, w1);
          if (r >= lo) add(r - lo, v);
          r = next_col<MK>(r, key.next(), dv, ds, negd);
        }
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; lo + i < hi; i += kJitThreads) {
      if constexpr (DIST == BPX_DIST_HOMO) out[lo + i] = repeat_sum_f(s_acc[i], w0);
      else if constexpr (FX) out[lo + i] = (float)((double)(long long)s_acc64[i] * fxi);
      else out[lo + i] = __uint_as_float(s_acc[i]);
    }
  }
  // rows past every stream's range (32 * step <= row < num_row) get no
  // connection: the last block stores their zeros
  if (blockIdx.x == gridDim.x - 1)
    for (int64_t row = (int64_t)step * 32 + threadIdx.x; row < num_row; row += kJitThreads) out[row] = 0.f;
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

static int jit_cus() {
  static int c = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) return 256;
    return v;
  }();
  return c;
}

static bool bits_fit(int64_t nwords) {
  static int shm = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 65536;
    return v;
  }();
  return nwords * 4 <= shm;
}

// workspace of the scatter form's event list: a count, then one u32 per column
static size_t scatter_ws_bytes(int64_t num_col) { return 256 + (size_t)num_col * 4; }

// Geometry and workspace layout of the two-pass scatter (k_jit_emit,
// k_jit_sum_lists): K row tiles of T rows per stream range (as the one-pass
// tiles), slices of 1024 * span columns, list counts at counts_off, the
// lists' tile-local rows (u16) at rows_off and, for uniform and normal
// weights, their weights (f32) at w_off.  A list holds `cap` connections:
// those of 12.5 % active columns with a 25 % margin.  Matrices whose lists
// would pass 1 GiB in all keep the one-pass tiles.
struct EmitPlan {
  int64_t K, T, span;
  uint64_t lists;
  uint32_t cap;
  size_t counts_off, rows_off, w_off, bytes;
};

static bool emit_plan(int64_t num_row, int64_t num_col, int32_t clen, int dist, EmitPlan* ep) {
  if (num_row <= 0 || num_col <= 0 || clen < 1 || num_row >= 0xFFFFFFFFll || num_col >= (1ll << 27)) return false;
  const int64_t step = ((num_row + 1) >> 5) > 1 ? ((num_row + 1) >> 5) : 1;
  int64_t K = (step + 32767) / 32768;
  if (K < 8) K = 8;
  if (K > step) K = step;
  const int64_t T = (step + K - 1) / K;
  const int64_t cw = (num_col + kEmitSlices - 1) / kEmitSlices;
  int64_t span = (cw + kJitThreads - 1) / kJitThreads;     // columns per thread
  if (span >= 16) span = (span + 15) / 16 * 16;               // whole 16-byte loads
  // connections of one list with every column of its slice active: columns
  // x T rows / the mean increment (clen + 1) / 2
  const double full = (double)span * kJitThreads * (double)T * 2.0 / ((double)clen + 1.0);
  const double capd = full * 0.125 * 1.25 + 64.0;
  if (capd > 2.0e9) return false;
  ep->K = K;
  ep->T = T;
  ep->span = span;
  ep->lists = (uint64_t)32 * kEmitSlices * (uint64_t)K;
  ep->cap = (uint32_t)capd;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  ep->counts_off = 256;
  ep->rows_off = al(ep->counts_off + ep->lists * 4);
  ep->w_off = al(ep->rows_off + ep->lists * ep->cap * 2);
  ep->bytes = ep->w_off + (dist == BPX_DIST_HOMO ? 0 : ep->lists * ep->cap * 4);
  return ep->bytes <= ((size_t)1 << 30);
}

template <int DIST, int VM>
static int run_jit(hipStream_t s, const void* vec, float w0, float w1, int32_t clen, uint32_t seed, int64_t num_row,
                   int64_t num_col, int outdim, int layout, float* out, void* ws, size_t ws_bytes) {
  const DivU32 dv = DivU32::make((uint32_t)clen);
  const uint32_t step = (uint32_t)(((num_row + 1) >> 5) > 1 ? ((num_row + 1) >> 5) : 1);
  EmitPlan ep{};
  if (outdim) {
    const int64_t nwords = (num_col + 31) / 32;
    if (layout == BPX_LAYOUT_LANE32 && VM == VM_BOOL && bits_fit(nwords + 1) && ws &&
        ws_bytes >= (size_t)nwords * 4 && num_col < 0xFFFFFFC0ll && num_row < 0x7FFFFFFFll &&
        (uint64_t)num_col + 64u + 2u * (uint64_t)clen < (1ull << 32)) {
      // (a lane's column counter, 32 above the column, passes its stream's
      // end by at most two increments of up to clen: no u32 wrap)
      uint32_t* bits = (uint32_t*)ws;
      hipLaunchKernelGGL(k_pack_bits, dim3((unsigned)((num_col + 255) / 256)), dim3(256), 0, s, (const uint8_t*)vec,
                         num_col, bits);
      int64_t blocks = (num_row + kJitThreads - 1) / kJitThreads;
      const int64_t maxb = (int64_t)jit_cus();
      if (blocks > maxb) blocks = maxb;
      // finite weights, and every draw finite too (lanes past their stream's
      // end draw a weight they multiply by a zero event bit): uniform draws
      // lie within |w0| + |w1 - w0|, normal ones within |w0| + 6.8 |w1|
      // (Box-Muller with u1 >= 1e-10 bounds |z| by sqrt(-2 ln 1e-10) < 6.8)
      bool fin = isfinite(w0) && (DIST == BPX_DIST_HOMO || isfinite(w1));
      if (DIST == BPX_DIST_UNIFORM)
        fin = fin && isfinite(w1 - w0) && fabs((double)w0) + fabs((double)(w1 - w0)) < 3.0e38;
      else if (DIST == BPX_DIST_NORMAL)
        fin = fin && fabs((double)w0) + 8.0 * fabs((double)w1) < 3.0e38;
      DivShort ds = {1u, 0u, 0u};
      const bool shrt = DivShort::make((uint32_t)clen, &ds) && DivShort::mad24_ok((uint32_t)clen);
      const bool clip = (uint64_t)step * 32u > (uint64_t)num_col;
#define JRB(MKV, F, CL)                                                                                       \
  BPX_LAUNCH(KID_JIT_OUTDIM, (k_jit_rows_bits<DIST, MKV, F, CL>), dim3((unsigned)blocks), dim3(kJitThreads),    \
             (size_t)(nwords + 1) * 4, s, (const uint32_t*)bits, nwords, w0, w1, dv, ds, seed, num_row, num_col, step, \
             out)
#define JRB_F(MKV)                                 \
  do {                                             \
    if (clip) {                                    \
      if (fin) JRB(MKV, true, true);               \
      else JRB(MKV, false, true);                  \
    } else {                                       \
      if (fin) JRB(MKV, true, false);              \
      else JRB(MKV, false, false);                 \
    }                                              \
  } while (0)
      if (shrt) JRB_F(MOD_SHORT24);
      else if (dv.small_quotient()) JRB_F(MOD_GM24);
      else JRB_F(MOD_GM);
#undef JRB_F
#undef JRB
    } else if (layout == BPX_LAYOUT_LANE32 && VM == VM_BOOL && DIST != BPX_DIST_NORMAL && bits_fit(nwords) && ws &&
        ws_bytes >= (size_t)nwords * 4) {
      uint32_t* bits = (uint32_t*)ws;
      hipLaunchKernelGGL(k_pack_bits, dim3((unsigned)((num_col + 255) / 256)), dim3(256), 0, s, (const uint8_t*)vec,
                         num_col, bits);
      int64_t blocks = (num_row + kRowsPerPass - 1) / kRowsPerPass;
      const int64_t maxb = (int64_t)jit_cus() * 1;
      if (blocks > maxb) blocks = maxb;
      if (dv.small_quotient())
        BPX_LAUNCH(KID_JIT_OUTDIM, (k_jit_lane32_outdim_bits<DIST, true>), dim3((unsigned)blocks), dim3(kJitThreads),
                   (size_t)nwords * 4, s, (const uint32_t*)bits, nwords, w0, w1, dv, seed, num_row, num_col, step, out);
      else
        BPX_LAUNCH(KID_JIT_OUTDIM, (k_jit_lane32_outdim_bits<DIST, false>), dim3((unsigned)blocks), dim3(kJitThreads),
                   (size_t)nwords * 4, s, (const uint32_t*)bits, nwords, w0, w1, dv, seed, num_row, num_col, step, out);
    } else if (layout == BPX_LAYOUT_LANE32) {
      const int64_t blocks = (num_row * 32 + 255) / 256;
      BPX_LAUNCH(KID_JIT_OUTDIM, (k_jit_lane32_outdim<DIST, VM>), dim3((unsigned)blocks), dim3(256), 0, s, vec, w0,
                 w1, dv, seed, num_row, num_col, step, out);
    } else {
      const int64_t blocks = (num_row + 255) / 256;
      BPX_LAUNCH(KID_JIT_OUTDIM, (k_jit_row_outdim<DIST, VM>), dim3((unsigned)blocks), dim3(256), 0, s, vec, w0, w1,
                 dv, seed, num_row, num_col, out);
    }
  } else if (layout == BPX_LAYOUT_LANE32 && VM == VM_BOOL && ws && emit_plan(num_row, num_col, clen, DIST, &ep) &&
             ws_bytes >= ep.bytes) {
    // two passes: each active stream generated once into per-tile lists
    // (grid: 32 stream ranges x kEmitSlices column slices), then each tile
    // summed from its lists (grid: 32 x K tiles)
    uint32_t* counts = (uint32_t*)((char*)ws + ep.counts_off);
    uint16_t* lrow = (uint16_t*)((char*)ws + ep.rows_off);
    float* lw = DIST == BPX_DIST_HOMO ? nullptr : (float*)((char*)ws + ep.w_off);
    const DivU32 dT = DivU32::make((uint32_t)ep.T);
    DivShort ds = {1u, 0u, 0u};
    const bool shrt = DivShort::make((uint32_t)clen, &ds) && DivShort::mad24_ok((uint32_t)clen);
    // fixed-point sums for uniform and normal weights when every draw is
    // finite and bounded by bnd (uniform |w0| + |w1 - w0|; normal |w0| + 8|w1|,
    // Box-Muller with u1 >= 1e-10 bounding |z| by 6.8) and 64-bit sums of a
    // tile fit the LDS: bnd * 2^shift * num_col < 2^62, terms below 2^51
    double bnd = 0.0;
    if (DIST == BPX_DIST_UNIFORM) bnd = fabs((double)w0) + fabs((double)w1 - (double)w0);
    else if (DIST == BPX_DIST_NORMAL) bnd = fabs((double)w0) + 8.0 * fabs((double)w1);
    const bool fx = DIST != BPX_DIST_HOMO && isfinite(w0) && isfinite(w1) && bnd < 3.0e38 && ep.T <= 16384;
    int eb = 0;
    if (bnd > 0.0) frexp(bnd, &eb);                       // bnd < 2^eb
    int ec = 0;
    frexp((double)num_col, &ec);                          // num_col < 2^ec
    int sh = 61 - ec - eb;
    if (sh > 50 - eb) sh = 50 - eb;
    const double fxs = ldexp(1.0, sh), fxi = ldexp(1.0, -sh);
#define JE(MKV, FXV)                                                                                          \
  do {                                                                                                        \
    BPX_LAUNCH(KID_JIT_EMIT, (k_jit_emit<DIST, MKV>), dim3((unsigned)(32 * kEmitSlices)), dim3(kJitThreads),  \
               (size_t)ep.K * 4, s, (const uint8_t*)vec, num_col, w0, w1, dv, ds, seed, num_row, step, (int)ep.K, \
               (int)ep.T, dT, ep.span, ep.cap, counts, lrow, lw);                                              \
    BPX_LAUNCH(KID_JIT_SCATTER, (k_jit_sum_lists<DIST, MKV, FXV>), dim3((unsigned)(32 * ep.K)),               \
               dim3(kJitThreads), (size_t)ep.T * (FXV ? 8 : 4), s, (const uint8_t*)vec, num_col, w0, w1, dv, ds, \
               seed, num_row, step, (int)ep.K, (int)ep.T, ep.cap, (const uint32_t*)counts,                    \
               (const uint16_t*)lrow, (const float*)lw, fxs, fxi, out);                                       \
  } while (0)
#define JE_F(MKV)                                \
  do {                                           \
    if (fx) JE(MKV, DIST != BPX_DIST_HOMO);      \
    else JE(MKV, false);                         \
  } while (0)
    if (shrt) JE_F(MOD_SHORT24);
    else if (dv.small_quotient()) JE_F(MOD_GM24);
    else JE_F(MOD_GM);
#undef JE_F
#undef JE
  } else if (layout == BPX_LAYOUT_LANE32 && VM == VM_BOOL && ws && num_col > 0 && num_row > 0 &&
             ws_bytes >= scatter_ws_bytes(num_col) && num_row < 0xFFFFFFFFll && num_col < (1ll << 27)) {
    // exclusive row tiles: K tiles per stream range (at least 8, so the 32
    // ranges give ~256 blocks), each at most 32768 rows of LDS accumulators;
    // the active columns listed first (count at ws, list at ws + 256)
    uint32_t* cnt = (uint32_t*)ws;
    uint32_t* list = (uint32_t*)((char*)ws + 256);
    hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(64), 0, s, cnt);
    hipLaunchKernelGGL(k_list_events, dim3((unsigned)((num_col + 4095) / 4096)), dim3(1024), 0, s,
                       (const uint8_t*)vec, num_col, cnt, list);
    int64_t K = ((int64_t)step + 32767) / 32768;
    if (K < 8) K = 8;
    if (K > (int64_t)step) K = step;
    const int64_t T = ((int64_t)step + K - 1) / K;
    DivShort ds = {1u, 0u, 0u};
    const bool shrt = DivShort::make((uint32_t)clen, &ds) && DivShort::mad24_ok((uint32_t)clen);
#define JST(MKV)                                                                                              \
  BPX_LAUNCH(KID_JIT_SCATTER, (k_jit_scatter_tiles<DIST, MKV>), dim3((unsigned)(32 * K)), dim3(kJitThreads),      \
             (size_t)T * 4, s, (const uint32_t*)cnt, (const uint32_t*)list, w0, w1, dv, ds, seed, num_row, step,  \
             (int)K, (int)T, out)
    if (shrt) JST(MOD_SHORT24);
    else if (dv.small_quotient()) JST(MOD_GM24);
    else JST(MOD_GM);
#undef JST
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
                      int64_t num_row, int64_t num_col, int outdim, int layout, float* out, void* ws,
                      size_t wsb) {
  if (vm == VM_DENSE)
    return run_jit<DIST, VM_DENSE>(s, vec, w0, w1, clen, seed, num_row, num_col, outdim, layout, out, ws, wsb);
  if (vm == VM_BOOL)
    return run_jit<DIST, VM_BOOL>(s, vec, w0, w1, clen, seed, num_row, num_col, outdim, layout, out, ws, wsb);
  return run_jit<DIST, VM_FLOAT>(s, vec, w0, w1, clen, seed, num_row, num_col, outdim, layout, out, ws, wsb);
}

}  // namespace bpx

using namespace bpx;

extern "C" size_t bpx_jitconn_workspace(int64_t num_row, int64_t num_col) {
  (void)num_row;
  if (num_col <= 0) return 0;
  const size_t bits = (size_t)((num_col + 31) / 32) * 4;
  const size_t lst = scatter_ws_bytes(num_col);
  return bits > lst ? bits : lst;
}

extern "C" size_t bpx_jitconn_workspace_ex(int64_t num_row, int64_t num_col, int32_t clen, int dist) {
  const size_t base = bpx_jitconn_workspace(num_row, num_col);
  EmitPlan ep{};
  if (dist < BPX_DIST_HOMO || dist > BPX_DIST_NORMAL || !emit_plan(num_row, num_col, clen, dist, &ep)) return base;
  return ep.bytes > base ? ep.bytes : base;
}

extern "C" int bpx_jitconn_mv(void* stream, const void* vec, int vec_dtype, int event, int dist, float w0, float w1,
                              int32_t clen, uint32_t seed, int64_t num_row, int64_t num_col, int outdim_parallel,
                              int layout, float* out, void* workspace, size_t workspace_bytes) {
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
  if (!out || (num_col > 0 && !vec)) return BPX_ERR_ARG;
  const int o = outdim_parallel ? 1 : 0;
  switch (dist) {
    case BPX_DIST_HOMO: return run_jit_vm<BPX_DIST_HOMO>(vm, s, vec, w0, w1, clen, seed, num_row, num_col, o, layout, out, workspace, workspace_bytes);
    case BPX_DIST_UNIFORM: return run_jit_vm<BPX_DIST_UNIFORM>(vm, s, vec, w0, w1, clen, seed, num_row, num_col, o, layout, out, workspace, workspace_bytes);
    default: return run_jit_vm<BPX_DIST_NORMAL>(vm, s, vec, w0, w1, clen, seed, num_row, num_col, o, layout, out, workspace, workspace_bytes);
  }
}
