/*
 * bpx_oracle.c -- CPU restatement of BrainPy's event-driven synaptic hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity reference for the HIP
 * library in brainpy_amd/csrc.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it; the product path never does.
 *
 * Every function restates a loop of the reference (chaoming0625/BrainPy
 * 2.5.0) in plain C, in the reference's own loop order and arithmetic type,
 * and cites the file:line it follows (paths relative to the reference root).
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math, so every
 * float operation is one IEEE-754 rounding in source order).
 *
 * Parity pins (what checks each function beyond re-reading it):
 *   lfsr88_*         full-period check of each Tausworthe component
 *                    (2^31-1, 2^29-1, 2^28-1) + independent numpy restatement
 *   orc_jitconn_*    independent numpy restatement (tiny shapes, brute force),
 *                    one-hot probes, transpose/shape-swap identity, linearity,
 *                    lane32 skipped-index quirk, mean in-degree
 *   orc_event_csrmv  dense FP64 closed form (scipy), homo == heter*ones,
 *                    bool == float 0/1
 *   orc_csrmv        dense FP64 closed form
 *   orc_event_info   numpy flatnonzero
 *   orc_lif_*        closed forms: single-neuron threshold time tau*ln2,
 *                    refractory length, g*exp(-dt/tau) decay; FP64 twin
 *   bpx_expf_ref     checked against libm exp in FP64 (<= 2 ulp)
 * The normal-distribution weights (log/sqrt/sin) are pinned to tolerance only.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

float orc_expon_coef(float tau, float dt);
int orc_lif_coba_step(int64_t n, int ltc, int n_g, float* V, float* t_last, uint8_t* spike, float** g,
                      const float* tau_g, const float* E_g, const float* c_g, float I_ext, const float* I_vec,
                      float V_rest, float V_th, float V_reset, float tau, float tau_ref, float R, float t,
                      float dt);

/* dtype codes shared with include/bpx.h */
enum {
  DT_BOOL = 0, DT_F32 = 1, DT_F64 = 2, DT_I32 = 3, DT_I64 = 4, DT_BF16 = 5,
  DT_F16 = 6, DT_U32 = 7, DT_U64 = 8, DT_I8 = 9, DT_U8 = 10, DT_I16 = 11, DT_U16 = 12
};
/* float-event truthiness rules */
enum { RULE_POSITIVE = 0, RULE_NONZERO = 1 };

/* ------------------------------------------------------------------------ */
/* element access helpers                                                    */
/* ------------------------------------------------------------------------ */
static float bf16_to_f32(uint16_t h) {
  uint32_t u = ((uint32_t)h) << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static float f16_to_f32(uint16_t h) {
  uint32_t s = (h >> 15) & 1u, e = (h >> 10) & 31u, m = h & 1023u, u;
  if (e == 0) {
    if (m == 0) { u = s << 31; }
    else { /* subnormal */
      float f = ldexpf((float)m, -24);
      return s ? -f : f;
    }
  } else if (e == 31) {
    u = (s << 31) | 0x7F800000u | (m << 13);
  } else {
    u = (s << 31) | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static double get_f(const void* p, int dt, int64_t i) {
  switch (dt) {
    case DT_BOOL: case DT_U8: return (double)((const uint8_t*)p)[i];
    case DT_I8: return (double)((const int8_t*)p)[i];
    case DT_F32: return (double)((const float*)p)[i];
    case DT_F64: return ((const double*)p)[i];
    case DT_BF16: return (double)bf16_to_f32(((const uint16_t*)p)[i]);
    case DT_F16: return (double)f16_to_f32(((const uint16_t*)p)[i]);
    case DT_I32: return (double)((const int32_t*)p)[i];
    case DT_I64: return (double)((const int64_t*)p)[i];
    case DT_U32: return (double)((const uint32_t*)p)[i];
    case DT_U64: return (double)((const uint64_t*)p)[i];
    case DT_I16: return (double)((const int16_t*)p)[i];
    case DT_U16: return (double)((const uint16_t*)p)[i];
  }
  return 0.0;
}

static int64_t get_i(const void* p, int dt, int64_t i) {
  switch (dt) {
    case DT_I32: return ((const int32_t*)p)[i];
    case DT_I64: return ((const int64_t*)p)[i];
    case DT_U32: return (int64_t)((const uint32_t*)p)[i];
    case DT_U64: return (int64_t)((const uint64_t*)p)[i];
    case DT_I8: return ((const int8_t*)p)[i];
    case DT_U8: return ((const uint8_t*)p)[i];
    case DT_I16: return ((const int16_t*)p)[i];
    case DT_U16: return ((const uint16_t*)p)[i];
  }
  return 0;
}

/* Truthiness of an event.  Bool: non-zero byte.  Float: the Numba kernels use
 * `event > 0.` (event/_csr_matvec.py:349,357,400,410); the Taichi kernels use
 * `events[i] != 0.` (event/_csr_matvec_taichi.py:70,130). */
static int ev_true(const void* ev, int dt, int64_t i, int rule) {
  if (dt == DT_BOOL) return ((const uint8_t*)ev)[i] != 0;
  double v = get_f(ev, dt, i);
  return rule == RULE_POSITIVE ? (v > 0.0) : (v != 0.0);
}

/* ------------------------------------------------------------------------ */
/* event.info  (event/_info_collection.py:143-154, batched :54-66)           */
/* ------------------------------------------------------------------------ */
int orc_event_info(const void* ev, int ev_dt, int64_t n, int32_t* ids, int32_t* num) {
  int32_t k = 0;
  for (int64_t i = 0; i < n; ++i) ids[i] = -1;
  for (int64_t i = 0; i < n; ++i) {
    if (ev_true(ev, ev_dt, i, RULE_NONZERO)) ids[k++] = (int32_t)i;
  }
  num[0] = k;
  return 0;
}

int orc_event_info_batch(const void* ev, int ev_dt, int64_t b, int64_t n, int32_t* ids, int32_t* num) {
  size_t esz = (ev_dt == DT_BOOL || ev_dt == DT_U8 || ev_dt == DT_I8) ? 1 :
               (ev_dt == DT_F64 || ev_dt == DT_I64 || ev_dt == DT_U64) ? 8 :
               (ev_dt == DT_BF16 || ev_dt == DT_F16 || ev_dt == DT_I16 || ev_dt == DT_U16) ? 2 : 4;
  for (int64_t r = 0; r < b; ++r)
    orc_event_info((const char*)ev + (size_t)(r * n) * esz, ev_dt, n, ids + r * n, num + r);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* event csrmv                                                               */
/*  transpose=True : serial scatter in ascending row order, accumulating in  */
/*                   the output type (event/_csr_matvec.py:321-361; Taichi   */
/*                   serialize=True event/_csr_matvec_taichi.py:34-58).      */
/*  transpose=False: per-row gather; the Numba kernel's `r = 0.` is a        */
/*                   float64 accumulator rounded once on store               */
/*                   (event/_csr_matvec.py:364-413).                         */
/* Output type: data type (f32/f64); bf16/f16 data accumulate in f32.        */
/* ------------------------------------------------------------------------ */
int orc_event_csrmv(const void* data, int data_dt, int64_t data_len,
                    const void* indices, int idx_dt, const void* indptr, int ptr_dt,
                    const void* ev, int ev_dt, int64_t m, int64_t n,
                    int transpose, int rule, void* out) {
  int homo = data_len == 1;
  if (data_dt == DT_F64) {
    double* o = (double*)out;
    int64_t nout = transpose ? n : m;
    for (int64_t i = 0; i < nout; ++i) o[i] = 0.0;
    if (transpose) {
      for (int64_t r = 0; r < m; ++r) {
        if (!ev_true(ev, ev_dt, r, rule)) continue;
        for (int64_t j = get_i(indptr, ptr_dt, r); j < get_i(indptr, ptr_dt, r + 1); ++j)
          o[get_i(indices, idx_dt, j)] += get_f(data, data_dt, homo ? 0 : j);
      }
    } else {
      for (int64_t r = 0; r < m; ++r) {
        double acc = 0.0;
        for (int64_t j = get_i(indptr, ptr_dt, r); j < get_i(indptr, ptr_dt, r + 1); ++j)
          if (ev_true(ev, ev_dt, get_i(indices, idx_dt, j), rule)) acc += get_f(data, data_dt, homo ? 0 : j);
        o[r] = acc;
      }
    }
  } else {
    float* o = (float*)out;
    int64_t nout = transpose ? n : m;
    for (int64_t i = 0; i < nout; ++i) o[i] = 0.0f;
    if (transpose) {
      for (int64_t r = 0; r < m; ++r) {
        if (!ev_true(ev, ev_dt, r, rule)) continue;
        for (int64_t j = get_i(indptr, ptr_dt, r); j < get_i(indptr, ptr_dt, r + 1); ++j)
          o[get_i(indices, idx_dt, j)] += (float)get_f(data, data_dt, homo ? 0 : j);
      }
    } else {
      for (int64_t r = 0; r < m; ++r) {
        double acc = 0.0;
        for (int64_t j = get_i(indptr, ptr_dt, r); j < get_i(indptr, ptr_dt, r + 1); ++j)
          if (ev_true(ev, ev_dt, get_i(indices, idx_dt, j), rule)) acc += (double)(float)get_f(data, data_dt, homo ? 0 : j);
        o[r] = (float)acc;
      }
    }
  }
  return 0;
}

/* Specialised serial scatter (int32 indices, bool events, f32 data): the
 * CPU-baseline form of the loop above (same order, same f32 accumulation),
 * without the per-element dtype dispatch.  event/_csr_matvec.py:321-340. */
int orc_event_csrmv_scatter_f32(const float* data, int64_t data_len, const int32_t* indices,
                                const int32_t* indptr, const uint8_t* ev, int64_t m, int64_t n,
                                float* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = 0.0f;
  if (data_len == 1) {
    const float w = data[0];
    for (int64_t r = 0; r < m; ++r) {
      if (!ev[r]) continue;
      for (int32_t j = indptr[r]; j < indptr[r + 1]; ++j) out[indices[j]] += w;
    }
  } else {
    for (int64_t r = 0; r < m; ++r) {
      if (!ev[r]) continue;
      for (int32_t j = indptr[r]; j < indptr[r + 1]; ++j) out[indices[j]] += data[j];
    }
  }
  return 0;
}

/* The same scatter spread over threads (the CPU baseline's all-thread
 * figure; the reference's own CPU transpose scatter is serial): thread t
 * scatters a contiguous block of rows into a private output, then the
 * private outputs are summed column by column.  Sums of a column are taken
 * in another order than the serial loop's, so results agree within f32
 * rounding, not bit for bit. */
#ifdef _OPENMP
#include <omp.h>
#endif
int orc_event_csrmv_scatter_f32_threads(const float* data, int64_t data_len, const int32_t* indices,
                                        const int32_t* indptr, const uint8_t* ev, int64_t m, int64_t n,
                                        float* out, int threads, float* scratch) {
  /* scratch: threads * n floats */
#ifdef _OPENMP
  if (threads < 1) threads = 1;
#pragma omp parallel num_threads(threads)
  {
    const int t = omp_get_thread_num(), nt = omp_get_num_threads();
    float* acc = scratch + (int64_t)t * n;
    for (int64_t i = 0; i < n; ++i) acc[i] = 0.0f;
    const int64_t r0 = m * t / nt, r1 = m * (t + 1) / nt;
    for (int64_t r = r0; r < r1; ++r) {
      if (!ev[r]) continue;
      if (data_len == 1) {
        const float w = data[0];
        for (int32_t j = indptr[r]; j < indptr[r + 1]; ++j) acc[indices[j]] += w;
      } else {
        for (int32_t j = indptr[r]; j < indptr[r + 1]; ++j) acc[indices[j]] += data[j];
      }
    }
#pragma omp barrier
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      float v = 0.0f;
      for (int k = 0; k < nt; ++k) v += scratch[(int64_t)k * n + i];
      out[i] = v;
    }
  }
  return 0;
#else
  (void)threads;
  (void)scratch;
  return orc_event_csrmv_scatter_f32(data, data_len, indices, indptr, ev, m, n, out);
#endif
}

/* ------------------------------------------------------------------------ */
/* sparse csrmv (sparse/_csr_mv.py:126-165)                                  */
/*  transpose=True : serial scatter `out[col] += w * v[row]` (f32 accum,     */
/*                   homo: values * v, heter: v * values[j])                 */
/*  transpose=False: gather with float64 accumulator of f32 products         */
/* ------------------------------------------------------------------------ */
int orc_csrmv(const void* data, int data_dt, int64_t data_len,
              const void* indices, int idx_dt, const void* indptr, int ptr_dt,
              const void* vec, int vec_dt, int64_t m, int64_t n, int transpose, void* out) {
  int homo = data_len == 1;
  int64_t nout = transpose ? n : m;
  if (data_dt == DT_F64) {
    double* o = (double*)out;
    for (int64_t i = 0; i < nout; ++i) o[i] = 0.0;
    if (transpose) {
      for (int64_t r = 0; r < m; ++r) {
        double v = get_f(vec, vec_dt, r);
        for (int64_t j = get_i(indptr, ptr_dt, r); j < get_i(indptr, ptr_dt, r + 1); ++j)
          o[get_i(indices, idx_dt, j)] += homo ? get_f(data, data_dt, 0) * v : v * get_f(data, data_dt, j);
      }
    } else {
      for (int64_t r = 0; r < m; ++r) {
        double acc = 0.0;
        for (int64_t j = get_i(indptr, ptr_dt, r); j < get_i(indptr, ptr_dt, r + 1); ++j)
          acc += get_f(data, data_dt, homo ? 0 : j) * get_f(vec, vec_dt, get_i(indices, idx_dt, j));
        o[r] = acc;
      }
    }
  } else {
    float* o = (float*)out;
    for (int64_t i = 0; i < nout; ++i) o[i] = 0.0f;
    if (transpose) {
      for (int64_t r = 0; r < m; ++r) {
        float v = (float)get_f(vec, vec_dt, r);
        for (int64_t j = get_i(indptr, ptr_dt, r); j < get_i(indptr, ptr_dt, r + 1); ++j) {
          float w = (float)get_f(data, data_dt, homo ? 0 : j);
          o[get_i(indices, idx_dt, j)] += homo ? w * v : v * w;
        }
      }
    } else {
      for (int64_t r = 0; r < m; ++r) {
        double acc = 0.0;
        for (int64_t j = get_i(indptr, ptr_dt, r); j < get_i(indptr, ptr_dt, r + 1); ++j) {
          float w = (float)get_f(data, data_dt, homo ? 0 : j);
          float x = (float)get_f(vec, vec_dt, get_i(indices, idx_dt, j));
          acc += (double)(w * x);
        }
        o[r] = (float)acc;
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* LFSR88 (math/tifunc.py:47-180), L'Ecuyer's taus88 as written there.        */
/* ------------------------------------------------------------------------ */
typedef struct { uint32_t s0, s1, s2, s3; } key4;

static inline key4 lfsr88_key(uint32_t seed) {           /* tifunc.py:67 */
  key4 k = {seed + 1u, seed + 7u, seed + 15u, 0u};
  return k;
}

static inline key4 lfsr88_next(key4 k) {                  /* tifunc.py:80-86 */
  key4 o;
  uint32_t b = ((k.s0 << 13) ^ k.s0) >> 19;
  o.s0 = ((k.s0 & 4294967294u) << 12) ^ b;
  b = ((k.s1 << 2) ^ k.s1) >> 25;
  o.s1 = ((k.s1 & 4294967288u) << 4) ^ b;
  b = ((k.s2 << 3) ^ k.s2) >> 11;
  o.s2 = ((k.s2 & 4294967280u) << 17) ^ b;
  o.s3 = b;
  return o;
}

static inline uint32_t lfsr88_bits(key4 k) { return k.s0 ^ k.s1 ^ k.s2; }

/* lfsr88_random_integers(key, low, high) (tifunc.py:136-147): u32 modulo by
 * (high + 1 - low), plus low, cast to i32.  Every call site in the jitconn
 * kernels uses (0, clen-1) or (1, clen), so the divisor is always clen. */
static inline int32_t lfsr88_randint(key4* k, int32_t low, int32_t high) {
  *k = lfsr88_next(*k);
  return (int32_t)(lfsr88_bits(*k) % (uint32_t)(high + 1 - low) + (uint32_t)low);
}

/* lfsr88_uniform (tifunc.py:156-168): f32(bits) * f32(2^-32), then
 * r*(high-low)+low as two f32 roundings (no contraction). */
static inline float lfsr88_uniform(key4* k, float low, float high) {
  *k = lfsr88_next(*k);
  float r = (float)lfsr88_bits(*k) * 2.3283064365386963e-10f;
  float span = high - low;
  float t = r * span;
  return t + low;
}

static inline float lfsr88_rand(key4* k) {                /* tifunc.py:171-180 */
  *k = lfsr88_next(*k);
  return (float)lfsr88_bits(*k) * 2.3283064365386963e-10f;
}

/* lfsr88_normal / lfsr88_randn (tifunc.py:89-133): Box-Muller sine branch in
 * f32.  log/sqrt/sin are libm's; the GPU uses its own device math, so this
 * value is pinned only to a tolerance. */
static inline float lfsr88_normal(key4* k, float mu, float sigma) {
  float u1 = lfsr88_rand(k);
  float u2 = lfsr88_rand(k);
  u1 = u1 > 1e-10f ? u1 : 1e-10f;
  float mag = sqrtf(-2.0f * logf(u1));
  float z = mag * sinf(6.283185307179586f * u2);
  float t = sigma * z;
  return mu + t;
}

/* raw taus88 output stream for a seed: bits after each next() */
int orc_lfsr88_stream(uint32_t seed, int64_t count, uint32_t* out) {
  key4 k = lfsr88_key(seed);
  for (int64_t i = 0; i < count; ++i) { k = lfsr88_next(k); out[i] = lfsr88_bits(k); }
  return 0;
}

int orc_lfsr88_randint_stream(uint32_t seed, int32_t low, int32_t high, int64_t count, int32_t* out) {
  key4 k = lfsr88_key(seed);
  for (int64_t i = 0; i < count; ++i) out[i] = lfsr88_randint(&k, low, high);
  return 0;
}

int orc_lfsr88_uniform_stream(uint32_t seed, float low, float high, int64_t count, float* out) {
  key4 k = lfsr88_key(seed);
  for (int64_t i = 0; i < count; ++i) out[i] = lfsr88_uniform(&k, low, high);
  return 0;
}

int orc_lfsr88_normal_stream(uint32_t seed, float mu, float sigma, int64_t count, float* out) {
  key4 k = lfsr88_key(seed);
  for (int64_t i = 0; i < count; ++i) out[i] = lfsr88_normal(&k, mu, sigma);
  return 0;
}

/* Period of one Tausworthe component started from `start` (component 0,1,2).
 * Returns the number of next() steps until the state repeats, or 0 if not
 * found within `limit` steps.  Used to pin the shift/mask constants: taus88's
 * components have periods 2^31-1, 2^29-1, 2^28-1 for any valid start. */
int64_t orc_lfsr88_component_period(int comp, uint32_t start, int64_t limit) {
  key4 k = {start, start, start, 0};
  uint32_t s0 = comp == 0 ? k.s0 : comp == 1 ? k.s1 : k.s2;
  uint32_t mask = comp == 0 ? 4294967294u : comp == 1 ? 4294967288u : 4294967280u;
  uint32_t first = s0 & mask;
  for (int64_t i = 1; i <= limit; ++i) {
    k = lfsr88_next(k);
    uint32_t s = comp == 0 ? k.s0 : comp == 1 ? k.s1 : k.s2;
    if ((s & mask) == first) return i;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Just-in-time connectivity matvec.                                          */
/*   dist : 0 homo (w0), 1 uniform [w0, w1), 2 normal (mu=w0, sigma=w1)      */
/*   vmode: 0 dense vector (mv_prob_*), 1 bool events, 2 float events        */
/*   layout: 0 "lane32" = Taichi GPU kernels, 1 "row" = Taichi CPU kernels   */
/* The matrix has shape (num_row, num_col) = (len(out), len(vec)) after the   */
/* transpose shape swap of _general_checking (jitconn/_matvec_taichi.py:     */
/* 174-207); the caller passes the already-swapped shape.                     */
/* Accumulators are f32 (Taichi's `r = 0.` with default_fp = f32).            */
/* ------------------------------------------------------------------------ */
static inline float draw_w(key4* k, int dist, float w0, float w1) {
  if (dist == 1) return lfsr88_uniform(k, w0, w1);
  if (dist == 2) return lfsr88_normal(k, w0, w1);
  return w0;
}

static inline float vec_at(const void* v, int dt, int64_t i) { return (float)get_f(v, dt, i); }

static inline int vec_on(const void* v, int dt, int64_t i, int vmode) {
  if (vmode == 1) return ((const uint8_t*)v)[i] != 0;
  return get_f(v, dt, i) != 0.0;                /* Taichi `!= 0.` */
}

/* Row layout (CPU kernels): _event_matvec_taichi.py:38-85 (homo),
 * :446-509 (uniform), :863-916 (normal); _matvec_taichi.py:30-76, 337-386,
 * 627-676. */
static void jit_row_layout(int dist, int vmode, int outdim_parallel, const void* vec, int vdt,
                           int64_t num_row, int64_t num_col, float w0, float w1,
                           int32_t clen, uint32_t seed, float* out) {
  if (outdim_parallel) {
    for (int64_t r = 0; r < num_row; ++r) {
      float acc = 0.0f;
      key4 k = lfsr88_key(seed + (uint32_t)r);
      int32_t c = lfsr88_randint(&k, 0, clen - 1);
      while (c < num_col) {
        if (vmode == 0) {
          if (dist == 0) acc += vec_at(vec, vdt, c);
          else { float w = draw_w(&k, dist, w0, w1); acc += vec_at(vec, vdt, c) * w; }
        } else {
          float w = draw_w(&k, dist, w0, w1);
          if (vec_on(vec, vdt, c, vmode)) acc += w;
        }
        c += lfsr88_randint(&k, 1, clen);
      }
      out[r] = (vmode == 0 && dist == 0) ? acc * w0 : acc;
    }
  } else {
    for (int64_t r = 0; r < num_row; ++r) out[r] = 0.0f;
    for (int64_t c = 0; c < num_col; ++c) {
      if (vmode != 0 && !vec_on(vec, vdt, c, vmode)) continue;
      float cv = vmode == 0 ? vec_at(vec, vdt, c) : 0.0f;
      float hv = (vmode == 0 && dist == 0) ? cv * w0 : 0.0f;
      key4 k = lfsr88_key(seed + (uint32_t)c);
      int32_t r = lfsr88_randint(&k, 0, clen - 1);
      while (r < num_row) {
        if (vmode == 0) {
          if (dist == 0) out[r] += hv;
          else { float w = draw_w(&k, dist, w0, w1); out[r] += cv * w; }
        } else {
          out[r] += draw_w(&k, dist, w0, w1);
        }
        r += lfsr88_randint(&k, 1, clen);
      }
    }
  }
}

/* lane32 layout (GPU kernels): _event_matvec_taichi.py:92-154 (homo),
 * :512-573 (uniform), :920-983 (normal); _matvec_taichi.py:79-137, 389-456,
 * 679-746.  32 independent streams per row (or column), keyed seed+i*32+l;
 * lane l walks (step*l-1, min(step*(l+1)-1, bound)) in u32 with
 * step = max((num_row+1)>>5, 1) taken from num_row in both variants.
 * Lane partials are added into out in ascending lane order. */
static void jit_lane32_layout(int dist, int vmode, int outdim_parallel, const void* vec, int vdt,
                              int64_t num_row, int64_t num_col, float w0, float w1,
                              int32_t clen, uint32_t seed, float* out) {
  uint32_t step = (uint32_t)(((num_row + 1) >> 5) > 1 ? ((num_row + 1) >> 5) : 1);
  for (int64_t r = 0; r < num_row; ++r) out[r] = 0.0f;
  if (outdim_parallel) {
    for (int64_t r = 0; r < num_row; ++r) {
      for (uint32_t l = 0; l < 32; ++l) {
        uint32_t c = step * l - 1u;
        uint32_t end = c + step;
        if ((int64_t)end > num_col) end = (uint32_t)num_col;
        float acc = 0.0f;
        key4 k = lfsr88_key(seed + (uint32_t)(r * 32 + l));
        c += (uint32_t)lfsr88_randint(&k, 1, clen);
        while (c < end) {
          if (vmode == 0) {
            if (dist == 0) acc += vec_at(vec, vdt, c);
            else { float w = draw_w(&k, dist, w0, w1); acc += vec_at(vec, vdt, c) * w; }
          } else {
            float w = draw_w(&k, dist, w0, w1);
            float e = vmode == 1 ? (float)(((const uint8_t*)vec)[c] != 0) : vec_at(vec, vdt, c);
            acc += w * e;                         /* Q4: GPU multiplies by the event */
          }
          c += (uint32_t)lfsr88_randint(&k, 1, clen);
        }
        out[r] += (vmode == 0 && dist == 0) ? w0 * acc : acc;
      }
    }
  } else {
    for (int64_t col = 0; col < num_col; ++col) {
      if (vmode != 0 && !vec_on(vec, vdt, col, vmode)) continue;
      float cv = vmode == 0 ? vec_at(vec, vdt, col) : 0.0f;
      for (uint32_t l = 0; l < 32; ++l) {
        uint32_t r = step * l - 1u;
        uint32_t end = r + step;
        if ((int64_t)end > num_row) end = (uint32_t)num_row;
        key4 k = lfsr88_key(seed + (uint32_t)(col * 32 + l));
        r += (uint32_t)lfsr88_randint(&k, 1, clen);
        while (r < end) {
          if (vmode == 0) {
            if (dist == 0) out[r] += w0 * cv;
            else { float w = draw_w(&k, dist, w0, w1); out[r] += w * cv; }
          } else {
            out[r] += draw_w(&k, dist, w0, w1);
          }
          r += (uint32_t)lfsr88_randint(&k, 1, clen);
        }
      }
    }
  }
}

int orc_jitconn_mv(int dist, int vmode, int outdim_parallel, int layout,
                   const void* vec, int vdt, int64_t num_row, int64_t num_col,
                   float w0, float w1, int32_t clen, uint32_t seed, float* out) {
  if (clen < 1) return -1;
  if (layout == 1) jit_row_layout(dist, vmode, outdim_parallel, vec, vdt, num_row, num_col, w0, w1, clen, seed, out);
  else jit_lane32_layout(dist, vmode, outdim_parallel, vec, vdt, num_row, num_col, w0, w1, clen, seed, out);
  return 0;
}

/* Enumerate the generated connections (row, col, weight) of the (num_row,
 * num_col) matrix, in generation order, for brute-force checks.  Returns the
 * number of connections (writing at most cap of them). */
int64_t orc_jitconn_enumerate(int dist, int outdim_parallel, int layout, int64_t num_row, int64_t num_col,
                              float w0, float w1, int32_t clen, uint32_t seed,
                              int32_t* rows, int32_t* cols, float* ws, int64_t cap) {
  int64_t cnt = 0;
#define EMIT(R, C, W) do { if (cnt < cap) { rows[cnt] = (int32_t)(R); cols[cnt] = (int32_t)(C); ws[cnt] = (W); } ++cnt; } while (0)
  if (layout == 1) {
    int64_t outer = outdim_parallel ? num_row : num_col, inner = outdim_parallel ? num_col : num_row;
    for (int64_t a = 0; a < outer; ++a) {
      key4 k = lfsr88_key(seed + (uint32_t)a);
      int32_t b = lfsr88_randint(&k, 0, clen - 1);
      while (b < inner) {
        float w = draw_w(&k, dist, w0, w1);
        if (outdim_parallel) EMIT(a, b, w); else EMIT(b, a, w);
        b += lfsr88_randint(&k, 1, clen);
      }
    }
  } else {
    uint32_t step = (uint32_t)(((num_row + 1) >> 5) > 1 ? ((num_row + 1) >> 5) : 1);
    int64_t outer = outdim_parallel ? num_row : num_col, inner = outdim_parallel ? num_col : num_row;
    for (int64_t a = 0; a < outer; ++a) {
      for (uint32_t l = 0; l < 32; ++l) {
        uint32_t b = step * l - 1u;
        uint32_t end = b + step;
        if ((int64_t)end > inner) end = (uint32_t)inner;
        key4 k = lfsr88_key(seed + (uint32_t)(a * 32 + l));
        b += (uint32_t)lfsr88_randint(&k, 1, clen);
        while (b < end) {
          float w = draw_w(&k, dist, w0, w1);
          if (outdim_parallel) EMIT(a, b, w); else EMIT(b, a, w);
          b += (uint32_t)lfsr88_randint(&k, 1, clen);
        }
      }
    }
  }
#undef EMIT
  return cnt;
}

/* ------------------------------------------------------------------------ */
/* exp(x) in f32 by Cody-Waite reduction and a degree-7 Taylor polynomial in */
/* Horner form with explicit fmaf.  The HIP library uses the same algorithm  */
/* (brainpy_amd/csrc/bpx_math.h) so the neuron update is bit-identical on    */
/* host and device; correctly rounded fmaf/rintf/ldexpf make it portable.    */
/* ------------------------------------------------------------------------ */
float bpx_expf_ref(float x) {
  if (x > 88.7228393554688f) return INFINITY;
  if (x < -103.972084045410f) return 0.0f;
  float n = rintf(x * 1.44269504088896341f);
  float r = fmaf(-n, 0.693145751953125f, x);
  r = fmaf(-n, 1.428606765330187045e-06f, r);
  float p = 1.98412698412698413e-04f;                  /* 1/7! */
  p = fmaf(p, r, 1.38888888888888889e-03f);            /* 1/6! */
  p = fmaf(p, r, 8.33333333333333333e-03f);            /* 1/5! */
  p = fmaf(p, r, 4.16666666666666667e-02f);            /* 1/4! */
  p = fmaf(p, r, 1.66666666666666667e-01f);            /* 1/3! */
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)n);
}

/* exprel (math/others.py:103-126): threshold 1e-5 for f32 */
static inline float exprel_f(float z) {
  float az = z < 0 ? -z : z;
  if (az <= 1e-5f) {
    float h = z / 2.0f;
    float s = 1.0f + h;
    float q = (z * z) / 6.0f;
    return s + q;
  }
  return (bpx_expf_ref(z) - 1.0f) / z;
}

static inline double exprel_d(double z) {
  if (fabs(z) <= 1e-8) return 1.0 + z / 2.0 + z * z / 6.0;
  return expm1(z) / z;
}

/* Expon decay coefficient c = dt * exprel(dt * (-(1/tau))) in f32, the factor
 * of exp_euler's `g + dt*phi*(-g/tau)` (integrators/ode/exponential.py:357-362,
 * dyn/synapses/abstract_models.py:121-131). */
float orc_expon_coef(float tau, float dt) {
  float lin = -(1.0f / tau);
  float z = dt * lin;
  return dt * exprel_f(z);
}

/* ------------------------------------------------------------------------ */
/* Fused LIF + Expon + COBA step for one population (f32), in the order      */
/* of the reference step (A.3 of the parity spec):                           */
/*   for each conductance k (registration order):                            */
/*     g_k = g_k + c_k * ((-g_k) / tau_k)           abstract_models.py:127-131 */
/*   I(V)  = I_ext + sum_k g_k*(E_k - V)           mixin.py:427-447,         */
/*                                                 outputs.py:53-54          */
/*   f     = ((-V + V_rest) + R*I(V)) / tau        lif.py:254-256            */
/*   L     = -(1 + R*sum_k g_k) / tau  (LTC)  or  -(1/tau)  (LifRef)         */
/*   V1    = V + (dt*exprel(dt*L)) * f             exponential.py:357-362    */
/*   ref   = (t - t_last) <= tau_ref; V1 = ref ? V : V1      lif.py:470-472  */
/*   spike = V1 >= V_th; V = spike ? V_reset : V1; t_last = spike ? t : t_last */
/*                                                 lif.py:487-497            */
/* ltc=1: LifRefLTC (currents inside the derivative);                         */
/* ltc=0: LifRef (currents summed once with the old V, lif.py:550-556).       */
/* ------------------------------------------------------------------------ */
int orc_lif_coba_step(int64_t n, int ltc, int n_g,
                      float* V, float* t_last, uint8_t* spike, float** g,
                      const float* tau_g, const float* E_g, const float* c_g,
                      float I_ext, const float* I_vec,
                      float V_rest, float V_th, float V_reset, float tau, float tau_ref, float R,
                      float t, float dt) {
  float cV = ltc ? 0.0f : dt * exprel_f(dt * (-(1.0f / tau)));
  for (int64_t i = 0; i < n; ++i) {
    float gk[8];
    float gsum = 0.0f;
    for (int k = 0; k < n_g; ++k) {
      float gv = g[k][i];
      gv = gv + c_g[k] * ((-gv) / tau_g[k]);
      g[k][i] = gv;
      gk[k] = gv;
      gsum = gsum + gv;
    }
    float v = V[i];
    float I = I_vec ? I_vec[i] : I_ext;
    for (int k = 0; k < n_g; ++k) I = I + gk[k] * (E_g[k] - v);
    float f = ((-v + V_rest) + R * I) / tau;
    float c;
    if (ltc) {
      float L = (-(1.0f + R * gsum)) / tau;
      c = dt * exprel_f(dt * L);
    } else {
      c = cV;
    }
    float v1 = v + c * f;
    int ref = (t - t_last[i]) <= tau_ref;
    if (ref) v1 = v;
    int sp = v1 >= V_th;
    V[i] = sp ? V_reset : v1;
    if (sp) t_last[i] = t;
    spike[i] = (uint8_t)sp;
  }
  return 0;
}

/* FP64 twin of orc_lif_coba_step (error reporting only). */
int orc_lif_coba_step_f64(int64_t n, int ltc, int n_g,
                          double* V, double* t_last, uint8_t* spike, double** g,
                          const double* tau_g, const double* E_g,
                          double I_ext, const double* I_vec,
                          double V_rest, double V_th, double V_reset, double tau, double tau_ref, double R,
                          double t, double dt) {
  for (int64_t i = 0; i < n; ++i) {
    double gk[8], gsum = 0.0;
    for (int k = 0; k < n_g; ++k) {
      double gv = g[k][i];
      gv = gv + dt * exprel_d(-dt / tau_g[k]) * (-gv / tau_g[k]);
      g[k][i] = gv;
      gk[k] = gv;
      gsum += gv;
    }
    double v = V[i];
    double I = I_vec ? I_vec[i] : I_ext;
    for (int k = 0; k < n_g; ++k) I += gk[k] * (E_g[k] - v);
    double f = (-v + V_rest + R * I) / tau;
    double L = ltc ? -(1.0 + R * gsum) / tau : -1.0 / tau;
    double v1 = v + dt * exprel_d(dt * L) * f;
    if ((t - t_last[i]) <= tau_ref) v1 = v;
    int sp = v1 >= V_th;
    V[i] = sp ? V_reset : v1;
    if (sp) t_last[i] = t;
    spike[i] = (uint8_t)sp;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Whole E/I COBA network run (f32), the CPU restatement of the reference    */
/* step loop (examples/dynamics_simulation/COBA.py:58-109 with DSRunner,     */
/* runners.py:619-643): per step, every projection's serial event csrmv      */
/* scatter into a zeroed output (event/_csr_matvec.py:321-340), g += out,   */
/* then the E and I neuron updates.  The CSR is the per-neuron combined      */
/* form (row p lists network-local targets; targets of E neurons feed g_exc, */
/* of I neurons g_inh).  Homogeneous weights: the scatter output of a target */
/* receiving c spikes is ((0 + w) + w) + ... (c terms), computed as such.    */
/* raster (nullable) receives steps x N spike flags, V_out (nullable) the    */
/* final potentials; returns total spikes.                                   */
/* ------------------------------------------------------------------------ */
int64_t orc_coba_net_run(int32_t n_exc, int32_t n_inh, const int32_t* indptr, const int32_t* indices,
                         float w_exc, float w_inh, float I_ext, const float* V0, int64_t steps, int ltc,
                         float V_rest, float V_th, float V_reset, float tau, float tau_ref, float R,
                         float tau_e, float tau_i, float E_e, float E_i, float dt, uint8_t* raster,
                         const int32_t* delay4, float* V_out) {
  /* delay4: steps of delay of the projections E2E, E2I, I2E, I2I (NULL: all
   * 1).  The projection with delay d reads the spikes of step i - d at step
   * i (delay.py:305-349, rotation update with the delay data written after
   * the pre update); d = 0 reads the pre's current spikes, those of step
   * i - 1, as d = 1 does. */
  const int32_t N = n_exc + n_inh;
  int32_t D[4];
  int32_t H = 2;
  for (int k = 0; k < 4; ++k) {
    D[k] = delay4 ? (delay4[k] < 1 ? 1 : delay4[k]) : 1;
    if (D[k] + 1 > H) H = D[k] + 1;
  }
  float* V = (float*)malloc(sizeof(float) * N);
  float* tl = (float*)malloc(sizeof(float) * N);
  float* ge = (float*)calloc(N, sizeof(float));
  float* gi = (float*)calloc(N, sizeof(float));
  float* oute = (float*)calloc(N, sizeof(float));
  float* outi = (float*)calloc(N, sizeof(float));
  uint8_t* sp = (uint8_t*)calloc(N, 1);
  uint8_t* hist = (uint8_t*)calloc((size_t)H * N, 1);     /* spikes of step i at hist[i % H] */
  int64_t total = 0;
  const float tau_g[2] = {tau_e, tau_i}, E_g[2] = {E_e, E_i};
  const float c_g[2] = {orc_expon_coef(tau_e, dt), orc_expon_coef(tau_i, dt)};
  for (int32_t j = 0; j < N; ++j) { V[j] = V0[j]; tl[j] = -1e7f; }
  for (int64_t i = 0; i < steps; ++i) {
    const float t = (float)i * dt;
    for (int32_t j = 0; j < N; ++j) { oute[j] = 0.f; outi[j] = 0.f; }
    for (int32_t p = 0; p < N; ++p) {
      const int exc = p < n_exc;
      const int ka = exc ? 0 : 2, kb = ka + 1;          /* projections to E and to I */
      const int64_t sa = i - D[ka], sb = i - D[kb];
      const int on_a = sa >= 0 && hist[(sa % H) * N + p], on_b = sb >= 0 && hist[(sb % H) * N + p];
      if (!on_a && !on_b) continue;
      float* o = exc ? oute : outi;
      const float w = exc ? w_exc : w_inh;
      for (int32_t q = indptr[p]; q < indptr[p + 1]; ++q) {
        const int32_t post = indices[q];
        if (post < n_exc ? on_a : on_b) o[post] += w;
      }
    }
    for (int32_t j = 0; j < N; ++j) { ge[j] = ge[j] + oute[j]; gi[j] = gi[j] + outi[j]; }
    float* gp[2] = {ge, gi};
    orc_lif_coba_step(N, ltc, 2, V, tl, sp, gp, tau_g, E_g, c_g, I_ext, NULL, V_rest, V_th, V_reset, tau, tau_ref,
                      R, t, dt);
    memcpy(hist + (i % H) * N, sp, N);
    for (int32_t j = 0; j < N; ++j) total += sp[j];
    if (raster) memcpy(raster + i * N, sp, N);
  }
  if (V_out) memcpy(V_out, V, sizeof(float) * N);
  free(V); free(tl); free(ge); free(gi); free(oute); free(outi); free(sp); free(hist);
  return total;
}

/* ---------------------------------------------------------------------------
 * STDP weight updates (dnn/linear.py:238-275, 650-700), f32.
 * np.maximum / np.minimum propagate NaN (fmaxf/fminf would not).
 * ------------------------------------------------------------------------- */
static float np_max_f(float a, float b) { return (a != a || b != b) ? NAN : (a > b ? a : b); }
static float np_min_f(float a, float b) { return (a != a || b != b) ? NAN : (a < b ? a : b); }

/* _cpu_csr_on_pre_update (dnn/linear.py:650-660) */
int orc_csr_on_pre_update(const float* w, const int32_t* indices, const int32_t* indptr, int64_t nnz,
                          const uint8_t* spike, int64_t num_pre, const float* trace, int64_t num_post,
                          float w_min, float w_max, float* out) {
  for (int64_t k = 0; k < nnz; ++k) out[k] = w[k];
  for (int64_t i = 0; i < num_pre; ++i) {
    if (!spike[i]) continue;
    for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k) {
      int64_t j = indices[k];
      if (j < 0 || j >= num_post) return -1;
      out[k] = np_min_f(np_max_f(out[k] + trace[j], w_min), w_max);
    }
  }
  return 0;
}

/* _cpu_csc_on_pre_update (dnn/linear.py:679-689), used by csc_on_post_update */
int orc_csc_on_post_update(const float* w, int64_t nnz, const int32_t* pre_ids, const int32_t* colptr,
                           const int32_t* w_ids, const uint8_t* spike, int64_t num_post, const float* trace,
                           int64_t num_pre, float w_min, float w_max, float* out) {
  for (int64_t k = 0; k < nnz; ++k) out[k] = w[k];
  for (int64_t i = 0; i < num_post; ++i) {
    if (!spike[i]) continue;
    for (int64_t k = colptr[i]; k < colptr[i + 1]; ++k) {
      int64_t j = pre_ids[k], l = w_ids[k];
      if (j < 0 || j >= num_pre || l < 0 || l >= nnz) return -1;
      out[l] = np_min_f(np_max_f(out[l] + trace[j], w_min), w_max);
    }
  }
  return 0;
}

/* _cpu_dense_on_pre / _cpu_dense_on_post (dnn/linear.py:238-244, 257-263):
 * np.clip of a row (pre spike) or column (post spike) of [num_pre, num_post] */
int orc_dense_on_spike(int rows, const float* w, const uint8_t* spike, int64_t num_pre, int64_t num_post,
                       const float* trace, float w_min, float w_max, float* out) {
  for (int64_t k = 0; k < num_pre * num_post; ++k) out[k] = w[k];
  for (int64_t r = 0; r < num_pre; ++r)
    for (int64_t c = 0; c < num_post; ++c) {
      if (!spike[rows ? r : c]) continue;
      float x = out[r * num_post + c] + trace[rows ? c : r];
      out[r * num_post + c] = np_min_f(np_max_f(x, w_min), w_max);
    }
  return 0;
}
