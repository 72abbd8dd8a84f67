// bpx_net.hip -- batched E/I COBA network step: many independent networks
// advanced by one launch per step.
//
// Per launch (step i), every block of T threads owns T neurons of one
// network (a tile from the host-built tile table; T = 64 ... 1024, chosen
// per batch by the host-side autotuner):
//   1. update: for each neuron, g_exc += S(c_exc), g_inh += S(c_inh), where
//      c_* are the numbers of pre-synaptic spikes of step i-1 delivered to it
//      and S(c) = ((0 + w) + w) + ... (c terms) is exactly the homogeneous
//      event csrmv / jitconn output of the reference (`out[row] += w` into a
//      zeroed output, then `g += out`: dyn/synapses/abstract_models.py:
//      133-134); the counts are zeroed for reuse; then the fused LIF/Expon/
//      COBA update (bpx_lif.h) in the reference order.
//   2. deliver: every neuron that spiked in step i regenerates or reads its
//      targets and adds 1 to their count for step i+D, D the projection's
//      delay in steps (>= 1; integer atomics into that step's count buffer).  Connectivity is either just-in-time (the
//      lane32 scatter form of jitconn/_event_matvec_taichi.py:92-123,
//      outdim_parallel=False: 32 lanes per spiking neuron, stream key
//      seed + pre*32 + lane) or a stored CSR per network.
// Integer counts make every conductance update exact and independent of
// scheduling, so a batched run, a network run alone, and the serial CPU
// restatement give the same spike raster.
// A ring of n_buf count buffers removes the grid-wide barrier between the
// two phases: step i reads/zeroes buffer i % n_buf and fills buffers
// (i + D) % n_buf, D < n_buf.
#include "bpx_common.h"
#include "bpx_lif.h"
#include "bpx_rng.h"

namespace bpx {

__device__ __forceinline__ float repeat_sum(uint32_t c, float w) {
  float s = 0.f;
  for (uint32_t k = 0; k < c; ++k) s = s + w;
  return s;
}

struct NetArgs {
  const bpx_net_desc* nets;
  const int32_t* tiles;        // pairs (net, first local neuron)
  int64_t total;               // neurons in the pools
  float* V;
  float* t_last;
  float* g_exc;
  float* g_inh;
  uint32_t* counts;            // [n_buf buffers][2 kinds][total]
  int n_buf;                   // count buffers: step i reads buffer i % n_buf
  uint8_t* spikes;             // [total] spike flags of this step (nullable)
  unsigned long long* events;  // [n_nets] delivered synaptic events (nullable)
  unsigned int* spike_count;   // [n_nets] spikes (nullable)
  LifParams p;
  float tau_e, tau_i, E_e, E_i, c_e, c_i;
  float dt;
  int64_t step;                // step index, or offset from *d_step
  const int64_t* d_step;       // device base step (graph replays) or null
};

template <int CONN, int T>
__global__ __launch_bounds__(T) void k_net_step(NetArgs a) {
  __shared__ int s_spk[T];
  __shared__ int s_nspk;
  __shared__ unsigned long long s_events;
  const int net = a.tiles[2 * blockIdx.x];
  const int start = a.tiles[2 * blockIdx.x + 1];
  const bpx_net_desc d = a.nets[net];
  const int N = d.n_exc + d.n_inh;
  if (threadIdx.x == 0) {
    s_nspk = 0;
    s_events = 0ull;
  }
  __syncthreads();
  const int j = start + (int)threadIdx.x;
  const int64_t total = a.total;
  const int64_t step = a.d_step ? *a.d_step + a.step : a.step;
  const float t = (float)(step - d.t0_step) * a.dt;      // network-local time
  const int64_t nb = a.n_buf;
  uint32_t* cur = a.counts + (step % nb) * 2 * total;
  // spikes of projection k (E2E, E2I, I2E, I2I) with delay D_k steps are
  // counted into the buffer step i + D_k reads (1 <= D_k < n_buf)
  auto buf_of = [&](int k) -> uint32_t* {
    int64_t dk = d.delay[k];
    dk = dk < 1 ? 1 : (dk > nb - 1 ? nb - 1 : dk);
    return a.counts + ((step + dk) % nb) * 2 * total;
  };
  if (j < N) {
    const int64_t gid = d.off + j;
    const uint32_t ce = cur[gid], ci = cur[total + gid];
    cur[gid] = 0u;
    cur[total + gid] = 0u;
    float g[2];
    g[0] = a.g_exc[gid];
    g[1] = a.g_inh[gid];
    if (ce) g[0] = g[0] + repeat_sum(ce, d.w_exc);
    if (ci) g[1] = g[1] + repeat_sum(ci, d.w_inh);
    const float tau_g[2] = {a.tau_e, a.tau_i};
    const float E_g[2] = {a.E_e, a.E_i};
    const float c_g[2] = {a.c_e, a.c_i};
    float V = a.V[gid], tl = a.t_last[gid];
    const bool sp = lif_update<2>(V, tl, g, tau_g, E_g, c_g, d.I_ext, a.p, t, a.dt);
    a.V[gid] = V;
    a.t_last[gid] = tl;
    a.g_exc[gid] = g[0];
    a.g_inh[gid] = g[1];
    if (a.spikes) a.spikes[gid] = sp ? 1 : 0;
    if (sp) s_spk[atomicAdd(&s_nspk, 1)] = j;
  }
  __syncthreads();
  const int nspk = s_nspk;
  // deliver: T/32 groups of 32 lanes, one spiking neuron per group
  const int grp = threadIdx.x >> 5;
  const uint32_t l = threadIdx.x & 31;
  unsigned long long ev = 0ull;
  for (int e = grp; e < nspk; e += T / 32) {
    const int p = s_spk[e];
    const bool exc = p < d.n_exc;
    const int64_t kind = exc ? 0 : total;                // excitatory or inhibitory conductance
    if constexpr (CONN == BPX_NET_JIT) {
      const int col = exc ? p : p - d.n_exc;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        // projections from this population: to E (post offset 0), to I (offset n_exc)
        const int k = exc ? q : 2 + q;                   // E2E, E2I, I2E, I2I
        uint32_t* cnt = buf_of(k) + kind + d.off;        // post ids are network-local
        const int64_t num_row = q == 0 ? d.n_exc : d.n_inh;
        const int post_off = q == 0 ? 0 : d.n_exc;
        if (num_row <= 0) continue;
        const DivU32 dv = {d.div[k][0], d.div[k][1], d.div[k][2], d.div[k][3]};
        const uint32_t step = (uint32_t)(((num_row + 1) >> 5) > 1 ? ((num_row + 1) >> 5) : 1);
        uint32_t r = step * l - 1u;
        uint32_t end = r + step;
        if ((int64_t)end > num_row) end = (uint32_t)num_row;
        Lfsr88 key(d.seed[k] + (uint32_t)(col * 32) + l);
        r += (uint32_t)rand_inc(key, dv);
        while (r < end) {
          atomicAdd(cnt + post_off + r, 1u);
          ++ev;
          r += (uint32_t)rand_inc(key, dv);
        }
      }
    } else {
      // targets in E belong to E2E / I2E, targets in I to E2I / I2I
      uint32_t* cnt_e = buf_of(exc ? 0 : 2) + kind + d.off;
      uint32_t* cnt_i = buf_of(exc ? 1 : 3) + kind + d.off;
      const int32_t s = d.indptr[p], t2 = d.indptr[p + 1];
      for (int32_t q = s + (int32_t)l; q < t2; q += 32) {
        const int32_t tgt = d.indices[q];
        if ((uint32_t)tgt < (uint32_t)N) {
          atomicAdd((tgt < d.n_exc ? cnt_e : cnt_i) + tgt, 1u);
          ++ev;
        }
      }
    }
  }
  if (nspk && (a.events || a.spike_count)) {
    // block totals, one global atomic each
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ev += __shfl_xor(ev, o, 64);
    if ((threadIdx.x & 63) == 0 && ev) atomicAdd(&s_events, ev);
    __syncthreads();
    if (threadIdx.x == 0) {
      if (a.events) atomicAdd(a.events + net, s_events);
      if (a.spike_count) atomicAdd(a.spike_count + net, (unsigned int)nspk);
    }
  }
}

// Multi-step form: one block of kRunThreads threads per network advances it
// K steps in one launch.  Each thread keeps kRunNpt neurons' state
// (V, t_last, g_exc, g_inh) in registers; the count ring lives in LDS with
// the excitatory and inhibitory counts of a neuron packed in one word (low
// and high 16 bits; exact while no neuron receives 65536 synapses of one
// kind, which the host checks), so delivery is LDS atomics and the two
// phases of a step are separated by block barriers instead of launches.
// The same integer counts and the same update as k_net_step: the spike
// raster, state and counters equal K launches of the step kernel.
constexpr int kRunThreads = 1024;
constexpr int kRunNpt = 8;                 // networks up to 8192 neurons

template <int CONN>
__global__ __launch_bounds__(kRunThreads) void k_net_run(NetArgs a, uint8_t* raster, int64_t n_steps, int cap,
                                                        int32_t* status) {
  extern __shared__ uint32_t s_mem[];
  __shared__ int s_nspk[2];
  __shared__ unsigned long long s_events;
  const int nb = a.n_buf;
  uint32_t* s_cnt = s_mem;                               // [n_buf][cap] packed counts
  int* s_spk = (int*)(s_mem + (int64_t)nb * cap);        // [cap] this step's spiking neurons
  const int net = blockIdx.x;
  const bpx_net_desc d = a.nets[net];
  const int N = d.n_exc + d.n_inh;
  if (N > cap) {                                         // more neurons than the caller's max_neurons:
    if (status && threadIdx.x == 0) atomicExch(status, (int32_t)BPX_ERR_SHAPE);   // left unchanged, reported
    return;
  }
  const int64_t total = a.total;
  const int tid = threadIdx.x;
  for (int b = 0; b < nb; ++b) {
    const uint32_t* src = a.counts + (int64_t)b * 2 * total + d.off;
    for (int j = tid; j < N; j += kRunThreads) s_cnt[b * cap + j] = src[j] | (src[total + j] << 16);
  }
  if (tid == 0) {
    s_nspk[0] = 0;
    s_nspk[1] = 0;
    s_events = 0ull;
  }
  float V[kRunNpt], tl[kRunNpt], ge[kRunNpt], gi[kRunNpt];
#pragma unroll
  for (int u = 0; u < kRunNpt; ++u) {
    const int j = tid + u * kRunThreads;
    if (j < N) {
      V[u] = a.V[d.off + j];
      tl[u] = a.t_last[d.off + j];
      ge[u] = a.g_exc[d.off + j];
      gi[u] = a.g_inh[d.off + j];
    }
  }
  int dly[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) dly[k] = d.delay[k] < 1 ? 1 : (d.delay[k] > nb - 1 ? nb - 1 : d.delay[k]);
  const float tau_g[2] = {a.tau_e, a.tau_i};
  const float E_g[2] = {a.E_e, a.E_i};
  const float c_g[2] = {a.c_e, a.c_i};
  const int64_t step0 = a.d_step ? *a.d_step + a.step : a.step;
  const int grp = tid >> 5;
  const uint32_t l = tid & 31;
  unsigned long long ev = 0ull;
  unsigned int spk_total = 0u;
  __syncthreads();
  for (int64_t k = 0; k < n_steps; ++k) {
    const int64_t step = step0 + k;
    const float t = (float)(step - d.t0_step) * a.dt;
    uint32_t* cur = s_cnt + (step % nb) * cap;
    int* nsp = s_nspk + (k & 1);
    uint8_t* rs = raster ? raster + k * total + d.off : nullptr;
    const bool last = k == n_steps - 1;
#pragma unroll
    for (int u = 0; u < kRunNpt; ++u) {
      const int j = tid + u * kRunThreads;
      if (j < N) {
        const uint32_t c = cur[j];
        cur[j] = 0u;
        float g[2] = {ge[u], gi[u]};
        if (c & 0xffffu) g[0] = g[0] + repeat_sum(c & 0xffffu, d.w_exc);
        if (c >> 16) g[1] = g[1] + repeat_sum(c >> 16, d.w_inh);
        const bool sp = lif_update<2>(V[u], tl[u], g, tau_g, E_g, c_g, d.I_ext, a.p, t, a.dt);
        ge[u] = g[0];
        gi[u] = g[1];
        if (rs) rs[j] = sp ? 1 : 0;
        if (last && a.spikes) a.spikes[d.off + j] = sp ? 1 : 0;
        if (sp) s_spk[atomicAdd(nsp, 1)] = j;
      }
    }
    __syncthreads();
    const int nspk = *nsp;
    if (tid == 0) {
      s_nspk[(k + 1) & 1] = 0;                           // last read before this step's first barrier
      spk_total += (unsigned int)nspk;
    }
    for (int e = grp; e < nspk; e += kRunThreads / 32) {
      const int p = s_spk[e];
      const bool exc = p < d.n_exc;
      const uint32_t inc = exc ? 1u : 0x10000u;          // excitatory or inhibitory conductance
      if constexpr (CONN == BPX_NET_JIT) {
        const int col = exc ? p : p - d.n_exc;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int kk = exc ? q : 2 + q;                // E2E, E2I, I2E, I2I
          const int64_t num_row = q == 0 ? d.n_exc : d.n_inh;
          if (num_row <= 0) continue;
          uint32_t* cnt = s_cnt + ((step + dly[kk]) % nb) * cap + (q == 0 ? 0 : d.n_exc);
          const DivU32 dv = {d.div[kk][0], d.div[kk][1], d.div[kk][2], d.div[kk][3]};
          const uint32_t rstep = (uint32_t)(((num_row + 1) >> 5) > 1 ? ((num_row + 1) >> 5) : 1);
          uint32_t r = rstep * l - 1u;
          uint32_t end = r + rstep;
          if ((int64_t)end > num_row) end = (uint32_t)num_row;
          Lfsr88 key(d.seed[kk] + (uint32_t)(col * 32) + l);
          r += (uint32_t)rand_inc(key, dv);
          while (r < end) {
            atomicAdd(cnt + r, inc);
            ++ev;
            r += (uint32_t)rand_inc(key, dv);
          }
        }
      } else {
        uint32_t* cnt_e = s_cnt + ((step + dly[exc ? 0 : 2]) % nb) * cap;
        uint32_t* cnt_i = s_cnt + ((step + dly[exc ? 1 : 3]) % nb) * cap;
        const int32_t s = d.indptr[p], t2 = d.indptr[p + 1];
        for (int32_t q = s + (int32_t)l; q < t2; q += 32) {
          const int32_t tgt = d.indices[q];
          if ((uint32_t)tgt < (uint32_t)N) {
            atomicAdd((tgt < d.n_exc ? cnt_e : cnt_i) + tgt, inc);
            ++ev;
          }
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < kRunNpt; ++u) {
    const int j = tid + u * kRunThreads;
    if (j < N) {
      a.V[d.off + j] = V[u];
      a.t_last[d.off + j] = tl[u];
      a.g_exc[d.off + j] = ge[u];
      a.g_inh[d.off + j] = gi[u];
    }
  }
  for (int b = 0; b < nb; ++b) {
    uint32_t* dst = a.counts + (int64_t)b * 2 * total + d.off;
    for (int j = tid; j < N; j += kRunThreads) {
      const uint32_t c = s_cnt[b * cap + j];
      dst[j] = c & 0xffffu;
      dst[total + j] = c >> 16;
    }
  }
  if (a.events || a.spike_count) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ev += __shfl_xor(ev, o, 64);
    if ((tid & 63) == 0 && ev) atomicAdd(&s_events, ev);
    __syncthreads();
    if (tid == 0) {
      if (a.events && s_events) atomicAdd(a.events + net, s_events);
      if (a.spike_count && spk_total) atomicAdd(a.spike_count + net, spk_total);
    }
  }
}

__global__ void k_net_advance(int64_t* d_step, int64_t delta) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *d_step += delta;
}

}  // namespace bpx

using namespace bpx;

static NetArgs net_args(const bpx_net_desc* nets, const int32_t* tiles, int64_t total, float* V, float* t_last,
                        float* g_exc, float* g_inh, uint32_t* counts, int n_buf, uint8_t* spikes,
                        unsigned long long* events, unsigned int* spike_count, const bpx_net_params* prm,
                        int64_t step, const int64_t* d_step) {
  NetArgs a;
  a.nets = nets;
  a.tiles = tiles;
  a.total = total;
  a.V = V;
  a.t_last = t_last;
  a.g_exc = g_exc;
  a.g_inh = g_inh;
  a.counts = counts;
  a.n_buf = n_buf;
  a.spikes = spikes;
  a.events = events;
  a.spike_count = spike_count;
  a.p.V_rest = prm->V_rest;
  a.p.V_th = prm->V_th;
  a.p.V_reset = prm->V_reset;
  a.p.tau = prm->tau;
  a.p.tau_ref = prm->tau_ref;
  a.p.R = prm->R;
  a.p.ltc = prm->ltc;
  a.p.cV = prm->dt * exprel_det(prm->dt * (-(1.0f / prm->tau)));
  a.tau_e = prm->tau_e;
  a.tau_i = prm->tau_i;
  a.E_e = prm->E_e;
  a.E_i = prm->E_i;
  a.c_e = prm->c_e;
  a.c_i = prm->c_i;
  a.dt = prm->dt;
  a.step = step;
  a.d_step = d_step;
  return a;
}

// LDS words of the multi-step kernel per network neuron: n_buf packed count
// words and one spike-list entry; the block stays within 64 KiB
static int64_t run_max_neurons(int n_buf) {
  const int64_t lim = (65536 - 64) / (4 * ((int64_t)n_buf + 1));
  return lim < (int64_t)kRunNpt * kRunThreads ? lim : (int64_t)kRunNpt * kRunThreads;
}

extern "C" {

void bpx_net_prepare(bpx_net_desc* d) {
  // divisor magic of each projection's clen (jitconn)
  for (int k = 0; k < 4; ++k) {
    const uint32_t c = d->clen[k] > 0 ? (uint32_t)d->clen[k] : 1u;
    const DivU32 dv = DivU32::make(c);
    d->div[k][0] = dv.d;
    d->div[k][1] = dv.m;
    d->div[k][2] = dv.s1;
    d->div[k][3] = dv.s2;
  }
}

int bpx_net_step(void* stream, const bpx_net_desc* nets, const int32_t* tiles, int n_tiles, int tile_size,
                 int conn_kind, int64_t total, float* V, float* t_last, float* g_exc, float* g_inh, uint32_t* counts, int n_buf,
                 uint8_t* spikes, unsigned long long* events, unsigned int* spike_count,
                 const bpx_net_params* prm, int64_t step, const int64_t* d_step) {
  if (!nets || !tiles || !prm) return BPX_ERR_ARG;
  if (n_buf < 2 || (!d_step && step < 0)) return BPX_ERR_ARG;
  if (n_tiles < 0 || total < 0) return BPX_ERR_SHAPE;
  if (conn_kind != BPX_NET_JIT && conn_kind != BPX_NET_CSR) return BPX_ERR_ARG;
  if (tile_size != 64 && tile_size != 128 && tile_size != 256 && tile_size != 512 && tile_size != 1024)
    return BPX_ERR_ARG;
  if (n_tiles == 0) return BPX_OK;
  if (!V || !t_last || !g_exc || !g_inh || !counts) return BPX_ERR_ARG;
  const NetArgs a = net_args(nets, tiles, total, V, t_last, g_exc, g_inh, counts, n_buf, spikes, events,
                             spike_count, prm, step, d_step);
  hipStream_t s = (hipStream_t)stream;
#define NET_GO(T_)                                                                                     \
  do {                                                                                                 \
    if (conn_kind == BPX_NET_JIT)                                                                      \
      BPX_LAUNCH(KID_NET, (k_net_step<BPX_NET_JIT, T_>), dim3((unsigned)n_tiles), dim3(T_), 0, s, a);  \
    else                                                                                               \
      BPX_LAUNCH(KID_NET, (k_net_step<BPX_NET_CSR, T_>), dim3((unsigned)n_tiles), dim3(T_), 0, s, a);  \
  } while (0)
  switch (tile_size) {
    case 64: NET_GO(64); break;
    case 128: NET_GO(128); break;
    case 256: NET_GO(256); break;
    case 512: NET_GO(512); break;
    default: NET_GO(1024); break;
  }
#undef NET_GO
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

int64_t bpx_net_run_max_neurons(int n_buf) { return n_buf < 2 ? 0 : run_max_neurons(n_buf); }

int bpx_net_run(void* stream, const bpx_net_desc* nets, int n_nets, int max_neurons, int max_in_degree,
                int conn_kind, int64_t total, float* V, float* t_last, float* g_exc, float* g_inh, uint32_t* counts,
                int n_buf, uint8_t* raster, uint8_t* spikes, unsigned long long* events, unsigned int* spike_count,
                const bpx_net_params* prm, int64_t step, const int64_t* d_step, int64_t n_steps,
                int32_t* status) {
  if (!nets || !prm) return BPX_ERR_ARG;
  if (n_buf < 2 || (!d_step && step < 0) || n_steps < 0) return BPX_ERR_ARG;
  if (n_nets < 0 || total < 0) return BPX_ERR_SHAPE;
  if (conn_kind != BPX_NET_JIT && conn_kind != BPX_NET_CSR) return BPX_ERR_ARG;
  if (n_nets == 0 || n_steps == 0) return BPX_OK;
  if (!V || !t_last || !g_exc || !g_inh || !counts) return BPX_ERR_ARG;
  if (max_neurons < 1 || max_neurons > run_max_neurons(n_buf)) return BPX_ERR_SHAPE;
  if (max_in_degree < 0 || max_in_degree > 65535) return BPX_ERR_SHAPE;
  const NetArgs a = net_args(nets, nullptr, total, V, t_last, g_exc, g_inh, counts, n_buf, spikes, events,
                             spike_count, prm, step, d_step);
  const size_t shm = (size_t)4 * ((size_t)n_buf + 1) * (size_t)max_neurons;
  hipStream_t s = (hipStream_t)stream;
  if (conn_kind == BPX_NET_JIT)
    BPX_LAUNCH(KID_NET_RUN, (k_net_run<BPX_NET_JIT>), dim3((unsigned)n_nets), dim3(kRunThreads), shm, s, a, raster,
               n_steps, max_neurons, status);
  else
    BPX_LAUNCH(KID_NET_RUN, (k_net_run<BPX_NET_CSR>), dim3((unsigned)n_nets), dim3(kRunThreads), shm, s, a, raster,
               n_steps, max_neurons, status);
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

int bpx_net_advance(void* stream, int64_t* d_step, int64_t delta) {
  if (!d_step) return BPX_ERR_ARG;
  hipLaunchKernelGGL(k_net_advance, dim3(1), dim3(64), 0, (hipStream_t)stream, d_step, delta);
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

}  // extern "C"
