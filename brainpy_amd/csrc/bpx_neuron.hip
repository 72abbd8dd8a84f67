// bpx_neuron.hip -- fused LIF + Expon + COBA exponential-Euler step (f32).
//
// One thread per neuron reads V, t_last and its n_g conductances once,
// decays the conductances, integrates V, applies the refractory hold,
// thresholds, resets and writes V, t_last, spike and g back: 4*(2+n_g)+1
// bytes read and written per neuron, one pass instead of the several XLA
// elementwise passes of the reference step.
//
// Arithmetic order (one IEEE rounding per operation, no contraction):
//   g_k  = g_k + c_k * ((-g_k) / tau_k)                 Expon, exp_euler
//   I    = I_ext + sum_k g_k * (E_k - V)                COBA currents
//   f    = ((-V + V_rest) + R*I) / tau                  LIF derivative
//   L    = (-(1 + R*sum_k g_k)) / tau   (LifRefLTC)     d f / d V
//   V1   = V + (dt*exprel(dt*L)) * f                    exp_euler
//   V1   = (t - t_last) <= tau_ref ? V : V1             refractory hold
//   spike = V1 >= V_th; V = spike ? V_reset : V1; t_last = spike ? t : t_last
// (dyn/neurons/lif.py:254-256, 461-500, 550-556; dyn/synapses/
// abstract_models.py:121-134; dyn/outs/outputs.py:53-54;
// integrators/ode/exponential.py:357-362)
#include "bpx_common.h"
#include "bpx_math.h"

namespace bpx {

constexpr int kMaxG = 4;

struct LifArgs {
  int64_t n;
  int ltc;
  int n_g;
  float* V;
  float* t_last;
  uint8_t* spike;
  float* g[kMaxG];
  float tau_g[kMaxG];
  float E_g[kMaxG];
  float c_g[kMaxG];
  float I_ext;
  const float* I_vec;
  float V_rest, V_th, V_reset, tau, tau_ref, R, t, dt, cV;
  const int64_t* d_step;   // non-null: t = f32(*d_step + step) * dt, read on the device
  int64_t step;
};

__global__ __launch_bounds__(256) void k_lif_coba(LifArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  float gk[kMaxG];
  float gsum = 0.f;
#pragma unroll
  for (int k = 0; k < kMaxG; ++k) {
    if (k < a.n_g) {
      float gv = a.g[k][i];
      gv = gv + a.c_g[k] * ((-gv) / a.tau_g[k]);
      a.g[k][i] = gv;
      gk[k] = gv;
      gsum = gsum + gv;
    }
  }
  const float v = a.V[i];
  float I = a.I_vec ? a.I_vec[i] : a.I_ext;
#pragma unroll
  for (int k = 0; k < kMaxG; ++k)
    if (k < a.n_g) I = I + gk[k] * (a.E_g[k] - v);
  const float f = ((-v + a.V_rest) + a.R * I) / a.tau;
  float c;
  if (a.ltc) {
    const float L = (-(1.0f + a.R * gsum)) / a.tau;
    c = a.dt * exprel_det(a.dt * L);
  } else {
    c = a.cV;
  }
  float v1 = v + c * f;
  const float t = a.d_step ? (float)(*a.d_step + a.step) * a.dt : a.t;
  const float tl = a.t_last[i];
  if ((t - tl) <= a.tau_ref) v1 = v;
  const bool sp = v1 >= a.V_th;
  a.V[i] = sp ? a.V_reset : v1;
  if (sp) a.t_last[i] = t;
  a.spike[i] = sp ? 1 : 0;
}

}  // namespace bpx

using namespace bpx;

static int lif_step(void* stream, int64_t n, int ltc, int n_g, float* V, float* t_last, uint8_t* spike,
                    float* const* g, const float* tau_g, const float* E_g, const float* c_g, float I_ext,
                    const float* I_vec, float V_rest, float V_th, float V_reset, float tau, float tau_ref, float R,
                    float t, const int64_t* d_step, int64_t step, float dt) {
  if (n < 0) return BPX_ERR_SHAPE;
  if (n_g < 0 || n_g > kMaxG) return BPX_ERR_ARG;
  if (n_g > 0 && (!g || !tau_g || !E_g || !c_g)) return BPX_ERR_ARG;
  if (n == 0) return BPX_OK;
  if (!V || !t_last || !spike) return BPX_ERR_ARG;
  LifArgs a = {};
  a.n = n;
  a.ltc = ltc;
  a.n_g = n_g;
  a.V = V;
  a.t_last = t_last;
  a.spike = spike;
  for (int k = 0; k < n_g; ++k) {
    if (!g[k]) return BPX_ERR_ARG;
    a.g[k] = g[k];
    a.tau_g[k] = tau_g[k];
    a.E_g[k] = E_g[k];
    a.c_g[k] = c_g[k];
  }
  a.I_ext = I_ext;
  a.I_vec = I_vec;
  a.V_rest = V_rest;
  a.V_th = V_th;
  a.V_reset = V_reset;
  a.tau = tau;
  a.tau_ref = tau_ref;
  a.R = R;
  a.t = t;
  a.dt = dt;
  a.d_step = d_step;
  a.step = step;
  a.cV = dt * exprel_det(dt * (-(1.0f / tau)));
  hipStream_t s = (hipStream_t)stream;
  BPX_LAUNCH(KID_LIF, k_lif_coba, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
  BPX_CHECK_LAUNCH();
  return BPX_OK;
}

extern "C" int bpx_lif_coba_step(void* stream, int64_t n, int ltc, int n_g, float* V, float* t_last, uint8_t* spike,
                                 float* const* g, const float* tau_g, const float* E_g, const float* c_g, float I_ext,
                                 const float* I_vec, float V_rest, float V_th, float V_reset, float tau, float tau_ref,
                                 float R, float t, float dt) {
  return lif_step(stream, n, ltc, n_g, V, t_last, spike, g, tau_g, E_g, c_g, I_ext, I_vec, V_rest, V_th, V_reset,
                  tau, tau_ref, R, t, nullptr, 0, dt);
}

extern "C" int bpx_lif_coba_step_dev(void* stream, int64_t n, int ltc, int n_g, float* V, float* t_last,
                                     uint8_t* spike, float* const* g, const float* tau_g, const float* E_g,
                                     const float* c_g, float I_ext, const float* I_vec, float V_rest, float V_th,
                                     float V_reset, float tau, float tau_ref, float R, const int64_t* d_step,
                                     int64_t step, float dt) {
  if (!d_step) return BPX_ERR_ARG;
  return lif_step(stream, n, ltc, n_g, V, t_last, spike, g, tau_g, E_g, c_g, I_ext, I_vec, V_rest, V_th, V_reset,
                  tau, tau_ref, R, 0.f, d_step, step, dt);
}

extern "C" float bpx_expon_coef(float tau, float dt) {
  const float lin = -(1.0f / tau);
  return dt * exprel_det(dt * lin);
}
