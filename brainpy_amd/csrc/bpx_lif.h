// bpx_lif.h -- the LIF + Expon + COBA exponential-Euler update of one neuron
// (f32, one IEEE rounding per operation, no contraction), shared by the
// per-population step and the batched network step.
//
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
#pragma once
#include "bpx_math.h"

namespace bpx {

struct LifParams {
  float V_rest, V_th, V_reset, tau, tau_ref, R;
  float cV;   // dt*exprel(-dt/tau), used when ltc == 0 (LifRef)
  int ltc;
};

// g[0..NG) are updated in place; returns the spike
template <int NG>
__device__ __forceinline__ bool lif_update(float& V, float& t_last, float* g, const float* tau_g, const float* E_g,
                                           const float* c_g, float I_ext, const LifParams& p, float t, float dt) {
  float gsum = 0.f;
#pragma unroll
  for (int k = 0; k < NG; ++k) {
    g[k] = g[k] + c_g[k] * ((-g[k]) / tau_g[k]);
    gsum = gsum + g[k];
  }
  const float v = V;
  float I = I_ext;
#pragma unroll
  for (int k = 0; k < NG; ++k) I = I + g[k] * (E_g[k] - v);
  const float f = ((-v + p.V_rest) + p.R * I) / p.tau;
  float c;
  if (p.ltc) {
    const float L = (-(1.0f + p.R * gsum)) / p.tau;
    c = dt * exprel_det(dt * L);
  } else {
    c = p.cV;
  }
  float v1 = v + c * f;
  if ((t - t_last) <= p.tau_ref) v1 = v;
  const bool sp = v1 >= p.V_th;
  V = sp ? p.V_reset : v1;
  if (sp) t_last = t;
  return sp;
}

}  // namespace bpx
