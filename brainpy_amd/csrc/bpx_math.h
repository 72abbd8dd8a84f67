// bpx_math.h -- deterministic float32 exp/exprel for the neuron update.
//
// exp(x) by Cody-Waite reduction (ln2 split into an 11-bit head and a tail)
// and the degree-7 Taylor polynomial of e^r in Horner form with explicit
// fmaf.  fmaf, rintf and ldexpf are correctly rounded on host and device, so
// the same source gives bit-identical results in the CPU check and on gfx950.
// Max error on the neuron-update range is below 2 ulp.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define BPX_MHD __host__ __device__ __forceinline__
#else
#define BPX_MHD inline
#endif

namespace bpx {

BPX_MHD float expf_det(float x) {
  if (x > 88.7228393554688f) return INFINITY;
  if (x < -103.972084045410f) return 0.0f;
  float n = rintf(x * 1.44269504088896341f);
  float r = fmaf(-n, 0.693145751953125f, x);
  r = fmaf(-n, 1.428606765330187045e-06f, r);
  float p = 1.98412698412698413e-04f;
  p = fmaf(p, r, 1.38888888888888889e-03f);
  p = fmaf(p, r, 8.33333333333333333e-03f);
  p = fmaf(p, r, 4.16666666666666667e-02f);
  p = fmaf(p, r, 1.66666666666666667e-01f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)n);
}

// exprel(z) = (e^z - 1)/z with the series 1 + z/2 + z^2/6 for |z| <= 1e-5
// (brainpy/_src/math/others.py:103-126, f32 threshold)
BPX_MHD float exprel_det(float z) {
  float az = z < 0.f ? -z : z;
  if (az <= 1e-5f) {
    float h = z / 2.0f;
    float s = 1.0f + h;
    float q = (z * z) / 6.0f;
    return s + q;
  }
  return (expf_det(z) - 1.0f) / z;
}

}  // namespace bpx
