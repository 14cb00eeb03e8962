// Device math shared by the beamformer (rca_das.hip) and the simulator
// (rca_sim.hip).  Header-only; every function is internal to its
// translation unit.
#ifndef RCA_MATH_H
#define RCA_MATH_H

#include <hip/hip_runtime.h>

namespace rca_math {

// sqrt in FP64 from the hardware reciprocal square root and a Newton pair
// (Goldschmidt form); x > 0.
__device__ __forceinline__ double sqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  const double d = fma(-g, g, x);
  return fma(d, h, g);
}

// sin/cos of 2*pi*p for p in [-1/2, 1/2], branch-free: sin/cos of the half
// angle pi*p (|pi p| <= pi/2, Taylor to t^13 / t^14, |error| < 1e-9), then the
// double-angle identities.
__device__ __forceinline__ void sincos2pi(float p, float& s, float& c) {
  const float t = p * 3.14159265358979f;
  const float t2 = t * t;
  float sh = fmaf(t2, -1.60590438e-10f, 2.50521084e-8f);
  sh = fmaf(t2, sh, -2.75573192e-6f);
  sh = fmaf(t2, sh, 1.98412698e-4f);
  sh = fmaf(t2, sh, -8.33333333e-3f);
  sh = fmaf(t2, sh, 1.66666667e-1f);
  sh = t * fmaf(-t2, sh, 1.0f);
  float ch = fmaf(t2, 4.77947733e-14f, -1.14707456e-11f);
  ch = fmaf(t2, ch, 2.08767570e-9f);
  ch = fmaf(t2, ch, -2.75573192e-7f);
  ch = fmaf(t2, ch, 2.48015873e-5f);
  ch = fmaf(t2, ch, -1.38888889e-3f);
  ch = fmaf(t2, ch, 4.16666667e-2f);
  ch = fmaf(t2, ch, -0.5f);
  ch = fmaf(t2, ch, 1.0f);
  s = 2.0f * sh * ch;
  c = fmaf(-2.0f * sh, sh, 1.0f);
}

// The same with the half-angle series one term shorter (sine to t^11, cosine
// to t^12): truncation below 6e-8 (sine) and 7e-9 (cosine) at |t| = pi/2,
// about one FP32 ulp after the double angle.  Three FMAs fewer, for kernels
// that evaluate it per term (the simulator).
__device__ __forceinline__ void sincos2pi_lean(float p, float& s, float& c) {
  const float t = p * 3.14159265358979f;
  const float t2 = t * t;
  float sh = fmaf(t2, 2.50521084e-8f, -2.75573192e-6f);
  sh = fmaf(t2, sh, 1.98412698e-4f);
  sh = fmaf(t2, sh, -8.33333333e-3f);
  sh = fmaf(t2, sh, 1.66666667e-1f);
  sh = t * fmaf(-t2, sh, 1.0f);
  float ch = fmaf(t2, 2.08767570e-9f, -2.75573192e-7f);
  ch = fmaf(t2, ch, 2.48015873e-5f);
  ch = fmaf(t2, ch, -1.38888889e-3f);
  ch = fmaf(t2, ch, 4.16666667e-2f);
  ch = fmaf(t2, ch, -0.5f);
  ch = fmaf(t2, ch, 1.0f);
  s = 2.0f * sh * ch;
  c = fmaf(-2.0f * sh, sh, 1.0f);
}

}  // namespace rca_math

#endif  // RCA_MATH_H
