// sin and cos of a phase-screen argument without quadrant selects.  Cody-Waite reduction by
// pi with a three-part float split of pi (fl(pi), then the float rounding of each remainder;
// k fl(pi) is exact for |k| < 2^21, so the first step is exact), r = x - k pi in [-pi/2, pi/2];
// then odd and even polynomials of degrees 9 and 10 in r, and the sign (-1)^k applied to both
// as one exclusive-or of the sign bit.  The coefficients are a weighted least-squares
// (near-minimax) fit on [-pi/2, pi/2] with the leading terms fixed at r and 1 - r^2/2; their
// truncation error there is below 5e-9, and evaluated in float with fused multiply-adds the
// result is within 1.5e-7 of sin and cos (rms 3.3e-8) over |x| < 2^17.  That is the domain
// (kScreenPhaseLimit); the callers' host code rejects screens that reach it.  The polynomial
// form costs 19 VALU operations per argument against 31 for a pi/2 reduction with its
// quadrant selects.
#pragma once
#include <hip/hip_runtime.h>

namespace trig {

constexpr float kScreenPhaseLimit = 131072.0f;

// Below this magnitude the nearest multiple of pi is 0 (x / pi rounds below 0.5), the reduction
// leaves x unchanged and the sign is +: sincos_reduced(x) is then bit for bit sincos_screen(x).
constexpr float kNoReduction = 1.5707f;

// The polynomials of sincos_screen on a reduced argument r in [-pi/2, pi/2].
__host__ __device__ __forceinline__ void sincos_reduced(float r, float* s, float* c) {
  const float r2 = r * r;
  float ps = fmaf(r2, 2.6000243e-06f, -1.9806602e-04f);
  ps = fmaf(r2, ps, 8.333017e-03f);
  ps = fmaf(r2, ps, -1.6666657e-01f);
  *s = fmaf(r * r2, ps, r);
  float pc = fmaf(r2, -2.619355e-07f, 2.4769291e-05f);
  pc = fmaf(r2, pc, -1.3888569e-03f);
  pc = fmaf(r2, pc, 4.1666657e-02f);
  pc = fmaf(r2, pc, -0.5f);
  *c = fmaf(r2, pc, 1.0f);
}

__host__ __device__ __forceinline__ void sincos_screen(float x, float* s, float* c) {
  const float k = rintf(x * 3.183098733e-01f);  // nearest multiple of pi
  float r = fmaf(-k, 3.141592741e+00f, x);
  r = fmaf(-k, -8.742277657e-08f, r);
  r = fmaf(-k, -3.552713679e-15f, r);
  float sr, cr;
  sincos_reduced(r, &sr, &cr);
  // sin(r + k pi) = (-1)^k sin r, cos(r + k pi) = (-1)^k cos r
  const unsigned sign = static_cast<unsigned>(static_cast<int>(k)) << 31;
  *s = __uint_as_float(__float_as_uint(sr) ^ sign);
  *c = __uint_as_float(__float_as_uint(cr) ^ sign);
}

}  // namespace trig
