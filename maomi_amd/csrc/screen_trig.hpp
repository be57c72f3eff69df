// sin and cos of a phase-screen argument.  Cody-Waite reduction by pi/2 with a
// three-part float split of pi/2 (about 72 bits, exact products through fmaf),
// then the Taylor series of sin (to r^9) and cos (to r^10) on [-pi/4, pi/4],
// whose truncation error there is below 3e-9; about 1.2 ulp of 1 overall.
// Domain |x| < 2^17 (kScreenPhaseLimit); the callers' host code rejects screens
// that reach it.  Without the general-argument reduction of sincosf the 16
// evaluations a thread makes per slice stay cheap in registers and code.
#pragma once
#include <hip/hip_runtime.h>

namespace trig {

constexpr float kScreenPhaseLimit = 131072.0f;

__host__ __device__ __forceinline__ void sincos_screen(float x, float* s, float* c) {
  const float k = rintf(x * 6.366197467e-01f);  // nearest multiple of pi/2
  float r = fmaf(-k, 1.570796371e+00f, x);
  r = fmaf(-k, -4.371138829e-08f, r);
  r = fmaf(-k, -1.715124510e-15f, r);
  const float r2 = r * r;
  // 1/3!, 1/5!, 1/7!, 1/9! and 1/2!, 1/4!, ..., 1/10!
  float ps = fmaf(r2, 2.755731884e-06f, -1.984127011e-04f);
  ps = fmaf(r2, ps, 8.333333768e-03f);
  ps = fmaf(r2, ps, -1.666666716e-01f);
  const float sr = fmaf(r * r2, ps, r);
  float pc = fmaf(r2, -2.755731998e-07f, 2.480158764e-05f);
  pc = fmaf(r2, pc, -1.388888923e-03f);
  pc = fmaf(r2, pc, 4.166666791e-02f);
  pc = fmaf(r2, pc, -0.5f);
  const float cr = fmaf(r2, pc, 1.0f);
  // x = q pi/2 + r: sin x = (sin r, cos r, -sin r, -cos r)[q], cos x = (cos r, -sin r, -cos r, sin r)[q]
  const int q = static_cast<int>(k) & 3;
  const float sv = (q & 1) ? cr : sr, cv = (q & 1) ? sr : cr;
  *s = (q & 2) ? -sv : sv;
  *c = ((q + 1) & 2) ? -cv : cv;
}

}  // namespace trig
