// lh_math.h -- scalar math of the liquid step, compiled for gfx950 device code.
//
// Every function restates one reference routine with the reference's FP32
// operation order; the library is compiled with -ffp-contract=off so no FMA is
// formed (the reference build, MSVC /fp:precise, forms none).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LH_HD __host__ __device__ __forceinline__

namespace lh {

// Kernel.h:12-17 (quadratic B-spline weight)
// Both pieces are evaluated and one is selected: the same operations as the branches, without
// the divergent branch (and its exec-mask bookkeeping) per call in the particle loops.  1.5 - d
// is positive exactly when d < 1.5, so clamping it at +0 gives the outer piece below 1.5 and
// +0 from 1.5 on, as the reference's third branch does.  The empty asm keeps the compiler from
// sinking either piece back into a branch.
LH_HD float quad_w(float d) {
  d = fabsf(d);
  float inner = 0.75f - d * d;
  const float t = fmaxf(1.5f - d, 0.0f);
  float outer = 0.5f * t * t;
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(inner), "+v"(outer));
#endif
  return d < 0.5f ? inner : outer;
}
// Kernel.h:26-31 (derivative w.r.t. |d|, unsigned as in the reference); d - 1.5 is negative
// exactly when d < 1.5
LH_HD float quad_gw(float d) {
  d = fabsf(d);
  return d < 0.5f ? -2.f * d : fminf(d - 1.5f, 0.0f);
}

constexpr float kEps = 1e-8f;  // Definition.h:93 EPSILON
LH_HD bool is_small(float v) { return fabsf(v) < kEps; }  // MathUtils.h:8
// MathUtils.h:10 with the windows.h min/max macros
LH_HD float clampf(float v, float lo, float hi) {
  float t = (lo > v) ? lo : v;
  return (hi < t) ? hi : t;
}
// MathUtils.h:12-34
LH_HD float lerp(float a, float b, float f) { return a + clampf(f, 0.0f, 1.0f) * (b - a); }
LH_HD float bilerp(float v00, float v10, float v01, float v11, float f1, float f2) {
  float a = lerp(v00, v10, f1);
  float b = lerp(v01, v11, f1);
  return lerp(a, b, clampf(f2, 0.0f, 1.0f));
}
LH_HD float trilerp(float v000, float v100, float v010, float v110, float v001, float v101, float v011, float v111,
                    float f1, float f2, float f3) {
  return lerp(bilerp(v000, v100, v010, v110, f1, f2), bilerp(v001, v101, v011, v111, f1, f2),
              clampf(f3, 0.0f, 1.0f));
}
// MathUtils.cpp:29-35
LH_HD float fraction_in_region(float l, float r) {
  float ans = 0.0f;
  if (l < 0.0f && r < 0.0f) ans = 1.0f;
  else if (l < 0.0f && 0.0f < r) ans = l / (l - r);
  else if (l > 0.0f && 0.0f > r) ans = r / (r - l);
  return ans;
}
// max(0.1, frac) through the windows.h macro: the comparison happens in double
// and the result is stored in a float (PressureSolver.cpp:569, LiquidModel.cpp:845).
LH_HD float max_tenth(float frac) { return (0.1 > (double)frac) ? 0.1f : frac; }
// MathUtils.cpp:75-88
LH_HD void barycentric(float x, int lo, int hi, int& idx, float& frac) {
  int i = (int)floorf(x);
  float f = x - (float)i;
  if (i < lo) {
    i = lo;
    f = 0.0f;
  } else if (i > hi - 2) {
    i = hi - 2;
    f = 1.0f;
  }
  idx = i;
  frac = f;
}
// Eigen fixed-size length-3 reductions: x0 + (x1 + x2) (Redux.h splits 3 as 1 + 2).
LH_HD float sum3(float a, float b, float c) { return a + (b + c); }
LH_HD float norm3(float a, float b, float c) { return sqrtf(a * a + (b * b + c * c)); }
LH_HD float max3(float a, float b, float c) {
  float bc = (b < c) ? c : b;
  return (a < bc) ? bc : a;
}

// Solid SDF program: postfix list of nodes, 12 floats each
//   [kind, cx, cy, cz, hx, hy, hz, radius, useInside, -, -, -]
// kind 0 RectBoundary (Boundary.h:77-94), 1 sphere, 10 Union (min), 11
// Intersection (max), 12 Subtract (max(a,-b)), 13 None (pass-through)
// (OperatorBoundary::Compute, Boundary.h:38-56).
constexpr int kMaxSdfNodes = 32;
LH_HD float sdf_eval(const float* prog, int n, float px, float py, float pz) {
  float st[16];
  int sp = 0;
  for (int k = 0; k < n; ++k) {
    const float* q = prog + 12 * k;
    int kind = (int)q[0];
    if (kind == 0) {
      float ax = fabsf(px - q[1]) - q[4];
      float ay = fabsf(py - q[2]) - q[5];
      float az = fabsf(pz - q[3]) - q[6];
      float mx = ax > 0.f ? ax : 0.f, my = ay > 0.f ? ay : 0.f, mz = az > 0.f ? az : 0.f;
      float outside = norm3(mx, my, mz);
      float mc = max3(ax, ay, az);
      float inside = (mc < 0.f) ? mc : 0.f;
      st[sp++] = (q[8] != 0.f) ? -outside - inside : outside + inside;
    } else if (kind == 1) {
      float d = norm3(px - q[1], py - q[2], pz - q[3]) - q[7];
      st[sp++] = (q[8] != 0.f) ? -d : d;
    } else if (kind == 13) {
    } else {
      float b = st[--sp];
      float a = st[--sp];
      float r;
      if (kind == 10) r = (a < b) ? a : b;
      else if (kind == 11) r = (a > b) ? a : b;
      else {
        float nb = -b;
        r = (a > nb) ? a : nb;
      }
      st[sp++] = r;
    }
  }
  return sp > 0 ? st[sp - 1] : 1.0f;
}

}  // namespace lh
