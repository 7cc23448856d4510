// lh2.hip -- LiquidModel<2> / CGSolver<2> (LiquidModel.cpp:10-474, PressureSolver.cpp:9-418) on
// gfx950 behind the lh2_* C ABI (include/lh_api.h).
//
// The 2D model is the reference's other instantiation of the same classes.  Its grids are small
// (a 512^2 grid is 1 MB per array), so the kernels here are plain one-thread-per-element
// kernels; the points where the result depends on an order follow the same rules as the 3D path:
//   * P2G is a per-face gather over the particles of the 3 x 3 cells around the face (a particle
//     of radius dx reaches exactly the faces around its cell), in cell order and, inside a cell,
//     in particle-id order: particles are sorted by cell with a stable radix sort;
//   * the PCG dot products accumulate in FP64 (block partials summed in a fixed order) and are
//     rounded to float, as the oracle's dot_mode 1;
//   * Extrapolate's in-place validity scan (x outer, y inner) is reproduced by one workgroup that
//     walks the anti-diagonals t = 2x + y: the faces a face can see validated earlier in the scan
//     ((x-1, y-1..y+1) and (x, y-1)) all have a smaller t, the ones after it a larger t.
// Compiled with -ffp-contract=off like the rest of the library (no FMA contraction).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/lh_api.h"
#include "lh_json.h"
#include "lh_math.h"

namespace lh2 {
using lh::bilerp;
using lh::clampf;
using lh::fraction_in_region;
using lh::is_small;
using lh::Json;
using lh::kEps;
using lh::max_tenth;
using lh::quad_gw;
using lh::quad_w;

constexpr int kB = 256;
constexpr int kMaxNodes = 32;

inline size_t div_up(size_t a, size_t b) { return (a + b - 1) / b; }
inline unsigned grid_for(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>(div_up(n, kB), 65535)); }

struct P2 {  // grid geometry passed to kernels
  int N0, N1;
  float mn0, mn1, mx0, mx1, dx0, dx1;
};

struct Level {
  int d0 = 0, d1 = 0;
  float *adiag = nullptr, *acoef = nullptr, *r = nullptr, *z = nullptr;  // acoef: 4 per cell
};

struct Ctx2 {
  // configuration (SceneParameter.h / SolverParameter.h, 2D)
  P2 g;
  float gravity = -9.81f, density = 1000.f;
  int max_iter = 100, tol_mode = 0, precond = 2, num_level = 1, num_pre = 2, num_post = 2, num_final = 10;
  float tol = 1e-6f, omega = 0.5f;
  int smoother = 0;  // 0 reference, 1 symmetric
  int correct_cycle = 8;
  long long num_particles = 0;
  unsigned long long seed = 0;
  std::vector<float> sdf;
  std::vector<float> seed_pos;  // optional FluidBlocks spawn
  // device state
  hipStream_t s = nullptr;
  float *vx = nullptr, *vy = nullptr, *tvx = nullptr, *tvy = nullptr, *lwx = nullptr, *lwy = nullptr;
  float *pressure = nullptr, *phi = nullptr, *rhs = nullptr, *solid = nullptr, *d_sdf = nullptr;
  uint8_t* valid = nullptr;
  int64_t P = 0, cap = 0;
  float *pos = nullptr, *vel = nullptr, *C = nullptr, *cvbuf = nullptr;
  int *key = nullptr, *key_s = nullptr, *idx = nullptr, *idx_s = nullptr, *start = nullptr, *count = nullptr;
  void* cub_tmp = nullptr;
  size_t cub_bytes = 0;
  std::vector<Level> lv;
  float *p = nullptr, *x = nullptr, *q = nullptr;
  double* partial = nullptr;  // block partials
  double* dres = nullptr;     // reduced sums
  double* h_pin = nullptr;
  unsigned* dmaxv = nullptr;
  double time = 0.0;
  int correct_step = 0;
  unsigned long long correct_count = 0;
  float last_maxv = 0.f;
  std::string err;
};

// ------------------------------------------------------------------ device helpers
__device__ __forceinline__ float norm2(float a, float b) { return sqrtf(a * a + b * b); }

// Boundary.h:27-94 on a 2D point (postfix program, 12 floats per node; z ignored)
__device__ float sdf_eval2(const float* __restrict__ prog, int n, float px, float py) {
  float st[kMaxNodes];
  int sp = 0;
  for (int k = 0; k < n; ++k) {
    const float* q = prog + 12 * k;
    const int kind = (int)q[0];
    if (kind == 0) {
      const float ax = fabsf(px - q[1]) - q[4], ay = fabsf(py - q[2]) - q[5];
      const float outside = norm2(ax > 0.f ? ax : 0.f, ay > 0.f ? ay : 0.f);
      const float mc = (ax < ay) ? ay : ax;
      const float inside = (mc < 0.f) ? mc : 0.f;
      st[sp++] = (q[8] != 0.f) ? -outside - inside : outside + inside;
    } else if (kind == 1) {
      const float d = norm2(px - q[1], py - q[2]) - q[7];
      st[sp++] = (q[8] != 0.f) ? -d : d;
    } else if (kind != 13) {
      const float b = st[--sp], a = st[--sp];
      float r;
      if (kind == 10) r = (a < b) ? a : b;
      else if (kind == 11) r = (a > b) ? a : b;
      else {
        const float nb = -b;
        r = (a > nb) ? a : nb;
      }
      st[sp++] = r;
    }
  }
  return sp > 0 ? st[sp - 1] : 1.0f;
}

// Grid<2>::GetSolidPhi / GetSolidGradient, Grid.cpp:37-50, 67-80
__device__ __forceinline__ bool in_grid(const P2& g, float x, float y) {
  return g.mn0 <= x && g.mn1 <= y && x <= g.mx0 && y <= g.mx1;
}
__device__ __forceinline__ void bary(float x, int lo, int hi, int& idx, float& frac) {
  int i = (int)floorf(x);
  float f = x - (float)i;
  if (i < lo) {
    i = lo;
    f = 0.f;
  } else if (i > hi - 2) {
    i = hi - 2;
    f = 1.f;
  }
  idx = i;
  frac = f;
}
__device__ float solid_at_in(const float* __restrict__ s, const P2& g, float x, float y) {
  const float lx = (x - g.mn0) / g.dx0, ly = (y - g.mn1) / g.dx1;
  int xi, yi;
  float fx, fy;
  bary(lx, 0, g.N0 + 1, xi, fx);
  bary(ly, 0, g.N1 + 1, yi, fy);
  const int w = g.N1 + 1;
  return bilerp(s[xi * w + yi], s[(xi + 1) * w + yi], s[xi * w + yi + 1], s[(xi + 1) * w + yi + 1], fx, fy);
}
__device__ void grad_in(const float* __restrict__ s, const P2& g, float x, float y, float gr[2]) {
  const float eps = 1e-4f;
  const float left = solid_at_in(s, g, x - eps, y), right = solid_at_in(s, g, x + eps, y);
  const float bottom = solid_at_in(s, g, x, y - eps), top = solid_at_in(s, g, x, y + eps);
  const float two_eps = (float)(2.0 * (double)eps);
  float a = (right - left) / two_eps, b = (top - bottom) / two_eps;
  const float n2 = a * a + b * b;
  if (n2 > 0.f) {
    const float n = sqrtf(n2);
    a = a / n;
    b = b / n;
  }
  gr[0] = a;
  gr[1] = b;
}
// the samples around an in-grid point stay in the grid's [min, max] box only up to eps; like
// the reference, the central differences of an in-grid point sample the in-grid branch (points
// within eps outside the box take the out-of-grid branch there): reproduced by recursion depth 1
__device__ float solid_at(const float* __restrict__ s, const P2& g, float x, float y);
__device__ void solid_grad(const float* __restrict__ s, const P2& g, float x, float y, float gr[2]) {
  if (!in_grid(g, x, y)) {
    x = clampf(x, g.mn0 + g.dx0, g.mx0 - g.dx0);
    y = clampf(y, g.mn1 + g.dx1, g.mx1 - g.dx1);
  }
  const float eps = 1e-4f;
  const float left = solid_at(s, g, x - eps, y), right = solid_at(s, g, x + eps, y);
  const float bottom = solid_at(s, g, x, y - eps), top = solid_at(s, g, x, y + eps);
  const float two_eps = (float)(2.0 * (double)eps);
  float a = (right - left) / two_eps, b = (top - bottom) / two_eps;
  const float n2 = a * a + b * b;
  if (n2 > 0.f) {
    const float n = sqrtf(n2);
    a = a / n;
    b = b / n;
  }
  gr[0] = a;
  gr[1] = b;
}
__device__ float solid_at(const float* __restrict__ s, const P2& g, float x, float y) {
  if (in_grid(g, x, y)) return solid_at_in(s, g, x, y);
  const float vx = clampf(x, g.mn0 + g.dx0, g.mx0 - g.dx0), vy = clampf(y, g.mn1 + g.dx1, g.mx1 - g.dx1);
  // (vx, vy) is at least dx inside the box: its +-eps samples are in the grid
  float gr[2];
  grad_in(s, g, vx, vy, gr);
  return solid_at_in(s, g, vx, vy) + gr[0] * (x - vx) + gr[1] * (y - vy);
}

// ------------------------------------------------------------------ kernels
__global__ void k_fill(float* __restrict__ a, size_t n, float v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}
__global__ void k_fill_i(int* __restrict__ a, size_t n, int v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}

// LiquidModel<2>::ComputeSolidPhi, LiquidModel.cpp:100-125
__global__ void k_solid_phi(float* __restrict__ s, const float* __restrict__ prog, int n, P2 g) {
  const int w = g.N1 + 1;
  const size_t tot = (size_t)(g.N0 + 1) * w;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
    const int xi = (int)(i / w), yi = (int)(i % w);
    s[i] = sdf_eval2(prog, n, (float)xi * g.dx0 + g.mn0, (float)yi * g.dx1 + g.mn1);
  }
}
__global__ void k_face_w(float* __restrict__ lwx, float* __restrict__ lwy, const float* __restrict__ s, P2 g) {
  const int w = g.N1 + 1;
  const size_t nx = (size_t)(g.N0 + 1) * g.N1, ny = (size_t)g.N0 * (g.N1 + 1);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nx + ny; i += (size_t)gridDim.x * blockDim.x) {
    if (i < nx) {
      const int xi = (int)(i / g.N1), yi = (int)(i % g.N1);
      lwx[i] = clampf(1.0f - fraction_in_region(s[xi * w + yi + 1], s[xi * w + yi]), 0.f, 1.f);
    } else {
      const size_t j = i - nx;
      const int xi = (int)(j / (g.N1 + 1)), yi = (int)(j % (g.N1 + 1));
      lwy[j] = clampf(1.0f - fraction_in_region(s[(xi + 1) * w + yi], s[xi * w + yi]), 0.f, 1.f);
    }
  }
}

// cell key of each particle (Point2Index: truncation, clamped into the grid)
__global__ void k_bin_key(const float* __restrict__ pos, int64_t P, P2 g, int* __restrict__ key, int* __restrict__ idx,
                          int* __restrict__ count) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < P; i += (int64_t)gridDim.x * blockDim.x) {
    int cx = (int)((pos[2 * i] - g.mn0) / g.dx0), cy = (int)((pos[2 * i + 1] - g.mn1) / g.dx1);
    cx = min(max(cx, 0), g.N0 - 1);
    cy = min(max(cy, 0), g.N1 - 1);
    const int k = cx * g.N1 + cy;
    key[i] = k;
    idx[i] = (int)i;
    atomicAdd(&count[k], 1);
  }
}

// LiquidModel<2>::ComputeLiquidPhi, LiquidModel.cpp:78-98: the particles of the 3 x 3 cells
// around a cell reach it (radius dx); min is order-free
__global__ void k_liquid_phi(float* __restrict__ phi, const float* __restrict__ pos, const int* __restrict__ order,
                             const int* __restrict__ start, P2 g, float radius) {
  const size_t n = (size_t)g.N0 * g.N1;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int xi = (int)(i / g.N1), yi = (int)(i % g.N1);
    const float px = ((float)xi + 0.5f) * g.dx0 + g.mn0, py = ((float)yi + 0.5f) * g.dx1 + g.mn1;
    float best = 3 * g.dx0;
    for (int cx = max(0, xi - 1); cx <= min(g.N0 - 1, xi + 1); ++cx)
      for (int cy = max(0, yi - 1); cy <= min(g.N1 - 1, yi + 1); ++cy) {
        const int k = cx * g.N1 + cy;
        for (int s = start[k]; s < start[k + 1]; ++s) {
          const int pi = order[s];
          const float ph = norm2(px - pos[2 * pi], py - pos[2 * pi + 1]) - radius;
          best = (best < ph) ? best : ph;
        }
      }
    phi[i] = best;
  }
}

// LiquidModel<2>::Particle2Grid, LiquidModel.cpp:127-186, gathered per face; normalised; plus
// ApplyGravity (LiquidModel.cpp:306-313) on the vy faces of x in [1, N0-2], y in [1, N1-1] when gdt
// is not 0
__global__ void k_p2g(float* __restrict__ vx, float* __restrict__ vy, const float* __restrict__ pos,
                      const float* __restrict__ vel, const float* __restrict__ C, const int* __restrict__ order,
                      const int* __restrict__ start, P2 g, float gdt) {
  const size_t nx = (size_t)(g.N0 + 1) * g.N1, ny = (size_t)g.N0 * (g.N1 + 1);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nx + ny; i += (size_t)gridDim.x * blockDim.x) {
    const int a = i < nx ? 0 : 1;
    const size_t j = a == 0 ? i : i - nx;
    const int d1 = g.N1 + a;
    const int xi = (int)(j / d1), yi = (int)(j % d1);
    const float ccx = (a == 0) ? (float)xi : (float)xi + 0.5f;
    const float ccy = (a == 1) ? (float)yi : (float)yi + 0.5f;
    const float fx = ccx * g.dx0 + g.mn0, fy = ccy * g.dx1 + g.mn1;
    float ws = 0.f, vs = 0.f;
    for (int cx = max(0, xi - 1); cx <= min(g.N0 - 1, xi + 1); ++cx)
      for (int cy = max(0, yi - 1); cy <= min(g.N1 - 1, yi + 1); ++cy) {
        const int k = cx * g.N1 + cy;
        for (int s = start[k]; s < start[k + 1]; ++s) {
          const int pi = order[s];
          const float p0 = pos[2 * pi], p1 = pos[2 * pi + 1];
          const float gx = (p0 - g.mn0) / g.dx0, gy = (p1 - g.mn1) / g.dx1;
          const float w = quad_w(gx - ccx) * quad_w(gy - ccy);
          const float d0 = p0 - fx, d1v = p1 - fy;
          const float dot = C[4 * pi + 2 * a] * d0 + C[4 * pi + 2 * a + 1] * d1v;
          ws += w;
          vs += w * (vel[2 * pi + a] + dot);
        }
      }
    float v = !is_small(ws) ? vs / ws : 0.0f;
    if (a == 1 && gdt != 0.f && xi >= 1 && xi <= g.N0 - 2 && yi >= 1 && yi <= g.N1 - 1) v += gdt;
    if (a == 0) vx[j] = v;
    else vy[j] = v;
  }
}
__global__ void k_gravity(float* __restrict__ vy, P2 g, float gdt) {
  const size_t n = (size_t)g.N0 * (g.N1 + 1);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int xi = (int)(i / (g.N1 + 1)), yi = (int)(i % (g.N1 + 1));
    if (xi >= 1 && xi <= g.N0 - 2 && yi >= 1 && yi <= g.N1 - 1) vy[i] += gdt;
  }
}

// CGSolver<2>::BuildRHS, PressureSolver.cpp:194-206
__global__ void k_build_rhs(float* __restrict__ b, const float* __restrict__ phi, const float* __restrict__ lwx,
                            const float* __restrict__ lwy, const float* __restrict__ vx, const float* __restrict__ vy,
                            P2 g) {
  const size_t n = (size_t)g.N0 * g.N1;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int xi = (int)(i / g.N1), yi = (int)(i % g.N1);
    float r = 0.0f;
    if (!(phi[i] > 0.0f)) {
      const size_t fx0 = (size_t)xi * g.N1 + yi, fx1 = fx0 + g.N1;
      const size_t fy0 = (size_t)xi * (g.N1 + 1) + yi, fy1 = fy0 + 1;
      r -= lwx[fx1] * vx[fx1] / g.dx0;
      r += lwx[fx0] * vx[fx0] / g.dx0;
      r -= lwy[fy1] * vy[fy1] / g.dx1;
      r += lwy[fy0] * vy[fy0] / g.dx1;
    }
    b[i] = r;
  }
}

// CGSolver<2>::GetApprox*, PressureSolver.cpp:69-94
__device__ __forceinline__ float approx_phi(const float* __restrict__ phi, const P2& g, int xi, int yi, int l) {
  if (l == 0) return phi[(size_t)xi * g.N1 + yi];
  const int f = 1 << l;
  const int bx = f * xi + (1 << (l - 1)) - 1, by = f * yi + (1 << (l - 1)) - 1;
  const size_t b = (size_t)bx * g.N1 + by;
  return bilerp(phi[b], phi[b + g.N1], phi[b + 1], phi[b + g.N1 + 1], 0.5f, 0.5f);
}
__device__ __forceinline__ float approx_w(const float* __restrict__ lwx, const float* __restrict__ lwy, const P2& g,
                                          int a, int xi, int yi, int l) {
  if (l == 0) return a == 0 ? lwx[(size_t)xi * g.N1 + yi] : lwy[(size_t)xi * (g.N1 + 1) + yi];
  const int f = 1 << l;
  if (a == 0) {
    const int bx = f * xi, by = f * yi + (1 << (l - 1)) - 1;
    const size_t b = (size_t)bx * g.N1 + by;
    return (float)(0.5 * (double)(lwx[b] + lwx[b + 1]));
  }
  const int bx = f * xi + (1 << (l - 1)) - 1, by = f * yi;
  const size_t b = (size_t)bx * (g.N1 + 1) + by;
  return (float)(0.5 * (double)(lwy[b] + lwy[b + g.N1 + 1]));
}

// CGSolver<2>::BuildLHS (one level), PressureSolver.cpp:96-175
__global__ void k_build_lhs(float* __restrict__ adiag, float* __restrict__ acoef, const float* __restrict__ phi,
                            const float* __restrict__ lwx, const float* __restrict__ lwy, P2 g, int l, int d0, int d1,
                            float dt, float rho) {
  const float f = (float)(1 << l);
  const float dxl[2] = {f * g.dx0, f * g.dx1};
  const size_t n = (size_t)d0 * d1;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int xi = (int)(i / d1), yi = (int)(i % d1);
    float diag = 0.f, co[4] = {0.f, 0.f, 0.f, 0.f};
    const float cphi = approx_phi(phi, g, xi, yi, l);
    if (!(cphi > 0.0f)) {
      const int ax[4] = {0, 0, 1, 1}, sg[4] = {1, -1, 1, -1};
      const int dd[2] = {d0, d1};
      for (int k = 0; k < 4; ++k) {
        const int a = ax[k], s = sg[k];
        const int c[2] = {xi, yi};
        if (s > 0 ? !(c[a] + 1 < dd[a]) : !(c[a] - 1 >= 0)) continue;
        int fc[2] = {xi, yi};
        if (s > 0) fc[a] += 1;
        const float wv = approx_w(lwx, lwy, g, a, fc[0], fc[1], l);
        if (!(wv > 0.0f)) continue;
        const float term = wv * dt / (rho * dxl[a] * dxl[a]);
        int nb[2] = {xi, yi};
        nb[a] += s;
        const float ph = approx_phi(phi, g, nb[0], nb[1], l);
        if (ph < 0.0f) {
          co[k] -= term;
          diag += term;
        } else {
          diag += term / max_tenth(fraction_in_region(cphi, ph));
        }
      }
    }
    adiag[i] = diag;
    for (int k = 0; k < 4; ++k) acoef[4 * i + k] = co[k];
  }
}

// CGSolver<2>::GetAx, PressureSolver.cpp:208-223 (order: diag, right, left, top, bottom)
__device__ __forceinline__ float get_ax(const float* __restrict__ adiag, const float* __restrict__ acoef,
                                        const float* __restrict__ z, int d0, int d1, int xi, int yi) {
  const size_t c = (size_t)xi * d1 + yi;
  float res = z[c] * adiag[c];
  if (xi + 1 < d0) res += z[c + d1] * acoef[4 * c + 0];
  if (xi - 1 >= 0) res += z[c - d1] * acoef[4 * c + 1];
  if (yi + 1 < d1) res += z[c + 1] * acoef[4 * c + 2];
  if (yi - 1 >= 0) res += z[c - 1] * acoef[4 * c + 3];
  return res;
}

// CGSolver<2>::Smooth, PressureSolver.cpp:312-327 (mode 1: z + w term, the symmetric option)
__global__ void k_smooth(float* __restrict__ z, const float* __restrict__ r, const float* __restrict__ adiag,
                         const float* __restrict__ acoef, int d0, int d1, int phase, float w, int mode) {
  const size_t n = (size_t)d0 * d1;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int xi = (int)(i / d1), yi = (int)(i % d1);
    if (((xi + yi) & 1) != phase) continue;
    const float ad = adiag[i];
    if (is_small(ad)) continue;
    const float term = (r[i] - get_ax(adiag, acoef, z, d0, d1, xi, yi)) / ad;
    if (mode == 0) {
      const float wt = w * term;
      z[i] = (float)((1.0 - (double)w) * (double)z[i] + (double)wt);
    } else {
      z[i] = z[i] + w * term;
    }
  }
}
// CGSolver<2>::Restrict / Prolong, PressureSolver.cpp:280-310 (children (0,0), (1,0), (0,1), (1,1))
__global__ void k_restrict(float* __restrict__ rc, float* __restrict__ zc, const float* __restrict__ rf,
                           const float* __restrict__ zf, const float* __restrict__ adf, const float* __restrict__ acf,
                           int fd0, int fd1, int cd0, int cd1) {
  const size_t n = (size_t)cd0 * cd1;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int xi = (int)(i / cd1), yi = (int)(i % cd1);
    float acc = 0.f;
    const int ox[4] = {0, 1, 0, 1}, oy[4] = {0, 0, 1, 1};
    for (int k = 0; k < 4; ++k) {
      const int x = 2 * xi + ox[k], y = 2 * yi + oy[k];
      const float d = rf[(size_t)x * fd1 + y] - get_ax(adf, acf, zf, fd0, fd1, x, y);
      acc = (float)((double)acc + 0.25 * (double)d);
    }
    rc[i] = acc;
    zc[i] = 0.f;
  }
}
__global__ void k_prolong(float* __restrict__ zf, const float* __restrict__ zc, int fd1, int cd0, int cd1) {
  const size_t n = (size_t)cd0 * cd1;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int xi = (int)(i / cd1), yi = (int)(i % cd1);
    const float v = zc[i];
    zf[(size_t)(2 * xi) * fd1 + 2 * yi] += v;
    zf[(size_t)(2 * xi + 1) * fd1 + 2 * yi] += v;
    zf[(size_t)(2 * xi) * fd1 + 2 * yi + 1] += v;
    zf[(size_t)(2 * xi + 1) * fd1 + 2 * yi + 1] += v;
  }
}
// PrecondType::None / Jacobi, PressureSolver.cpp:227-246
__global__ void k_precond_simple(float* __restrict__ z, const float* __restrict__ r, const float* __restrict__ adiag,
                                 const float* __restrict__ phi, size_t n, int jacobi) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    if (jacobi) {
      if (!is_small(adiag[i])) z[i] = r[i] / adiag[i];
    } else if (!(phi[i] > 0.0f)) {
      z[i] = r[i];
    }
  }
}

// FP64 block partials of up to 2 sums, written to part[k * gridDim.x + block]
template <int NV>
__device__ void block_sum(double (&v)[NV], double* __restrict__ part) {
  __shared__ double s[NV][kB];
  for (int k = 0; k < NV; ++k) s[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int st = kB / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int k = 0; k < NV; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < NV; ++k) part[(size_t)k * gridDim.x + blockIdx.x] = s[k][0];
}
// sums of the partials: one block, thread t adds partials t, t + 256, ... in order, then a
// fixed tree (deterministic)
__global__ void __launch_bounds__(kB) k_sum_partials(const double* __restrict__ part, int nb, int nv,
                                                     double* __restrict__ out) {
  __shared__ double s[kB];
  for (int k = 0; k < nv; ++k) {
    double a = 0.0;
    for (int b = threadIdx.x; b < nb; b += kB) a += part[(size_t)k * nb + b];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int st = kB / 2; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = s[0];
    __syncthreads();
  }
}

// r = b on liquid cells; sums r.r and the liquid count (Solve, :331-340)
__global__ void k_cg_init(float* __restrict__ r, const float* __restrict__ b, const float* __restrict__ phi, size_t n,
                          double* __restrict__ part) {
  double v[2] = {0.0, 0.0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    if (!(phi[i] > 0.0f)) {
      r[i] = b[i];
      v[1] += 1.0;
    }
  block_sum<2>(v, part);
}
// p = z on liquid cells; sums r.z and r.r (:341-349)
__global__ void k_cg_first(float* __restrict__ p, const float* __restrict__ z, const float* __restrict__ r,
                           const float* __restrict__ phi, size_t n, double* __restrict__ part) {
  double v[2] = {0.0, 0.0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    if (!(phi[i] > 0.0f)) p[i] = z[i];
    v[0] += (double)r[i] * (double)z[i];
    v[1] += (double)r[i] * (double)r[i];
  }
  block_sum<2>(v, part);
}
// q = A p and sum p.q over interior liquid cells (:355-362)
__global__ void k_cg_spmv(float* __restrict__ q, const float* __restrict__ p, const float* __restrict__ adiag,
                          const float* __restrict__ acoef, const float* __restrict__ phi, int d0, int d1,
                          double* __restrict__ part) {
  double v[1] = {0.0};
  const size_t n = (size_t)d0 * d1;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int xi = (int)(i / d1), yi = (int)(i % d1);
    if (xi < 1 || xi > d0 - 2 || yi < 1 || yi > d1 - 2 || phi[i] > 0.0f) continue;
    const float ap = get_ax(adiag, acoef, p, d0, d1, xi, yi);
    q[i] = ap;
    v[0] += (double)p[i] * (double)ap;
  }
  block_sum<1>(v, part);
}
// x += alpha p, r -= alpha q on interior liquid cells; sums r.r and their count (:370-384)
__global__ void k_cg_update_xr(float* __restrict__ x, float* __restrict__ r, const float* __restrict__ p,
                               const float* __restrict__ q, const float* __restrict__ phi, int d0, int d1, float alpha,
                               double* __restrict__ part) {
  double v[2] = {0.0, 0.0};
  const size_t n = (size_t)d0 * d1;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int xi = (int)(i / d1), yi = (int)(i % d1);
    if (xi < 1 || xi > d0 - 2 || yi < 1 || yi > d1 - 2 || phi[i] > 0.0f) continue;
    x[i] += alpha * p[i];
    const float rn = r[i] - alpha * q[i];
    r[i] = rn;
    v[0] += (double)rn * (double)rn;
    v[1] += 1.0;
  }
  block_sum<2>(v, part);
}
__global__ void k_cg_dot(const float* __restrict__ a, const float* __restrict__ b, size_t n, double* __restrict__ part) {
  double v[1] = {0.0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    v[0] += (double)a[i] * (double)b[i];
  block_sum<1>(v, part);
}
// p = z + beta p on liquid cells (:394-400)
__global__ void k_cg_update_p(float* __restrict__ p, const float* __restrict__ z, const float* __restrict__ phi, size_t n,
                              float beta) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    if (!(phi[i] > 0.0f)) p[i] = z[i] + beta * p[i];
}

// LiquidModel<2>::SolvePressure velocity update, LiquidModel.cpp:246-291
__global__ void k_velocity_update(float* __restrict__ vx, float* __restrict__ vy, const float* __restrict__ p,
                                  const float* __restrict__ phi, const float* __restrict__ lwx,
                                  const float* __restrict__ lwy, P2 g, float dt, float rho) {
  const size_t n = (size_t)g.N0 * g.N1;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int xi = (int)(i / g.N1), yi = (int)(i % g.N1);
    if (xi < 1 || xi > g.N0 - 2 || yi < 1 || yi > g.N1 - 2) continue;
    const float cphi = phi[i];
    for (int a = 0; a < 2; ++a) {
      const size_t f = a == 0 ? (size_t)xi * g.N1 + yi : (size_t)xi * (g.N1 + 1) + yi;
      float* v = a == 0 ? vx : vy;
      const float w = a == 0 ? lwx[f] : lwy[f];
      if (w > 0.0f) {
        const size_t nb = a == 0 ? i - g.N1 : i - 1;
        const float ph = phi[nb];
        if (cphi < 0.0f || ph < 0.0f) {
          float frac = 1.0f;
          if (cphi >= 0.0f || ph >= 0.0f) frac = max_tenth(fraction_in_region(cphi, ph));
          const float gp = (p[i] - p[nb]) / ((a == 0 ? g.dx0 : g.dx1) * frac);
          v[f] -= gp * dt / rho;
        }
      } else {
        v[f] = 0.0f;
      }
    }
  }
}

// LiquidModel<2>::Extrapolate for one component, LiquidModel.cpp:341-409: one workgroup walks
// the anti-diagonals t = 2x + y of the range x in [1, xhi], y in [1, yhi] five times; sweep k
// reads src (v, tv, v, ...) and writes the other buffer, as std::swap(temp, v) does.  With
// kLds the validity of every node lives in LDS as bits (row x: words [x * wpr, (x + 1) * wpr)),
// loaded at the start and written back at the end; a step touches one face per row, so no two
// threads of a step write the same word
template <bool kLds>
__global__ void __launch_bounds__(1024) k_extrapolate(float* __restrict__ v, float* __restrict__ tv,
                                                      uint8_t* __restrict__ valid, const float* __restrict__ lw,
                                                      int d1, int vw, int vrows, int xhi, int yhi) {
  // d1: the component's face array y-size; vw: isValid row length (N1 + 1); vrows: N0 + 1
  extern __shared__ uint32_t bits[];
  const int wpr = (vw + 31) >> 5;
  auto get = [&](int x, int y) -> bool {
    if (kLds) return (bits[x * wpr + (y >> 5)] >> (y & 31)) & 1u;
    return valid[(size_t)x * vw + y] != 0;
  };
  auto set = [&](int x, int y) {
    if (kLds) bits[x * wpr + (y >> 5)] |= 1u << (y & 31);
    else valid[(size_t)x * vw + y] = 1;
  };
  if (kLds) {
    for (int w = threadIdx.x; w < vrows * wpr; w += blockDim.x) {
      const int x = w / wpr, y0 = (w - x * wpr) * 32;
      uint32_t b = 0;
      for (int k = 0; k < 32 && y0 + k < vw; ++k) {
        const int y = y0 + k;
        bool val = valid[(size_t)x * vw + y] != 0;
        if (x >= 1 && x <= xhi && y >= 1 && y <= yhi) val = lw[(size_t)x * d1 + y] > 0.0f;
        b |= (val ? 1u : 0u) << k;
      }
      bits[w] = b;
    }
  } else {
    for (int x = 1 + (int)threadIdx.x; x <= xhi; x += blockDim.x)
      for (int y = 1; y <= yhi; ++y) valid[(size_t)x * vw + y] = lw[(size_t)x * d1 + y] > 0.0f ? 1 : 0;
  }
  __syncthreads();
  if (kLds && xhi <= (int)blockDim.x) {
    // thread x - 1 owns row x: at step t it handles y = t - 2x, so its faces come in y order,
    // and the ten source values of its next face (the face and its 3 x 3 block; the source is
    // read-only during a sweep) are loaded before the barrier and used after it
    const int x = 1 + (int)threadIdx.x;
    const bool row = x <= xhi;
    for (int it = 0; it < 5; ++it) {
      const float* __restrict__ src = (it & 1) ? tv : v;
      float* __restrict__ dst = (it & 1) ? v : tv;
      float pf[3][3];
      int ynext = 1;
      if (row)
        for (int dx = -1; dx <= 1; ++dx)
          for (int dy = -1; dy <= 1; ++dy) pf[dx + 1][dy + 1] = src[(size_t)(x + dx) * d1 + (ynext + dy)];
      for (int t = 3; t <= 2 * xhi + yhi; ++t) {
        const int y = t - 2 * x;
        if (row && y == ynext && y <= yhi) {
          float out = pf[1][1];
          if (!get(x, y)) {
            float sum = 0.0f;
            int cnt = 0;
            for (int dx = -1; dx <= 1; ++dx)
              for (int dy = -1; dy <= 1; ++dy)
                if (get(x + dx, y + dy)) {
                  sum += pf[dx + 1][dy + 1];
                  ++cnt;
                }
            if (cnt > 0) {
              out = sum / (float)cnt;
              set(x, y);
            }
          }
          dst[(size_t)x * d1 + y] = out;
          ynext = y + 1;
          if (ynext <= yhi)
            for (int dx = -1; dx <= 1; ++dx)
              for (int dy = -1; dy <= 1; ++dy) pf[dx + 1][dy + 1] = src[(size_t)(x + dx) * d1 + (ynext + dy)];
        }
        __syncthreads();
      }
    }
  } else
  for (int it = 0; it < 5; ++it) {
    const float* __restrict__ src = (it & 1) ? tv : v;
    float* __restrict__ dst = (it & 1) ? v : tv;
    for (int t = 3; t <= 2 * xhi + yhi; ++t) {
      const int xlo = max(1, (t - yhi + 1) / 2), xtop = min(xhi, (t - 1) / 2);
      for (int x = xlo + (int)threadIdx.x; x <= xtop; x += blockDim.x) {
        const int y = t - 2 * x;
        const size_t f = (size_t)x * d1 + y;
        float out = src[f];
        if (!get(x, y)) {
          float sum = 0.0f;
          int cnt = 0;
          for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
              if (get(x + dx, y + dy)) {
                sum += src[(size_t)(x + dx) * d1 + (y + dy)];
                ++cnt;
              }
          if (cnt > 0) {
            out = sum / (float)cnt;
            set(x, y);
          }
        }
        dst[f] = out;
      }
      __syncthreads();
    }
  }
  if (kLds)
    for (int w = threadIdx.x; w < vrows * wpr; w += blockDim.x) {
      const int x = w / wpr, y0 = (w - x * wpr) * 32;
      const uint32_t b = bits[w];
      for (int k = 0; k < 32 && y0 + k < vw; ++k) valid[(size_t)x * vw + y0 + k] = (uint8_t)((b >> k) & 1u);
    }
}

// LiquidModel<2>::Grid2Particle, LiquidModel.cpp:188-236, and AdvectParticle (:315-339), and the
// max |v| for GetCFL
__global__ void k_g2p(float* __restrict__ vel, float* __restrict__ C, const float* __restrict__ pos,
                      const float* __restrict__ vx, const float* __restrict__ vy, int64_t P, P2 g, float radius) {
  for (int64_t pi = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; pi < P; pi += (int64_t)gridDim.x * blockDim.x) {
    const float gx = (pos[2 * pi] - g.mn0) / g.dx0, gy = (pos[2 * pi + 1] - g.mn1) / g.dx1;
    const int cx = (int)gx, cy = (int)gy;
    const int dX = (int)ceilf(radius / g.dx0), dY = (int)ceilf(radius / g.dx1);
    float vv[2] = {0.f, 0.f}, Cm[4] = {0.f, 0.f, 0.f, 0.f};
    for (int a = 0; a < 2; ++a) {
      const float* V = a == 0 ? vx : vy;
      const int d1 = g.N1 + a;
      for (int ddx = -dX; ddx <= dX; ++ddx)
        for (int ddy = -dY; ddy <= dY; ++ddy) {
          const int xi = cx + ddx, yi = cy + ddy;
          if (xi < 0 || yi < 0 || xi > g.N0 - (a == 0 ? 0 : 1) || yi > g.N1 - (a == 1 ? 0 : 1)) continue;
          const float ccx = (a == 0) ? (float)xi : (float)xi + 0.5f;
          const float ccy = (a == 1) ? (float)yi : (float)yi + 0.5f;
          const float wX = quad_w(gx - ccx), wY = quad_w(gy - ccy);
          const float gX = quad_gw(gx - ccx), gY = quad_gw(gy - ccy);
          const float fv = V[(size_t)xi * d1 + yi];
          vv[a] += fv * (wX * wY);
          Cm[2 * a + 0] += fv * (wY * gX);
          Cm[2 * a + 1] += fv * (wX * gY);
        }
    }
    vel[2 * pi] = vv[0];
    vel[2 * pi + 1] = vv[1];
    for (int k = 0; k < 4; ++k) C[4 * pi + k] = Cm[k];
  }
}
__global__ void k_advect(float* __restrict__ pos, const float* __restrict__ vel, const float* __restrict__ solid,
                         int64_t P, P2 g, float dt, unsigned* __restrict__ maxv) {
  for (int64_t pi = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; pi < P; pi += (int64_t)gridDim.x * blockDim.x) {
    float p0 = pos[2 * pi] + vel[2 * pi] * dt, p1 = pos[2 * pi + 1] + vel[2 * pi + 1] * dt;
    const float sp = solid_at(solid, g, p0, p1);
    if (sp < 0.0f) {
      float gr[2];
      solid_grad(solid, g, p0, p1, gr);
      const float s = sp + kEps;
      p0 -= s * gr[0];
      p1 -= s * gr[1];
    }
    pos[2 * pi] = p0;
    pos[2 * pi + 1] = p1;
    const float a = fabsf(vel[2 * pi]), b = fabsf(vel[2 * pi + 1]);
    const float m = (a < b) ? b : a;
    atomicMax(maxv, __float_as_uint(m));  // non-negative floats order as their bits
  }
}

// LiquidModel<2>::CorrectVolume, LiquidModel.cpp:428-468 (see the oracle's correct_volume2): the
// neighbours within the radius among the particles of the cells within 2, each visited once per
// pair of region cells whose hash slots match; positions into out, committed by k_cv_commit
__device__ __forceinline__ int ns_slot2(int x, int y, int m) {
  const uint32_t h = (73856093u * (uint32_t)x) ^ (19349663u * (uint32_t)y);
  const int hs = (int)h;
  return ((hs % m) + m) % m;
}
__device__ __forceinline__ uint64_t splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ float cv_rand(uint64_t seed, uint64_t count, int64_t pi, int64_t npi, int v, int k) {
  uint64_t h = splitmix(seed ^ 0x436F7272566F6Cull);
  h = splitmix(h ^ count);
  h = splitmix(h ^ (uint64_t)pi);
  h = splitmix(h ^ ((uint64_t)npi * 128u + (uint64_t)v * 4u + (uint64_t)k));
  return (float)(h >> 40) * (1.0f / 16777216.0f);
}
__global__ void k_correct_volume(float* __restrict__ out, const float* __restrict__ pos, const int* __restrict__ order,
                                 const int* __restrict__ start, const float* __restrict__ solid, int64_t P, P2 g,
                                 float radius, float dt, int cycle, int step, unsigned long long seed,
                                 unsigned long long count) {
  const int nslot = (int)(2 * P);
  const float wf = sqrtf(g.dx0 * g.dx1);
  for (int64_t pi = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; pi < P; pi += (int64_t)gridDim.x * blockDim.x) {
    if (pi % cycle != step) continue;
    const float p0 = pos[2 * pi], p1 = pos[2 * pi + 1];
    const float r = radius, r2 = r * r;
    const int cx = (int)((p0 - g.mn0) / g.dx0), cy = (int)((p1 - g.mn1) / g.dx1);
    const int lo0 = min(max(cx - 1, 0), g.N0 - 1), hi0 = min(max(cx + 1, 0), g.N0 - 1);
    const int lo1 = min(max(cy - 1, 0), g.N1 - 1), hi1 = min(max(cy + 1, 0), g.N1 - 1);
    int reg[9], nreg = 0;
    for (int x1 = lo0; x1 <= hi0; ++x1)
      for (int y1 = lo1; y1 <= hi1; ++y1) reg[nreg++] = ns_slot2(x1, y1, nslot);
    float rep0 = 0.f, rep1 = 0.f;
    for (int bx = max(0, cx - 2); bx <= min(g.N0 - 1, cx + 2); ++bx)
      for (int by = max(0, cy - 2); by <= min(g.N1 - 1, cy + 2); ++by) {
        const int k = bx * g.N1 + by;
        for (int s = start[k]; s < start[k + 1]; ++s) {
          const int npi = order[s];
          if (npi == pi) continue;
          const float q0 = pos[2 * npi], q1 = pos[2 * npi + 1];
          const float d0 = p0 - q0, d1 = p1 - q1;
          const float sq = d0 * d0 + d1 * d1;
          if (sq >= r2) continue;
          const int c2x = (int)((q0 - g.mn0) / g.dx0), c2y = (int)((q1 - g.mn1) / g.dx1);
          const int nl0 = min(max(c2x - 1, 0), g.N0 - 1), nh0 = min(max(c2x + 1, 0), g.N0 - 1);
          const int nl1 = min(max(c2y - 1, 0), g.N1 - 1), nh1 = min(max(c2y + 1, 0), g.N1 - 1);
          int mult = 0;
          for (int x2 = nl0; x2 <= nh0; ++x2)
            for (int y2 = nl1; y2 <= nh1; ++y2) {
              const int s2 = ns_slot2(x2, y2, nslot);
              for (int kk = 0; kk < nreg; ++kk) mult += reg[kk] == s2 ? 1 : 0;
            }
          const float dist = sqrtf(sq);
          const float w = wf * quad_w(dist / r);
          for (int vv = 0; vv < mult; ++vv) {
            if (dist > 1e-4f * r) {
              const float wr = w * r;
              rep0 += wr * (d0 / dist);
              rep1 += wr * (d1 / dist);
            } else {
              rep0 += r / dt * cv_rand(seed, count, pi, npi, vv, 0);
              rep1 += r / dt * cv_rand(seed, count, pi, npi, vv, 1);
            }
          }
        }
      }
    float n0 = p0 + rep0 * dt, n1 = p1 + rep1 * dt;
    const float sp = solid_at(solid, g, n0, n1);
    if (sp < 0.0f) {
      float gr[2];
      solid_grad(solid, g, n0, n1, gr);
      n0 -= sp * gr[0];
      n1 -= sp * gr[1];
    }
    out[2 * pi] = n0;
    out[2 * pi + 1] = n1;
  }
}
__global__ void k_cv_commit(float* __restrict__ pos, const float* __restrict__ out, int64_t P, int cycle, int step) {
  for (int64_t pi = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; pi < P; pi += (int64_t)gridDim.x * blockDim.x)
    if (pi % cycle == step) {
      pos[2 * pi] = out[2 * pi];
      pos[2 * pi + 1] = out[2 * pi + 1];
    }
}

// LiquidModel<2>::ComputeDivergence, LiquidModel.cpp:411-426: mean (FP64) and L-inf
__global__ void k_divergence(const float* __restrict__ phi, const float* __restrict__ lwx, const float* __restrict__ lwy,
                             const float* __restrict__ vx, const float* __restrict__ vy, P2 g,
                             double* __restrict__ part, unsigned* __restrict__ linf) {
  double v[2] = {0.0, 0.0};
  const size_t n = (size_t)g.N0 * g.N1;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    if (phi[i] > 0.0f) continue;
    const int xi = (int)(i / g.N1), yi = (int)(i % g.N1);
    const size_t fx0 = (size_t)xi * g.N1 + yi, fx1 = fx0 + g.N1;
    const size_t fy0 = (size_t)xi * (g.N1 + 1) + yi, fy1 = fy0 + 1;
    float d = lwx[fx1] * vx[fx1] - lwx[fx0] * vx[fx0];
    d += lwy[fy1] * vy[fy1] - lwy[fy0] * vy[fy0];
    d /= g.dx0;
    v[0] += (double)d;
    v[1] += 1.0;
    atomicMax(linf, __float_as_uint(fabsf(d)));
  }
  block_sum<2>(v, part);
}

// ------------------------------------------------------------------ host
bool ok(Ctx2& c, hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  c.err = std::string(what) + ": " + hipGetErrorString(e);
  return false;
}
template <class T>
bool dalloc(Ctx2& c, T*& p, size_t n) {
  p = nullptr;
  return ok(c, hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
}
size_t nfx(const Ctx2& c) { return (size_t)(c.g.N0 + 1) * c.g.N1; }
size_t nfy(const Ctx2& c) { return (size_t)c.g.N0 * (c.g.N1 + 1); }
size_t ncell(const Ctx2& c) { return (size_t)c.g.N0 * c.g.N1; }
size_t nnode(const Ctx2& c) { return (size_t)(c.g.N0 + 1) * (c.g.N1 + 1); }
void fill(Ctx2& c, float* a, size_t n, float v) {
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kB), 0, c.s, a, n, v);
}

void emit_sdf(const Json& n, std::vector<float>& prog) {
  const std::string& t = n.at("type").as_string();
  if (t == "Rect" || t == "Sphere") {
    float node[12] = {t == "Rect" ? 0.f : 1.f, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (n.has("center"))
      for (size_t a = 0; a < n.at("center").size() && a < 3; ++a) node[1 + a] = n.at("center").at(a).as_float();
    if (n.has("halfExtent"))
      for (size_t a = 0; a < n.at("halfExtent").size() && a < 3; ++a) node[4 + a] = n.at("halfExtent").at(a).as_float();
    node[7] = n.has("radius") ? n.at("radius").as_float() : 0.f;
    node[8] = (n.has("useInside") && n.at("useInside").as_bool()) ? 1.f : 0.f;
    prog.insert(prog.end(), node, node + 12);
    return;
  }
  float kind;
  if (t == "Union") kind = 10;
  else if (t == "Intersection") kind = 11;
  else if (t == "Subtract" || t == "Substract") kind = 12;
  else if (t == "None") kind = 13;
  else throw std::runtime_error("Failed to recognize OperatorType");  // Boundary.h:52-54
  if (kind == 13) {
    emit_sdf(n.at("child"), prog);
  } else {
    emit_sdf(n.at("left"), prog);
    emit_sdf(n.at("right"), prog);
  }
  float node[12] = {kind, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  prog.insert(prog.end(), node, node + 12);
}

// SceneParameter.h / SolverParameter.h / LiquidParameter (2D): the reference's keys, and the
// build's optional Boundary (2D tree), toleranceMode, smoother, seed
void parse_config(Ctx2& c, const std::string& text) {
  const Json j = Json::parse(text);
  const Json& gd = j.at("gridDimension");
  if (gd.size() != 2) throw std::runtime_error("lh2: gridDimension must have 2 entries");
  c.g.N0 = gd.at((size_t)0).as_int();
  c.g.N1 = gd.at((size_t)1).as_int();
  if (c.g.N0 < 3 || c.g.N1 < 3) throw std::runtime_error("gridDimension must be >= 3 on every axis");
  c.g.mn0 = j.at("minCoord").at((size_t)0).as_float();
  c.g.mn1 = j.at("minCoord").at((size_t)1).as_float();
  c.g.mx0 = j.at("maxCoord").at((size_t)0).as_float();
  c.g.mx1 = j.at("maxCoord").at((size_t)1).as_float();
  c.g.dx0 = (float)((c.g.mx0 - c.g.mn0) / (float)c.g.N0);  // Grid<2> ctor, Grid.cpp:15-16
  c.g.dx1 = (float)((c.g.mx1 - c.g.mn1) / (float)c.g.N1);
  c.gravity = j.at("gravity").as_float();
  const Json& s = j.at("SolverParameter");
  c.max_iter = s.at("maxIterations").as_int();
  c.tol = s.at("tolerance").as_float();
  const std::string pt = s.at("preconditionerType").as_string();
  if (pt == "None") c.precond = 0;
  else if (pt == "Jacobi") c.precond = 1;
  else if (pt == "Multigrid") c.precond = 2;
  else throw std::runtime_error("Failed to recognize the preconditioner type");
  c.num_level = s.at("numLevel").as_int();
  c.num_pre = s.at("numPreSmooth").as_int();
  c.num_post = s.at("numPostSmooth").as_int();
  c.num_final = s.at("numFinalSmooth").as_int();
  c.omega = s.at("smoothingFactor").as_float();
  c.tol_mode = s.get("toleranceMode", "absolute") == "relative" ? 1 : 0;
  c.smoother = s.get("smoother", "reference") == "symmetric" ? 1 : 0;
  const Json& lq = j.at("LiquidParameter");
  c.density = lq.at("density").as_float();
  c.num_particles = lq.at("numParticles").as_int();
  c.correct_cycle = std::max(1, lq.at("correctCycle").as_int());
  c.seed = (unsigned long long)j.get("seed", 0.0);
  if (j.has("Boundary")) {
    emit_sdf(j.at("Boundary"), c.sdf);
  } else {  // Scene.cpp:18: OperatorBoundary(None) over RectBoundary(0, 0.9, useInside)
    const float rect[12] = {0, 0, 0, 0, 0.9f, 0.9f, 0, 0, 1.f, 0, 0, 0};
    const float none[12] = {13, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    c.sdf.assign(rect, rect + 12);
    c.sdf.insert(c.sdf.end(), none, none + 12);
  }
  if (c.sdf.size() / 12 > (size_t)kMaxNodes) throw std::runtime_error("Boundary tree too large");
}

int ensure_particles(Ctx2& c, int64_t n) {
  if (n <= c.cap && c.pos) return LH_OK;
  float* old[4] = {c.pos, c.vel, c.C, c.cvbuf};
  int* oldi[5] = {c.key, c.key_s, c.idx, c.idx_s, nullptr};
  for (auto p : old)
    if (p) hipFree(p);
  for (auto p : oldi)
    if (p) hipFree(p);
  if (c.cub_tmp) hipFree(c.cub_tmp);
  c.cub_tmp = nullptr;
  const int64_t cap = std::max<int64_t>(n, 1);
  if (!(dalloc(c, c.pos, 2 * cap) && dalloc(c, c.vel, 2 * cap) && dalloc(c, c.C, 4 * cap) &&
        dalloc(c, c.cvbuf, 2 * cap) && dalloc(c, c.key, cap) && dalloc(c, c.key_s, cap) && dalloc(c, c.idx, cap) &&
        dalloc(c, c.idx_s, cap)))
    return LH_ERR_DEVICE;
  c.cub_bytes = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, c.cub_bytes, c.key, c.key_s, c.idx, c.idx_s, (int)cap, 0, 32, c.s);
  size_t scan_bytes = 0;
  hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, c.count, c.start, (int)(ncell(c) + 1), c.s);
  c.cub_bytes = std::max(c.cub_bytes, scan_bytes);
  if (!ok(c, hipMalloc(&c.cub_tmp, std::max<size_t>(c.cub_bytes, 1)), "hipMalloc")) return LH_ERR_DEVICE;
  c.cap = cap;
  return LH_OK;
}

int alloc(Ctx2& c) {
  const size_t nc = ncell(c);
  if (!(ok(c, hipStreamCreate(&c.s), "stream") && dalloc(c, c.vx, nfx(c)) && dalloc(c, c.vy, nfy(c)) &&
        dalloc(c, c.tvx, nfx(c)) && dalloc(c, c.tvy, nfy(c)) && dalloc(c, c.lwx, nfx(c)) && dalloc(c, c.lwy, nfy(c)) &&
        dalloc(c, c.pressure, nc) && dalloc(c, c.phi, nc) && dalloc(c, c.rhs, nc) && dalloc(c, c.solid, nnode(c)) &&
        dalloc(c, c.valid, nnode(c)) && dalloc(c, c.d_sdf, (size_t)12 * kMaxNodes) && dalloc(c, c.start, nc + 1) &&
        dalloc(c, c.count, nc + 1) && dalloc(c, c.p, nc) && dalloc(c, c.x, nc) && dalloc(c, c.q, nc) &&
        dalloc(c, c.dres, 4) && dalloc(c, c.dmaxv, 2)))
    return LH_ERR_DEVICE;
  const int nb = (int)grid_for(nc);
  if (!dalloc(c, c.partial, (size_t)2 * nb)) return LH_ERR_DEVICE;
  if (!ok(c, hipHostMalloc((void**)&c.h_pin, 8 * sizeof(double)), "hipHostMalloc")) return LH_ERR_DEVICE;
  int L = std::max(1, c.num_level);
  if (c.precond != 2) L = 1;
  c.lv.resize(L);
  for (int l = 0; l < L; ++l) {
    Level& q = c.lv[l];
    q.d0 = c.g.N0 / (1 << l);  // CGSolver<2> ctor, PressureSolver.cpp:26-33
    q.d1 = c.g.N1 / (1 << l);
    const size_t n = (size_t)q.d0 * q.d1;
    if (q.d0 < 1 || q.d1 < 1) throw std::runtime_error("numLevel too large for the grid");
    if (!(dalloc(c, q.adiag, n) && dalloc(c, q.acoef, 4 * n) && dalloc(c, q.r, n) && dalloc(c, q.z, n)))
      return LH_ERR_DEVICE;
  }
  hipMemcpyAsync(c.d_sdf, c.sdf.data(), c.sdf.size() * sizeof(float), hipMemcpyHostToDevice, c.s);
  for (float* a : {c.vx, c.tvx, c.lwx}) fill(c, a, nfx(c), 0.f);
  for (float* a : {c.vy, c.tvy, c.lwy}) fill(c, a, nfy(c), 0.f);
  for (float* a : {c.pressure, c.phi, c.rhs, c.p, c.x, c.q}) fill(c, a, nc, 0.f);
  hipMemsetAsync(c.valid, 0, nnode(c), c.s);
  hipLaunchKernelGGL(k_fill_i, dim3(grid_for(nc + 1)), dim3(kB), 0, c.s, c.count, nc + 1, 0);
  return ok(c, hipStreamSynchronize(c.s), "init") ? LH_OK : LH_ERR_DEVICE;
}

void free_all(Ctx2& c) {
  auto F = [](void* p) {
    if (p) hipFree(p);
  };
  for (void* p : {(void*)c.vx, (void*)c.vy, (void*)c.tvx, (void*)c.tvy, (void*)c.lwx, (void*)c.lwy, (void*)c.pressure,
                  (void*)c.phi, (void*)c.rhs, (void*)c.solid, (void*)c.d_sdf, (void*)c.valid, (void*)c.pos,
                  (void*)c.vel, (void*)c.C, (void*)c.cvbuf, (void*)c.key, (void*)c.key_s, (void*)c.idx,
                  (void*)c.idx_s, (void*)c.start, (void*)c.count, c.cub_tmp, (void*)c.p, (void*)c.x, (void*)c.q,
                  (void*)c.partial, (void*)c.dres, (void*)c.dmaxv})
    F(p);
  for (auto& q : c.lv) {
    F(q.adiag);
    F(q.acoef);
    F(q.r);
    F(q.z);
  }
  if (c.h_pin) hipHostFree(c.h_pin);
  if (c.s) hipStreamDestroy(c.s);
}

// particles sorted by cell, id order inside a cell: start[] (ncell + 1), idx_s (the order)
int bin(Ctx2& c) {
  const size_t nc = ncell(c);
  hipLaunchKernelGGL(k_fill_i, dim3(grid_for(nc + 1)), dim3(kB), 0, c.s, c.count, nc + 1, 0);
  if (c.P == 0) {
    hipLaunchKernelGGL(k_fill_i, dim3(grid_for(nc + 1)), dim3(kB), 0, c.s, c.start, nc + 1, 0);
    return LH_OK;
  }
  hipLaunchKernelGGL(k_bin_key, dim3(grid_for((size_t)c.P)), dim3(kB), 0, c.s, c.pos, c.P, c.g, c.key, c.idx, c.count);
  size_t bytes = c.cub_bytes;
  if (!ok(c, hipcub::DeviceRadixSort::SortPairs(c.cub_tmp, bytes, c.key, c.key_s, c.idx, c.idx_s, (int)c.P, 0, 32, c.s),
          "radix sort"))
    return LH_ERR_DEVICE;
  bytes = c.cub_bytes;
  if (!ok(c, hipcub::DeviceScan::ExclusiveSum(c.cub_tmp, bytes, c.count, c.start, (int)(nc + 1), c.s), "scan"))
    return LH_ERR_DEVICE;
  return LH_OK;
}

double reduce(Ctx2& c, int nv, int nb, bool& good) {
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kB), 0, c.s, c.partial, nb, nv, c.dres);
  hipMemcpyAsync(c.h_pin, c.dres, nv * sizeof(double), hipMemcpyDeviceToHost, c.s);
  good = ok(c, hipStreamSynchronize(c.s), "reduce");
  return c.h_pin[0];
}

void smooth(Ctx2& c, int l, int phase) {
  Level& q = c.lv[l];
  hipLaunchKernelGGL(k_smooth, dim3(grid_for((size_t)q.d0 * q.d1)), dim3(kB), 0, c.s, q.z, q.r, q.adiag, q.acoef, q.d0,
                     q.d1, phase, c.omega, c.smoother);
}
// CGSolver<2>::ApplyPreconditioner, PressureSolver.cpp:225-278 (symmetric: z[0] = 0 first,
// post-smoothing and the second half of the coarsest sweeps in reverse colour order)
void precond(Ctx2& c) {
  Level& q0 = c.lv[0];
  const size_t n0 = ncell(c);
  if (c.precond != 2) {
    hipLaunchKernelGGL(k_precond_simple, dim3(grid_for(n0)), dim3(kB), 0, c.s, q0.z, q0.r, q0.adiag, c.phi, n0,
                       c.precond == 1 ? 1 : 0);
    return;
  }
  const int L = (int)c.lv.size();
  const bool sym = c.smoother == 1;
  if (sym) fill(c, q0.z, n0, 0.f);
  for (int l = 1; l < L; ++l) {
    for (int it = 0; it < c.num_pre; ++it) {
      smooth(c, l - 1, 0);
      smooth(c, l - 1, 1);
    }
    Level& f = c.lv[l - 1];
    Level& k = c.lv[l];
    hipLaunchKernelGGL(k_restrict, dim3(grid_for((size_t)k.d0 * k.d1)), dim3(kB), 0, c.s, k.r, k.z, f.r, f.z, f.adiag,
                       f.acoef, f.d0, f.d1, k.d0, k.d1);
  }
  for (int it = 0; it < c.num_final; ++it) {
    smooth(c, L - 1, 0);
    smooth(c, L - 1, 1);
  }
  if (sym)
    for (int it = 0; it < c.num_final; ++it) {
      smooth(c, L - 1, 1);
      smooth(c, L - 1, 0);
    }
  for (int l = L - 2; l >= 0; --l) {
    Level& f = c.lv[l];
    Level& k = c.lv[l + 1];
    hipLaunchKernelGGL(k_prolong, dim3(grid_for((size_t)k.d0 * k.d1)), dim3(kB), 0, c.s, f.z, k.z, f.d1, k.d0, k.d1);
    for (int it = 0; it < c.num_post; ++it) {
      smooth(c, l, sym ? 1 : 0);
      smooth(c, l, sym ? 0 : 1);
    }
  }
}

// CGSolver<2>::Solve, PressureSolver.cpp:329-416
int solve(Ctx2& c, float dt, lh_solve_stats* st) {
  const size_t n0 = ncell(c);
  const unsigned nb = grid_for(n0);
  hipLaunchKernelGGL(k_build_rhs, dim3(nb), dim3(kB), 0, c.s, c.rhs, c.phi, c.lwx, c.lwy, c.vx, c.vy, c.g);
  for (size_t l = 0; l < c.lv.size(); ++l) {
    Level& q = c.lv[l];
    hipLaunchKernelGGL(k_build_lhs, dim3(grid_for((size_t)q.d0 * q.d1)), dim3(kB), 0, c.s, q.adiag, q.acoef, c.phi,
                       c.lwx, c.lwy, c.g, (int)l, q.d0, q.d1, dt, c.density);
    fill(c, q.r, (size_t)q.d0 * q.d1, 0.f);
    fill(c, q.z, (size_t)q.d0 * q.d1, 0.f);
  }
  fill(c, c.p, n0, 0.f);
  fill(c, c.x, n0, 0.f);
  Level& q0 = c.lv[0];
  bool good = true;
  hipLaunchKernelGGL(k_cg_init, dim3(nb), dim3(kB), 0, c.s, q0.r, c.rhs, c.phi, n0, c.partial);
  reduce(c, 2, nb, good);
  if (!good) return LH_ERR_DEVICE;
  const int nliquid = (int)c.h_pin[1];
  precond(c);
  hipLaunchKernelGGL(k_cg_first, dim3(nb), dim3(kB), 0, c.s, c.p, q0.z, q0.r, c.phi, n0, c.partial);
  reduce(c, 2, nb, good);
  if (!good) return LH_ERR_DEVICE;
  float oldRZ = (float)c.h_pin[0];
  float residual = sqrtf((float)c.h_pin[1]) / (float)nliquid;
  const float res0 = residual;
  auto done = [&](float res) { return c.tol_mode == 0 ? res < c.tol : (double)res <= (double)c.tol * (double)res0; };
  int it = 0;
  bool conv = false, nan = false;
  for (it = 0; it < c.max_iter; ++it) {
    if (done(residual)) {
      conv = true;
      break;
    }
    hipLaunchKernelGGL(k_cg_spmv, dim3(nb), dim3(kB), 0, c.s, c.q, c.p, q0.adiag, q0.acoef, c.phi, q0.d0, q0.d1,
                       c.partial);
    const float denom = (float)reduce(c, 1, nb, good);
    if (!good) return LH_ERR_DEVICE;
    const float alpha = oldRZ / denom;
    if (std::isnan(alpha) || std::isinf(alpha)) {
      nan = true;
      break;
    }
    hipLaunchKernelGGL(k_cg_update_xr, dim3(nb), dim3(kB), 0, c.s, c.x, q0.r, c.p, c.q, c.phi, q0.d0, q0.d1, alpha,
                       c.partial);
    reduce(c, 2, nb, good);
    if (!good) return LH_ERR_DEVICE;
    residual = sqrtf((float)c.h_pin[0]) / (float)(int)c.h_pin[1];
    precond(c);
    hipLaunchKernelGGL(k_cg_dot, dim3(nb), dim3(kB), 0, c.s, q0.r, q0.z, n0, c.partial);
    const float newRZ = (float)reduce(c, 1, nb, good);
    if (!good) return LH_ERR_DEVICE;
    const float beta = newRZ / oldRZ;
    hipLaunchKernelGGL(k_cg_update_p, dim3(nb), dim3(kB), 0, c.s, c.p, q0.z, c.phi, n0, beta);
    oldRZ = newRZ;
  }
  if (!conv && !nan && it >= c.max_iter) conv = done(residual);
  hipMemcpyAsync(c.pressure, c.x, n0 * sizeof(float), hipMemcpyDeviceToDevice, c.s);
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = it;
    st->converged = conv ? 1 : 0;
    st->initial_residual = res0;
    st->final_residual = residual;
  }
  if (!ok(c, hipStreamSynchronize(c.s), "solve")) return LH_ERR_DEVICE;
  if (nan) {
    c.err = "PCG produced a NaN/Inf step length";
    return LH_ERR_NAN;
  }
  return LH_OK;
}

void extrapolate(Ctx2& c) {
  if (c.P == 0) return;
  const int vw = c.g.N1 + 1, vrows = c.g.N0 + 1;
  const size_t lds = (size_t)vrows * ((vw + 31) / 32) * sizeof(uint32_t);
  const bool in_lds = lds <= 64 * 1024;  // validity bits in LDS up to 64 KiB (grids to ~1400^2)
  // x component: x in [1, N0 - 1], y in [1, N1 - 2]; y component: x in [1, N0 - 2], y in [1, N1 - 1]
  for (int a = 0; a < 2; ++a) {
    float* v = a == 0 ? c.vx : c.vy;
    float* tv = a == 0 ? c.tvx : c.tvy;
    const float* lw = a == 0 ? c.lwx : c.lwy;
    const int d1 = a == 0 ? c.g.N1 : c.g.N1 + 1;
    const int xhi = a == 0 ? c.g.N0 - 1 : c.g.N0 - 2, yhi = a == 0 ? c.g.N1 - 2 : c.g.N1 - 1;
    if (in_lds)
      hipLaunchKernelGGL(k_extrapolate<true>, dim3(1), dim3(1024), lds, c.s, v, tv, c.valid, lw, d1, vw, vrows, xhi,
                         yhi);
    else
      hipLaunchKernelGGL(k_extrapolate<false>, dim3(1), dim3(1024), 0, c.s, v, tv, c.valid, lw, d1, vw, vrows, xhi,
                         yhi);
    if (a == 0) std::swap(c.vx, c.tvx);  // five sweeps: the result is in the other buffer
    else std::swap(c.vy, c.tvy);
  }
}

void clear(Ctx2& c) {  // Grid<2>::Clear, Grid.cpp:52-61
  fill(c, c.pressure, ncell(c), 0.f);
  fill(c, c.phi, ncell(c), 3 * c.g.dx0);
  for (float* a : {c.vx, c.tvx}) fill(c, a, nfx(c), 0.f);
  for (float* a : {c.vy, c.tvy}) fill(c, a, nfy(c), 0.f);
}

int liquid_phi(Ctx2& c) {
  int rc = bin(c);
  if (rc) return rc;
  hipLaunchKernelGGL(k_liquid_phi, dim3(grid_for(ncell(c))), dim3(kB), 0, c.s, c.phi, c.pos, c.idx_s, c.start, c.g,
                     c.g.dx0);
  return LH_OK;
}
int p2g(Ctx2& c, float gdt) {
  int rc = bin(c);
  if (rc) return rc;
  hipLaunchKernelGGL(k_p2g, dim3(grid_for(nfx(c) + nfy(c))), dim3(kB), 0, c.s, c.vx, c.vy, c.pos, c.vel, c.C, c.idx_s,
                     c.start, c.g, gdt);
  return LH_OK;
}
int advect(Ctx2& c, float dt) {
  hipMemsetAsync(c.dmaxv, 0, sizeof(unsigned), c.s);
  if (c.P > 0)
    hipLaunchKernelGGL(k_advect, dim3(grid_for((size_t)c.P)), dim3(kB), 0, c.s, c.pos, c.vel, c.solid, c.P, c.g, dt,
                       c.dmaxv);
  unsigned mv = 0;
  hipMemcpyAsync(c.h_pin, c.dmaxv, sizeof(unsigned), hipMemcpyDeviceToHost, c.s);
  if (!ok(c, hipStreamSynchronize(c.s), "advect")) return LH_ERR_DEVICE;
  std::memcpy(&mv, c.h_pin, sizeof(unsigned));
  std::memcpy(&c.last_maxv, &mv, sizeof(float));
  return LH_OK;
}
int correct_volume(Ctx2& c, float dt) {
  if (c.P == 0) {  // no particles: only the cycle advances (LiquidModel.cpp:464)
    c.correct_step = (c.correct_step + 1) % c.correct_cycle;
    c.correct_count++;
    return LH_OK;
  }
  if (c.P > (int64_t)(INT32_MAX / 2)) {
    c.err = "CorrectVolume: the hash table of 2 x NumParticles slots exceeds a 32-bit index";
    return LH_ERR_DOMAIN;
  }
  int rc = bin(c);  // the hash table of the advected positions (AdvectParticle rebuilds it)
  if (rc) return rc;
  const int step = c.correct_step;
  hipLaunchKernelGGL(k_correct_volume, dim3(grid_for((size_t)c.P)), dim3(kB), 0, c.s, c.cvbuf, c.pos, c.idx_s, c.start,
                     c.solid, c.P, c.g, c.g.dx0, dt, c.correct_cycle, step, c.seed, c.correct_count);
  hipLaunchKernelGGL(k_cv_commit, dim3(grid_for((size_t)c.P)), dim3(kB), 0, c.s, c.pos, c.cvbuf, c.P, c.correct_cycle,
                     step);
  c.correct_step = (step + 1) % c.correct_cycle;
  c.correct_count++;
  return LH_OK;
}

int set_particles(Ctx2& c, const float* pos, const float* vel, const float* C, int64_t n) {
  int rc = ensure_particles(c, n);
  if (rc) return rc;
  c.P = n;
  if (n > 0) {
    hipMemcpyAsync(c.pos, pos, 2 * n * sizeof(float), hipMemcpyHostToDevice, c.s);
    if (vel) hipMemcpyAsync(c.vel, vel, 2 * n * sizeof(float), hipMemcpyHostToDevice, c.s);
    else fill(c, c.vel, 2 * n, 0.f);
    if (C) hipMemcpyAsync(c.C, C, 4 * n * sizeof(float), hipMemcpyHostToDevice, c.s);
    else fill(c, c.C, 4 * n, 0.f);
  }
  return ok(c, hipStreamSynchronize(c.s), "set particles") ? LH_OK : LH_ERR_DEVICE;
}

// ParticleSystem<2>::SpawnParticles (ParticleSystem.cpp:23-42) at LiquidModel<2>::Init's box
// (:44-49): spacing dx / 2 from min + 0.3 (max - min), at most numParticles (the break leaves
// only the y loop)
void spawn(const Ctx2& c, std::vector<float>& pos) {
  const float rad = c.g.dx0;
  const float lo0 = c.g.mn0 + 0.3f * (c.g.mx0 - c.g.mn0), hi0 = c.g.mx0 - 0.3f * (c.g.mx0 - c.g.mn0);
  const float lo1 = c.g.mn1 + 0.3f * (c.g.mx1 - c.g.mn1), hi1 = c.g.mx1 - 0.3f * (c.g.mx1 - c.g.mn1);
  const float interval = (float)(0.5 * (double)rad);
  const int nx = (int)((hi0 - lo0) / interval), ny = (int)((hi1 - lo1) / interval);
  const int64_t n = std::min<int64_t>((int64_t)nx * ny, c.num_particles);
  pos.assign(2 * n, 0.f);
  for (int xi = 0; xi < nx; ++xi)
    for (int yi = 0; yi < ny; ++yi) {
      const int64_t index = (int64_t)xi * ny + yi;
      if (index >= n) break;
      pos[2 * index] = (float)xi * interval + lo0;
      pos[2 * index + 1] = (float)yi * interval + lo1;
    }
}

}  // namespace lh2

struct lh2_ctx {
  lh2::Ctx2 c;
};

namespace {
template <class F>
int guarded2(lh2_ctx* ctx, F&& f) {
  if (!ctx) return LH_ERR_ARG;
  try {
    return f(ctx->c);
  } catch (const std::exception& e) {
    ctx->c.err = e.what();
    return LH_ERR_CONFIG;
  }
}
size_t field_count(const lh2::Ctx2& c, int f) {
  switch (f) {
    case LH_FIELD_VX: case LH_FIELD_TEMPVX: case LH_FIELD_LIQUID_WX: return lh2::nfx(c);
    case LH_FIELD_VY: case LH_FIELD_TEMPVY: case LH_FIELD_LIQUID_WY: return lh2::nfy(c);
    case LH_FIELD_PRESSURE: case LH_FIELD_LIQUID_PHI: case LH_FIELD_RHS: return lh2::ncell(c);
    case LH_FIELD_SOLID_PHI: case LH_FIELD_IS_VALID: return lh2::nnode(c);
    default: return 0;
  }
}
void* field_ptr(lh2::Ctx2& c, int f) {
  switch (f) {
    case LH_FIELD_VX: return c.vx;
    case LH_FIELD_VY: return c.vy;
    case LH_FIELD_TEMPVX: return c.tvx;
    case LH_FIELD_TEMPVY: return c.tvy;
    case LH_FIELD_LIQUID_WX: return c.lwx;
    case LH_FIELD_LIQUID_WY: return c.lwy;
    case LH_FIELD_PRESSURE: return c.pressure;
    case LH_FIELD_LIQUID_PHI: return c.phi;
    case LH_FIELD_RHS: return c.rhs;
    case LH_FIELD_SOLID_PHI: return c.solid;
    case LH_FIELD_IS_VALID: return c.valid;
    default: return nullptr;
  }
}
}  // namespace

extern "C" {

int lh2_create_from_string(const char* json_text, lh2_ctx** out) {
  if (!out || !json_text) return LH_ERR_ARG;
  *out = nullptr;
  lh2_ctx* ctx = new lh2_ctx();
  int rc = guarded2(ctx, [&](lh2::Ctx2& c) {
    lh2::parse_config(c, json_text);
    return lh2::alloc(c);
  });
  if (rc) {
    lh2::free_all(ctx->c);
    delete ctx;
    return rc;
  }
  *out = ctx;
  return LH_OK;
}
int lh2_create(const char* json_path, lh2_ctx** out) {
  if (!json_path) return LH_ERR_ARG;
  std::ifstream f(json_path);
  if (!f) return LH_ERR_CONFIG;
  std::stringstream ss;
  ss << f.rdbuf();
  return lh2_create_from_string(ss.str().c_str(), out);
}
void lh2_destroy(lh2_ctx* ctx) {
  if (!ctx) return;
  lh2::free_all(ctx->c);
  delete ctx;
}
const char* lh2_last_error(const lh2_ctx* ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

int lh2_compute_solid_phi(lh2_ctx* ctx) {
  return guarded2(ctx, [](lh2::Ctx2& c) {
    hipLaunchKernelGGL(lh2::k_solid_phi, dim3(lh2::grid_for(lh2::nnode(c))), dim3(lh2::kB), 0, c.s, c.solid, c.d_sdf,
                       (int)(c.sdf.size() / 12), c.g);
    hipLaunchKernelGGL(lh2::k_face_w, dim3(lh2::grid_for(lh2::nfx(c) + lh2::nfy(c))), dim3(lh2::kB), 0, c.s, c.lwx,
                       c.lwy, c.solid, c.g);
    return lh2::ok(c, hipStreamSynchronize(c.s), "solid phi") ? LH_OK : LH_ERR_DEVICE;
  });
}
int lh2_set_boundary(lh2_ctx* ctx, const char* json) {
  return guarded2(ctx, [json](lh2::Ctx2& c) {
    if (!json) return LH_ERR_ARG;
    std::vector<float> prog;
    lh2::emit_sdf(lh::Json::parse(json), prog);
    if (prog.empty() || prog.size() / 12 > (size_t)lh2::kMaxNodes) {
      c.err = "Boundary tree too large";
      return LH_ERR_CONFIG;
    }
    c.sdf = prog;
    hipMemcpyAsync(c.d_sdf, c.sdf.data(), c.sdf.size() * sizeof(float), hipMemcpyHostToDevice, c.s);
    return lh2::ok(c, hipStreamSynchronize(c.s), "boundary") ? LH_OK : LH_ERR_DEVICE;
  });
}
// LiquidModel<2>::Init, LiquidModel.cpp:44-49
int lh2_init(lh2_ctx* ctx) {
  int rc = lh2_compute_solid_phi(ctx);
  if (rc) return rc;
  return guarded2(ctx, [](lh2::Ctx2& c) {
    std::vector<float> pos;
    lh2::spawn(c, pos);
    c.time = 0.0;
    return lh2::set_particles(c, pos.data(), nullptr, nullptr, (int64_t)(pos.size() / 2));
  });
}
int lh2_clear_grid(lh2_ctx* ctx) {
  return guarded2(ctx, [](lh2::Ctx2& c) {
    lh2::clear(c);
    return LH_OK;
  });
}
int lh2_compute_liquid_phi(lh2_ctx* ctx) {
  return guarded2(ctx, [](lh2::Ctx2& c) { return lh2::liquid_phi(c); });
}
// LiquidModel<2>::PreStep, LiquidModel.cpp:51-55
int lh2_pre_step(lh2_ctx* ctx, float dt) {
  (void)dt;
  return guarded2(ctx, [](lh2::Ctx2& c) {
    lh2::clear(c);
    return lh2::liquid_phi(c);
  });
}
int lh2_p2g(lh2_ctx* ctx) {
  return guarded2(ctx, [](lh2::Ctx2& c) { return lh2::p2g(c, 0.f); });
}
int lh2_apply_gravity(lh2_ctx* ctx, float dt) {
  return guarded2(ctx, [dt](lh2::Ctx2& c) {
    hipLaunchKernelGGL(lh2::k_gravity, dim3(lh2::grid_for(lh2::nfy(c))), dim3(lh2::kB), 0, c.s, c.vy, c.g,
                       c.gravity * dt);
    return LH_OK;
  });
}
int lh2_solve_pressure(lh2_ctx* ctx, float dt, lh_solve_stats* st) {
  return guarded2(ctx, [=](lh2::Ctx2& c) { return lh2::solve(c, dt, st); });
}
int lh2_velocity_update(lh2_ctx* ctx, float dt) {
  return guarded2(ctx, [dt](lh2::Ctx2& c) {
    hipLaunchKernelGGL(lh2::k_velocity_update, dim3(lh2::grid_for(lh2::ncell(c))), dim3(lh2::kB), 0, c.s, c.vx, c.vy,
                       c.pressure, c.phi, c.lwx, c.lwy, c.g, dt, c.density);
    return LH_OK;
  });
}
int lh2_extrapolate(lh2_ctx* ctx) {
  return guarded2(ctx, [](lh2::Ctx2& c) {
    lh2::extrapolate(c);
    return LH_OK;
  });
}
int lh2_g2p(lh2_ctx* ctx) {
  return guarded2(ctx, [](lh2::Ctx2& c) {
    if (c.P > 0)
      hipLaunchKernelGGL(lh2::k_g2p, dim3(lh2::grid_for((size_t)c.P)), dim3(lh2::kB), 0, c.s, c.vel, c.C, c.pos, c.vx,
                         c.vy, c.P, c.g, c.g.dx0);
    return LH_OK;
  });
}
int lh2_advect(lh2_ctx* ctx, float dt) {
  return guarded2(ctx, [dt](lh2::Ctx2& c) { return lh2::advect(c, dt); });
}
int lh2_correct_volume(lh2_ctx* ctx, float dt) {
  return guarded2(ctx, [dt](lh2::Ctx2& c) { return lh2::correct_volume(c, dt); });
}
// LiquidModel<2>::Step, LiquidModel.cpp:57-66: P2G (+ gravity), projection, Extrapolate, G2P,
// AdvectParticle, CorrectVolume; the clock advances by dt
int lh2_step(lh2_ctx* ctx, float dt, lh_solve_stats* st) {
  return guarded2(ctx, [=](lh2::Ctx2& c) {
    if (st) std::memset(st, 0, sizeof(*st));
    int rc = lh2::p2g(c, c.gravity * dt);
    if (rc) return rc;
    if (c.P > 0) {  // SolvePressure returns early without particles (:240)
      rc = lh2::solve(c, dt, st);
      if (rc) return rc;
      hipLaunchKernelGGL(lh2::k_velocity_update, dim3(lh2::grid_for(lh2::ncell(c))), dim3(lh2::kB), 0, c.s, c.vx, c.vy,
                         c.pressure, c.phi, c.lwx, c.lwy, c.g, dt, c.density);
    }
    lh2::extrapolate(c);
    if (c.P > 0)
      hipLaunchKernelGGL(lh2::k_g2p, dim3(lh2::grid_for((size_t)c.P)), dim3(lh2::kB), 0, c.s, c.vel, c.C, c.pos, c.vx,
                         c.vy, c.P, c.g, c.g.dx0);
    rc = lh2::advect(c, dt);
    if (rc) return rc;
    rc = lh2::correct_volume(c, dt);
    if (rc) return rc;
    c.time += (double)dt;
    return lh2::ok(c, hipStreamSynchronize(c.s), "step") ? LH_OK : LH_ERR_DEVICE;
  });
}
// GetCFL, LiquidModel.h:49-55
int lh2_get_cfl(lh2_ctx* ctx, float* dt) {
  return guarded2(ctx, [dt](lh2::Ctx2& c) {
    if (!dt) return LH_ERR_ARG;
    *dt = lh::clampf(c.g.dx0 / c.last_maxv, 0.001f, 0.1f);
    return LH_OK;
  });
}
// n x {PreStep, Step, PostStep, dt = GetCFL()}; dt in/out; the last step's solve statistics
int lh2_run_steps(lh2_ctx* ctx, int n, float* dt, lh_solve_stats* st) {
  if (!ctx || !dt) return LH_ERR_ARG;
  for (int k = 0; k < n; ++k) {
    int rc = lh2_pre_step(ctx, *dt);
    if (rc) return rc;
    rc = lh2_step(ctx, *dt, st);
    if (rc) return rc;
    rc = lh2_get_cfl(ctx, dt);
    if (rc) return rc;
  }
  return LH_OK;
}
int lh2_divergence(lh2_ctx* ctx, double* mean, double* linf, int* count) {
  return guarded2(ctx, [=](lh2::Ctx2& c) {
    const unsigned nb = lh2::grid_for(lh2::ncell(c));
    hipMemsetAsync(c.dmaxv + 1, 0, sizeof(unsigned), c.s);
    hipLaunchKernelGGL(lh2::k_divergence, dim3(nb), dim3(lh2::kB), 0, c.s, c.phi, c.lwx, c.lwy, c.vx, c.vy, c.g,
                       c.partial, c.dmaxv + 1);
    bool good = true;
    lh2::reduce(c, 2, nb, good);
    if (!good) return LH_ERR_DEVICE;
    const double sum = c.h_pin[0], cnt = c.h_pin[1];
    unsigned m = 0;
    hipMemcpy(&m, c.dmaxv + 1, sizeof(unsigned), hipMemcpyDeviceToHost);
    float mf;
    std::memcpy(&mf, &m, sizeof(float));
    if (mean) *mean = cnt > 0 ? sum / cnt : 0.0;
    if (linf) *linf = mf;
    if (count) *count = (int)cnt;
    return LH_OK;
  });
}
int lh2_num_particles(const lh2_ctx* ctx, int64_t* n) {
  if (!ctx || !n) return LH_ERR_ARG;
  *n = ctx->c.P;
  return LH_OK;
}
int lh2_set_particles(lh2_ctx* ctx, const float* pos, const float* vel, const float* C, int64_t n) {
  return guarded2(ctx, [=](lh2::Ctx2& c) {
    if (n < 0 || (n > 0 && !pos)) return LH_ERR_ARG;
    return lh2::set_particles(c, pos, vel, C, n);
  });
}
int lh2_get_particles(lh2_ctx* ctx, float* pos, float* vel, float* C, int64_t n) {
  return guarded2(ctx, [=](lh2::Ctx2& c) {
    if (n != c.P) {
      c.err = "lh2_get_particles: n does not match the particle count";
      return LH_ERR_ARG;
    }
    if (n == 0) return LH_OK;
    if (pos) hipMemcpyAsync(pos, c.pos, 2 * n * sizeof(float), hipMemcpyDeviceToHost, c.s);
    if (vel) hipMemcpyAsync(vel, c.vel, 2 * n * sizeof(float), hipMemcpyDeviceToHost, c.s);
    if (C) hipMemcpyAsync(C, c.C, 4 * n * sizeof(float), hipMemcpyDeviceToHost, c.s);
    return lh2::ok(c, hipStreamSynchronize(c.s), "get particles") ? LH_OK : LH_ERR_DEVICE;
  });
}
int lh2_get_field(lh2_ctx* ctx, int field, void* host, size_t count) {
  return guarded2(ctx, [=](lh2::Ctx2& c) {
    const size_t n = field_count(c, field);
    if (n == 0 || n != count || !host) {
      c.err = "lh2_get_field: bad field id or element count";
      return LH_ERR_ARG;
    }
    const size_t bytes = field == LH_FIELD_IS_VALID ? n : n * sizeof(float);
    hipMemcpyAsync(host, field_ptr(c, field), bytes, hipMemcpyDeviceToHost, c.s);
    return lh2::ok(c, hipStreamSynchronize(c.s), "get field") ? LH_OK : LH_ERR_DEVICE;
  });
}
int lh2_set_field(lh2_ctx* ctx, int field, const void* host, size_t count) {
  return guarded2(ctx, [=](lh2::Ctx2& c) {
    const size_t n = field_count(c, field);
    if (n == 0 || n != count || !host) {
      c.err = "lh2_set_field: bad field id or element count";
      return LH_ERR_ARG;
    }
    const size_t bytes = field == LH_FIELD_IS_VALID ? n : n * sizeof(float);
    hipMemcpyAsync(field_ptr(c, field), host, bytes, hipMemcpyHostToDevice, c.s);
    return lh2::ok(c, hipStreamSynchronize(c.s), "set field") ? LH_OK : LH_ERR_DEVICE;
  });
}
int lh2_set_time(lh2_ctx* ctx, double t) {
  if (!ctx) return LH_ERR_ARG;
  ctx->c.time = t;
  return LH_OK;
}
int lh2_get_time(const lh2_ctx* ctx, double* t) {
  if (!ctx || !t) return LH_ERR_ARG;
  *t = ctx->c.time;
  return LH_OK;
}
}  // extern "C"
