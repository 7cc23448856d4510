// lh_oracle.cpp -- CPU oracle for the liquid step.  TEST INFRASTRUCTURE ONLY.
//
// A line-by-line FP32 restatement of the reference's 3D liquid path
// (LiquidModel.cpp:478-1083, PressureSolver.cpp:422-934, Grid.cpp:86-192,
// Array.h:67-126, Kernel.h:12-31, MathUtils.h:8-34, MathUtils.cpp:29-35,75-88,
// Boundary.h:27-94, ParticleSystem.cpp:6-64, LiquidModel.h:49-55).  It is the
// checker of the HIP library, never the thing measured or shipped: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Determinism choices (the reference's OpenMP code is racy or order-free):
//  * CGSolver denom / residual sums get a proper reduction (the reference has
//    no reduction clause, PressureSolver.cpp:874-898).
//  * Reduce / denom / residual: dot_mode 0 sums in FP32 in index order
//    (a serial run of Array.h:111-122); dot_mode 1 sums in FP64 and rounds the
//    total to float, which is what the HIP library does.
//  * P2G: p2g_order 0 adds particle contributions in particle-index order
//    (a serial run of the atomics at LiquidModel.cpp:640-688); p2g_order 1 adds
//    them per face in half-cell-bin order (bins x-major, z fastest; particles in
//    a bin by ascending index), the order of the HIP gather.
//  * Extrapolate runs sequentially in (x,y,z) order with in-place validity
//    (LiquidModel.cpp:953-978), which fixes the std::vector<bool> race.
//  * Eigen fixed-size reductions of length 3 evaluate as x0 + (x1 + x2)
//    (Eigen Redux.h redux_novec_unroller splits Length=3 as 1 + 2); this is
//    used for norm(), dot() and maxCoeff().
//
// Parity status per function is listed in the header comment of each section
// ("pinned by:").  Built with -ffp-contract=off: MSVC /fp:precise forms no FMA.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using i64 = long long;

// ---------------------------------------------------------------- params
}  // namespace

extern "C" {
struct orc_params {
  int N[3];
  float mn[3], mx[3];
  float gravity, density;
  int max_iter;
  float tol;
  int tol_mode;   // 0 absolute (reference), 1 relative to ||b||
  int precond;    // 0 None, 1 Jacobi, 2 Multigrid
  int num_level, num_pre, num_post, num_final;
  float omega;
  int dot_mode;   // 0 fp32 serial, 1 fp64
  int smoother;   // 0 reference (Q4/Q5), 1 symmetric
  int p2g_order;  // 0 particle order, 1 half-cell bin order
  float tank_amp, tank_freq;
  int tank_axis;  // -1: none
  int nsdf;       // number of SDF program nodes (12 floats each)
  const float* sdf;
  int correct_volume;  // 1: CorrectVolume after AdvectParticle (commented out of the reference's 3D step)
  int correct_cycle;   // LiquidParameter.correctCycle
  unsigned long long seed;
  long long num_particles;  // LiquidParameter.numParticles (the 2D lattice spawn's cap)
  int coarsest_fp64;        // 1: the multigrid coarsest level's sweeps in FP64 (refined mode's default)
  int coarsest_cheb;        // > 0: the coarsest level solved by that many Chebyshev steps instead of sweeps
  float cheb_lo;            // the Chebyshev interval's lower end (the upper end is 2)
};
struct orc_solve_stats {
  int iterations, converged;
  double initial_residual, final_residual, rhs_norm;
};
}

namespace {

// ------------------------------------------------------------ Array3
// Array.h:67-126: x-major, z fastest.
struct A3 {
  int d[3] = {1, 1, 1};
  std::vector<float> v;
  void resize(int x, int y, int z) { d[0] = x; d[1] = y; d[2] = z; v.assign((size_t)x * y * z, 0.f); }
  inline size_t id(int x, int y, int z) const { return ((size_t)x * d[1] + y) * d[2] + z; }
  inline float& operator()(int x, int y, int z) { return v[id(x, y, z)]; }
  inline float operator()(int x, int y, int z) const { return v[id(x, y, z)]; }
  void fill(float a) { std::fill(v.begin(), v.end(), a); }
};

// ------------------------------------------------------------ L0 numerics
// Kernel.h:12-17 / 26-31.  pinned by: closed forms (w(0)=.75, w(.5)=.5,
// w(1)=.125, w(1.5)=0, partition of unity, first moment) in tests.
inline float quad_w(float d) {
  d = std::fabs(d);
  if (d < 0.5f) return 0.75f - d * d;
  else if (d < 1.5f) return 0.5f * (1.5f - d) * (1.5f - d);
  return 0.0f;
}
inline float quad_gw(float d) {
  d = std::fabs(d);
  if (d < 0.5f) return -2.f * d;
  else if (d < 1.5f) return d - 1.5f;
  return 0.f;
}
// MathUtils.h:8-34
const float kEps = 1e-8f;  // Definition.h:93
inline bool is_small(float v) { return std::fabs(v) < kEps; }
inline float clampf(float v, float lo, float hi) {  // min(maxVal, max(minVal, val)) with the windows.h macros
  float t = (lo > v) ? lo : v;
  return (hi < t) ? hi : t;
}
inline float lerp(float a, float b, float f) { return a + clampf(f, 0.0f, 1.0f) * (b - a); }
inline float bilerp(float v00, float v10, float v01, float v11, float f1, float f2) {
  float a = lerp(v00, v10, f1);
  float b = lerp(v01, v11, f1);
  return lerp(a, b, clampf(f2, 0.0f, 1.0f));
}
inline float trilerp(float v000, float v100, float v010, float v110, float v001, float v101, float v011,
                     float v111, float f1, float f2, float f3) {
  return lerp(bilerp(v000, v100, v010, v110, f1, f2), bilerp(v001, v101, v011, v111, f1, f2),
              clampf(f3, 0.0f, 1.0f));
}
// MathUtils.cpp:29-35.  pinned by: closed-form table in tests.
inline float fraction_in_region(float l, float r) {
  float ans = 0.0f;
  if (l < 0.0f && r < 0.0f) ans = 1.0f;
  else if (l < 0.0f && 0.0f < r) ans = l / (l - r);
  else if (l > 0.0f && 0.0f > r) ans = r / (r - l);
  return ans;
}
// max(0.1, frac) with the windows.h macro: compared in double, result stored as float.
inline float max_tenth(float frac) { return (0.1 > (double)frac) ? (float)0.1 : frac; }
// MathUtils.cpp:75-88
inline void barycentric(float x, int lo, int hi, int& idx, float& frac) {
  int i = (int)std::floor(x);
  float f = x - (float)i;
  if (i < lo) { i = lo; f = 0.0f; }
  else if (i > hi - 2) { i = hi - 2; f = 1.0f; }
  idx = i; frac = f;
}
// Eigen length-3 reductions (see file header).
inline float sum3(float a, float b, float c) { return a + (b + c); }
inline float norm3(float a, float b, float c) { return std::sqrt(a * a + (b * b + c * c)); }
inline float max3(float a, float b, float c) {
  float bc = (b < c) ? c : b;
  return (a < bc) ? bc : a;
}

// ------------------------------------------------------------ SDF program
// Boundary.h:27-94 as a postfix program (12 floats per node):
//   [kind, cx, cy, cz, hx, hy, hz, radius, useInside, 0, 0, 0]
//   kind 0 Rect leaf, 1 Sphere leaf, 10 Union(min), 11 Intersection(max),
//   12 Subtract max(a,-b), 13 None (identity on the single child).
// pinned by: box-SDF closed forms in tests.
float sdf_eval(const std::vector<float>& prog, float px, float py, float pz) {
  float st[32];
  int sp = 0;
  size_t n = prog.size() / 12;
  for (size_t k = 0; k < n; ++k) {
    const float* q = &prog[12 * k];
    int kind = (int)q[0];
    if (kind == 0) {  // RectBoundary::Compute, Boundary.h:83-90
      float ax = std::fabs(px - q[1]) - q[4];
      float ay = std::fabs(py - q[2]) - q[5];
      float az = std::fabs(pz - q[3]) - q[6];
      float mx = ax > 0.f ? ax : 0.f, my = ay > 0.f ? ay : 0.f, mz = az > 0.f ? az : 0.f;
      float outside = norm3(mx, my, mz);
      float mc = max3(ax, ay, az);
      float inside = (mc < 0.f) ? mc : 0.f;
      st[sp++] = (q[8] != 0.f) ? -outside - inside : outside + inside;
    } else if (kind == 1) {  // sphere: |p - c| - r (negated with useInside)
      float d = norm3(px - q[1], py - q[2], pz - q[3]) - q[7];
      st[sp++] = (q[8] != 0.f) ? -d : d;
    } else if (kind == 13) {
      // None: returns the (single) child value unchanged
    } else {
      float b = st[--sp];
      float a = st[--sp];
      float r;
      if (kind == 10) r = (a < b) ? a : b;
      else if (kind == 11) r = (a > b) ? a : b;
      else { float nb = -b; r = (a > nb) ? a : nb; }
      st[sp++] = r;
    }
  }
  return sp > 0 ? st[sp - 1] : 1.0f;
}

// ------------------------------------------------------------ model state
struct Level {
  int d[3];
  A3 r, z, adiag;
  std::vector<float> acoef;  // 6 per cell: right, left, top, bottom, front, back
};

struct Model {
  orc_params prm;
  std::vector<float> sdf;
  int N[3];
  float mn[3], mx[3], dx[3];
  A3 vx, vy, vz, tvx, tvy, tvz, wx, wy, wz, lwx, lwy, lwz, pressure, liquidPhi, solidPhi, rhs, p, x;
  std::vector<uint8_t> isValid;  // (N+1)^3, Grid.h:160, never cleared by Clear()
  std::vector<float> pos, vel, C, radius;
  i64 P = 0;
  std::vector<Level> lv;
  int max_level = 1;
  double time = 0.0;
  int correct_step = 0;      // m_correctStep (LiquidModel.h:73)
  uint64_t correct_count = 0;  // CorrectVolume calls so far (random-stream counter)
  std::string err;
};

inline size_t vid(const Model& m, int x, int y, int z) {
  return ((size_t)x * (m.N[1] + 1) + y) * (m.N[2] + 1) + z;
}

// Grid<3> ctor, Grid.cpp:86-98; CGSolver<3> ctor, PressureSolver.cpp:422-447.
void model_alloc(Model& m) {
  const int* N = m.N;
  m.vx.resize(N[0] + 1, N[1], N[2]); m.vy.resize(N[0], N[1] + 1, N[2]); m.vz.resize(N[0], N[1], N[2] + 1);
  m.tvx.resize(N[0] + 1, N[1], N[2]); m.tvy.resize(N[0], N[1] + 1, N[2]); m.tvz.resize(N[0], N[1], N[2] + 1);
  m.wx.resize(N[0] + 1, N[1], N[2]); m.wy.resize(N[0], N[1] + 1, N[2]); m.wz.resize(N[0], N[1], N[2] + 1);
  m.lwx.resize(N[0] + 1, N[1], N[2]); m.lwy.resize(N[0], N[1] + 1, N[2]); m.lwz.resize(N[0], N[1], N[2] + 1);
  m.pressure.resize(N[0], N[1], N[2]); m.liquidPhi.resize(N[0], N[1], N[2]);
  m.solidPhi.resize(N[0] + 1, N[1] + 1, N[2] + 1);
  m.rhs.resize(N[0], N[1], N[2]); m.p.resize(N[0], N[1], N[2]); m.x.resize(N[0], N[1], N[2]);
  m.isValid.assign((size_t)(N[0] + 1) * (N[1] + 1) * (N[2] + 1), 0);
  for (int a = 0; a < 3; ++a) {
    float interval = m.mx[a] - m.mn[a];
    m.dx[a] = (float)(interval / (float)N[a]);
  }
  int L = std::max(1, m.prm.num_level);
  if (m.prm.precond != 2) L = 1;
  m.max_level = L;
  m.lv.resize(L);
  for (int l = 0; l < L; ++l) {
    Level& q = m.lv[l];
    for (int a = 0; a < 3; ++a) q.d[a] = N[a] / (1 << l);
    q.r.resize(q.d[0], q.d[1], q.d[2]); q.z.resize(q.d[0], q.d[1], q.d[2]);
    q.adiag.resize(q.d[0], q.d[1], q.d[2]);
    q.acoef.assign((size_t)q.d[0] * q.d[1] * q.d[2] * 6, 0.f);
  }
}

// ------------------------------------------------------------ Grid sampling
// Grid<3>::GetSolidPhi / GetSolidGradient, Grid.cpp:128-155, 175-192.
// pinned by: linear-field exactness and edge-clamp cases in tests.
inline bool in_grid(const Model& m, float x, float y, float z) {
  return m.mn[0] <= x && m.mn[1] <= y && m.mn[2] <= z && x <= m.mx[0] && y <= m.mx[1] && z <= m.mx[2];
}
void solid_gradient(const Model& m, float x, float y, float z, float g[3]);
float solid_phi_at(const Model& m, float x, float y, float z) {
  if (!in_grid(m, x, y, z)) {
    float vx = clampf(x, m.mn[0] + m.dx[0], m.mx[0] - m.dx[0]);
    float vy = clampf(y, m.mn[1] + m.dx[1], m.mx[1] - m.dx[1]);
    float vz = clampf(z, m.mn[2] + m.dx[2], m.mx[2] - m.dx[2]);
    float g[3];
    solid_gradient(m, vx, vy, vz, g);
    return solid_phi_at(m, vx, vy, vz) + g[0] * (x - vx) + g[1] * (y - vy) + g[2] * (z - vz);
  }
  float lx = (x - m.mn[0]) / m.dx[0];
  float ly = (y - m.mn[1]) / m.dx[1];
  float lz = (z - m.mn[2]) / m.dx[2];
  int xi, yi, zi;
  float fx, fy, fz;
  barycentric(lx, 0, m.N[0] + 1, xi, fx);
  barycentric(ly, 0, m.N[1] + 1, yi, fy);
  barycentric(lz, 0, m.N[2] + 1, zi, fz);
  const A3& s = m.solidPhi;
  return trilerp(s(xi, yi, zi), s(xi + 1, yi, zi), s(xi, yi + 1, zi), s(xi + 1, yi + 1, zi), s(xi, yi, zi + 1),
                 s(xi + 1, yi, zi + 1), s(xi, yi + 1, zi + 1), s(xi + 1, yi + 1, zi + 1), fx, fy, fz);
}
void solid_gradient(const Model& m, float x, float y, float z, float g[3]) {
  if (!in_grid(m, x, y, z)) {
    float vx = clampf(x, m.mn[0] + m.dx[0], m.mx[0] - m.dx[0]);
    float vy = clampf(y, m.mn[1] + m.dx[1], m.mx[1] - m.dx[1]);
    float vz = clampf(z, m.mn[2] + m.dx[2], m.mx[2] - m.dx[2]);
    solid_gradient(m, vx, vy, vz, g);
    return;
  }
  const float eps = 1e-4f;
  float left = solid_phi_at(m, x - eps, y, z);
  float right = solid_phi_at(m, x + eps, y, z);
  float bottom = solid_phi_at(m, x, y - eps, z);
  float top = solid_phi_at(m, x, y + eps, z);
  float front = solid_phi_at(m, x, y, z + eps);
  float back = solid_phi_at(m, x, y, z - eps);
  const float two_eps = (float)(2.0 * (double)eps);
  float a = (right - left) / two_eps, b = (top - bottom) / two_eps, c = (front - back) / two_eps;
  float n2 = a * a + (b * b + c * c);  // Eigen normalized(): v / sqrt(squaredNorm) if > 0
  if (n2 > 0.f) {
    float n = std::sqrt(n2);
    a = a / n; b = b / n; c = c / n;
  }
  g[0] = a; g[1] = b; g[2] = c;
}

// ------------------------------------------------------------ stages
// LiquidModel<3>::ComputeSolidPhi, LiquidModel.cpp:562-611 (edge-based face
// fractions, Q2).  pinned by: box SDF closed forms; W=0 inside walls, 1 in the
// open region, fractional on cut edges, checked in tests.
void compute_solid_phi(Model& m) {
  const int* N = m.N;
#pragma omp parallel for
  for (int xi = 0; xi < N[0] + 1; ++xi)
    for (int yi = 0; yi < N[1] + 1; ++yi)
      for (int zi = 0; zi < N[2] + 1; ++zi) {
        float px = (float)xi * m.dx[0] + m.mn[0];
        float py = (float)yi * m.dx[1] + m.mn[1];
        float pz = (float)zi * m.dx[2] + m.mn[2];
        m.solidPhi(xi, yi, zi) = sdf_eval(m.sdf, px, py, pz);
      }
  for (int xi = 0; xi < N[0] + 1; ++xi)
    for (int yi = 0; yi < N[1]; ++yi)
      for (int zi = 0; zi < N[2]; ++zi)
        m.lwx(xi, yi, zi) = clampf(1.0f - fraction_in_region(m.solidPhi(xi, yi, zi), m.solidPhi(xi, yi + 1, zi)), 0.f, 1.f);
  for (int xi = 0; xi < N[0]; ++xi)
    for (int yi = 0; yi < N[1] + 1; ++yi)
      for (int zi = 0; zi < N[2]; ++zi)
        m.lwy(xi, yi, zi) = clampf(1.0f - fraction_in_region(m.solidPhi(xi, yi, zi), m.solidPhi(xi + 1, yi, zi)), 0.f, 1.f);
  for (int xi = 0; xi < N[0]; ++xi)
    for (int yi = 0; yi < N[1]; ++yi)
      for (int zi = 0; zi < N[2] + 1; ++zi)
        m.lwz(xi, yi, zi) = clampf(1.0f - fraction_in_region(m.solidPhi(xi, yi, zi), m.solidPhi(xi + 1, yi, zi)), 0.f, 1.f);
}

// Grid<3>::Clear, Grid.cpp:157-169 (isValid, liquidW, solidPhi untouched).
void clear_grid(Model& m) {
  m.pressure.fill(0.f);
  m.liquidPhi.fill(3 * m.dx[0]);
  m.vx.fill(0.f); m.wx.fill(0.f); m.tvx.fill(0.f);
  m.vy.fill(0.f); m.wy.fill(0.f); m.tvy.fill(0.f);
  m.vz.fill(0.f); m.wz.fill(0.f); m.tvz.fill(0.f);
}

// LiquidModel<3>::ComputeLiquidPhi, LiquidModel.cpp:539-560.  The min is
// order-free, so a serial loop gives the reference's result bit for bit.
// pinned by: an FP64 brute-force min over the particles whose cell window holds
// the cell (tests/test_oracle_pins_3d.py), the LiquidData frame-1 known answer.
void compute_liquid_phi(Model& m) {
  if (m.P == 0) return;
  const int* N = m.N;
  for (i64 pi = 0; pi < m.P; ++pi) {
    const float px = m.pos[3 * pi], py = m.pos[3 * pi + 1], pz = m.pos[3 * pi + 2];
    const float radius = m.radius[pi];
    int cx = (int)((px - m.mn[0]) / m.dx[0]);
    int cy = (int)((py - m.mn[1]) / m.dx[1]);
    int cz = (int)((pz - m.mn[2]) / m.dx[2]);
    int dX = (int)std::ceil(radius / m.dx[0]);
    int dY = (int)std::ceil(radius / m.dx[1]);
    int dZ = (int)std::ceil(radius / m.dx[2]);
    for (int xi = std::max(0, cx - dX); xi <= std::min(N[0] - 1, cx + dX); ++xi)
      for (int yi = std::max(0, cy - dY); yi <= std::min(N[1] - 1, cy + dY); ++yi)
        for (int zi = std::max(0, cz - dZ); zi <= std::min(N[2] - 1, cz + dZ); ++zi) {
          float cpx = ((float)xi + 0.5f) * m.dx[0] + m.mn[0];
          float cpy = ((float)yi + 0.5f) * m.dx[1] + m.mn[1];
          float cpz = ((float)zi + 0.5f) * m.dx[2] + m.mn[2];
          float phi = norm3(cpx - px, cpy - py, cpz - pz) - radius;
          float& cur = m.liquidPhi(xi, yi, zi);
          cur = (cur < phi) ? cur : phi;  // windows.h min(a,b) = (a < b) ? a : b
        }
  }
}

// ---- P2G, LiquidModel.cpp:613-721.
// pinned by (tests/test_oracle_pins_3d.py): both orders against an FP64 brute
// force over every particle-face pair (random particles, particles 0.05 cells
// inside half-cell boundaries, boundary half-cells), the orders against each
// other, uniform-field reproduction; the LiquidData frame-1 known answer.
struct P2GContrib {
  float w, vw;
};
// Contribution of particle pi to face (xi,yi,zi) of component a, exactly as
// LiquidModel.cpp:628-636 (a=0), 657-665 (a=1), 686-694 (a=2).
inline P2GContrib p2g_term(const Model& m, i64 pi, int a, int xi, int yi, int zi) {
  const float* pp = &m.pos[3 * pi];
  float gx = (pp[0] - m.mn[0]) / m.dx[0];
  float gy = (pp[1] - m.mn[1]) / m.dx[1];
  float gz = (pp[2] - m.mn[2]) / m.dx[2];
  float ccx = (a == 0) ? (float)xi : (float)xi + 0.5f;
  float ccy = (a == 1) ? (float)yi : (float)yi + 0.5f;
  float ccz = (a == 2) ? (float)zi : (float)zi + 0.5f;
  float weight = quad_w(gx - ccx) * quad_w(gy - ccy) * quad_w(gz - ccz);
  float d0 = pp[0] - (ccx * m.dx[0] + m.mn[0]);
  float d1 = pp[1] - (ccy * m.dx[1] + m.mn[1]);
  float d2 = pp[2] - (ccz * m.dx[2] + m.mn[2]);
  const float* col = &m.C[9 * pi + 3 * a];
  float dot = sum3(col[0] * d0, col[1] * d1, col[2] * d2);
  return {weight, weight * (m.vel[3 * pi + a] + dot)};
}

void p2g_particle_order(Model& m) {
  const int* N = m.N;
  A3* V[3] = {&m.vx, &m.vy, &m.vz};
  A3* W[3] = {&m.wx, &m.wy, &m.wz};
  for (i64 pi = 0; pi < m.P; ++pi) {
    const float* pp = &m.pos[3 * pi];
    float gx = (pp[0] - m.mn[0]) / m.dx[0];
    float gy = (pp[1] - m.mn[1]) / m.dx[1];
    float gz = (pp[2] - m.mn[2]) / m.dx[2];
    int cx = (int)gx, cy = (int)gy, cz = (int)gz;
    float r = m.radius[pi];
    int dX = (int)std::ceil(r / m.dx[0]), dY = (int)std::ceil(r / m.dx[1]), dZ = (int)std::ceil(r / m.dx[2]);
    for (int a = 0; a < 3; ++a) {
      int hi[3] = {N[0] - 1, N[1] - 1, N[2] - 1};
      hi[a] += 1;
      for (int ddx = -dX; ddx <= dX; ++ddx)
        for (int ddy = -dY; ddy <= dY; ++ddy)
          for (int ddz = -dZ; ddz <= dZ; ++ddz) {
            int xi = cx + ddx, yi = cy + ddy, zi = cz + ddz;
            if (xi < 0 || xi > hi[0] || yi < 0 || yi > hi[1] || zi < 0 || zi > hi[2]) continue;
            P2GContrib t = p2g_term(m, pi, a, xi, yi, zi);
            (*W[a])(xi, yi, zi) += t.w;
            (*V[a])(xi, yi, zi) += t.vw;
          }
    }
  }
}

// Half-cell binning: key of a particle = ((hx*2Ny)+hy)*2Nz+hz with h = floor(2*gridPos)
// (gridPos = (pos-min)/dx).  Counting sort, stable in particle index.
struct Bins {
  int H[3];
  std::vector<i64> start;  // size H0*H1*H2 + 1
  std::vector<i64> order;  // particle ids in bin order
};
bool build_bins(const Model& m, Bins& b) {
  for (int a = 0; a < 3; ++a) b.H[a] = 2 * m.N[a];
  size_t nb = (size_t)b.H[0] * b.H[1] * b.H[2];
  std::vector<i64> key(m.P);
  b.start.assign(nb + 1, 0);
  for (i64 pi = 0; pi < m.P; ++pi) {
    int h[3];
    for (int a = 0; a < 3; ++a) {
      float g = (m.pos[3 * pi + a] - m.mn[a]) / m.dx[a];
      float g2 = 2.0f * g;
      if (!(g2 >= 0.f) || !(g2 < (float)b.H[a])) return false;
      h[a] = (int)g2;
    }
    key[pi] = ((i64)h[0] * b.H[1] + h[1]) * b.H[2] + h[2];
    b.start[key[pi] + 1]++;
  }
  for (size_t k = 0; k < nb; ++k) b.start[k + 1] += b.start[k];
  std::vector<i64> cur(b.start.begin(), b.start.end() - 1);
  b.order.resize(m.P);
  for (i64 pi = 0; pi < m.P; ++pi) b.order[cur[key[pi]]++] = pi;
  return true;
}

// Per-face gather in bin order.  Face (xi,yi,zi) of component a receives
// weight from particles whose gridPos lies within 1.5 of the face centre on
// every axis, i.e. half-cells [2c-3, 2c+2] on the normal axis and [2c-2, 2c+3]
// on the other two.  Zero-weight taps add exact zeros in the reference, so
// skipping them changes nothing.
bool p2g_bin_order(Model& m) {
  Bins b;
  if (!build_bins(m, b)) return false;
  const int* N = m.N;
  A3* V[3] = {&m.vx, &m.vy, &m.vz};
  A3* W[3] = {&m.wx, &m.wy, &m.wz};
  for (int a = 0; a < 3; ++a) {
    int dim[3] = {N[0], N[1], N[2]};
    dim[a] += 1;
#pragma omp parallel for
    for (int xi = 0; xi < dim[0]; ++xi)
      for (int yi = 0; yi < dim[1]; ++yi)
        for (int zi = 0; zi < dim[2]; ++zi) {
          int c[3] = {xi, yi, zi};
          int lo[3], hi[3];
          for (int k = 0; k < 3; ++k) {
            lo[k] = (k == a) ? 2 * c[k] - 3 : 2 * c[k] - 2;
            hi[k] = lo[k] + 5;
            lo[k] = std::max(lo[k], 0);
            hi[k] = std::min(hi[k], b.H[k] - 1);
          }
          float ws = 0.f, vs = 0.f;
          for (int hx = lo[0]; hx <= hi[0]; ++hx)
            for (int hy = lo[1]; hy <= hi[1]; ++hy) {
              size_t k0 = ((size_t)hx * b.H[1] + hy) * b.H[2];
              for (i64 s = b.start[k0 + lo[2]]; s < b.start[k0 + hi[2] + 1]; ++s) {
                i64 pi = b.order[s];
                // the reference also bounds taps to cellIndex +- ceil(r/dx) (LiquidModel.cpp:625)
                float r = m.radius[pi];
                int dd[3] = {(int)std::ceil(r / m.dx[0]), (int)std::ceil(r / m.dx[1]), (int)std::ceil(r / m.dx[2])};
                bool inwin = true;
                for (int k = 0; k < 3; ++k) {
                  int ck = (int)((m.pos[3 * pi + k] - m.mn[k]) / m.dx[k]);
                  if (c[k] < ck - dd[k] || c[k] > ck + dd[k]) inwin = false;
                }
                if (!inwin) continue;
                P2GContrib t = p2g_term(m, pi, a, xi, yi, zi);
                if (t.w == 0.f) continue;
                ws += t.w;
                vs += t.vw;
              }
            }
          (*W[a])(xi, yi, zi) = ws;
          (*V[a])(xi, yi, zi) = vs;
        }
  }
  return true;
}

void p2g_normalize(Model& m) {
  A3* V[3] = {&m.vx, &m.vy, &m.vz};
  A3* W[3] = {&m.wx, &m.wy, &m.wz};
  for (int a = 0; a < 3; ++a) {
    std::vector<float>& v = V[a]->v;
    const std::vector<float>& w = W[a]->v;
    for (size_t i = 0; i < v.size(); ++i) v[i] = !is_small(w[i]) ? v[i] / w[i] : 0.0f;
  }
}

// LiquidModel<3>::ApplyGravity, LiquidModel.cpp:898-910 (all faces, Q8), plus
// the optional horizontal tank acceleration a(t) = A sin(2 pi f t) on the
// interior faces of the chosen axis (same index ranges rotated to that axis).
// pinned by: the LiquidData frame-1 known answer (gravity); a floating block's
// exact one-step displacement and the rotated face mask (tank acceleration,
// tests/test_oracle_pins_3d.py).
void apply_gravity(Model& m, float dt) {
  const int* N = m.N;
  float gdt = m.prm.gravity * dt;
  for (int xi = 1; xi < N[0] - 1; ++xi)
    for (int yi = 1; yi < N[1]; ++yi)
      for (int zi = 1; zi < N[2] - 1; ++zi) m.vy(xi, yi, zi) += gdt;
  int ax = m.prm.tank_axis;
  if (ax >= 0 && ax <= 2 && m.prm.tank_amp != 0.f) {
    double acc = (double)m.prm.tank_amp * std::sin(2.0 * M_PI * (double)m.prm.tank_freq * m.time);
    float adt = (float)acc * dt;
    A3* V[3] = {&m.vx, &m.vy, &m.vz};
    int lo[3] = {1, 1, 1}, hi[3] = {N[0] - 2, N[1] - 2, N[2] - 2};
    hi[ax] = N[ax] - 1;
    for (int xi = lo[0]; xi <= hi[0]; ++xi)
      for (int yi = lo[1]; yi <= hi[1]; ++yi)
        for (int zi = lo[2]; zi <= hi[2]; ++zi) (*V[ax])(xi, yi, zi) += adt;
  }
}

// ------------------------------------------------------------ solver
// CGSolver<3>, PressureSolver.cpp:476-934.
// pinned by: symmetry / row sums / SPD of A (tests), scipy FP64 direct solves
// on small grids, and the projection invariants.
float approx_phi(const Model& m, int xi, int yi, int zi, int level) {  // :476-494
  if (level == 0) return m.liquidPhi(xi, yi, zi);
  int f = 1 << level;
  int bx = f * xi + (1 << (level - 1)) - 1, by = f * yi + (1 << (level - 1)) - 1, bz = f * zi + (1 << (level - 1)) - 1;
  const A3& p = m.liquidPhi;
  return trilerp(p(bx, by, bz), p(bx + 1, by, bz), p(bx, by + 1, bz), p(bx + 1, by + 1, bz), p(bx, by, bz + 1),
                 p(bx + 1, by, bz + 1), p(bx, by + 1, bz + 1), p(bx + 1, by + 1, bz + 1), 0.5f, 0.5f, 0.5f);
}
float approx_w(const Model& m, int a, int xi, int yi, int zi, int level) {  // :496-542
  const A3& w = (a == 0) ? m.lwx : (a == 1) ? m.lwy : m.lwz;
  if (level == 0) return w(xi, yi, zi);
  int f = 1 << level, h = (1 << (level - 1)) - 1;
  if (a == 0) {
    int bx = f * xi, by = f * yi + h, bz = f * zi + h;
    return bilerp(w(bx, by, bz), w(bx, by + 1, bz), w(bx, by, bz + 1), w(bx, by + 1, bz + 1), 0.5f, 0.5f);
  } else if (a == 1) {
    int bx = f * xi + h, by = f * yi, bz = f * zi + h;
    return bilerp(w(bx, by, bz), w(bx + 1, by, bz), w(bx, by, bz + 1), w(bx + 1, by, bz + 1), 0.5f, 0.5f);
  }
  int bx = f * xi + h, by = f * yi + h, bz = f * zi;
  return bilerp(w(bx, by, bz), w(bx + 1, by, bz), w(bx, by + 1, bz), w(bx + 1, by + 1, bz), 0.5f, 0.5f);
}

// BuildLHS, :544-664
void build_lhs(Model& m, float rho, float dt) {
  for (int l = 0; l < m.max_level; ++l) {
    Level& q = m.lv[l];
    int f = 1 << l;
    float dxl[3] = {(float)f * m.dx[0], (float)f * m.dx[1], (float)f * m.dx[2]};
#pragma omp parallel for
    for (int xi = 0; xi < q.d[0]; ++xi)
      for (int yi = 0; yi < q.d[1]; ++yi)
        for (int zi = 0; zi < q.d[2]; ++zi) {
          float cphi = approx_phi(m, xi, yi, zi, l);
          if (cphi > 0.0f) continue;
          float& diag = q.adiag(xi, yi, zi);
          float* co = &q.acoef[q.adiag.id(xi, yi, zi) * 6];
          // (face index k, axis, neighbour offset, face index for W)
          struct F { int axis, s; };
          const F faces[6] = {{0, +1}, {0, -1}, {1, +1}, {1, -1}, {2, +1}, {2, -1}};
          for (int k = 0; k < 6; ++k) {
            int ax = faces[k].axis, s = faces[k].s;
            int c[3] = {xi, yi, zi};
            if (s > 0 ? !(c[ax] + 1 < q.d[ax]) : !(c[ax] - 1 >= 0)) continue;
            int fc[3] = {xi, yi, zi};
            if (s > 0) fc[ax] += 1;
            float wv = approx_w(m, ax, fc[0], fc[1], fc[2], l);
            if (!(wv > 0.0f)) continue;
            float term = wv * dt / (rho * dxl[ax] * dxl[ax]);
            int nb[3] = {xi, yi, zi};
            nb[ax] += s;
            float phi = approx_phi(m, nb[0], nb[1], nb[2], l);
            if (phi < 0.0f) {
              co[k] -= term;
              diag += term;
            } else {
              float frac = max_tenth(fraction_in_region(cphi, phi));
              diag += term / frac;
            }
          }
        }
  }
}

// BuildRHS, :682-698
void build_rhs(Model& m) {
  const int* N = m.N;
#pragma omp parallel for
  for (int xi = 0; xi < N[0]; ++xi)
    for (int yi = 0; yi < N[1]; ++yi)
      for (int zi = 0; zi < N[2]; ++zi) {
        float& b = m.rhs(xi, yi, zi);
        b = 0.0f;
        if (m.liquidPhi(xi, yi, zi) > 0.0f) continue;
        b -= m.lwx(xi + 1, yi, zi) * m.vx(xi + 1, yi, zi) / m.dx[0];
        b += m.lwx(xi, yi, zi) * m.vx(xi, yi, zi) / m.dx[0];
        b -= m.lwy(xi, yi + 1, zi) * m.vy(xi, yi + 1, zi) / m.dx[1];
        b += m.lwy(xi, yi, zi) * m.vy(xi, yi, zi) / m.dx[1];
        b -= m.lwz(xi, yi, zi + 1) * m.vz(xi, yi, zi + 1) / m.dx[2];
        b += m.lwz(xi, yi, zi) * m.vz(xi, yi, zi) / m.dx[2];
      }
}

// GetAx, :700-717 (summation order: diag, +x, -x, +y, -y, +z, -z)
template <class T>
inline T get_ax(const Level& q, const std::vector<T>& g, int xi, int yi, int zi) {
  auto I = [&](int x, int y, int z) { return ((size_t)x * q.d[1] + y) * q.d[2] + z; };
  size_t c = I(xi, yi, zi);
  const float* co = &q.acoef[c * 6];
  T res = g[c] * (T)q.adiag.v[c];
  if (xi + 1 < q.d[0]) res += g[I(xi + 1, yi, zi)] * (T)co[0];
  if (xi - 1 >= 0) res += g[I(xi - 1, yi, zi)] * (T)co[1];
  if (yi + 1 < q.d[1]) res += g[I(xi, yi + 1, zi)] * (T)co[2];
  if (yi - 1 >= 0) res += g[I(xi, yi - 1, zi)] * (T)co[3];
  if (zi + 1 < q.d[2]) res += g[I(xi, yi, zi + 1)] * (T)co[4];
  if (zi - 1 >= 0) res += g[I(xi, yi, zi - 1)] * (T)co[5];
  return res;
}

// Solver working vectors in precision T (float: the reference path; double:
// the FP64 oracle of the refined mode, on the same FP32 coefficients).
template <class T>
struct SolveVecs {
  std::vector<std::vector<T>> r, z;
  std::vector<T> p, x, q;
};

template <class T>
void smooth(const Model& m, SolveVecs<T>& v, int l, int phase, int mode) {  // :821-837
  const Level& q = m.lv[l];
  const double w = (double)m.prm.omega;
  std::vector<T>& z = v.z[l];
  const std::vector<T>& r = v.r[l];
#pragma omp parallel for
  for (int xi = 0; xi < q.d[0]; ++xi)
    for (int yi = 0; yi < q.d[1]; ++yi)
      for (int zi = 0; zi < q.d[2]; ++zi) {
        if (((xi + yi + zi) & 1) != phase) continue;
        size_t c = ((size_t)xi * q.d[1] + yi) * q.d[2] + zi;
        float ad = q.adiag.v[c];
        if (is_small(ad)) continue;
        T term = (r[c] - get_ax(q, z, xi, yi, zi)) / (T)ad;
        if (mode == 0) {
          // (1. - w) * z + w * term: double * double + (float w * term) promoted
          T wt = (T)m.prm.omega * term;
          z[c] = (T)((1.0 - w) * (double)z[c] + (double)wt);
        } else {
          z[c] = z[c] + (T)m.prm.omega * term;
        }
      }
}

template <class T>
void restrict_(const Model& m, SolveVecs<T>& v, int to) {  // :796-819
  const Level& qc = m.lv[to];
  const Level& qf = m.lv[to - 1];
  std::vector<T>& rc = v.r[to];
  std::vector<T>& zc = v.z[to];
  const std::vector<T>& rf = v.r[to - 1];
  const std::vector<T>& zf = v.z[to - 1];
#pragma omp parallel for
  for (int xi = 0; xi < qc.d[0]; ++xi)
    for (int yi = 0; yi < qc.d[1]; ++yi)
      for (int zi = 0; zi < qc.d[2]; ++zi) {
        size_t c = ((size_t)xi * qc.d[1] + yi) * qc.d[2] + zi;
        T acc = 0;
        int bx = xi << 1, by = yi << 1, bz = zi << 1;
        const int ord[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1}, {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};
        for (int k = 0; k < 8; ++k) {
          int x = bx + ord[k][0], y = by + ord[k][1], z = bz + ord[k][2];
          size_t f = ((size_t)x * qf.d[1] + y) * qf.d[2] + z;
          // 0.125 (double) * float difference, added to the float accumulator
          T d = rf[f] - get_ax(qf, zf, x, y, z);
          acc = (T)((double)acc + 0.125 * (double)d);
        }
        rc[c] = acc;
        zc[c] = 0;
      }
}

template <class T>
void prolong(const Model& m, SolveVecs<T>& v, int to) {  // :774-794
  const Level& qc = m.lv[to + 1];
  const Level& qf = m.lv[to];
  std::vector<T>& zf = v.z[to];
  const std::vector<T>& zc = v.z[to + 1];
#pragma omp parallel for
  for (int xi = 0; xi < qc.d[0]; ++xi)
    for (int yi = 0; yi < qc.d[1]; ++yi)
      for (int zi = 0; zi < qc.d[2]; ++zi) {
        T val = zc[((size_t)xi * qc.d[1] + yi) * qc.d[2] + zi];
        for (int k = 0; k < 8; ++k) {
          int x = 2 * xi + ((k >> 2) & 1), y = 2 * yi + ((k >> 1) & 1), z = 2 * zi + (k & 1);
          zf[((size_t)x * qf.d[1] + y) * qf.d[2] + z] += val;
        }
      }
}

// The coarsest level of the V-cycle, PressureSolver.cpp:754-757: num_final red-black sweeps
// (symmetric mode: then num_final in reverse colour order).  With coarsest_fp64 the sweeps run in
// FP64: r and z (zero after the restriction) promoted, the same formulas and order on the FP32
// coefficients promoted, and z rounded back to T for the prolongation.
// Chebyshev iteration on the Jacobi-scaled coarsest system D^-1 A z = D^-1 r from z = 0, for the
// eigenvalue interval [lo, 2] (2 bounds the spectrum of D^-1 A for a diagonally dominant A by
// Gershgorin).  Saad, Iterative Methods for Sparse Linear Systems, Alg. 12.1: theta = (hi + lo) / 2,
// delta = (hi - lo) / 2, sigma = theta / delta; d_0 = D^-1 r / theta; rho_k = 1 / (2 sigma -
// rho_{k-1}), d_k = rho_k rho_{k-1} d_{k-1} + 2 rho_k / delta D^-1 (r - A z_k); z_{k+1} = z_k + d_k.
// The result is p(D^-1 A) D^-1 r for a fixed polynomial p positive on (0, 2]: a symmetric positive
// definite preconditioner.  Cells with a negligible diagonal keep z = 0.
template <class T>
void coarsest_chebyshev(const Model& m, SolveVecs<T>& v, int steps, double lo) {
  const int L = m.max_level;
  const Level& q = m.lv[L - 1];
  std::vector<T>& z = v.z[L - 1];
  const std::vector<T>& r = v.r[L - 1];
  const double hi = 2.0, theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
  std::vector<T> d(z.size(), (T)0), zn(z.size(), (T)0);
  double rho_prev = 1.0 / sigma;
  for (int k = 0; k < steps; ++k) {
    double a = 0.0, b = 1.0 / theta;
    if (k > 0) {
      const double rho = 1.0 / (2.0 * sigma - rho_prev);
      a = rho * rho_prev;
      b = 2.0 * rho / delta;
      rho_prev = rho;
    }
#pragma omp parallel for
    for (int xi = 0; xi < q.d[0]; ++xi)
      for (int yi = 0; yi < q.d[1]; ++yi)
        for (int zi = 0; zi < q.d[2]; ++zi) {
          const size_t c = ((size_t)xi * q.d[1] + yi) * q.d[2] + zi;
          const float ad = q.adiag.v[c];
          if (is_small(ad)) {
            zn[c] = z[c];
            continue;
          }
          const T res = r[c] - get_ax(q, z, xi, yi, zi);
          // step 0 has no previous direction (no 0 * d term, whose sign of zero could differ)
          d[c] = k == 0 ? (T)b * (res / (T)ad) : (T)a * d[c] + (T)b * (res / (T)ad);
          zn[c] = z[c] + d[c];
        }
    z.swap(zn);
  }
}

template <class T>
void coarsest_sweeps(const Model& m, SolveVecs<T>& v, int smoother) {
  const int L = m.max_level;
  auto run = [&](auto& vv) {
    if (m.prm.coarsest_cheb > 0) {
      coarsest_chebyshev(m, vv, m.prm.coarsest_cheb, (double)m.prm.cheb_lo);
      return;
    }
    for (int it = 0; it < m.prm.num_final; ++it) {
      smooth(m, vv, L - 1, 0, smoother);
      smooth(m, vv, L - 1, 1, smoother);
    }
    if (smoother == 1) {
      for (int it = 0; it < m.prm.num_final; ++it) {
        smooth(m, vv, L - 1, 1, smoother);
        smooth(m, vv, L - 1, 0, smoother);
      }
    }
  };
  if (!m.prm.coarsest_fp64) {
    run(v);
    return;
  }
  SolveVecs<double> d;
  d.r.resize(L);
  d.z.resize(L);
  d.r[L - 1].assign(v.r[L - 1].begin(), v.r[L - 1].end());
  d.z[L - 1].assign(v.z[L - 1].begin(), v.z[L - 1].end());
  run(d);
  std::vector<T>& z = v.z[L - 1];
  for (size_t c = 0; c < z.size(); ++c) z[c] = (T)d.z[L - 1][c];
}

template <class T>
void apply_precond(const Model& m, SolveVecs<T>& v, int smoother) {  // :719-772
  const A3& phi = m.liquidPhi;
  if (m.prm.precond == 0) {
    for (size_t c = 0; c < phi.v.size(); ++c)
      if (!(phi.v[c] > 0.0f)) v.z[0][c] = v.r[0][c];
    return;
  }
  if (m.prm.precond == 1) {
    const Level& q = m.lv[0];
    for (size_t c = 0; c < phi.v.size(); ++c) {
      float dn = q.adiag.v[c];
      if (is_small(dn)) continue;
      v.z[0][c] = v.r[0][c] / (T)dn;
    }
    return;
  }
  const int L = m.max_level;
  if (smoother == 1) std::fill(v.z[0].begin(), v.z[0].end(), (T)0);
  for (int l = 1; l < L; ++l) {
    for (int it = 0; it < m.prm.num_pre; ++it) {
      smooth(m, v, l - 1, 0, smoother);
      smooth(m, v, l - 1, 1, smoother);
    }
    restrict_(m, v, l);
  }
  coarsest_sweeps(m, v, smoother);
  for (int l = L - 2; l >= 0; --l) {
    prolong(m, v, l);
    for (int it = 0; it < m.prm.num_post; ++it) {
      if (smoother == 1) {
        smooth(m, v, l, 1, smoother);
        smooth(m, v, l, 0, smoother);
      } else {
        smooth(m, v, l, 0, smoother);
        smooth(m, v, l, 1, smoother);
      }
    }
  }
}

// Array3::Reduce (Array.h:111-122) in the chosen accumulation precision.
template <class T>
T reduce(const std::vector<T>& a, const std::vector<T>& b, int dot_mode) {
  if (dot_mode == 0) {
    T s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
  }
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += (double)a[i] * (double)b[i];
  return (T)s;
}

// CGSolver<3>::Solve core loop, :839-934, on prebuilt A and m.rhs.
// x0 (optional) is a warm start in the same precision; the reference starts
// from x = 0 (Init, :666-680).
template <class T>
void pcg(Model& m, SolveVecs<T>& v, const std::vector<T>& b, std::vector<T>& xout, int smoother, int dot_mode,
         int tol_mode, double tol, int max_iter, orc_solve_stats* st) {
  const A3& phi = m.liquidPhi;
  const int* N = m.N;
  const int L = m.max_level;
  v.r.assign(L, {}); v.z.assign(L, {});
  for (int l = 0; l < L; ++l) {
    size_t n = m.lv[l].r.v.size();
    v.r[l].assign(n, (T)0); v.z[l].assign(n, (T)0);
  }
  size_t n0 = phi.v.size();
  v.p.assign(n0, (T)0);
  v.x.assign(n0, (T)0);
  v.q.assign(n0, (T)0);
  for (size_t c = 0; c < n0; ++c)
    if (!(phi.v[c] > 0.0f)) v.r[0][c] = b[c];
  apply_precond(m, v, smoother);
  for (size_t c = 0; c < n0; ++c)
    if (!(phi.v[c] > 0.0f)) v.p[c] = v.z[0][c];
  T oldRZ = reduce(v.r[0], v.z[0], dot_mode);
  T residual = std::sqrt(reduce(v.r[0], v.r[0], dot_mode));
  const T res0 = residual;
  const Level& q0 = m.lv[0];
  int it = 0;
  bool conv = false;
  auto done = [&](T res) {
    if (tol_mode == 0) return res < (T)tol;
    return (double)res <= tol * (double)res0;
  };
  for (it = 0; it < max_iter; ++it) {
    if (done(residual)) { conv = true; break; }
    // denom = sum p.Ap over interior liquid cells (reduction added)
    double dd = 0;
    T df = 0;
    for (int xi = 1; xi < N[0] - 1; ++xi)
      for (int yi = 1; yi < N[1] - 1; ++yi)
        for (int zi = 1; zi < N[2] - 1; ++zi) {
          size_t c = phi.id(xi, yi, zi);
          if (phi.v[c] > 0.0f) continue;
          T ap = get_ax(q0, v.p, xi, yi, zi);
          v.q[c] = ap;
          if (dot_mode == 0) df += v.p[c] * ap;
          else dd += (double)v.p[c] * (double)ap;
        }
    T denom = (dot_mode == 0) ? df : (T)dd;
    T alpha = oldRZ / denom;
    if (std::isnan((double)alpha) || std::isinf((double)alpha)) break;
    double rd = 0;
    T rf = 0;
    for (int xi = 1; xi < N[0] - 1; ++xi)
      for (int yi = 1; yi < N[1] - 1; ++yi)
        for (int zi = 1; zi < N[2] - 1; ++zi) {
          size_t c = phi.id(xi, yi, zi);
          if (phi.v[c] > 0.0f) continue;
          v.x[c] += alpha * v.p[c];
          v.r[0][c] -= alpha * v.q[c];
          if (dot_mode == 0) rf += v.r[0][c] * v.r[0][c];
          else rd += (double)v.r[0][c] * (double)v.r[0][c];
        }
    residual = std::sqrt((dot_mode == 0) ? rf : (T)rd);
    apply_precond(m, v, smoother);
    T newRZ = reduce(v.r[0], v.z[0], dot_mode);
    T beta = newRZ / oldRZ;
    if (std::isnan((double)beta) || std::isinf((double)beta)) { ++it; break; }
    for (size_t c = 0; c < n0; ++c)
      if (!(phi.v[c] > 0.0f)) v.p[c] = v.z[0][c] + beta * v.p[c];
    oldRZ = newRZ;
  }
  if (!conv && it >= max_iter) conv = done(residual);
  xout = v.x;
  if (st) {
    st->iterations = it;
    st->converged = conv ? 1 : 0;
    st->initial_residual = (double)res0;
    st->final_residual = (double)residual;
    st->rhs_norm = (double)res0;
  }
}

void init_solver(Model& m) {  // Init, :666-680
  for (int l = 0; l < m.max_level; ++l) {
    Level& q = m.lv[l];
    q.r.fill(0.f); q.z.fill(0.f); q.adiag.fill(0.f);
    std::fill(q.acoef.begin(), q.acoef.end(), 0.f);
  }
  m.p.fill(0.f); m.x.fill(0.f); m.rhs.fill(0.f);
}

// CGSolver<3>::Solve(liquidDensity, dt), PressureSolver.cpp:839-934
int solve_pressure_only(Model& m, float rho, float dt, orc_solve_stats* st) {
  init_solver(m);
  build_rhs(m);
  build_lhs(m, rho, dt);
  SolveVecs<float> v;
  std::vector<float> xo;
  pcg<float>(m, v, m.rhs.v, xo, m.prm.smoother, m.prm.dot_mode, m.prm.tol_mode, (double)m.prm.tol, m.prm.max_iter, st);
  m.pressure.v = xo;  // m_grid.pressure.CopyFrom(m_x), :933
  return 0;
}

// LiquidModel<3>::SolvePressure velocity update, LiquidModel.cpp:834-889
void velocity_update(Model& m, float dt) {
  const int* N = m.N;
  const float rho = m.prm.density;
#pragma omp parallel for
  for (int xi = 1; xi < N[0] - 1; ++xi)
    for (int yi = 1; yi < N[1] - 1; ++yi)
      for (int zi = 1; zi < N[2] - 1; ++zi) {
        float cphi = m.liquidPhi(xi, yi, zi);
        A3* V[3] = {&m.vx, &m.vy, &m.vz};
        const A3* W[3] = {&m.lwx, &m.lwy, &m.lwz};
        for (int a = 0; a < 3; ++a) {
          int nb[3] = {xi, yi, zi};
          nb[a] -= 1;
          if ((*W[a])(xi, yi, zi) > 0.0f) {
            float phi = m.liquidPhi(nb[0], nb[1], nb[2]);
            if (cphi < 0.0f || phi < 0.0f) {
              float frac = 1.0f;
              if (cphi >= 0.0f || phi >= 0.0f) frac = max_tenth(fraction_in_region(cphi, phi));
              float gp = (m.pressure(xi, yi, zi) - m.pressure(nb[0], nb[1], nb[2])) / (m.dx[a] * frac);
              (*V[a])(xi, yi, zi) -= gp * dt / rho;
            }
          } else {
            (*V[a])(xi, yi, zi) = 0.0f;
          }
        }
      }
}

// LiquidModel<3>::Extrapolate, LiquidModel.cpp:937-1063 (sequential, Q6).
// pinned by: hand-derived values of a validity chain along the scan order
// (tests/test_oracle_pins_3d.py).
void extrapolate(Model& m) {
  if (m.P == 0) return;
  const int* N = m.N;
  A3* V[3] = {&m.vx, &m.vy, &m.vz};
  A3* T[3] = {&m.tvx, &m.tvy, &m.tvz};
  const A3* W[3] = {&m.lwx, &m.lwy, &m.lwz};
  for (int a = 0; a < 3; ++a) {
    int lo[3] = {1, 1, 1}, hi[3] = {N[0] - 2, N[1] - 2, N[2] - 2};
    hi[a] = N[a] - 1;
    for (int xi = lo[0]; xi <= hi[0]; ++xi)
      for (int yi = lo[1]; yi <= hi[1]; ++yi)
        for (int zi = lo[2]; zi <= hi[2]; ++zi) m.isValid[vid(m, xi, yi, zi)] = ((*W[a])(xi, yi, zi) > 0.0f) ? 1 : 0;
    for (int iter = 0; iter < 5; ++iter) {
      A3& v = *V[a];
      A3& t = *T[a];
      for (int xi = lo[0]; xi <= hi[0]; ++xi)
        for (int yi = lo[1]; yi <= hi[1]; ++yi)
          for (int zi = lo[2]; zi <= hi[2]; ++zi) {
            t(xi, yi, zi) = v(xi, yi, zi);
            if (m.isValid[vid(m, xi, yi, zi)]) continue;
            float sum = 0.0f;
            int count = 0;
            for (int ddx = -1; ddx <= 1; ++ddx)
              for (int ddy = -1; ddy <= 1; ++ddy)
                for (int ddz = -1; ddz <= 1; ++ddz)
                  if (m.isValid[vid(m, xi + ddx, yi + ddy, zi + ddz)]) {
                    sum += v(xi + ddx, yi + ddy, zi + ddz);
                    ++count;
                  }
            if (count > 0) {
              t(xi, yi, zi) = sum / (float)count;
              m.isValid[vid(m, xi, yi, zi)] = 1;
            }
          }
      std::swap(V[a]->v, T[a]->v);  // std::swap(tempV, v)
    }
  }
}

// LiquidModel<3>::Grid2Particle, LiquidModel.cpp:723-823.
// pinned by: an FP64 brute force of v and C over every face
// (tests/test_oracle_pins_3d.py); uniform field -> exact v; frame-1 KAT.
void g2p(Model& m) {
  if (m.P == 0) return;
  const int* N = m.N;
#pragma omp parallel for
  for (i64 pi = 0; pi < m.P; ++pi) {
    const float* pp = &m.pos[3 * pi];
    float g[3] = {(pp[0] - m.mn[0]) / m.dx[0], (pp[1] - m.mn[1]) / m.dx[1], (pp[2] - m.mn[2]) / m.dx[2]};
    int ci[3] = {(int)g[0], (int)g[1], (int)g[2]};
    float r = m.radius[pi];
    int dd[3] = {(int)std::ceil(r / m.dx[0]), (int)std::ceil(r / m.dx[1]), (int)std::ceil(r / m.dx[2])};
    float vel[3] = {0.f, 0.f, 0.f};
    float Cm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const A3* V[3] = {&m.vx, &m.vy, &m.vz};
    for (int a = 0; a < 3; ++a) {
      int hi[3] = {N[0] - 1, N[1] - 1, N[2] - 1};
      hi[a] += 1;
      for (int ddx = -dd[0]; ddx <= dd[0]; ++ddx)
        for (int ddy = -dd[1]; ddy <= dd[1]; ++ddy)
          for (int ddz = -dd[2]; ddz <= dd[2]; ++ddz) {
            int xi = ci[0] + ddx, yi = ci[1] + ddy, zi = ci[2] + ddz;
            if (xi < 0 || xi > hi[0] || yi < 0 || yi > hi[1] || zi < 0 || zi > hi[2]) continue;
            float cc[3] = {(float)xi, (float)yi, (float)zi};
            for (int k = 0; k < 3; ++k)
              if (k != a) cc[k] += 0.5f;
            float wX = quad_w(g[0] - cc[0]), wY = quad_w(g[1] - cc[1]), wZ = quad_w(g[2] - cc[2]);
            float gX = quad_gw(g[0] - cc[0]), gY = quad_gw(g[1] - cc[1]), gZ = quad_gw(g[2] - cc[2]);
            float weight = wX * wY * wZ;
            float gw0 = gX * wY * wZ, gw1 = wX * gY * wZ, gw2 = wX * wY * gZ;
            float fv = (*V[a])(xi, yi, zi);
            vel[a] += fv * weight;
            Cm[3 * a + 0] += fv * gw0;
            Cm[3 * a + 1] += fv * gw1;
            Cm[3 * a + 2] += fv * gw2;
          }
    }
    for (int a = 0; a < 3; ++a) m.vel[3 * pi + a] = vel[a];
    for (int k = 0; k < 9; ++k) m.C[9 * pi + k] = Cm[k];
  }
}

// LiquidModel<3>::AdvectParticle, LiquidModel.cpp:912-935 (the velocity
// reflection is dead code in the reference, Q7).
void advect(Model& m, float dt) {
  if (m.P == 0) return;
#pragma omp parallel for
  for (i64 pi = 0; pi < m.P; ++pi) {
    float* pp = &m.pos[3 * pi];
    const float* v = &m.vel[3 * pi];
    float p0 = pp[0] + v[0] * dt, p1 = pp[1] + v[1] * dt, p2 = pp[2] + v[2] * dt;
    float sp = solid_phi_at(m, p0, p1, p2);
    if (sp < 0.0f) {
      float g[3];
      solid_gradient(m, p0, p1, p2, g);
      float s = sp + kEps;
      p0 -= s * g[0]; p1 -= s * g[1]; p2 -= s * g[2];
    }
    pp[0] = p0; pp[1] = p1; pp[2] = p2;
  }
}

inline uint64_t cv_splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// uniform [0, 1) for the k-th component of the v-th visit of neighbour npi by particle pi
// in the count-th CorrectVolume call (replaces the reference's random_device-seeded mt19937)
inline float cv_rand(uint64_t seed, uint64_t count, int64_t pi, int64_t npi, int v, int k) {
  uint64_t h = cv_splitmix(seed ^ 0x436F7272566F6Cull);
  h = cv_splitmix(h ^ count);
  h = cv_splitmix(h ^ (uint64_t)pi);
  h = cv_splitmix(h ^ ((uint64_t)npi * 128u + (uint64_t)v * 4u + (uint64_t)k));
  return (float)(h >> 40) * (1.0f / 16777216.0f);
}
inline int cv_cell(float p, float mn, float dx) { return (int)((p - mn) / dx); }  // cast<Integer>(): truncation
// hash slot of cell (x, y, z), NeighborSearch.h:45 with the constants of :134-136: Integer is
// a 32-bit int, so the products wrap; % truncates toward zero, then the result is made >= 0
inline int ns_slot(int x, int y, int z, int m) {
  const uint32_t h = (73856093u * (uint32_t)x) ^ (19349663u * (uint32_t)y) ^ (83492791u * (uint32_t)z);
  const int hs = (int)h;
  return ((hs % m) + m) % m;
}
inline int cv_clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// LiquidModel<3>::CorrectVolume, LiquidModel.cpp:1085-1132, on every particle with
// id % correctCycle == correctStep.  The reference's hash search (NeighborSearch.h:22-131)
// puts each particle into the slots of the 3x3x3 cells around its own cell (clamped to the
// grid) and visits the slots of the 3x3x3 cells around the querying particle, so a
// neighbour closer than the radius is visited once per cell the two clamped regions share:
// mult = prod_axes |[max(c1-1,0), min(c1+1,N-1)] & [max(c2-1,0), min(c2+1,N-1)]|.
// Two distinct cells whose slots collide add visits too, so the exact count is
//   visits = #{(C1, C2) in R(pi) x R(npi) : slot(C1) == slot(C2)},  m = 2 * NumParticles slots.
// Restated here with that count, the table built from the current positions (the recorded
// run LiquidData.csv behaves so; the current source rebuilds it only at the end of
// CorrectVolume), neighbours visited in half-cell bin order (the reference's order depends
// on its parallel insertion).  Positions are updated after all particles are processed.
bool correct_volume(Model& m, float dt) {
  if (m.P == 0) return true;
  Bins b;
  if (!build_bins(m, b)) return false;
  const int cyc = std::max(1, m.prm.correct_cycle), step = m.correct_step;
  const int nslot = (int)(2 * m.P);  // HashNeighborSearch m_size = scaleFactor (2) * NumParticles
  const float wf = std::sqrt(m.dx[0] * m.dx[1] * m.dx[2]);
  std::vector<float> buf(3 * (size_t)m.P);
#pragma omp parallel for schedule(dynamic, 256)
  for (i64 pi = 0; pi < m.P; ++pi) {
    if (pi % cyc != step) continue;
    const float* p = &m.pos[3 * pi];
    const float r = m.radius[pi];
    const float r2 = r * r;
    int lo[3], hi[3], hlo[3], hhi[3];
    for (int a = 0; a < 3; ++a) {
      const int c = cv_cell(p[a], m.mn[a], m.dx[a]);
      lo[a] = cv_clampi(c - 1, 0, m.N[a] - 1);
      hi[a] = cv_clampi(c + 1, 0, m.N[a] - 1);
      hlo[a] = std::max(0, 2 * (c - 2) - 1);
      hhi[a] = std::min(b.H[a] - 1, 2 * (c + 2) + 2);
    }
    int reg[27], nreg = 0;  // slots of the cells around pi, with repeats
    for (int x1 = lo[0]; x1 <= hi[0]; ++x1)
      for (int y1 = lo[1]; y1 <= hi[1]; ++y1)
        for (int z1 = lo[2]; z1 <= hi[2]; ++z1) reg[nreg++] = ns_slot(x1, y1, z1, nslot);
    float rep[3] = {0.f, 0.f, 0.f};
    for (int hx = hlo[0]; hx <= hhi[0]; ++hx)
      for (int hy = hlo[1]; hy <= hhi[1]; ++hy) {
        const size_t k0 = ((size_t)hx * b.H[1] + hy) * b.H[2];
        for (i64 s = b.start[k0 + hlo[2]]; s < b.start[k0 + hhi[2] + 1]; ++s) {
          const i64 npi = b.order[s];
          if (npi == pi) continue;
          const float* q = &m.pos[3 * npi];
          const float d0 = p[0] - q[0], d1 = p[1] - q[1], d2 = p[2] - q[2];
          const float sq = d0 * d0 + (d1 * d1 + d2 * d2);
          if (sq >= r2) continue;
          int nlo[3], nhi[3];
          for (int a = 0; a < 3; ++a) {
            const int c2 = cv_cell(q[a], m.mn[a], m.dx[a]);
            nlo[a] = cv_clampi(c2 - 1, 0, m.N[a] - 1);
            nhi[a] = cv_clampi(c2 + 1, 0, m.N[a] - 1);
          }
          int mult = 0;  // visits: slot matches between the two regions
          for (int x2 = nlo[0]; x2 <= nhi[0]; ++x2)
            for (int y2 = nlo[1]; y2 <= nhi[1]; ++y2)
              for (int z2 = nlo[2]; z2 <= nhi[2]; ++z2) {
                const int s2 = ns_slot(x2, y2, z2, nslot);
                for (int k = 0; k < nreg; ++k) mult += (reg[k] == s2) ? 1 : 0;
              }
          const float dist = std::sqrt(sq);
          const float w = wf * quad_w(dist / r);
          for (int v = 0; v < mult; ++v) {
            if (dist > 1e-4f * r) {
              const float wr = w * r;
              rep[0] += wr * (d0 / dist);
              rep[1] += wr * (d1 / dist);
              rep[2] += wr * (d2 / dist);
            } else {
              for (int k = 0; k < 3; ++k) rep[k] += r / dt * cv_rand(m.prm.seed, m.correct_count, pi, npi, v, k);
            }
          }
        }
      }
    float np[3] = {p[0] + rep[0] * dt, p[1] + rep[1] * dt, p[2] + rep[2] * dt};
    const float sp = solid_phi_at(m, np[0], np[1], np[2]);
    if (sp < 0.0f) {
      float g[3];
      solid_gradient(m, np[0], np[1], np[2], g);
      for (int k = 0; k < 3; ++k) np[k] -= sp * g[k];
    }
    for (int k = 0; k < 3; ++k) buf[3 * pi + k] = np[k];
  }
  for (i64 pi = 0; pi < m.P; ++pi)
    if (pi % cyc == step)
      for (int k = 0; k < 3; ++k) m.pos[3 * pi + k] = buf[3 * pi + k];
  m.correct_step = (step + 1) % cyc;
  m.correct_count++;
  return true;
}

// LiquidModel::GetCFL, LiquidModel.h:49-55
float get_cfl(const Model& m, float* maxv_out) {
  float maxVel = 0.0f;
  for (i64 pi = 0; pi < m.P; ++pi) {
    const float* v = &m.vel[3 * pi];
    float c = max3(std::fabs(v[0]), std::fabs(v[1]), std::fabs(v[2]));
    maxVel = (maxVel > c) ? maxVel : c;
  }
  if (maxv_out) *maxv_out = maxVel;
  return clampf(m.dx[0] / maxVel, 0.001f, 0.1f);
}

// LiquidModel<3>::ComputeDivergence, LiquidModel.cpp:1065-1083 (+ L-inf).  The
// mean's sum runs in FP32 in index order (dot_mode 0, the reference's serial loop)
// or in FP64 (dot_mode 1, the HIP library's reduction), like the solver's dots.
// pinned by: closed-form fields (linear velocities, a single face spike, liquid
// test phi <= 0, openness weights) in tests/test_oracle_pins_3d.py.
void divergence(const Model& m, double* mean, double* linf, int* count) {
  const int* N = m.N;
  float dv = 0.0f;
  double dv64 = 0.0;
  int cnt = 0;
  float mx = 0.0f;
  for (int xi = 0; xi < N[0]; ++xi)
    for (int yi = 0; yi < N[1]; ++yi)
      for (int zi = 0; zi < N[2]; ++zi) {
        if (m.liquidPhi(xi, yi, zi) > 0.0f) continue;
        ++cnt;
        float d = m.lwx(xi + 1, yi, zi) * m.vx(xi + 1, yi, zi) - m.lwx(xi, yi, zi) * m.vx(xi, yi, zi);
        d += m.lwy(xi, yi + 1, zi) * m.vy(xi, yi + 1, zi) - m.lwy(xi, yi, zi) * m.vy(xi, yi, zi);
        d += m.lwz(xi, yi, zi + 1) * m.vz(xi, yi, zi + 1) - m.lwz(xi, yi, zi) * m.vz(xi, yi, zi);
        d /= m.dx[0];
        dv += d;
        dv64 += (double)d;
        float ad = std::fabs(d);
        mx = (ad > mx) ? ad : mx;
      }
  if (mean) *mean = (m.prm.dot_mode == 1) ? (cnt ? dv64 / (double)cnt : 0.0) : (double)(dv / (float)cnt);
  if (linf) *linf = (double)mx;
  if (count) *count = cnt;
}

// ------------------------------------------------------------ seeding
// ParticleSystem<3>::SpawnParticles, ParticleSystem.cpp:44-64 (lattice at
// spacing 0.5*defaultRadius = dx, truncated at numParticles with the z-loop break).
void spawn_lattice(Model& m, const float lo[3], const float hi[3], i64 cap) {
  float interval = (float)(0.5 * (double)(2.0f * m.dx[0]));
  int d[3];
  for (int a = 0; a < 3; ++a) d[a] = (int)((hi[a] - lo[a]) / interval);
  i64 total = (i64)d[0] * d[1] * d[2];
  i64 n = std::min(cap, total);
  m.P = n;
  m.pos.assign(3 * n, 0.f); m.vel.assign(3 * n, 0.f); m.C.assign(9 * n, 0.f);
  m.radius.assign(n, 2.0f * m.dx[0]);
  for (int xi = 0; xi < d[0]; ++xi)
    for (int yi = 0; yi < d[1]; ++yi)
      for (int zi = 0; zi < d[2]; ++zi) {
        i64 idx = (i64)xi * d[1] * d[2] + (i64)yi * d[2] + zi;
        if (idx >= n) break;
        m.pos[3 * idx] = (float)xi * interval + lo[0];
        m.pos[3 * idx + 1] = (float)yi * interval + lo[1];
        m.pos[3 * idx + 2] = (float)zi * interval + lo[2];
      }
}

// Jittered block seeding (no reference counterpart): each cell whose centre
// lies in the block gets k^3 particles, one per sub-cell, at
// min + ((k*i + s) + u) * (dx/k) with u from a SplitMix64 counter hash, and
// velocity v0 + sigma * N(0,1) (Box-Muller, FP64) per component.  Particles
// with SDF <= 0 (inside solids) are dropped.
inline uint64_t splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t rng_u64(uint64_t seed, uint64_t block, uint64_t cell, uint64_t sub, uint64_t stream) {
  uint64_t h = splitmix(seed);
  h = splitmix(h ^ block);
  h = splitmix(h ^ cell);
  h = splitmix(h ^ (sub * 16 + stream));
  return h;
}
inline float rng_unit24(uint64_t h) { return (float)(h >> 40) * (1.0f / 16777216.0f); }
inline double rng_normal(uint64_t h1, uint64_t h2) {
  double u1 = ((double)(h1 >> 11) + 1.0) * (1.0 / 9007199254740992.0);
  double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

}  // namespace

// ====================================================================== C API
extern "C" {

void* orc_create(const orc_params* prm) {
  Model* m = new Model();
  m->prm = *prm;
  if (prm->nsdf > 0 && prm->sdf) m->sdf.assign(prm->sdf, prm->sdf + 12 * prm->nsdf);
  m->prm.sdf = nullptr;
  for (int a = 0; a < 3; ++a) { m->N[a] = prm->N[a]; m->mn[a] = prm->mn[a]; m->mx[a] = prm->mx[a]; }
  model_alloc(*m);
  return m;
}
void orc_destroy(void* h) { delete (Model*)h; }
// LiquidModel<3>::SetBoundary (LiquidModel.h:57): a new SDF program; solidPhi is recomputed by
// orc_compute_solid_phi, as the reference recomputes it in ComputeSolidPhi
void orc_set_sdf(void* h, const float* prog, int nnodes) {
  Model* m = (Model*)h;
  m->sdf.assign(prog, prog + 12 * (size_t)nnodes);
}
void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}
int orc_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void orc_grid_info(void* h, int dims[3], float mn[3], float dx[3]) {
  Model* m = (Model*)h;
  for (int a = 0; a < 3; ++a) { dims[a] = m->N[a]; mn[a] = m->mn[a]; dx[a] = m->dx[a]; }
}

static A3* field_ptr(Model* m, int f) {
  switch (f) {
    case 0: return &m->vx; case 1: return &m->vy; case 2: return &m->vz;
    case 3: return &m->tvx; case 4: return &m->tvy; case 5: return &m->tvz;
    case 6: return &m->lwx; case 7: return &m->lwy; case 8: return &m->lwz;
    case 9: return &m->pressure; case 10: return &m->liquidPhi; case 11: return &m->solidPhi;
    case 13: return &m->rhs;
    case 100: return &m->wx; case 101: return &m->wy; case 102: return &m->wz;
    default: return nullptr;
  }
}
int orc_field(void* h, int f, void* buf, long long count, int set) {
  Model* m = (Model*)h;
  if (f == 12) {
    if ((size_t)count != m->isValid.size()) return 6;
    if (set) std::memcpy(m->isValid.data(), buf, count);
    else std::memcpy(buf, m->isValid.data(), count);
    return 0;
  }
  A3* a = field_ptr(m, f);
  if (!a || (size_t)count != a->v.size()) return 6;
  if (set) std::memcpy(a->v.data(), buf, count * 4);
  else std::memcpy(buf, a->v.data(), count * 4);
  return 0;
}
int orc_set_particles(void* h, const float* pos, const float* vel, const float* C, long long n) {
  Model* m = (Model*)h;
  m->P = n;
  m->pos.assign(pos, pos + 3 * n);
  if (vel) m->vel.assign(vel, vel + 3 * n); else m->vel.assign(3 * n, 0.f);
  if (C) m->C.assign(C, C + 9 * n); else m->C.assign(9 * n, 0.f);
  m->radius.assign(n, 2.0f * m->dx[0]);  // m_defaultRadius = 2 * dx (ParticleSystem.cpp:6-8)
  return 0;
}
long long orc_num_particles(void* h) { return ((Model*)h)->P; }
int orc_get_particles(void* h, float* pos, float* vel, float* C, long long n) {
  Model* m = (Model*)h;
  if (n != m->P) return 6;
  if (pos) std::memcpy(pos, m->pos.data(), 12 * n);
  if (vel) std::memcpy(vel, m->vel.data(), 12 * n);
  if (C) std::memcpy(C, m->C.data(), 36 * n);
  return 0;
}
void orc_set_time(void* h, double t) { ((Model*)h)->time = t; }
double orc_get_time(void* h) { return ((Model*)h)->time; }

void orc_spawn_lattice(void* h, const float lo[3], const float hi[3], long long cap) {
  spawn_lattice(*(Model*)h, lo, hi, cap);
}
// blocks: 10 floats each: min[3], max[3], vel[3], sigma
long long orc_seed_blocks(void* h, const float* blocks, int nblocks, int k, unsigned long long seed) {
  Model* m = (Model*)h;
  m->pos.clear(); m->vel.clear();
  for (int b = 0; b < nblocks; ++b) {
    const float* B = blocks + 10 * b;
    // cells whose centre lies in [min, max] on every axis
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = m->N[a]; hi[a] = -1;
      for (int i = 0; i < m->N[a]; ++i) {
        float c = ((float)i + 0.5f) * m->dx[a] + m->mn[a];
        if (c >= B[a] && c <= B[3 + a]) { lo[a] = std::min(lo[a], i); hi[a] = i; }
      }
    }
    for (int xi = lo[0]; xi <= hi[0]; ++xi)
      for (int yi = lo[1]; yi <= hi[1]; ++yi)
        for (int zi = lo[2]; zi <= hi[2]; ++zi) {
          uint64_t cell = ((uint64_t)xi * m->N[1] + yi) * m->N[2] + zi;
          for (int s = 0; s < k * k * k; ++s) {
            int sx = s / (k * k), sy = (s / k) % k, sz = s % k;
            int cc[3] = {xi, yi, zi}, ss[3] = {sx, sy, sz};
            float p[3];
            for (int a = 0; a < 3; ++a) {
              float u = rng_unit24(rng_u64(seed, b, cell, s, a));
              p[a] = m->mn[a] + ((float)(k * cc[a] + ss[a]) + u) * (m->dx[a] / (float)k);
            }
            if (!(sdf_eval(m->sdf, p[0], p[1], p[2]) > 0.0f)) continue;
            float v[3];
            for (int a = 0; a < 3; ++a) {
              v[a] = B[6 + a];
              if (B[9] > 0.f) {
                double nrm = rng_normal(rng_u64(seed, b, cell, s, 3 + 2 * a), rng_u64(seed, b, cell, s, 4 + 2 * a));
                v[a] = (float)((double)B[6 + a] + (double)B[9] * nrm);
              }
            }
            for (int a = 0; a < 3; ++a) { m->pos.push_back(p[a]); m->vel.push_back(v[a]); }
          }
        }
  }
  m->P = (i64)(m->pos.size() / 3);
  m->C.assign(9 * m->P, 0.f);
  m->radius.assign(m->P, 2.0f * m->dx[0]);
  return m->P;
}

int orc_compute_solid_phi(void* h) { compute_solid_phi(*(Model*)h); return 0; }
int orc_clear_grid(void* h) { clear_grid(*(Model*)h); return 0; }
int orc_compute_liquid_phi(void* h) { compute_liquid_phi(*(Model*)h); return 0; }
int orc_p2g(void* h, int order) {
  Model& m = *(Model*)h;
  if (m.P == 0) return 0;
  if (order == 0) p2g_particle_order(m);
  else if (!p2g_bin_order(m)) return 7;
  p2g_normalize(m);
  return 0;
}
int orc_apply_gravity(void* h, float dt) { apply_gravity(*(Model*)h, dt); return 0; }
int orc_solve_pressure(void* h, float rho, float dt, orc_solve_stats* st) {
  Model& m = *(Model*)h;
  if (m.P == 0) return 0;
  return solve_pressure_only(m, rho, dt, st);
}
int orc_velocity_update(void* h, float dt) {
  Model& m = *(Model*)h;
  if (m.P == 0) return 0;
  velocity_update(m, dt);
  return 0;
}
int orc_extrapolate(void* h) { extrapolate(*(Model*)h); return 0; }
int orc_g2p(void* h) { g2p(*(Model*)h); return 0; }
int orc_advect(void* h, float dt) { advect(*(Model*)h, dt); return 0; }
int orc_correct_volume(void* h, float dt) { return correct_volume(*(Model*)h, dt) ? 0 : 7; }
float orc_get_cfl(void* h, float* maxv) { return get_cfl(*(Model*)h, maxv); }
void orc_divergence(void* h, double* mean, double* linf, int* count) { divergence(*(Model*)h, mean, linf, count); }
float orc_solid_phi_at(void* h, float x, float y, float z) { return solid_phi_at(*(Model*)h, x, y, z); }
void orc_solid_gradient(void* h, float x, float y, float z, float g[3]) { solid_gradient(*(Model*)h, x, y, z, g); }

// one liquid step: PreStep + Step + PostStep (main.cpp:50-52)
int orc_step(void* h, float dt, orc_solve_stats* st, double* div_mean, double* div_linf) {
  Model& m = *(Model*)h;
  clear_grid(m);
  compute_liquid_phi(m);
  int rc = orc_p2g(h, m.prm.p2g_order);
  if (rc) return rc;
  apply_gravity(m, dt);
  if (m.P > 0) {
    solve_pressure_only(m, m.prm.density, dt, st);  // the scene's density, LiquidModel.cpp:830-831
    velocity_update(m, dt);
  }
  extrapolate(m);
  g2p(m);
  advect(m, dt);
  if (m.prm.correct_volume && !correct_volume(m, dt)) return 7;
  divergence(m, div_mean, div_linf, nullptr);
  m.time += (double)dt;
  return 0;
}

// Solver-level access for the refined / FP64 checks: build A from the
// current liquidPhi / liquidW, then solve A x = b (b given) in FP32 (reference
// loop) or FP64 (mode 1: symmetric MG, FP64 vectors, FP64 dots).
int orc_build_system(void* h, float rho, float dt) {
  Model& m = *(Model*)h;
  init_solver(m);
  build_rhs(m);
  build_lhs(m, rho, dt);
  return 0;
}
int orc_get_level(void* h, int level, float* adiag, float* acoef6, int dims[3]) {
  Model& m = *(Model*)h;
  if (level < 0 || level >= m.max_level) return 6;
  Level& q = m.lv[level];
  for (int a = 0; a < 3; ++a) dims[a] = q.d[a];
  if (adiag) std::memcpy(adiag, q.adiag.v.data(), q.adiag.v.size() * 4);
  if (acoef6) std::memcpy(acoef6, q.acoef.data(), q.acoef.size() * 4);
  return 0;
}
int orc_num_levels(void* h) { return ((Model*)h)->max_level; }
int orc_solve_fp32(void* h, const float* b, float* x, orc_solve_stats* st) {
  Model& m = *(Model*)h;
  SolveVecs<float> v;
  std::vector<float> bb(b, b + m.liquidPhi.v.size()), xo;
  pcg<float>(m, v, bb, xo, m.prm.smoother, m.prm.dot_mode, m.prm.tol_mode, (double)m.prm.tol, m.prm.max_iter, st);
  std::memcpy(x, xo.data(), xo.size() * 4);
  return 0;
}
int orc_solve_fp64(void* h, const double* b, double* x, double rel_tol, int max_iter, orc_solve_stats* st) {
  Model& m = *(Model*)h;
  SolveVecs<double> v;
  std::vector<double> bb(b, b + m.liquidPhi.v.size()), xo;
  pcg<double>(m, v, bb, xo, 1, 1, 1, rel_tol, max_iter, st);
  std::memcpy(x, xo.data(), xo.size() * 8);
  return 0;
}
// z = the coarsest level's solve of r (sweeps or Chebyshev, as configured) in FP64 from z = 0, on
// the system of the last build; r and z hold the coarsest level's cells
int orc_coarsest_solve64(void* h, const double* r, double* z) {
  Model& m = *(Model*)h;
  const int L = m.max_level;
  if (L < 1 || m.lv.size() < (size_t)L) return 6;
  const size_t n = m.lv[L - 1].adiag.v.size();
  SolveVecs<double> v;
  v.r.resize(L);
  v.z.resize(L);
  v.r[L - 1].assign(r, r + n);
  v.z[L - 1].assign(n, 0.0);
  coarsest_sweeps(m, v, m.prm.smoother);
  std::memcpy(z, v.z[L - 1].data(), n * 8);
  return 0;
}
// y = A x at level 0 (GetAx per cell), FP64 arithmetic on the FP32 coefficients
void orc_apply_a64(void* h, const double* x, double* y) {
  Model& m = *(Model*)h;
  const Level& q = m.lv[0];
  std::vector<double> xv(x, x + q.adiag.v.size());
#pragma omp parallel for
  for (int xi = 0; xi < q.d[0]; ++xi)
    for (int yi = 0; yi < q.d[1]; ++yi)
      for (int zi = 0; zi < q.d[2]; ++zi) y[q.adiag.id(xi, yi, zi)] = get_ax(q, xv, xi, yi, zi);
}

}  // extern "C"

// ---- scalar helpers exported for closed-form tests
extern "C" {
float orc_quad_w(float d) { return quad_w(d); }
float orc_quad_gw(float d) { return quad_gw(d); }
float orc_fraction_in_region(float l, float r) { return fraction_in_region(l, r); }
float orc_trilerp(const float* v, float f1, float f2, float f3) {
  return trilerp(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], f1, f2, f3);
}
float orc_sdf_eval(const float* prog, int nnodes, float x, float y, float z) {
  std::vector<float> p(prog, prog + 12 * nnodes);
  return sdf_eval(p, x, y, z);
}
void orc_barycentric(float x, int lo, int hi, int* idx, float* frac) { barycentric(x, lo, hi, *idx, *frac); }
}

// ======================================================================== 2D
// LiquidModel<2> (LiquidModel.cpp:10-474) and CGSolver<2> (PressureSolver.cpp:9-418):
// the same restatement rules as the 3D model above (serial where the reference races,
// dot products in FP64 rounded to float with dot_mode 1).  Array2 is x-major, y fastest
// (Array.h:5-60).  Differences from 3D that the reference has: particle radius = dx (so P2G,
// G2P and ComputeLiquidPhi visit the 3 x 3 cells around a particle's cell, which cuts the
// kernel's support on the face-normal axis), lattice spawn with spacing dx / 2 inside the box
// shrunk by 30 % per side, Restrict weight 0.25, the PCG residual divided by the liquid cell
// count, CorrectVolume always in the step.
namespace {

struct A2 {
  int d[2] = {1, 1};
  std::vector<float> v;
  void resize(int x, int y) { d[0] = x; d[1] = y; v.assign((size_t)x * y, 0.f); }
  inline size_t id(int x, int y) const { return (size_t)x * d[1] + y; }
  inline float& operator()(int x, int y) { return v[id(x, y)]; }
  inline float operator()(int x, int y) const { return v[id(x, y)]; }
  void fill(float a) { std::fill(v.begin(), v.end(), a); }
};
inline float norm2(float a, float b) { return std::sqrt(a * a + b * b); }
inline float max2(float a, float b) { return (a < b) ? b : a; }

// Boundary.h:27-94 on 2D points (the same program format; z ignored; Sphere is a circle)
float sdf_eval2(const std::vector<float>& prog, float px, float py) {
  float st[32];
  int sp = 0;
  const size_t n = prog.size() / 12;
  for (size_t k = 0; k < n; ++k) {
    const float* q = &prog[12 * k];
    const int kind = (int)q[0];
    if (kind == 0) {
      float ax = std::fabs(px - q[1]) - q[4];
      float ay = std::fabs(py - q[2]) - q[5];
      float outside = norm2(ax > 0.f ? ax : 0.f, ay > 0.f ? ay : 0.f);
      float mc = max2(ax, ay);
      float inside = (mc < 0.f) ? mc : 0.f;
      st[sp++] = (q[8] != 0.f) ? -outside - inside : outside + inside;
    } else if (kind == 1) {
      float d = norm2(px - q[1], py - q[2]) - q[7];
      st[sp++] = (q[8] != 0.f) ? -d : d;
    } else if (kind == 13) {
    } else {
      float b = st[--sp];
      float a = st[--sp];
      float r;
      if (kind == 10) r = (a < b) ? a : b;
      else if (kind == 11) r = (a > b) ? a : b;
      else { float nb = -b; r = (a > nb) ? a : nb; }
      st[sp++] = r;
    }
  }
  return sp > 0 ? st[sp - 1] : 1.0f;
}

struct Level2 {
  int d[2];
  A2 adiag;
  std::vector<float> acoef;  // 4 per cell: right, left, top, bottom
};

struct Model2 {
  orc_params prm;
  std::vector<float> sdf;
  int N[2];
  float mn[2], mx[2], dx[2];
  A2 vx, vy, tvx, tvy, wx, wy, lwx, lwy, pressure, liquidPhi, solidPhi, rhs;
  std::vector<uint8_t> isValid;  // (N0+1) x (N1+1)
  std::vector<float> pos, vel, C, radius;  // 2, 2, 4 (2x2 column-major) per particle
  i64 P = 0;
  std::vector<Level2> lv;
  int max_level = 1;
  double time = 0.0;
  int correct_step = 0;
  uint64_t correct_count = 0;
  int last_iterations = 0;
};

inline size_t vid2(const Model2& m, int x, int y) { return (size_t)x * (m.N[1] + 1) + y; }

// Grid<2> ctor (Grid.cpp:9-17), CGSolver<2> ctor (PressureSolver.cpp:11-35)
void model2_alloc(Model2& m) {
  const int* N = m.N;
  m.vx.resize(N[0] + 1, N[1]); m.vy.resize(N[0], N[1] + 1);
  m.tvx.resize(N[0] + 1, N[1]); m.tvy.resize(N[0], N[1] + 1);
  m.wx.resize(N[0] + 1, N[1]); m.wy.resize(N[0], N[1] + 1);
  m.lwx.resize(N[0] + 1, N[1]); m.lwy.resize(N[0], N[1] + 1);
  m.pressure.resize(N[0], N[1]); m.liquidPhi.resize(N[0], N[1]); m.rhs.resize(N[0], N[1]);
  m.solidPhi.resize(N[0] + 1, N[1] + 1);
  m.isValid.assign((size_t)(N[0] + 1) * (N[1] + 1), 0);
  for (int a = 0; a < 2; ++a) m.dx[a] = (float)((m.mx[a] - m.mn[a]) / (float)N[a]);
  int L = std::max(1, m.prm.num_level);
  if (m.prm.precond != 2) L = 1;
  m.max_level = L;
  m.lv.resize(L);
  for (int l = 0; l < L; ++l) {
    Level2& q = m.lv[l];
    q.d[0] = N[0] / (1 << l);
    q.d[1] = N[1] / (1 << l);
    q.adiag.resize(q.d[0], q.d[1]);
    q.acoef.assign((size_t)q.d[0] * q.d[1] * 4, 0.f);
  }
}

// Grid<2>::GetSolidPhi / GetSolidGradient, Grid.cpp:33-50, 63-80
inline bool in_grid2(const Model2& m, float x, float y) {
  return m.mn[0] <= x && m.mn[1] <= y && x <= m.mx[0] && y <= m.mx[1];
}
void solid_gradient2(const Model2& m, float x, float y, float g[2]);
float solid_phi2(const Model2& m, float x, float y) {
  if (!in_grid2(m, x, y)) {
    float vx = clampf(x, m.mn[0] + m.dx[0], m.mx[0] - m.dx[0]);
    float vy = clampf(y, m.mn[1] + m.dx[1], m.mx[1] - m.dx[1]);
    float g[2];
    solid_gradient2(m, vx, vy, g);
    return solid_phi2(m, vx, vy) + g[0] * (x - vx) + g[1] * (y - vy);
  }
  float lx = (x - m.mn[0]) / m.dx[0];
  float ly = (y - m.mn[1]) / m.dx[1];
  int xi, yi;
  float fx, fy;
  barycentric(lx, 0, m.N[0] + 1, xi, fx);
  barycentric(ly, 0, m.N[1] + 1, yi, fy);
  const A2& s = m.solidPhi;
  return bilerp(s(xi, yi), s(xi + 1, yi), s(xi, yi + 1), s(xi + 1, yi + 1), fx, fy);
}
void solid_gradient2(const Model2& m, float x, float y, float g[2]) {
  if (!in_grid2(m, x, y)) {
    float vx = clampf(x, m.mn[0] + m.dx[0], m.mx[0] - m.dx[0]);
    float vy = clampf(y, m.mn[1] + m.dx[1], m.mx[1] - m.dx[1]);
    solid_gradient2(m, vx, vy, g);
    return;
  }
  const float eps = 1e-4f;
  float left = solid_phi2(m, x - eps, y), right = solid_phi2(m, x + eps, y);
  float bottom = solid_phi2(m, x, y - eps), top = solid_phi2(m, x, y + eps);
  const float two_eps = (float)(2.0 * (double)eps);
  float a = (right - left) / two_eps, b = (top - bottom) / two_eps;
  float n2 = a * a + b * b;
  if (n2 > 0.f) {
    float n = std::sqrt(n2);
    a = a / n; b = b / n;
  }
  g[0] = a; g[1] = b;
}

// LiquidModel<2>::ComputeSolidPhi, LiquidModel.cpp:100-125
void compute_solid_phi2(Model2& m) {
  const int* N = m.N;
  for (int xi = 0; xi < N[0] + 1; ++xi)
    for (int yi = 0; yi < N[1] + 1; ++yi)
      m.solidPhi(xi, yi) = sdf_eval2(m.sdf, (float)xi * m.dx[0] + m.mn[0], (float)yi * m.dx[1] + m.mn[1]);
  for (int xi = 0; xi < N[0] + 1; ++xi)
    for (int yi = 0; yi < N[1]; ++yi)
      m.lwx(xi, yi) = clampf(1.0f - fraction_in_region(m.solidPhi(xi, yi + 1), m.solidPhi(xi, yi)), 0.f, 1.f);
  for (int xi = 0; xi < N[0]; ++xi)
    for (int yi = 0; yi < N[1] + 1; ++yi)
      m.lwy(xi, yi) = clampf(1.0f - fraction_in_region(m.solidPhi(xi + 1, yi), m.solidPhi(xi, yi)), 0.f, 1.f);
}

// Grid<2>::Clear, Grid.cpp:52-61
void clear_grid2(Model2& m) {
  m.pressure.fill(0.f);
  m.liquidPhi.fill(3 * m.dx[0]);
  m.vx.fill(0.f); m.wx.fill(0.f); m.tvx.fill(0.f);
  m.vy.fill(0.f); m.wy.fill(0.f); m.tvy.fill(0.f);
}

// LiquidModel<2>::ComputeLiquidPhi, LiquidModel.cpp:78-98 (min: order-free)
void compute_liquid_phi2(Model2& m) {
  const int* N = m.N;
  for (i64 pi = 0; pi < m.P; ++pi) {
    const float* p = &m.pos[2 * pi];
    const float r = m.radius[pi];
    const int cx = (int)((p[0] - m.mn[0]) / m.dx[0]), cy = (int)((p[1] - m.mn[1]) / m.dx[1]);
    const int dX = (int)std::ceil(r / m.dx[0]), dY = (int)std::ceil(r / m.dx[1]);
    for (int xi = std::max(0, cx - dX); xi <= std::min(N[0] - 1, cx + dX); ++xi)
      for (int yi = std::max(0, cy - dY); yi <= std::min(N[1] - 1, cy + dY); ++yi) {
        float px = ((float)xi + 0.5f) * m.dx[0] + m.mn[0], py = ((float)yi + 0.5f) * m.dx[1] + m.mn[1];
        float phi = norm2(px - p[0], py - p[1]) - r;
        float& d = m.liquidPhi(xi, yi);
        d = (d < phi) ? d : phi;  // min(a, b) macro
      }
  }
}

// particles grouped by cell (x-major), by id inside a cell: the order the GPU gathers in
struct Bins2 {
  std::vector<i64> start, order;
};
void build_bins2(const Model2& m, Bins2& b) {
  const size_t nc = (size_t)m.N[0] * m.N[1];
  std::vector<i64> key(m.P);
  b.start.assign(nc + 1, 0);
  for (i64 pi = 0; pi < m.P; ++pi) {
    int cx = (int)((m.pos[2 * pi] - m.mn[0]) / m.dx[0]), cy = (int)((m.pos[2 * pi + 1] - m.mn[1]) / m.dx[1]);
    cx = std::min(std::max(cx, 0), m.N[0] - 1);
    cy = std::min(std::max(cy, 0), m.N[1] - 1);
    key[pi] = (i64)cx * m.N[1] + cy;
    b.start[key[pi] + 1]++;
  }
  for (size_t k = 0; k < nc; ++k) b.start[k + 1] += b.start[k];
  b.order.assign(m.P, 0);
  std::vector<i64> fill(b.start.begin(), b.start.end() - 1);
  for (i64 pi = 0; pi < m.P; ++pi) b.order[fill[key[pi]]++] = pi;
}

// LiquidModel<2>::Particle2Grid, LiquidModel.cpp:127-186.  A particle reaches the faces
// (c + d) for d in [-1, 1]^2 around its cell c (radius dx), so face f receives exactly the
// particles of the 3 x 3 cells around it; summed per face in cell order, id order inside a cell
// (order 1), or in particle order (order 0).
void p2g2(Model2& m, int order) {
  const int* N = m.N;
  A2* V[2] = {&m.vx, &m.vy};
  A2* W[2] = {&m.wx, &m.wy};
  auto term = [&](i64 pi, int a, int xi, int yi, float& w, float& vw) {
    const float* p = &m.pos[2 * pi];
    float gx = (p[0] - m.mn[0]) / m.dx[0], gy = (p[1] - m.mn[1]) / m.dx[1];
    float ccx = (a == 0) ? (float)xi : (float)xi + 0.5f;
    float ccy = (a == 1) ? (float)yi : (float)yi + 0.5f;
    w = quad_w(gx - ccx) * quad_w(gy - ccy);
    float d0 = p[0] - (ccx * m.dx[0] + m.mn[0]), d1 = p[1] - (ccy * m.dx[1] + m.mn[1]);
    const float* col = &m.C[4 * pi + 2 * a];
    vw = w * (m.vel[2 * pi + a] + (col[0] * d0 + col[1] * d1));
  };
  if (order == 0) {
    for (i64 pi = 0; pi < m.P; ++pi) {
      const float* p = &m.pos[2 * pi];
      const int cx = (int)((p[0] - m.mn[0]) / m.dx[0]), cy = (int)((p[1] - m.mn[1]) / m.dx[1]);
      const int dX = (int)std::ceil(m.radius[pi] / m.dx[0]), dY = (int)std::ceil(m.radius[pi] / m.dx[1]);
      for (int a = 0; a < 2; ++a)
        for (int ddx = -dX; ddx <= dX; ++ddx)
          for (int ddy = -dY; ddy <= dY; ++ddy) {
            int xi = cx + ddx, yi = cy + ddy;
            if (xi < 0 || yi < 0 || xi > N[0] - (a == 0 ? 0 : 1) || yi > N[1] - (a == 1 ? 0 : 1)) continue;
            float w, vw;
            term(pi, a, xi, yi, w, vw);
            (*W[a])(xi, yi) += w;
            (*V[a])(xi, yi) += vw;
          }
    }
  } else {
    Bins2 b;
    build_bins2(m, b);
    for (int a = 0; a < 2; ++a) {
      const int d0 = N[0] + (a == 0), d1 = N[1] + (a == 1);
      for (int xi = 0; xi < d0; ++xi)
        for (int yi = 0; yi < d1; ++yi) {
          float ws = 0.f, vs = 0.f;
          for (int cx = std::max(0, xi - 1); cx <= std::min(N[0] - 1, xi + 1); ++cx)
            for (int cy = std::max(0, yi - 1); cy <= std::min(N[1] - 1, yi + 1); ++cy) {
              const i64 k = (i64)cx * N[1] + cy;
              for (i64 s = b.start[k]; s < b.start[k + 1]; ++s) {
                float w, vw;
                term(b.order[s], a, xi, yi, w, vw);
                ws += w;
                vs += vw;
              }
            }
          (*W[a])(xi, yi) = ws;
          (*V[a])(xi, yi) = vs;
        }
    }
  }
  for (int a = 0; a < 2; ++a)
    for (size_t i = 0; i < V[a]->v.size(); ++i) {
      float w = W[a]->v[i];
      V[a]->v[i] = !is_small(w) ? V[a]->v[i] / w : 0.0f;
    }
}

// LiquidModel<2>::ApplyGravity, LiquidModel.cpp:306-313
void apply_gravity2(Model2& m, float dt) {
  for (int xi = 1; xi < m.N[0] - 1; ++xi)
    for (int yi = 1; yi < m.N[1]; ++yi) m.vy(xi, yi) += m.prm.gravity * dt;
}

// CGSolver<2>::GetApprox*, PressureSolver.cpp:69-94
float approx_phi2(const Model2& m, int xi, int yi, int l) {
  if (l == 0) return m.liquidPhi(xi, yi);
  const int f = 1 << l;
  const int bx = f * xi + (1 << (l - 1)) - 1, by = f * yi + (1 << (l - 1)) - 1;
  const A2& p = m.liquidPhi;
  return bilerp(p(bx, by), p(bx + 1, by), p(bx, by + 1), p(bx + 1, by + 1), 0.5f, 0.5f);
}
float approx_w2(const Model2& m, int a, int xi, int yi, int l) {
  if (l == 0) return a == 0 ? m.lwx(xi, yi) : m.lwy(xi, yi);
  const int f = 1 << l;
  if (a == 0) {
    const int bx = f * xi, by = f * yi + (1 << (l - 1)) - 1;
    return (float)(0.5 * (double)(m.lwx(bx, by) + m.lwx(bx, by + 1)));
  }
  const int bx = f * xi + (1 << (l - 1)) - 1, by = f * yi;
  return (float)(0.5 * (double)(m.lwy(bx, by) + m.lwy(bx + 1, by)));
}

// CGSolver<2>::BuildLHS, PressureSolver.cpp:96-175 (faces right, left, top, bottom)
void build_lhs2(Model2& m, float rho, float dt) {
  for (int l = 0; l < m.max_level; ++l) {
    Level2& q = m.lv[l];
    q.adiag.fill(0.f);
    std::fill(q.acoef.begin(), q.acoef.end(), 0.f);
    const float f = (float)(1 << l);
    const float dxl[2] = {f * m.dx[0], f * m.dx[1]};
    for (int xi = 0; xi < q.d[0]; ++xi)
      for (int yi = 0; yi < q.d[1]; ++yi) {
        const float cphi = approx_phi2(m, xi, yi, l);
        if (cphi > 0.0f) continue;
        float& diag = q.adiag(xi, yi);
        float* co = &q.acoef[q.adiag.id(xi, yi) * 4];
        const int ax[4] = {0, 0, 1, 1}, sg[4] = {1, -1, 1, -1};
        for (int k = 0; k < 4; ++k) {
          const int a = ax[k], s = sg[k];
          const int c[2] = {xi, yi};
          if (s > 0 ? !(c[a] + 1 < q.d[a]) : !(c[a] - 1 >= 0)) continue;
          int fc[2] = {xi, yi};
          if (s > 0) fc[a] += 1;
          const float wv = approx_w2(m, a, fc[0], fc[1], l);
          if (!(wv > 0.0f)) continue;
          const float term = wv * dt / (rho * dxl[a] * dxl[a]);
          int nb[2] = {xi, yi};
          nb[a] += s;
          const float phi = approx_phi2(m, nb[0], nb[1], l);
          if (phi < 0.0f) {
            co[k] -= term;
            diag += term;
          } else {
            diag += term / max_tenth(fraction_in_region(cphi, phi));
          }
        }
      }
  }
}

// CGSolver<2>::BuildRHS, PressureSolver.cpp:194-206
void build_rhs2(Model2& m) {
  for (int xi = 0; xi < m.N[0]; ++xi)
    for (int yi = 0; yi < m.N[1]; ++yi) {
      float& b = m.rhs(xi, yi);
      b = 0.0f;
      if (m.liquidPhi(xi, yi) > 0.0f) continue;
      b -= m.lwx(xi + 1, yi) * m.vx(xi + 1, yi) / m.dx[0];
      b += m.lwx(xi, yi) * m.vx(xi, yi) / m.dx[0];
      b -= m.lwy(xi, yi + 1) * m.vy(xi, yi + 1) / m.dx[1];
      b += m.lwy(xi, yi) * m.vy(xi, yi) / m.dx[1];
    }
}

// CGSolver<2>::GetAx, PressureSolver.cpp:208-223 (order: diag, right, left, top, bottom)
inline float get_ax2(const Level2& q, const std::vector<float>& g, int xi, int yi) {
  const size_t c = (size_t)xi * q.d[1] + yi;
  const float* co = &q.acoef[c * 4];
  float res = g[c] * q.adiag.v[c];
  if (xi + 1 < q.d[0]) res += g[c + q.d[1]] * co[0];
  if (xi - 1 >= 0) res += g[c - q.d[1]] * co[1];
  if (yi + 1 < q.d[1]) res += g[c + 1] * co[2];
  if (yi - 1 >= 0) res += g[c - 1] * co[3];
  return res;
}

struct Vecs2 {
  std::vector<std::vector<float>> r, z;
  std::vector<float> p, x;
};

// CGSolver<2>::Smooth, PressureSolver.cpp:312-327 (symmetric: z + w term, as the 3D build option)
void smooth2(const Model2& m, Vecs2& v, int l, int phase, int mode) {
  const Level2& q = m.lv[l];
  const double w = (double)m.prm.omega;
  std::vector<float>& z = v.z[l];
  const std::vector<float>& r = v.r[l];
  for (int xi = 0; xi < q.d[0]; ++xi)
    for (int yi = 0; yi < q.d[1]; ++yi) {
      if (((xi + yi) & 1) != phase) continue;
      const size_t c = (size_t)xi * q.d[1] + yi;
      const float ad = q.adiag.v[c];
      if (is_small(ad)) continue;
      const float term = (r[c] - get_ax2(q, z, xi, yi)) / ad;
      if (mode == 0) {
        const float wt = m.prm.omega * term;
        z[c] = (float)((1.0 - w) * (double)z[c] + (double)wt);
      } else {
        z[c] = z[c] + m.prm.omega * term;
      }
    }
}

// CGSolver<2>::Restrict, PressureSolver.cpp:296-310 (children (0,0), (1,0), (0,1), (1,1))
void restrict2(const Model2& m, Vecs2& v, int to) {
  const Level2& qc = m.lv[to];
  const Level2& qf = m.lv[to - 1];
  for (int xi = 0; xi < qc.d[0]; ++xi)
    for (int yi = 0; yi < qc.d[1]; ++yi) {
      const size_t c = (size_t)xi * qc.d[1] + yi;
      float acc = 0.f;
      const int ord[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
      for (int k = 0; k < 4; ++k) {
        const int x = 2 * xi + ord[k][0], y = 2 * yi + ord[k][1];
        const size_t f = (size_t)x * qf.d[1] + y;
        const float d = v.r[to - 1][f] - get_ax2(qf, v.z[to - 1], x, y);
        acc = (float)((double)acc + 0.25 * (double)d);
      }
      v.r[to][c] = acc;
      v.z[to][c] = 0.f;
    }
}

// CGSolver<2>::Prolong, PressureSolver.cpp:280-294
void prolong2(const Model2& m, Vecs2& v, int to) {
  const Level2& qc = m.lv[to + 1];
  const Level2& qf = m.lv[to];
  for (int xi = 0; xi < qc.d[0]; ++xi)
    for (int yi = 0; yi < qc.d[1]; ++yi) {
      const float val = v.z[to + 1][(size_t)xi * qc.d[1] + yi];
      const int ord[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
      for (int k = 0; k < 4; ++k)
        v.z[to][(size_t)(2 * xi + ord[k][0]) * qf.d[1] + (2 * yi + ord[k][1])] += val;
    }
}

// CGSolver<2>::ApplyPreconditioner, PressureSolver.cpp:225-278
void apply_precond2(const Model2& m, Vecs2& v, int smoother) {
  const A2& phi = m.liquidPhi;
  if (m.prm.precond == 0) {
    for (size_t c = 0; c < phi.v.size(); ++c)
      if (!(phi.v[c] > 0.0f)) v.z[0][c] = v.r[0][c];
    return;
  }
  if (m.prm.precond == 1) {
    for (size_t c = 0; c < phi.v.size(); ++c) {
      const float dn = m.lv[0].adiag.v[c];
      if (!is_small(dn)) v.z[0][c] = v.r[0][c] / dn;
    }
    return;
  }
  const int L = m.max_level;
  if (smoother == 1) std::fill(v.z[0].begin(), v.z[0].end(), 0.f);
  for (int l = 1; l < L; ++l) {
    for (int it = 0; it < m.prm.num_pre; ++it) {
      smooth2(m, v, l - 1, 0, smoother);
      smooth2(m, v, l - 1, 1, smoother);
    }
    restrict2(m, v, l);
  }
  for (int it = 0; it < m.prm.num_final; ++it) {
    smooth2(m, v, L - 1, 0, smoother);
    smooth2(m, v, L - 1, 1, smoother);
  }
  if (smoother == 1)
    for (int it = 0; it < m.prm.num_final; ++it) {
      smooth2(m, v, L - 1, 1, smoother);
      smooth2(m, v, L - 1, 0, smoother);
    }
  for (int l = L - 2; l >= 0; --l) {
    prolong2(m, v, l);
    for (int it = 0; it < m.prm.num_post; ++it) {
      smooth2(m, v, l, smoother == 1 ? 1 : 0, smoother);
      smooth2(m, v, l, smoother == 1 ? 0 : 1, smoother);
    }
  }
}

float reduce2(const std::vector<float>& a, const std::vector<float>& b, int dot_mode) {
  if (dot_mode == 0) {
    float s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
  }
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += (double)a[i] * (double)b[i];
  return (float)s;
}

// CGSolver<2>::Solve, PressureSolver.cpp:329-416: the residual is ||r|| divided by the number of
// liquid cells (all of them for the initial residual, the interior ones counted in the loop)
int solve2(Model2& m, float dt, orc_solve_stats* st) {
  const int* N = m.N;
  const A2& phi = m.liquidPhi;
  build_rhs2(m);
  build_lhs2(m, m.prm.density, dt);
  const int L = m.max_level;
  Vecs2 v;
  v.r.assign(L, {});
  v.z.assign(L, {});
  for (int l = 0; l < L; ++l) {
    const size_t n = (size_t)m.lv[l].d[0] * m.lv[l].d[1];
    v.r[l].assign(n, 0.f);
    v.z[l].assign(n, 0.f);
  }
  const size_t n0 = phi.v.size();
  v.p.assign(n0, 0.f);
  v.x.assign(n0, 0.f);
  int nliquid = 0;
  for (size_t c = 0; c < n0; ++c)
    if (!(phi.v[c] > 0.0f)) {
      v.r[0][c] = m.rhs.v[c];
      ++nliquid;
    }
  apply_precond2(m, v, m.prm.smoother);
  for (size_t c = 0; c < n0; ++c)
    if (!(phi.v[c] > 0.0f)) v.p[c] = v.z[0][c];
  const int dm = m.prm.dot_mode;
  float oldRZ = reduce2(v.r[0], v.z[0], dm);
  float residual = std::sqrt(reduce2(v.r[0], v.r[0], dm)) / (float)nliquid;
  const float res0 = residual;
  auto done = [&](float res) {
    if (m.prm.tol_mode == 0) return res < m.prm.tol;
    return (double)res <= (double)m.prm.tol * (double)res0;
  };
  const Level2& q0 = m.lv[0];
  int it = 0;
  bool conv = false;
  for (it = 0; it < m.prm.max_iter; ++it) {
    if (done(residual)) { conv = true; break; }
    double dd = 0;
    float df = 0;
    for (int xi = 1; xi < N[0] - 1; ++xi)
      for (int yi = 1; yi < N[1] - 1; ++yi) {
        if (phi(xi, yi) > 0.0f) continue;
        const float ap = get_ax2(q0, v.p, xi, yi);
        const size_t c = phi.id(xi, yi);
        if (dm == 0) df += v.p[c] * ap;
        else dd += (double)v.p[c] * (double)ap;
      }
    const float denom = (dm == 0) ? df : (float)dd;
    const float alpha = oldRZ / denom;
    if (std::isnan(alpha) || std::isinf(alpha)) break;
    double rd = 0;
    float rf = 0;
    int cnt = 0;
    for (int xi = 1; xi < N[0] - 1; ++xi)
      for (int yi = 1; yi < N[1] - 1; ++yi) {
        if (phi(xi, yi) > 0.0f) continue;
        const size_t c = phi.id(xi, yi);
        const float ap = get_ax2(q0, v.p, xi, yi);
        v.x[c] += alpha * v.p[c];
        v.r[0][c] -= alpha * ap;
        if (dm == 0) rf += v.r[0][c] * v.r[0][c];
        else rd += (double)v.r[0][c] * (double)v.r[0][c];
        ++cnt;
      }
    residual = std::sqrt(dm == 0 ? rf : (float)rd) / (float)cnt;
    apply_precond2(m, v, m.prm.smoother);
    const float newRZ = reduce2(v.r[0], v.z[0], dm);
    const float beta = newRZ / oldRZ;
    for (size_t c = 0; c < n0; ++c)
      if (!(phi.v[c] > 0.0f)) v.p[c] = v.z[0][c] + beta * v.p[c];
    oldRZ = newRZ;
  }
  if (!conv && it >= m.prm.max_iter) conv = done(residual);
  m.pressure.v = v.x;
  m.last_iterations = it;
  if (st) {
    st->iterations = it;
    st->converged = conv ? 1 : 0;
    st->initial_residual = res0;
    st->final_residual = residual;
    st->rhs_norm = res0;
  }
  return 0;
}

// LiquidModel<2>::SolvePressure velocity update, LiquidModel.cpp:238-291
void velocity_update2(Model2& m, float dt) {
  const float rho = m.prm.density;
  for (int xi = 1; xi < m.N[0] - 1; ++xi)
    for (int yi = 1; yi < m.N[1] - 1; ++yi) {
      const float cphi = m.liquidPhi(xi, yi);
      for (int a = 0; a < 2; ++a) {
        A2& v = a == 0 ? m.vx : m.vy;
        const float w = a == 0 ? m.lwx(xi, yi) : m.lwy(xi, yi);
        if (w > 0.0f) {
          const int nx = a == 0 ? xi - 1 : xi, ny = a == 0 ? yi : yi - 1;
          const float phi = m.liquidPhi(nx, ny);
          if (cphi < 0.0f || phi < 0.0f) {
            float frac = 1.0f;
            if (cphi >= 0.0f || phi >= 0.0f) frac = max_tenth(fraction_in_region(cphi, phi));
            const float gp = (m.pressure(xi, yi) - m.pressure(nx, ny)) / (m.dx[a] * frac);
            v(xi, yi) -= gp * dt / rho;
          }
        } else {
          v(xi, yi) = 0.0f;
        }
      }
    }
}

// LiquidModel<2>::Extrapolate, LiquidModel.cpp:341-409 (in-place validity, x then y)
void extrapolate2(Model2& m) {
  if (m.P == 0) return;
  A2* V[2] = {&m.vx, &m.vy};
  A2* T[2] = {&m.tvx, &m.tvy};
  const A2* W[2] = {&m.lwx, &m.lwy};
  for (int a = 0; a < 2; ++a) {
    const int xhi = a == 0 ? m.N[0] - 1 : m.N[0] - 2, yhi = a == 0 ? m.N[1] - 2 : m.N[1] - 1;
    for (int xi = 1; xi <= xhi; ++xi)
      for (int yi = 1; yi <= yhi; ++yi) m.isValid[vid2(m, xi, yi)] = ((*W[a])(xi, yi) > 0.0f) ? 1 : 0;
    for (int iter = 0; iter < 5; ++iter) {
      A2& v = *V[a];
      A2& t = *T[a];
      for (int xi = 1; xi <= xhi; ++xi)
        for (int yi = 1; yi <= yhi; ++yi) {
          t(xi, yi) = v(xi, yi);
          if (m.isValid[vid2(m, xi, yi)]) continue;
          float sum = 0.0f;
          int count = 0;
          for (int ddx = -1; ddx <= 1; ++ddx)
            for (int ddy = -1; ddy <= 1; ++ddy)
              if (m.isValid[vid2(m, xi + ddx, yi + ddy)]) {
                sum += v(xi + ddx, yi + ddy);
                ++count;
              }
          if (count > 0) {
            t(xi, yi) = sum / (float)count;
            m.isValid[vid2(m, xi, yi)] = 1;
          }
        }
      std::swap(V[a]->v, T[a]->v);
    }
  }
}

// LiquidModel<2>::Grid2Particle, LiquidModel.cpp:188-236 (not normalised; unsigned gradient)
void g2p2(Model2& m) {
  const int* N = m.N;
  for (i64 pi = 0; pi < m.P; ++pi) {
    const float* p = &m.pos[2 * pi];
    const float gx = (p[0] - m.mn[0]) / m.dx[0], gy = (p[1] - m.mn[1]) / m.dx[1];
    const int cx = (int)gx, cy = (int)gy;
    const int dX = (int)std::ceil(m.radius[pi] / m.dx[0]), dY = (int)std::ceil(m.radius[pi] / m.dx[1]);
    float vel[2] = {0.f, 0.f};
    float Cm[4] = {0.f, 0.f, 0.f, 0.f};
    for (int a = 0; a < 2; ++a) {
      const A2& V = a == 0 ? m.vx : m.vy;
      for (int ddx = -dX; ddx <= dX; ++ddx)
        for (int ddy = -dY; ddy <= dY; ++ddy) {
          const int xi = cx + ddx, yi = cy + ddy;
          if (xi < 0 || yi < 0 || xi > N[0] - (a == 0 ? 0 : 1) || yi > N[1] - (a == 1 ? 0 : 1)) continue;
          const float ccx = (a == 0) ? (float)xi : (float)xi + 0.5f;
          const float ccy = (a == 1) ? (float)yi : (float)yi + 0.5f;
          const float wX = quad_w(gx - ccx), wY = quad_w(gy - ccy);
          const float gX = quad_gw(gx - ccx), gY = quad_gw(gy - ccy);
          const float weight = wX * wY;
          const float fv = V(xi, yi);
          vel[a] += fv * weight;
          Cm[2 * a + 0] += fv * (wY * gX);
          Cm[2 * a + 1] += fv * (wX * gY);
        }
    }
    m.vel[2 * pi] = vel[0];
    m.vel[2 * pi + 1] = vel[1];
    for (int k = 0; k < 4; ++k) m.C[4 * pi + k] = Cm[k];
  }
}

// LiquidModel<2>::AdvectParticle, LiquidModel.cpp:315-339 (velocity reflection dead, Q7)
void advect2(Model2& m, float dt) {
  for (i64 pi = 0; pi < m.P; ++pi) {
    float* pp = &m.pos[2 * pi];
    const float* v = &m.vel[2 * pi];
    float p0 = pp[0] + v[0] * dt, p1 = pp[1] + v[1] * dt;
    const float sp = solid_phi2(m, p0, p1);
    if (sp < 0.0f) {
      float g[2];
      solid_gradient2(m, p0, p1, g);
      const float s = sp + kEps;
      p0 -= s * g[0];
      p1 -= s * g[1];
    }
    pp[0] = p0;
    pp[1] = p1;
  }
}

// hash slot of a 2D cell (NeighborSearch.h:28-31 with the constants of :134-135)
inline int ns_slot2(int x, int y, int mm) {
  const uint32_t h = (73856093u * (uint32_t)x) ^ (19349663u * (uint32_t)y);
  const int hs = (int)h;
  return ((hs % mm) + mm) % mm;
}

// LiquidModel<2>::CorrectVolume, LiquidModel.cpp:430-468, on the hash table of the advected
// positions (AdvectParticle rebuilds it, :338): neighbours closer than the radius (dx), each
// visited once per pair of region cells (3 x 3 around each particle) whose slots match; the
// coincident-particle kick from the counter generator; positions written after all particles.
void correct_volume2(Model2& m, float dt) {
  const int cyc = std::max(1, m.prm.correct_cycle), step = m.correct_step;
  const int nslot = (int)(2 * m.P);
  const float wf = std::sqrt(m.dx[0] * m.dx[1]);
  Bins2 b;
  build_bins2(m, b);
  std::vector<float> buf(2 * (size_t)m.P);
  for (i64 pi = 0; pi < m.P; ++pi) {
    if (pi % cyc != step) continue;
    const float* p = &m.pos[2 * pi];
    const float r = m.radius[pi];
    const float r2 = r * r;
    int lo[2], hi[2];
    for (int a = 0; a < 2; ++a) {
      const int c = cv_cell(p[a], m.mn[a], m.dx[a]);
      lo[a] = cv_clampi(c - 1, 0, m.N[a] - 1);
      hi[a] = cv_clampi(c + 1, 0, m.N[a] - 1);
    }
    int reg[9], nreg = 0;
    for (int x1 = lo[0]; x1 <= hi[0]; ++x1)
      for (int y1 = lo[1]; y1 <= hi[1]; ++y1) reg[nreg++] = ns_slot2(x1, y1, nslot);
    float rep[2] = {0.f, 0.f};
    // candidates: the particles of the cells within 2 of p's cell (every neighbour within dx)
    const int cx = cv_cell(p[0], m.mn[0], m.dx[0]), cy = cv_cell(p[1], m.mn[1], m.dx[1]);
    for (int bx = std::max(0, cx - 2); bx <= std::min(m.N[0] - 1, cx + 2); ++bx)
      for (int by = std::max(0, cy - 2); by <= std::min(m.N[1] - 1, cy + 2); ++by) {
        const i64 k = (i64)bx * m.N[1] + by;
        for (i64 s = b.start[k]; s < b.start[k + 1]; ++s) {
          const i64 npi = b.order[s];
          if (npi == pi) continue;
          const float* q = &m.pos[2 * npi];
          const float d0 = p[0] - q[0], d1 = p[1] - q[1];
          const float sq = d0 * d0 + d1 * d1;
          if (sq >= r2) continue;
          int nlo[2], nhi[2];
          for (int a = 0; a < 2; ++a) {
            const int c2 = cv_cell(q[a], m.mn[a], m.dx[a]);
            nlo[a] = cv_clampi(c2 - 1, 0, m.N[a] - 1);
            nhi[a] = cv_clampi(c2 + 1, 0, m.N[a] - 1);
          }
          int mult = 0;
          for (int x2 = nlo[0]; x2 <= nhi[0]; ++x2)
            for (int y2 = nlo[1]; y2 <= nhi[1]; ++y2) {
              const int s2 = ns_slot2(x2, y2, nslot);
              for (int kk = 0; kk < nreg; ++kk) mult += (reg[kk] == s2) ? 1 : 0;
            }
          const float dist = std::sqrt(sq);
          const float w = wf * quad_w(dist / r);
          for (int vv = 0; vv < mult; ++vv) {
            if (dist > 1e-4f * r) {
              const float wr = w * r;
              rep[0] += wr * (d0 / dist);
              rep[1] += wr * (d1 / dist);
            } else {
              for (int kk = 0; kk < 2; ++kk) rep[kk] += r / dt * cv_rand(m.prm.seed, m.correct_count, pi, npi, vv, kk);
            }
          }
        }
      }
    float np[2] = {p[0] + rep[0] * dt, p[1] + rep[1] * dt};
    const float sp = solid_phi2(m, np[0], np[1]);
    if (sp < 0.0f) {
      float g[2];
      solid_gradient2(m, np[0], np[1], g);
      np[0] -= sp * g[0];
      np[1] -= sp * g[1];
    }
    buf[2 * pi] = np[0];
    buf[2 * pi + 1] = np[1];
  }
  for (i64 pi = 0; pi < m.P; ++pi)
    if (pi % cyc == step) {
      m.pos[2 * pi] = buf[2 * pi];
      m.pos[2 * pi + 1] = buf[2 * pi + 1];
    }
  m.correct_step = (step + 1) % cyc;
  m.correct_count++;
}

// LiquidModel::GetCFL, LiquidModel.h:49-55
float get_cfl2(const Model2& m, float* maxv_out) {
  float maxVel = 0.0f;
  for (i64 pi = 0; pi < m.P; ++pi) {
    const float* v = &m.vel[2 * pi];
    const float c = max2(std::fabs(v[0]), std::fabs(v[1]));
    maxVel = (maxVel > c) ? maxVel : c;
  }
  if (maxv_out) *maxv_out = maxVel;
  return clampf(m.dx[0] / maxVel, 0.001f, 0.1f);
}

// LiquidModel<2>::ComputeDivergence, LiquidModel.cpp:411-426 (+ L-inf)
void divergence2(const Model2& m, double* mean, double* linf, int* count) {
  float dv = 0.0f, mx = 0.0f;
  int cnt = 0;
  for (int xi = 0; xi < m.N[0]; ++xi)
    for (int yi = 0; yi < m.N[1]; ++yi) {
      if (m.liquidPhi(xi, yi) > 0.0f) continue;
      ++cnt;
      float d = m.lwx(xi + 1, yi) * m.vx(xi + 1, yi) - m.lwx(xi, yi) * m.vx(xi, yi);
      d += m.lwy(xi, yi + 1) * m.vy(xi, yi + 1) - m.lwy(xi, yi) * m.vy(xi, yi);
      d /= m.dx[0];
      dv += d;
      mx = std::max(mx, std::fabs(d));
    }
  if (mean) *mean = cnt ? (double)(dv / (float)cnt) : 0.0;
  if (linf) *linf = mx;
  if (count) *count = cnt;
}

// ParticleSystem<2>::SpawnParticles (ParticleSystem.cpp:23-42) at LiquidModel<2>::Init's box
// (:44-49): spacing radius / 2 from min + 0.3 (max - min); at most numParticles (the break
// leaves only the y loop)
void spawn2(Model2& m) {
  const float rad = m.dx[0];
  float lo[2], hi[2];
  for (int a = 0; a < 2; ++a) {
    const float off = 0.3f * (m.mx[a] - m.mn[a]);
    lo[a] = m.mn[a] + off;
    hi[a] = m.mx[a] - off;
  }
  const float interval = (float)(0.5 * (double)rad);
  const int nx = (int)((hi[0] - lo[0]) / interval), ny = (int)((hi[1] - lo[1]) / interval);
  const i64 cap = m.prm.num_particles;
  const i64 n = std::min((i64)nx * ny, cap);
  m.P = n;
  m.pos.assign(2 * n, 0.f);
  m.vel.assign(2 * n, 0.f);
  m.C.assign(4 * n, 0.f);
  m.radius.assign(n, rad);
  for (int xi = 0; xi < nx; ++xi)
    for (int yi = 0; yi < ny; ++yi) {
      const i64 index = (i64)xi * ny + yi;
      if (index >= n) break;
      m.pos[2 * index] = (float)xi * interval + lo[0];
      m.pos[2 * index + 1] = (float)yi * interval + lo[1];
    }
}

}  // namespace

extern "C" {
void* orc2_create(const orc_params* prm) {
  Model2* m = new Model2();
  m->prm = *prm;
  if (prm->nsdf > 0 && prm->sdf) m->sdf.assign(prm->sdf, prm->sdf + 12 * prm->nsdf);
  m->prm.sdf = nullptr;
  for (int a = 0; a < 2; ++a) { m->N[a] = prm->N[a]; m->mn[a] = prm->mn[a]; m->mx[a] = prm->mx[a]; }
  model2_alloc(*m);
  return m;
}
void orc2_destroy(void* h) { delete (Model2*)h; }
void orc2_init(void* h) {
  Model2* m = (Model2*)h;
  compute_solid_phi2(*m);
  spawn2(*m);
  m->time = 0.0;
}
void orc2_set_sdf(void* h, const float* prog, int nnodes) {
  ((Model2*)h)->sdf.assign(prog, prog + 12 * (size_t)nnodes);
}
void orc2_compute_solid_phi(void* h) { compute_solid_phi2(*(Model2*)h); }
i64 orc2_num_particles(void* h) { return ((Model2*)h)->P; }
// particles: pos (2P), vel (2P), C (4P, column-major 2x2); dir 0 get, 1 set (resizes)
void orc2_particles(void* h, float* pos, float* vel, float* C, i64 n, int dir) {
  Model2* m = (Model2*)h;
  if (dir == 1) {
    m->P = n;
    m->pos.assign(pos, pos + 2 * n);
    m->vel.assign(2 * n, 0.f);
    m->C.assign(4 * n, 0.f);
    if (vel) m->vel.assign(vel, vel + 2 * n);
    if (C) m->C.assign(C, C + 4 * n);
    m->radius.assign(n, m->dx[0]);
    return;
  }
  std::copy(m->pos.begin(), m->pos.end(), pos);
  if (vel) std::copy(m->vel.begin(), m->vel.end(), vel);
  if (C) std::copy(m->C.begin(), m->C.end(), C);
}
// fields: 0 vx, 1 vy, 3 tempVx, 4 tempVy, 6 liquidWx, 7 liquidWy, 9 pressure, 10 liquidPhi,
// 11 solidPhi, 12 isValid (uint8), 13 rhs; dir 0 get, 1 set
int orc2_field(void* h, int f, void* data, size_t n, int dir) {
  Model2* m = (Model2*)h;
  if (f == 12) {
    if (n != m->isValid.size()) return -1;
    if (dir) std::memcpy(m->isValid.data(), data, n);
    else std::memcpy(data, m->isValid.data(), n);
    return 0;
  }
  A2* a = nullptr;
  switch (f) {
    case 0: a = &m->vx; break;
    case 1: a = &m->vy; break;
    case 3: a = &m->tvx; break;
    case 4: a = &m->tvy; break;
    case 6: a = &m->lwx; break;
    case 7: a = &m->lwy; break;
    case 9: a = &m->pressure; break;
    case 10: a = &m->liquidPhi; break;
    case 11: a = &m->solidPhi; break;
    case 13: a = &m->rhs; break;
    default: return -1;
  }
  if (n != a->v.size()) return -1;
  if (dir) std::memcpy(a->v.data(), data, n * sizeof(float));
  else std::memcpy(data, a->v.data(), n * sizeof(float));
  return 0;
}
void orc2_clear_grid(void* h) { clear_grid2(*(Model2*)h); }
void orc2_compute_liquid_phi(void* h) { compute_liquid_phi2(*(Model2*)h); }
void orc2_p2g(void* h, int order) { p2g2(*(Model2*)h, order); }
void orc2_apply_gravity(void* h, float dt) { apply_gravity2(*(Model2*)h, dt); }
int orc2_solve_pressure(void* h, float dt, orc_solve_stats* st) { return solve2(*(Model2*)h, dt, st); }
void orc2_velocity_update(void* h, float dt) { velocity_update2(*(Model2*)h, dt); }
void orc2_extrapolate(void* h) { extrapolate2(*(Model2*)h); }
void orc2_g2p(void* h) { g2p2(*(Model2*)h); }
void orc2_advect(void* h, float dt) { advect2(*(Model2*)h, dt); }
void orc2_correct_volume(void* h, float dt) { correct_volume2(*(Model2*)h, dt); }
float orc2_get_cfl(void* h, float* maxv) { return get_cfl2(*(Model2*)h, maxv); }
void orc2_divergence(void* h, double* mean, double* linf, int* count) { divergence2(*(Model2*)h, mean, linf, count); }
void orc2_set_time(void* h, double t) { ((Model2*)h)->time = t; }
double orc2_time(void* h) { return ((Model2*)h)->time; }
// LiquidModel<2>: PreStep (:51-55), Step (:57-66), no PostStep work without _DEBUG
int orc2_step(void* h, float dt, orc_solve_stats* st) {
  Model2* m = (Model2*)h;
  clear_grid2(*m);
  compute_liquid_phi2(*m);
  p2g2(*m, m->prm.p2g_order);
  apply_gravity2(*m, dt);
  if (m->P > 0) {
    solve2(*m, dt, st);
    velocity_update2(*m, dt);
  }
  extrapolate2(*m);
  g2p2(*m);
  advect2(*m, dt);
  correct_volume2(*m, dt);
  m->time += (double)dt;
  return 0;
}
}  // extern "C"
