// k_grid.hip -- MAC-grid stages of the liquid step on gfx950:
// solid SDF and face openness (LiquidModel.cpp:562-611), Grid<3>::Clear
// (Grid.cpp:157-169), gravity (LiquidModel.cpp:898-910), the pressure
// velocity update (LiquidModel.cpp:834-889), Extrapolate (LiquidModel.cpp:937-1063)
// and ComputeDivergence (LiquidModel.cpp:1065-1083).
//
// All arrays use the reference Array3 layout (z fastest), so consecutive
// lanes of a wavefront touch consecutive z and every access below coalesces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "lh_ctx.h"
#include "lh_math.h"

namespace lh {

namespace {

constexpr int kBlock = 256;

__global__ void k_fill_f32(float* __restrict__ a, size_t n, float v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}
__global__ void k_fill_u8(uint8_t* __restrict__ a, size_t n, uint8_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}

inline unsigned grid_for(size_t n) {
  size_t b = div_up(n, kBlock);
  if (b > 65536 * 8) b = 65536 * 8;
  return (unsigned)(b ? b : 1);
}

// nodal SDF: solidPhi(x,y,z) = boundary.Compute(node), LiquidModel.cpp:566-574
__global__ void k_solid_phi(float* __restrict__ sp, GridP g, const float* __restrict__ prog, int nprog) {
  __shared__ float sprog[12 * kMaxSdfNodes];
  for (int i = threadIdx.x; i < 12 * nprog; i += blockDim.x) sprog[i] = prog[i];
  __syncthreads();
  const int n0 = g.N[0] + 1, n1 = g.N[1] + 1, n2 = g.N[2] + 1;
  size_t total = (size_t)n0 * n1 * n2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    int zi = (int)(i % n2);
    int yi = (int)((i / n2) % n1);
    int xi = (int)(i / ((size_t)n1 * n2));
    float px = (float)xi * g.dx[0] + g.mn[0];
    float py = (float)yi * g.dx[1] + g.mn[1];
    float pz = (float)zi * g.dx[2] + g.mn[2];
    sp[i] = sdf_eval(sprog, nprog, px, py, pz);
  }
}

// liquidW of one face component from two nodal samples (Q2: the reference
// uses an edge: Wx pairs (x,y,z)/(x,y+1,z); Wy and Wz pair (x,y,z)/(x+1,y,z)),
// LiquidModel.cpp:577-610.
__global__ void k_face_open(float* __restrict__ w, const float* __restrict__ sp, GridP g, int a) {
  int d0 = g.N[0] + (a == 0), d1 = g.N[1] + (a == 1), d2 = g.N[2] + (a == 2);
  const int n1 = g.N[1] + 1, n2 = g.N[2] + 1;
  size_t total = (size_t)d0 * d1 * d2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    int zi = (int)(i % d2);
    int yi = (int)((i / d2) % d1);
    int xi = (int)(i / ((size_t)d1 * d2));
    float va = sp[((size_t)xi * n1 + yi) * n2 + zi];
    float vb = (a == 0) ? sp[((size_t)xi * n1 + yi + 1) * n2 + zi] : sp[((size_t)(xi + 1) * n1 + yi) * n2 + zi];
    w[i] = clampf(1.0f - fraction_in_region(va, vb), 0.f, 1.f);
  }
}

// vy += g dt on x in [1,N-2], y in [1,N-1], z in [1,N-2] (LiquidModel.cpp:903-909);
// optional tank acceleration on the chosen axis' interior faces.
__global__ void k_add_interior(float* __restrict__ v, GridP g, int a, float inc) {
  int d0 = g.N[0] + (a == 0), d1 = g.N[1] + (a == 1), d2 = g.N[2] + (a == 2);
  size_t total = (size_t)d0 * d1 * d2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    int c[3];
    c[2] = (int)(i % d2);
    c[1] = (int)((i / d2) % d1);
    c[0] = (int)(i / ((size_t)d1 * d2));
    bool in = true;
    for (int k = 0; k < 3; ++k) {
      int hi = (k == a) ? g.N[k] - 1 : g.N[k] - 2;
      if (c[k] < 1 || c[k] > hi) in = false;
    }
    if (in) v[i] += inc;
  }
}

// SolvePressure velocity update, LiquidModel.cpp:834-889: cells in [1,N-2]^3,
// each updates its lower face on every axis.
__global__ void k_velocity_update(float* __restrict__ vx, float* __restrict__ vy, float* __restrict__ vz,
                                  const float* __restrict__ lwx, const float* __restrict__ lwy,
                                  const float* __restrict__ lwz, const float* __restrict__ phi,
                                  const float* __restrict__ pr, GridP g, float dt, float rho) {
  const int N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
  const int m0 = N0 - 2, m1 = N1 - 2, m2 = N2 - 2;
  if (m0 <= 0 || m1 <= 0 || m2 <= 0) return;
  size_t total = (size_t)m0 * m1 * m2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    int zi = 1 + (int)(i % m2);
    int yi = 1 + (int)((i / m2) % m1);
    int xi = 1 + (int)(i / ((size_t)m1 * m2));
    size_t c = ((size_t)xi * N1 + yi) * N2 + zi;
    float cphi = phi[c];
    float pc = pr[c];
    // x face (x,y,z) of vx with dims (N0+1,N1,N2): same linear index pattern
    {
      size_t f = ((size_t)xi * N1 + yi) * N2 + zi;
      size_t nb = ((size_t)(xi - 1) * N1 + yi) * N2 + zi;
      if (lwx[f] > 0.0f) {
        float ph = phi[nb];
        if (cphi < 0.0f || ph < 0.0f) {
          float frac = 1.0f;
          if (cphi >= 0.0f || ph >= 0.0f) frac = max_tenth(fraction_in_region(cphi, ph));
          float gp = (pc - pr[nb]) / (g.dx[0] * frac);
          vx[f] -= gp * dt / rho;
        }
      } else {
        vx[f] = 0.0f;
      }
    }
    {
      size_t f = ((size_t)xi * (N1 + 1) + yi) * N2 + zi;
      size_t nb = ((size_t)xi * N1 + (yi - 1)) * N2 + zi;
      if (lwy[f] > 0.0f) {
        float ph = phi[nb];
        if (cphi < 0.0f || ph < 0.0f) {
          float frac = 1.0f;
          if (cphi >= 0.0f || ph >= 0.0f) frac = max_tenth(fraction_in_region(cphi, ph));
          float gp = (pc - pr[nb]) / (g.dx[1] * frac);
          vy[f] -= gp * dt / rho;
        }
      } else {
        vy[f] = 0.0f;
      }
    }
    {
      size_t f = ((size_t)xi * N1 + yi) * (N2 + 1) + zi;
      size_t nb = c - 1;
      if (lwz[f] > 0.0f) {
        float ph = phi[nb];
        if (cphi < 0.0f || ph < 0.0f) {
          float frac = 1.0f;
          if (cphi >= 0.0f || ph >= 0.0f) frac = max_tenth(fraction_in_region(cphi, ph));
          float gp = (pc - pr[nb]) / (g.dx[2] * frac);
          vz[f] -= gp * dt / rho;
        }
      } else {
        vz[f] = 0.0f;
      }
    }
  }
}

// ---------------------------------------------------------------- Extrapolate
// One sweep of LiquidModel.cpp:953-978 visits the faces of the component's
// range in (x,y,z) order; an invalid face becomes valid if a neighbour is valid
// at that moment, i.e. valid when the sweep began, or earlier in the order and
// validated during this sweep.  The 13 neighbours that precede a face in the
// order are the first 13 of the 27-offset loop.  The set of faces validated in
// a sweep is therefore the least fixpoint of
//   validated(f) = invalid0(f) & (any valid0 neighbour | any validated predecessor),
// which is unique (the predecessor relation is acyclic) and is reached by
// monotone relaxation in any order.  Values are then averages of the OLD field.
// xstate: 0 = valid at sweep start or outside the range, 1 = invalid,
// 2 = validated in this sweep.
struct XRange {
  int d[3];   // face array dims of the component
  int lo[3];  // inclusive range
  int hi[3];
  int vd[3];  // isValid dims (N+1)^3
};

__device__ __forceinline__ size_t vidx(const XRange& R, int x, int y, int z) {
  return ((size_t)x * R.vd[1] + y) * R.vd[2] + z;
}

__global__ void k_xvalid_init(uint8_t* __restrict__ valid, const float* __restrict__ lw, XRange R) {
  const int m0 = R.hi[0] - R.lo[0] + 1, m1 = R.hi[1] - R.lo[1] + 1, m2 = R.hi[2] - R.lo[2] + 1;
  size_t total = (size_t)m0 * m1 * m2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    int z = R.lo[2] + (int)(i % m2), y = R.lo[1] + (int)((i / m2) % m1), x = R.lo[0] + (int)(i / ((size_t)m1 * m2));
    valid[vidx(R, x, y, z)] = (lw[((size_t)x * R.d[1] + y) * R.d[2] + z] > 0.0f) ? 1 : 0;
  }
}

__global__ void k_xseed(uint8_t* __restrict__ st, const uint8_t* __restrict__ valid, XRange R) {
  const int m0 = R.hi[0] - R.lo[0] + 1, m1 = R.hi[1] - R.lo[1] + 1, m2 = R.hi[2] - R.lo[2] + 1;
  size_t total = (size_t)m0 * m1 * m2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    int z = R.lo[2] + (int)(i % m2), y = R.lo[1] + (int)((i / m2) % m1), x = R.lo[0] + (int)(i / ((size_t)m1 * m2));
    size_t f = ((size_t)x * R.d[1] + y) * R.d[2] + z;
    uint8_t s = 0;
    if (!valid[vidx(R, x, y, z)]) {
      s = 1;
      for (int dx = -1; dx <= 1 && s == 1; ++dx)
        for (int dy = -1; dy <= 1 && s == 1; ++dy)
          for (int dz = -1; dz <= 1; ++dz)
            if (valid[vidx(R, x + dx, y + dy, z + dz)]) { s = 2; break; }
    }
    st[f] = s;
  }
}

__global__ void k_xprop(uint8_t* __restrict__ st, XRange R, int* __restrict__ changed) {
  const int m0 = R.hi[0] - R.lo[0] + 1, m1 = R.hi[1] - R.lo[1] + 1, m2 = R.hi[2] - R.lo[2] + 1;
  size_t total = (size_t)m0 * m1 * m2;
  int any = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    int z = R.lo[2] + (int)(i % m2), y = R.lo[1] + (int)((i / m2) % m1), x = R.lo[0] + (int)(i / ((size_t)m1 * m2));
    size_t f = ((size_t)x * R.d[1] + y) * R.d[2] + z;
    if (st[f] != 1) continue;
    bool hit = false;
    for (int k = 0; k < 13 && !hit; ++k) {
      int dx = k / 9 - 1, dy = (k / 3) % 3 - 1, dz = k % 3 - 1;
      size_t g = ((size_t)(x + dx) * R.d[1] + (y + dy)) * R.d[2] + (z + dz);
      if (st[g] == 2) hit = true;
    }
    if (hit) {
      st[f] = 2;
      any = 1;
    }
  }
  if (any) atomicOr(changed, 1);
}

__global__ void k_xvalue(float* __restrict__ t, const float* __restrict__ v, const uint8_t* __restrict__ st,
                         const uint8_t* __restrict__ valid, XRange R) {
  const int m0 = R.hi[0] - R.lo[0] + 1, m1 = R.hi[1] - R.lo[1] + 1, m2 = R.hi[2] - R.lo[2] + 1;
  size_t total = (size_t)m0 * m1 * m2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    int z = R.lo[2] + (int)(i % m2), y = R.lo[1] + (int)((i / m2) % m1), x = R.lo[0] + (int)(i / ((size_t)m1 * m2));
    size_t f = ((size_t)x * R.d[1] + y) * R.d[2] + z;
    float out = v[f];
    if (st[f] == 2) {
      float sum = 0.0f;
      int count = 0;
      for (int k = 0; k < 27; ++k) {
        int dx = k / 9 - 1, dy = (k / 3) % 3 - 1, dz = k % 3 - 1;
        size_t g = ((size_t)(x + dx) * R.d[1] + (y + dy)) * R.d[2] + (z + dz);
        bool ok = valid[vidx(R, x + dx, y + dy, z + dz)] != 0;
        if (!ok && k < 13) ok = (st[g] == 2);
        if (ok) {
          sum += v[g];
          ++count;
        }
      }
      out = sum / (float)count;
    }
    t[f] = out;
  }
}

__global__ void k_xcommit(uint8_t* __restrict__ valid, const uint8_t* __restrict__ st, XRange R) {
  const int m0 = R.hi[0] - R.lo[0] + 1, m1 = R.hi[1] - R.lo[1] + 1, m2 = R.hi[2] - R.lo[2] + 1;
  size_t total = (size_t)m0 * m1 * m2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    int z = R.lo[2] + (int)(i % m2), y = R.lo[1] + (int)((i / m2) % m1), x = R.lo[0] + (int)(i / ((size_t)m1 * m2));
    if (st[((size_t)x * R.d[1] + y) * R.d[2] + z] == 2) valid[vidx(R, x, y, z)] = 1;
  }
}

// ---------------------------------------------------------------- divergence
__global__ void k_divergence(const float* __restrict__ vx, const float* __restrict__ vy,
                             const float* __restrict__ vz, const float* __restrict__ lwx,
                             const float* __restrict__ lwy, const float* __restrict__ lwz,
                             const float* __restrict__ phi, GridP g, double* __restrict__ partial,
                             unsigned* __restrict__ maxbits, int* __restrict__ count) {
  const int N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
  size_t total = (size_t)N0 * N1 * N2;
  double acc = 0.0;
  float mx = 0.f;
  int cnt = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    if (phi[i] > 0.0f) continue;
    int zi = (int)(i % N2), yi = (int)((i / N2) % N1), xi = (int)(i / ((size_t)N1 * N2));
    size_t fx0 = ((size_t)xi * N1 + yi) * N2 + zi, fx1 = fx0 + (size_t)N1 * N2;
    size_t fy0 = ((size_t)xi * (N1 + 1) + yi) * N2 + zi, fy1 = fy0 + N2;
    size_t fz0 = ((size_t)xi * N1 + yi) * (N2 + 1) + zi, fz1 = fz0 + 1;
    float d = lwx[fx1] * vx[fx1] - lwx[fx0] * vx[fx0];
    d += lwy[fy1] * vy[fy1] - lwy[fy0] * vy[fy0];
    d += lwz[fz1] * vz[fz1] - lwz[fz0] * vz[fz0];
    d /= g.dx[0];
    acc += (double)d;
    float ad = fabsf(d);
    mx = ad > mx ? ad : mx;
    ++cnt;
  }
  __shared__ double sacc[kBlock];
  __shared__ float smx[kBlock];
  __shared__ int scnt[kBlock];
  sacc[threadIdx.x] = acc;
  smx[threadIdx.x] = mx;
  scnt[threadIdx.x] = cnt;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sacc[threadIdx.x] += sacc[threadIdx.x + s];
      smx[threadIdx.x] = fmaxf(smx[threadIdx.x], smx[threadIdx.x + s]);
      scnt[threadIdx.x] += scnt[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = sacc[0];
    atomicMax(maxbits, __float_as_uint(smx[0]));
    atomicAdd(count, scnt[0]);
  }
}

__global__ void k_sum_partials(const double* __restrict__ partial, int n, double* __restrict__ out) {
  __shared__ double s[kBlock];
  double a = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) a += partial[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int k = kBlock / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

}  // namespace

bool hip_ok(Ctx& c, hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  c.err = std::string(what) + ": " + hipGetErrorString(e);
  return false;
}

void launch_fill_f32(Ctx& c, float* a, size_t n, float v) {
  hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n)), dim3(kBlock), 0, c.s, a, n, v);
}
void launch_fill_u8(Ctx& c, uint8_t* a, size_t n, uint8_t v) {
  hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(n)), dim3(kBlock), 0, c.s, a, n, v);
}

void launch_solid_phi(Ctx& c) {
  hipLaunchKernelGGL(k_solid_phi, dim3(grid_for(c.nnode)), dim3(kBlock), 0, c.s, c.solid_phi, c.g, c.d_sdf, c.nsdf);
  for (int a = 0; a < 3; ++a)
    hipLaunchKernelGGL(k_face_open, dim3(grid_for(c.nface[a])), dim3(kBlock), 0, c.s, c.lw[a], c.solid_phi, c.g, a);
}

void launch_clear(Ctx& c) {
  launch_fill_f32(c, c.pressure, c.ncell, 0.f);
  launch_fill_f32(c, c.phi, c.ncell, 3 * c.g.dx[0]);
  for (int a = 0; a < 3; ++a) {
    launch_fill_f32(c, c.v[a], c.nface[a], 0.f);
    launch_fill_f32(c, c.tv[a], c.nface[a], 0.f);
  }
}

void launch_gravity(Ctx& c, float dt) {
  float gdt = c.cfg.gravity * dt;
  hipLaunchKernelGGL(k_add_interior, dim3(grid_for(c.nface[1])), dim3(kBlock), 0, c.s, c.v[1], c.g, 1, gdt);
  int ax = c.cfg.tank_axis;
  if (ax >= 0 && ax <= 2 && c.cfg.tank_amp != 0.f) {
    double acc = (double)c.cfg.tank_amp * std::sin(2.0 * M_PI * (double)c.cfg.tank_freq * c.time);
    float adt = (float)acc * dt;
    hipLaunchKernelGGL(k_add_interior, dim3(grid_for(c.nface[ax])), dim3(kBlock), 0, c.s, c.v[ax], c.g, ax, adt);
  }
}

void launch_velocity_update(Ctx& c, float dt) {
  size_t n = (size_t)std::max(c.g.N[0] - 2, 0) * std::max(c.g.N[1] - 2, 0) * std::max(c.g.N[2] - 2, 0);
  if (!n) return;
  hipLaunchKernelGGL(k_velocity_update, dim3(grid_for(n)), dim3(kBlock), 0, c.s, c.v[0], c.v[1], c.v[2], c.lw[0],
                     c.lw[1], c.lw[2], c.phi, c.pressure, c.g, dt, c.cfg.density);
}

int run_extrapolate(Ctx& c) {
  if (c.P == 0) return 0;
  int rounds = 0;
  for (int a = 0; a < 3; ++a) {
    XRange R;
    for (int k = 0; k < 3; ++k) {
      R.d[k] = c.g.N[k] + (k == a);
      R.lo[k] = 1;
      R.hi[k] = (k == a) ? c.g.N[k] - 1 : c.g.N[k] - 2;
      R.vd[k] = c.g.N[k] + 1;
    }
    size_t m = (size_t)std::max(R.hi[0] - R.lo[0] + 1, 0) * std::max(R.hi[1] - R.lo[1] + 1, 0) *
               std::max(R.hi[2] - R.lo[2] + 1, 0);
    if (!m) continue;
    unsigned gb = grid_for(m);
    hipLaunchKernelGGL(k_xvalid_init, dim3(gb), dim3(kBlock), 0, c.s, c.is_valid, c.lw[a], R);
    for (int iter = 0; iter < 5; ++iter) {
      launch_fill_u8(c, c.xstate, c.nface[a], 0);
      hipLaunchKernelGGL(k_xseed, dim3(gb), dim3(kBlock), 0, c.s, c.xstate, c.is_valid, R);
      while (true) {
        *c.h_flag = 0;
        if (!hip_ok(c, hipMemsetAsync(c.dflag, 0, sizeof(int), c.s), "memset")) return -1;
        hipLaunchKernelGGL(k_xprop, dim3(gb), dim3(kBlock), 0, c.s, c.xstate, R, c.dflag);
        if (!hip_ok(c, hipMemcpyAsync(c.h_flag, c.dflag, sizeof(int), hipMemcpyDeviceToHost, c.s), "copy")) return -1;
        if (!hip_ok(c, hipStreamSynchronize(c.s), "extrapolate")) return -1;
        ++rounds;
        if (*c.h_flag == 0) break;
      }
      hipLaunchKernelGGL(k_xvalue, dim3(gb), dim3(kBlock), 0, c.s, c.tv[a], c.v[a], c.xstate, c.is_valid, R);
      hipLaunchKernelGGL(k_xcommit, dim3(gb), dim3(kBlock), 0, c.s, c.is_valid, c.xstate, R);
      std::swap(c.v[a], c.tv[a]);  // std::swap(m_grid.tempV, m_grid.v)
    }
  }
  return rounds;
}

void launch_divergence(Ctx& c, double* h_mean, double* h_linf, int* h_count) {
  const int nb = c.npartial;
  hipMemsetAsync(c.dmaxv, 0, sizeof(unsigned), c.s);
  hipMemsetAsync(c.dflag, 0, sizeof(int), c.s);
  hipLaunchKernelGGL(k_divergence, dim3(nb), dim3(kBlock), 0, c.s, c.v[0], c.v[1], c.v[2], c.lw[0], c.lw[1], c.lw[2],
                     c.phi, c.g, c.partial, c.dmaxv, c.dflag);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, c.s, c.partial, nb, c.dsc64 + 7);
  hipMemcpyAsync(c.h_pin + 7, c.dsc64 + 7, sizeof(double), hipMemcpyDeviceToHost, c.s);
  hipMemcpyAsync(c.h_flag, c.dflag, sizeof(int), hipMemcpyDeviceToHost, c.s);
  unsigned mb = 0;
  hipMemcpyAsync(c.h_pin + 8, c.dmaxv, sizeof(unsigned), hipMemcpyDeviceToHost, c.s);
  hipStreamSynchronize(c.s);
  std::memcpy(&mb, c.h_pin + 8, sizeof(unsigned));
  int cnt = *c.h_flag;
  if (h_count) *h_count = cnt;
  if (h_mean) *h_mean = cnt ? c.h_pin[7] / (double)cnt : 0.0;
  float mf;
  std::memcpy(&mf, &mb, sizeof(float));
  if (h_linf) *h_linf = (double)mf;
}

}  // namespace lh
