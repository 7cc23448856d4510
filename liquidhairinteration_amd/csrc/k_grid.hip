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
  // blocks over the interior (x, y) rows, threads over z: no per-cell division
  const int nrow = m0 * m1;
  for (int row = blockIdx.x; row < nrow; row += gridDim.x) {
    const int xi = 1 + row / m1, yi = 1 + row % m1;
    for (int zi = 1 + (int)threadIdx.x; zi <= m2; zi += (int)blockDim.x) {
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

// Row-scan propagation.  One wavefront owns one (x, y) row of the range and
// walks z in chunks of 64 lanes.  Inside a row the dependency on the z-1
// predecessor is a chain: a candidate face is validated iff some face at or
// before it in the row, reached without crossing a non-candidate face, has a
// "base" reason (a valid neighbour at sweep start, or a validated predecessor in
// another row).  That is a segmented OR-scan, done exactly with two ballots and
// a carry between chunks, so one pass settles every in-row chain and rounds are
// only needed for chains that cross rows (the 12 predecessors in rows x-1 and
// (x, y-1)).
__device__ __forceinline__ bool row_chain(bool in_s, bool base, bool& carry) {
  const int lane = __lane_id();
  const unsigned long long S = __ballot(in_s);
  const unsigned long long B = __ballot(in_s && base);
  const unsigned long long le = (lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1ull);
  const unsigned long long brk = ~S & le;
  bool v;
  if (brk) {
    int j = 63 - __clzll(brk);
    unsigned long long win = le & ~((j == 63) ? ~0ull : ((1ull << (j + 1)) - 1ull));
    v = in_s && (B & win) != 0ull;
  } else {
    v = in_s && (((B & le) != 0ull) || carry);
  }
  carry = __shfl(v ? 1 : 0, 63) != 0;
  return v;
}

// round 0: state from the validity at sweep start; base = a valid neighbour among the 26
__global__ void k_xrow_seed(uint8_t* __restrict__ st, const uint8_t* __restrict__ valid, XRange R,
                            int* __restrict__ changed) {
  const int m0 = R.hi[0] - R.lo[0] + 1, m1 = R.hi[1] - R.lo[1] + 1;
  const int wave = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
  if (wave >= m0 * m1) return;  // whole wavefront exits together
  const int x = R.lo[0] + wave / m1, y = R.lo[1] + wave % m1;
  bool carry = false;
  int any = 0;
  for (int z0 = R.lo[2]; z0 <= R.hi[2]; z0 += 64) {
    const int z = z0 + __lane_id();
    const bool inr = z <= R.hi[2];
    bool in_s = false, base = false;
    if (inr) {
      in_s = !valid[vidx(R, x, y, z)];
      if (in_s)
        for (int dx = -1; dx <= 1 && !base; ++dx)
          for (int dy = -1; dy <= 1 && !base; ++dy)
            for (int dz = -1; dz <= 1; ++dz)
              if (valid[vidx(R, x + dx, y + dy, z + dz)]) { base = true; break; }
    }
    bool v = row_chain(in_s, base, carry);
    if (inr) st[((size_t)x * R.d[1] + y) * R.d[2] + z] = v ? 2 : (in_s ? 1 : 0);
    any |= v ? 1 : 0;
  }
  if (__any(any) && __lane_id() == 0) atomicOr(changed, 1);
}

// later rounds: base = already validated, or a validated predecessor in another row
__global__ void k_xrow_prop(uint8_t* __restrict__ st, XRange R, int* __restrict__ changed) {
  const int m0 = R.hi[0] - R.lo[0] + 1, m1 = R.hi[1] - R.lo[1] + 1;
  const int wave = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
  if (wave >= m0 * m1) return;
  const int x = R.lo[0] + wave / m1, y = R.lo[1] + wave % m1;
  const size_t sx = (size_t)R.d[1] * R.d[2], sy = (size_t)R.d[2];
  const size_t row = (size_t)x * sx + (size_t)y * sy;
  bool carry = false;
  int grew = 0;
  for (int z0 = R.lo[2]; z0 <= R.hi[2]; z0 += 64) {
    const int z = z0 + __lane_id();
    const bool inr = z <= R.hi[2];
    uint8_t s = inr ? st[row + z] : 0;
    bool in_s = s != 0, base = (s == 2);
    if (inr && s == 1) {
      // predecessors outside this row: (x-1, y-1..y+1, z-1..z+1) and (x, y-1, z-1..z+1)
      for (int k = 0; k < 12 && !base; ++k) {
        int dx = (k < 9) ? -1 : 0;
        int dy = (k < 9) ? (k / 3 - 1) : -1;
        int dz = (k < 9) ? (k % 3 - 1) : (k - 9 - 1);
        if (st[row + (ptrdiff_t)dx * sx + (ptrdiff_t)dy * sy + z + dz] == 2) base = true;
      }
    }
    bool v = row_chain(in_s, base, carry);
    if (inr && v && s == 1) {
      st[row + z] = 2;
      grew = 1;
    }
  }
  if (__any(grew) && __lane_id() == 0) atomicOr(changed, 1);
}

// Plane rounds: one 64-lane workgroup owns one x-plane and visits its rows in
// increasing y, so a chain that moves through rows of the same plane (offsets
// (0,-1,-1..1)) is settled in a single pass, as in the sequential sweep.  The
// plane's previous row and the x-1 plane's rows y-1..y+1 are kept in LDS.
// Only chains that cross planes need further rounds.
constexpr int kPlaneRow = 2064;  // LDS row capacity (face rows up to 2063 long)

__global__ void __launch_bounds__(64) k_xplane_prop(uint8_t* __restrict__ st, XRange R, int* __restrict__ changed) {
  __shared__ uint8_t rows[5][kPlaneRow];  // [0..2] plane x-1 rows y-1, y, y+1; [3] prev row; [4] current row
  const int x = R.lo[0] + blockIdx.x;
  const int lane = threadIdx.x;
  const int D2 = R.d[2];
  const size_t sx = (size_t)R.d[1] * R.d[2], sy = (size_t)R.d[2];
  // LDS index = z + 1 (index 0 stands for z = -1, never a candidate)
  int ia = 0, ib = 1, ic = 2, ip = 3, iq = 4;
  for (int k = lane; k < 5 * kPlaneRow; k += 64) (&rows[0][0])[k] = 0;
  __syncthreads();
  auto load_row = [&](int buf, int xx, int yy) {
    const uint8_t* src = st + (size_t)xx * sx + (size_t)yy * sy;
    for (int z = lane; z < D2; z += 64) rows[buf][z + 1] = src[z];
  };
  load_row(ia, x - 1, R.lo[1] - 1);
  load_row(ib, x - 1, R.lo[1]);
  load_row(ip, x, R.lo[1] - 1);
  int grew = 0;
  for (int y = R.lo[1]; y <= R.hi[1]; ++y) {
    load_row(ic, x - 1, y + 1);
    __syncthreads();
    bool carry = false;
    const size_t row = (size_t)x * sx + (size_t)y * sy;
    for (int z0 = R.lo[2]; z0 <= R.hi[2]; z0 += 64) {
      const int z = z0 + lane;
      const bool inr = z <= R.hi[2];
      uint8_t s = inr ? st[row + z] : 0;
      bool in_s = s != 0, base = (s == 2);
      if (inr && s == 1) {
#pragma unroll
        for (int dz = 0; dz < 3; ++dz) {
          base = base || rows[ia][z + dz] == 2 || rows[ib][z + dz] == 2 || rows[ic][z + dz] == 2 ||
                 rows[ip][z + dz] == 2;
        }
      }
      bool v = row_chain(in_s, base, carry);
      uint8_t ns = inr ? (v ? (uint8_t)2 : s) : (uint8_t)0;
      if (z + 1 < kPlaneRow) rows[iq][z + 1] = ns;
      if (inr && v && s == 1) {
        st[row + z] = 2;
        grew = 1;
      }
    }
    __syncthreads();
    int t = ip;  // current row becomes the previous row; the x-1 window slides up
    ip = iq;
    iq = t;
    t = ia;
    ia = ib;
    ib = ic;
    ic = t;
  }
  if (__any(grew) && lane == 0) atomicOr(changed, 1);
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

// recording variant of k_xvalue: also appends (face, neighbour mask) of every
// validated face to the sweep's list (list order is irrelevant: faces are independent)
__global__ void k_xvalue_rec(float* __restrict__ t, const float* __restrict__ v, const uint8_t* __restrict__ st,
                             const uint8_t* __restrict__ valid, XRange R, uint32_t* __restrict__ list,
                             int* __restrict__ counter) {
  const int m0 = R.hi[0] - R.lo[0] + 1, m1 = R.hi[1] - R.lo[1] + 1, m2 = R.hi[2] - R.lo[2] + 1;
  size_t total = (size_t)m0 * m1 * m2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    int z = R.lo[2] + (int)(i % m2), y = R.lo[1] + (int)((i / m2) % m1), x = R.lo[0] + (int)(i / ((size_t)m1 * m2));
    size_t f = ((size_t)x * R.d[1] + y) * R.d[2] + z;
    float out = v[f];
    if (st[f] == 2) {
      float sum = 0.0f;
      int count = 0;
      uint32_t mask = 0;
      for (int k = 0; k < 27; ++k) {
        int dx = k / 9 - 1, dy = (k / 3) % 3 - 1, dz = k % 3 - 1;
        size_t g = ((size_t)(x + dx) * R.d[1] + (y + dy)) * R.d[2] + (z + dz);
        bool ok = valid[vidx(R, x + dx, y + dy, z + dz)] != 0;
        if (!ok && k < 13) ok = (st[g] == 2);
        if (ok) {
          sum += v[g];
          ++count;
          mask |= 1u << k;
        }
      }
      out = sum / (float)count;
      int slot = atomicAdd(counter, 1);
      list[2 * (size_t)slot] = (uint32_t)f;
      list[2 * (size_t)slot + 1] = mask;
    }
    t[f] = out;
  }
}

// replay: new value of each listed face from its recorded neighbour set, in the
// reference's 27-offset order
__global__ void k_xreplay(float* __restrict__ t, const float* __restrict__ v, const uint32_t* __restrict__ list,
                          int n, XRange R) {
  const size_t sx = (size_t)R.d[1] * R.d[2], sy = (size_t)R.d[2];
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    size_t f = list[2 * (size_t)e];
    uint32_t mask = list[2 * (size_t)e + 1];
    float sum = 0.0f;
    int count = 0;
    for (int k = 0; k < 27; ++k) {
      if (!(mask & (1u << k))) continue;
      ptrdiff_t off = (ptrdiff_t)(k / 9 - 1) * (ptrdiff_t)sx + (ptrdiff_t)((k / 3) % 3 - 1) * (ptrdiff_t)sy + (k % 3 - 1);
      sum += v[(ptrdiff_t)f + off];
      ++count;
    }
    t[f] = sum / (float)count;
  }
}

__global__ void k_xcopy_list(float* __restrict__ t, const float* __restrict__ v, const uint32_t* __restrict__ list,
                             int n) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    size_t f = list[2 * (size_t)e];
    t[f] = v[f];
  }
}

// t = v on the range: blocks over its (x, y) rows, threads over z (no per-element division)
__global__ void k_xcopy_range(float* __restrict__ t, const float* __restrict__ v, XRange R) {
  const int m1 = R.hi[1] - R.lo[1] + 1;
  const int nrow = (R.hi[0] - R.lo[0] + 1) * m1;
  for (int row = blockIdx.x; row < nrow; row += gridDim.x) {
    const int x = R.lo[0] + row / m1, y = R.lo[1] + row % m1;
    const size_t base = ((size_t)x * R.d[1] + y) * R.d[2];
    for (int z = R.lo[2] + (int)threadIdx.x; z <= R.hi[2]; z += (int)blockDim.x) t[base + z] = v[base + z];
  }
}

// *differ |= 1 if the byte arrays differ: 16 bytes per thread and step, one atomic per block
__global__ void k_cmp_u8(const uint8_t* __restrict__ a, const uint8_t* __restrict__ b, size_t n,
                         int* __restrict__ differ) {
  int d = 0;
  const size_t n16 = n / 16;
  const uint4* a4 = reinterpret_cast<const uint4*>(a);
  const uint4* b4 = reinterpret_cast<const uint4*>(b);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x) {
    const uint4 p = a4[i], q = b4[i];
    d |= ((p.x ^ q.x) | (p.y ^ q.y) | (p.z ^ q.z) | (p.w ^ q.w)) != 0u;
  }
  for (size_t i = 16 * n16 + blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    d |= (a[i] != b[i]) ? 1 : 0;
  if (__syncthreads_or(d) && threadIdx.x == 0) atomicOr(differ, 1);
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
                             const float* __restrict__ phi, GridP g, double* __restrict__ partial) {
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
  if (threadIdx.x == 0) {  // the partial layout of the row-list divergence's final reduction
    partial[blockIdx.x] = sacc[0];
    partial[gridDim.x + blockIdx.x] = (double)smx[0];
    partial[2 * gridDim.x + blockIdx.x] = (double)scnt[0];
  }
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

void launch_clear(Ctx& c, bool in_step) {
  // liquid phi writes every cell, P2G every face, the solve's result every pressure cell
  const bool full = !in_step || c.P == 0;
  if (full) {
    launch_fill_f32(c, c.pressure, c.ncell, 0.f);
    launch_fill_f32(c, c.phi, c.ncell, 3 * c.g.dx[0]);
  }
  for (int a = 0; a < 3; ++a) {
    if (full) launch_fill_f32(c, c.v[a], c.nface[a], 0.f);
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
  hipLaunchKernelGGL(k_velocity_update,
                     dim3((unsigned)std::max(1, std::min((c.g.N[0] - 2) * (c.g.N[1] - 2), 65535))), dim3(kBlock),
                     0, c.s, c.v[0], c.v[1], c.v[2], c.lw[0],
                     c.lw[1], c.lw[2], c.phi, c.pressure, c.g, dt, c.cfg.density);
}

namespace {

XRange xrange_of(const Ctx& c, int a) {
  XRange R;
  for (int k = 0; k < 3; ++k) {
    R.d[k] = c.g.N[k] + (k == a);
    R.lo[k] = 1;
    R.hi[k] = (k == a) ? c.g.N[k] - 1 : c.g.N[k] - 2;
    R.vd[k] = c.g.N[k] + 1;
  }
  return R;
}

size_t xrange_size(const XRange& R) {
  return (size_t)std::max(R.hi[0] - R.lo[0] + 1, 0) * std::max(R.hi[1] - R.lo[1] + 1, 0) *
         std::max(R.hi[2] - R.lo[2] + 1, 0);
}

bool ensure_xcache(Ctx& c) {
  if (c.xv_snap) return true;
  if (hipMalloc((void**)&c.xv_snap, c.nnode) != hipSuccess || hipMalloc((void**)&c.xv_final, c.nnode) != hipSuccess ||
      hipMalloc((void**)&c.xcount, 32 * sizeof(int)) != hipSuccess) {
    c.err = "hipMalloc failed (extrapolation schedule)";
    return false;
  }
  for (int a = 0; a < 3; ++a) {
    c.xlist_cap[a] = std::max<size_t>(xrange_size(xrange_of(c, a)), 1);
    if (hipMalloc((void**)&c.xlist[a], 2 * sizeof(uint32_t) * c.xlist_cap[a]) != hipSuccess) {
      c.err = "hipMalloc failed (extrapolation schedule)";
      return false;
    }
  }
  return true;
}

// the exact algorithm (seed, propagation rounds, values), recording the schedule
int extrapolate_record(Ctx& c) {
  int rounds = 0;
  constexpr int kBatch = 2;  // row-parallel rounds launched before the first convergence check
  hipMemcpyAsync(c.xv_snap, c.is_valid, c.nnode, hipMemcpyDeviceToDevice, c.s);
  hipMemsetAsync(c.xcount, 0, 32 * sizeof(int), c.s);
  for (int a = 0; a < 3; ++a) {
    XRange R = xrange_of(c, a);
    size_t m = xrange_size(R);
    if (!m) continue;
    const size_t rows = (size_t)(R.hi[0] - R.lo[0] + 1) * (R.hi[1] - R.lo[1] + 1);
    const unsigned gb = grid_for(m);
    const unsigned gr = (unsigned)div_up(rows * 64, kBlock);
    hipLaunchKernelGGL(k_xvalid_init, dim3(gb), dim3(kBlock), 0, c.s, c.is_valid, c.lw[a], R);
    // faces outside the range keep state 0 for the whole component
    launch_fill_u8(c, c.xstate, c.nface[a], 0);
    for (int iter = 0; iter < 5; ++iter) {
      hipLaunchKernelGGL(k_xrow_seed, dim3(gr), dim3(kBlock), 0, c.s, c.xstate, c.is_valid, R, c.dflag + 8);
      // a few cheap row-parallel rounds settle the common short cross-row chains;
      // if the last one still changed something, plane rounds (in-plane chains
      // settled per pass) run until a round changes nothing
      bool plane_mode = false;
      while (true) {
        if (!hip_ok(c, hipMemsetAsync(c.dflag + 8, 0, kBatch * sizeof(int), c.s), "memset")) return -1;
        int launched = 0;
        if (!plane_mode) {
          for (int r = 0; r < kBatch; ++r)
            hipLaunchKernelGGL(k_xrow_prop, dim3(gr), dim3(kBlock), 0, c.s, c.xstate, R, c.dflag + 8 + r);
          launched = kBatch;
        } else {
          hipLaunchKernelGGL(k_xplane_prop, dim3(R.hi[0] - R.lo[0] + 1), dim3(64), 0, c.s, c.xstate, R, c.dflag + 8);
          launched = 1;
        }
        if (!hip_ok(c, hipMemcpyAsync(c.h_flag + 8, c.dflag + 8, kBatch * sizeof(int), hipMemcpyDeviceToHost, c.s),
                    "copy"))
          return -1;
        if (!hip_ok(c, hipStreamSynchronize(c.s), "extrapolate")) return -1;
        rounds += launched;
        if (c.h_flag[8 + launched - 1] == 0) break;
        if (R.d[2] + 2 < kPlaneRow) plane_mode = true;
      }
      // one running counter per component; the value after each sweep is kept as
      // that sweep's end offset in the component's list
      hipLaunchKernelGGL(k_xvalue_rec, dim3(gb), dim3(kBlock), 0, c.s, c.tv[a], c.v[a], c.xstate, c.is_valid, R,
                         c.xlist[a], c.xcount + a);
      hipMemcpyAsync(c.xcount + 3 + a * 5 + iter, c.xcount + a, sizeof(int), hipMemcpyDeviceToDevice, c.s);
      hipLaunchKernelGGL(k_xcommit, dim3(gb), dim3(kBlock), 0, c.s, c.is_valid, c.xstate, R);
      std::swap(c.v[a], c.tv[a]);  // std::swap(m_grid.tempV, m_grid.v)
    }
  }
  hipMemcpyAsync(c.xv_final, c.is_valid, c.nnode, hipMemcpyDeviceToDevice, c.s);
  int ends[15];
  hipMemcpyAsync(c.h_flag, c.xcount + 3, 15 * sizeof(int), hipMemcpyDeviceToHost, c.s);
  if (!hip_ok(c, hipStreamSynchronize(c.s), "extrapolate")) return -1;
  std::memcpy(ends, c.h_flag, sizeof(ends));
  for (int k = 0; k < 15; ++k) c.xcnt[k] = ends[k] - ((k % 5) ? ends[k - 1] : 0);
  c.xcache_ok = true;
  return rounds;
}

// values only, from the recorded schedule: sweep k copies the faces sweep k-1
// changed (the buffers agree everywhere else in the range) and writes its own
void extrapolate_replay(Ctx& c) {
  for (int a = 0; a < 3; ++a) {
    XRange R = xrange_of(c, a);
    size_t m = xrange_size(R);
    if (!m) continue;
    size_t off = 0, prev_off = 0;
    int prev_n = 0;
    for (int iter = 0; iter < 5; ++iter) {
      int n = c.xcnt[a * 5 + iter];
      if (iter == 0) {
        hipLaunchKernelGGL(k_xcopy_range, dim3((unsigned)std::min(m / (size_t)(R.hi[2] - R.lo[2] + 1), (size_t)65535)),
                           dim3(kBlock), 0, c.s, c.tv[a], c.v[a], R);
      } else if (prev_n > 0) {
        hipLaunchKernelGGL(k_xcopy_list, dim3(grid_for(prev_n)), dim3(kBlock), 0, c.s, c.tv[a], c.v[a],
                           c.xlist[a] + 2 * prev_off, prev_n);
      }
      if (n > 0)
        hipLaunchKernelGGL(k_xreplay, dim3(grid_for(n)), dim3(kBlock), 0, c.s, c.tv[a], c.v[a], c.xlist[a] + 2 * off,
                           n, R);
      std::swap(c.v[a], c.tv[a]);
      prev_off = off;
      prev_n = n;
      off += (size_t)n;
    }
  }
  hipMemcpyAsync(c.is_valid, c.xv_final, c.nnode, hipMemcpyDeviceToDevice, c.s);
}

}  // namespace

int run_extrapolate(Ctx& c) {
  if (c.P == 0) return 0;
  if (!ensure_xcache(c)) return -1;
  if (c.xcache_ok) {
    hipMemsetAsync(c.dflag + 12, 0, sizeof(int), c.s);
    hipLaunchKernelGGL(k_cmp_u8, dim3((unsigned)std::min<size_t>(2048, div_up(c.nnode / 16 + 1, kBlock))), dim3(kBlock),
                       0, c.s, c.is_valid, c.xv_snap, c.nnode, c.dflag + 12);
    hipMemcpyAsync(c.h_flag + 12, c.dflag + 12, sizeof(int), hipMemcpyDeviceToHost, c.s);
    if (!hip_ok(c, hipStreamSynchronize(c.s), "extrapolate")) return -1;
    if (c.h_flag[12] == 0) {
      extrapolate_replay(c);
      return 0;
    }
  }
  return extrapolate_record(c);
}

void launch_divergence_arrays(hipStream_t s, const float* const v[3], const float* const lw[3], const float* phi,
                              const GridP& g, double* partial, int nb, double* out, int as_mean) {
  hipLaunchKernelGGL(k_divergence, dim3(nb), dim3(kBlock), 0, s, v[0], v[1], v[2], lw[0], lw[1], lw[2], phi, g,
                     partial);
  launch_div_final_on(s, partial, nb, out, as_mean);
}

void launch_divergence_dense(Ctx& c, double* out) {
  const float* v[3] = {c.v[0], c.v[1], c.v[2]};
  const float* lw[3] = {c.lw[0], c.lw[1], c.lw[2]};
  launch_divergence_arrays(c.s, v, lw, c.phi, c.g, c.partial, c.npartial, out, 0);
}

void launch_divergence(Ctx& c, double* h_mean, double* h_linf, int* h_count) {
  launch_divergence_dense(c, reinterpret_cast<double*>(c.h_pin_dev) + kPinDiv);
  hipStreamSynchronize(c.s);
  const int cnt = (int)c.h_pin[kPinDiv + 2];
  if (h_count) *h_count = cnt;
  if (h_mean) *h_mean = cnt ? c.h_pin[kPinDiv] / (double)cnt : 0.0;
  if (h_linf) *h_linf = c.h_pin[kPinDiv + 1];
}

}  // namespace lh
