// k_solver.hip -- CGSolver<3> (PressureSolver.cpp:422-934) on gfx950.
//
// Operator storage per multigrid level (17 B/cell instead of the reference's
// 28 B of Adiag + Acoef[6]):
//   adiag      Adiag exactly as BuildLHS accumulates it (PressureSolver.cpp:544-664)
//   t[a]       term = W*dt/(rho*dx^2) of the +a face of the cell (0 if W <= 0 or
//              the face leaves the level's box); the -a face of a cell is the
//              +a face of its -a neighbour
//   cls        0: phi > 0 (no row), 1: phi == 0 or NaN, 2: phi < 0
// The reference's off-diagonal Acoef[k] of row c towards neighbour n equals
// -term if phi_c <= 0 and phi_n < 0, else 0, so GetAx (PressureSolver.cpp:700-717)
// is  Adiag*z_c + sum_k [cls_c >= 1 and cls_n == 2] * z_n * (-term_k)
// in the reference's order (+x, -x, +y, -y, +z, -z).  Adding z_n * 0 in the
// reference is an exact no-op, so skipped terms leave the result unchanged.
//
// Dot products (Array3::Reduce, Array.h:111-122, and the denom / residual sums
// of Solve) accumulate in FP64 over fixed per-block partials and a fixed-order
// final sum, so they are deterministic; the totals are rounded to float like
// the reference's Scalar members m_oldRZ / m_newRZ / residual.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "lh_ctx.h"
#include "lh_nn.h"
#include "lh_math.h"

namespace lh {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(size_t n) {
  size_t b = div_up(n, kBlock);
  if (b > 65536 * 8) b = 65536 * 8;
  return (unsigned)(b ? b : 1);
}

struct LvP {
  int d[3];
  const float* adiag;
  const float* t0;
  const float* t1;
  const float* t2;
  const uint8_t* cls;
  const uint16_t* cmask;
};

LvP lvp(const Level& L) {
  LvP p;
  for (int k = 0; k < 3; ++k) p.d[k] = L.d[k];
  p.adiag = L.adiag;
  p.t0 = L.t[0];
  p.t1 = L.t[1];
  p.t2 = L.t[2];
  p.cls = L.cls;
  p.cmask = L.cmask;
  return p;
}

// (A z)[c], GetAx (PressureSolver.cpp:745-772): off-diagonal terms only towards neighbours of
// class 2, added in the order +x, -x, +y, -y, +z, -z.  Every load is issued before any test
// (out-of-range neighbours read cell c instead) and the terms are selected rather than
// branched around, so a launch waits for one round trip to memory, not one per test; the
// value is the same as adding only the present terms.  The classes come from the cell's
// coupling mask (one load instead of seven class bytes from five rows).
struct AxAt {
  float ax;   // (A z)[c]
  float ad;   // adiag[c]
  float zc;   // z[c]
  float v;    // v[c] of the extra array handed in (r, b, ...)
  int cls;    // 0 where cls[c] == 0, else 1
};
__device__ __forceinline__ AxAt get_ax_full(const LvP& L, const float* __restrict__ z, const float* __restrict__ v,
                                            int x, int y, int w, size_t c) {
  const size_t sx = (size_t)L.d[1] * L.d[2], sy = (size_t)L.d[2];
  const bool hxp = x + 1 < L.d[0], hxm = x >= 1, hyp = y + 1 < L.d[1], hym = y >= 1, hzp = w + 1 < L.d[2],
             hzm = w >= 1;
  const size_t cxp = hxp ? c + sx : c, cxm = hxm ? c - sx : c, cyp = hyp ? c + sy : c, cym = hym ? c - sy : c,
               czp = hzp ? c + 1 : c, czm = hzm ? c - 1 : c;
  float zc = z[c], ad = L.adiag[c], vc = v[c];
  int m = L.cmask[c];
  float zxp = z[cxp], zxm = z[cxm], zyp = z[cyp], zym = z[cym], zzp = z[czp], zzm = z[czm];
  float txp = L.t0[c], txm = L.t0[cxm], typ = L.t1[c], tym = L.t1[cym], tzp = L.t2[c], tzm = L.t2[czm];
  // an empty statement that takes every loaded value: the loads complete here, together,
  // and the compiler cannot move any of them behind the selects or the caller's tests
  asm volatile("" : "+v"(zc), "+v"(ad), "+v"(vc), "+v"(m));
  asm volatile("" : "+v"(zxp), "+v"(zxm), "+v"(zyp), "+v"(zym), "+v"(zzp), "+v"(zzm), "+v"(txp), "+v"(txm),
               "+v"(typ), "+v"(tym), "+v"(tzp), "+v"(tzm));
  const float res0 = zc * ad;
  float res = res0;
  res = (m & 1) ? res + zxp * (-txp) : res;
  res = (m & 2) ? res + zxm * (-txm) : res;
  res = (m & 4) ? res + zyp * (-typ) : res;
  res = (m & 8) ? res + zym * (-tym) : res;
  res = (m & 16) ? res + zzp * (-tzp) : res;
  res = (m & 32) ? res + zzm * (-tzm) : res;
  const bool air = (m & 64) != 0;
  return {air ? res0 : res, ad, zc, vc, air ? 0 : 1};
}
__device__ __forceinline__ float get_ax(const LvP& L, const float* __restrict__ z, int x, int y, int w, size_t c) {
  return get_ax_full(L, z, z, x, y, w, c).ax;
}

// cmask of a level from its classes (layout in Level::cmask)
__global__ void k_build_cmask(uint16_t* __restrict__ m, const uint8_t* __restrict__ cls, int d0, int d1, int d2) {
  const size_t total = (size_t)d0 * d1 * d2, sx = (size_t)d1 * d2, sy = (size_t)d2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const int z = (int)(i % d2), y = (int)((i / d2) % d1), x = (int)(i / sx);
    const bool has[6] = {x + 1 < d0, x >= 1, y + 1 < d1, y >= 1, z + 1 < d2, z >= 1};
    const size_t nb[6] = {i + sx, i - sx, i + sy, i - sy, i + 1, i - 1};
    int v = cls[i] == 0 ? 64 : 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int kn = has[k] ? cls[nb[k]] : 0;
      v |= (kn == 2 ? 1 : 0) << k;
      v |= (kn != 0 ? 1 : 0) << (8 + k);
    }
    m[i] = (uint16_t)v;
  }
}

// ---------------------------------------------------------------- active z-rows
// The kernels below visit only z-rows (x, y) that hold a liquid cell at their level:
// every cell they update has cls != 0, and every array they reduce is zero on cells
// with cls == 0 (r is zeroed there, adiag is 0 there), so skipping the other rows
// changes no value.  R = kBlock / d2 rows per block pass (one row, strided, if d2 >= kBlock).
struct RowIter {
  int R, sub, z0, zs;
  bool on;
};
__device__ __forceinline__ RowIter row_iter(int d2) {
  RowIter it;
  if (d2 >= kBlock) {
    it.R = 1;
    it.sub = 0;
    it.z0 = threadIdx.x;
    it.zs = kBlock;
  } else {
    it.R = kBlock / d2;
    it.sub = threadIdx.x / d2;
    it.z0 = threadIdx.x - it.sub * d2;
    it.zs = d2;
  }
  it.on = it.sub < it.R;
  return it;
}

// An item of a level-0 work list: a z-row (nseg == 0: the list holds rows, every cell of the row)
// or a z-segment (the list holds row * nseg + s: the kSegZ cells from kSegZ s).  Kernels map
// threads to an item's cells as RowIter does to a row of the item's length (list_iter).
struct WorkItem {
  int row, zlo, zhi;
};
__device__ __forceinline__ WorkItem work_item(int id, int nseg, int d2) {
  WorkItem w;
  if (nseg == 0) {
    w.row = id;
    w.zlo = 0;
    w.zhi = d2;
  } else {
    w.row = id / nseg;
    w.zlo = (id - w.row * nseg) * kSegZ;
    w.zhi = w.zlo + kSegZ;
  }
  return w;
}
__device__ __forceinline__ RowIter list_iter(int d2, int nseg) { return row_iter(nseg ? kSegZ : d2); }

// A row list split into 8 contiguous runs, one per XCD: workgroups are dispatched to the 8 XCDs
// round-robin, so workgroup b serves run b % 8, and a row's x/y neighbours (adjacent and nearby
// list entries) are mostly read through the same XCD's L2.  The grid is a multiple of 8; R list
// entries per workgroup pass, sub = the thread's entry among them.
struct XcdRun {
  int k, end, step;
};
__device__ __forceinline__ XcdRun xcd_run(int nr, int R, int sub) {
  const int xcd = (int)blockIdx.x & 7;
  XcdRun q;
  q.end = (int)(((long long)nr * (xcd + 1)) >> 3);
  q.step = ((int)gridDim.x >> 3) * R;
  q.k = (int)(((long long)nr * xcd) >> 3) + ((int)blockIdx.x >> 3) * R + sub;
  return q;
}

// flag[row] = 1 when the z-row holds a cell with cls != 0; one wave per row.  With smask (d2 a
// multiple of kSegZ == 64), also bit s of smask[row] when its z-segment s holds such a cell.
static_assert(kSegZ == 64, "one wave-wide load per segment");
__global__ void k_row_flags(const uint8_t* __restrict__ cls, int* __restrict__ flag, int nrow, int d2,
                            int* __restrict__ smask) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (row >= nrow) return;
  int mask = 0;
  for (int s = 0; s * 64 < d2; ++s) {  // uniform trip count
    const int z = s * 64 + lane;
    const bool v = z < d2 && cls[(size_t)row * d2 + z] != 0;
    if (__ballot(v) != 0) mask |= 1 << (s & 31);
  }
  if (lane == 0) {
    flag[row] = mask != 0 ? 1 : 0;
    if (smask) smask[row] = mask;
  }
}

// The finest level's segment list: for the rows of the (ascending) row list, the ids row * nseg +
// s of the segments that smask flags, ascending.  One block of 1024 threads; wave w owns a
// contiguous run of 64-entry chunks of the row list: pass 1 counts its segments, the block forms
// the waves' offsets, pass 2 writes each entry's segments at its offset plus the wave's inclusive
// scan of the counts.
__global__ __launch_bounds__(1024) void k_seg_compact(const int* __restrict__ rows, const int* __restrict__ nrows,
                                                      const int* __restrict__ smask, int nseg,
                                                      int* __restrict__ segs, int* __restrict__ nsegs) {
  __shared__ int s_w[16];
  const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
  const int nr = *nrows;
  const int nch = (nr + 63) >> 6;
  const int per = (nch + 15) >> 4;
  const int c0 = min(nch, w * per), c1 = min(nch, c0 + per);
  int cnt = 0;
  for (int ch = c0; ch < c1; ++ch) {
    const int i = ch * 64 + lane;
    if (i < nr) cnt += __popc(smask[rows[i]]);
  }
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
  if (lane == 0) s_w[w] = cnt;
  __syncthreads();
  int pos = 0;
  for (int k = 0; k < w; ++k) pos += s_w[k];
  for (int ch = c0; ch < c1; ++ch) {
    const int i = ch * 64 + lane;
    const int row = i < nr ? rows[i] : 0;
    const int m = i < nr ? smask[row] : 0;
    const int n = __popc(m);
    int inc = n;  // inclusive scan of the entries' counts over the wave
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(inc, o);
      if (lane >= o) inc += u;
    }
    int at = pos + inc - n;
    for (int sg = 0; sg < nseg; ++sg)
      if (m & (1 << sg)) segs[at++] = row * nseg + sg;
    pos += __shfl(inc, 63);
  }
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int k = 0; k < 16; ++k) tot += s_w[k];
    *nsegs = tot;
  }
}
// ascending list of the flagged rows (one block of 1024 threads; deterministic order).  Wave w
// owns a contiguous run of 64-flag segments; each segment is one coalesced load.  Pass 1 counts
// the wave's flags, the block forms the waves' offsets, pass 2 writes each flagged row at its
// offset plus its rank among the flagged lanes of the segment (ballot), loading 8 segments ahead.
__global__ __launch_bounds__(1024) void k_row_compact(const int* __restrict__ flag, int nrow, int* __restrict__ rows,
                                                      int* __restrict__ nrows) {
  __shared__ int s_w[16];
  const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
  const int nseg = (nrow + 63) >> 6;
  const int per = (nseg + 15) >> 4;
  const int s0 = min(nseg, w * per), s1 = min(nseg, s0 + per);
  int cnt = 0;
#pragma unroll 8
  for (int sg = s0; sg < s1; ++sg) {
    const int i = sg * 64 + lane;
    cnt += (i < nrow && flag[i] != 0) ? 1 : 0;
  }
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
  if (lane == 0) s_w[w] = cnt;
  __syncthreads();
  int pos = 0;
  for (int k = 0; k < w; ++k) pos += s_w[k];
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int sg = s0; sg < s1; sg += 8) {
    int fv[8];  // 8 segments' flags loaded before the ballots
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = (sg + u) * 64 + lane;
      fv[u] = (sg + u < s1 && i < nrow) ? flag[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const bool f = fv[u] != 0;
      const unsigned long long m = __ballot(f);
      if (f) rows[pos + __popcll(m & below)] = (sg + u) * 64 + lane;
      pos += __popcll(m);
    }
  }
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int k = 0; k < 16; ++k) tot += s_w[k];
    *nrows = tot;
  }
}

// qflag[X, Y] of a coarse level: its own liquid row, or a child row (2X+i, 2Y+j) that is a
// liquid or q row of the finer level (fqflag null on level 0, whose r is zero off liquid)
__global__ void k_qrow_flags(const int* __restrict__ caflag, const int* __restrict__ faflag,
                             const int* __restrict__ fqflag, int* __restrict__ cqflag, int c0, int c1, int f1) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= c0 * c1) return;
  const int X = row / c1, Y = row - X * c1;
  int f = caflag[row];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int fr = (2 * X + (k >> 1)) * f1 + 2 * Y + (k & 1);
    f |= faflag[fr] | (fqflag ? fqflag[fr] : 0);
  }
  cqflag[row] = f ? 1 : 0;
}

// zflag of a level: its own liquid row, or a parent row (x/2, y/2) in the coarser level's
// z-support (czflag null on the coarsest level)
__global__ void k_zrow_flags(const int* __restrict__ aflag, const int* __restrict__ czflag, int* __restrict__ zflag,
                             int d0, int d1, int c0, int c1) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= d0 * d1) return;
  const int x = row / d1, y = row - x * d1;
  int f = aflag[row];
  if (czflag && (x >> 1) < c0 && (y >> 1) < c1) f |= czflag[(x >> 1) * c1 + (y >> 1)];
  zflag[row] = f ? 1 : 0;
}
__global__ void k_or_flags(int* __restrict__ a, const int* __restrict__ b, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) a[i] |= b[i];
}

// ---------------------------------------------------------------- assembly
struct FineP {
  int N[3];
  const float* phi;
  const float* lw0;
  const float* lw1;
  const float* lw2;
};

// GetApproxLiquidPhi, PressureSolver.cpp:476-494
__device__ __forceinline__ float approx_phi(const FineP& F, int x, int y, int z, int l) {
  const size_t s1 = (size_t)F.N[2], s0 = (size_t)F.N[1] * F.N[2];
  if (l == 0) return F.phi[x * s0 + y * s1 + z];
  int f = 1 << l, h = (1 << (l - 1)) - 1;
  int bx = f * x + h, by = f * y + h, bz = f * z + h;
  const float* p = F.phi;
  auto P = [&](int a, int b, int c) { return p[(size_t)a * s0 + (size_t)b * s1 + c]; };
  return trilerp(P(bx, by, bz), P(bx + 1, by, bz), P(bx, by + 1, bz), P(bx + 1, by + 1, bz), P(bx, by, bz + 1),
                 P(bx + 1, by, bz + 1), P(bx, by + 1, bz + 1), P(bx + 1, by + 1, bz + 1), 0.5f, 0.5f, 0.5f);
}
// GetApproxWx/Wy/Wz, PressureSolver.cpp:496-542
__device__ __forceinline__ float approx_w(const FineP& F, int a, int x, int y, int z, int l) {
  int d1 = F.N[1] + (a == 1), d2 = F.N[2] + (a == 2);
  const float* w = (a == 0) ? F.lw0 : (a == 1) ? F.lw1 : F.lw2;
  auto W = [&](int i, int j, int k) { return w[((size_t)i * d1 + j) * d2 + k]; };
  if (l == 0) return W(x, y, z);
  int f = 1 << l, h = (1 << (l - 1)) - 1;
  if (a == 0) {
    int bx = f * x, by = f * y + h, bz = f * z + h;
    return bilerp(W(bx, by, bz), W(bx, by + 1, bz), W(bx, by, bz + 1), W(bx, by + 1, bz + 1), 0.5f, 0.5f);
  } else if (a == 1) {
    int bx = f * x + h, by = f * y, bz = f * z + h;
    return bilerp(W(bx, by, bz), W(bx + 1, by, bz), W(bx, by, bz + 1), W(bx + 1, by, bz + 1), 0.5f, 0.5f);
  }
  int bx = f * x + h, by = f * y + h, bz = f * z;
  return bilerp(W(bx, by, bz), W(bx + 1, by, bz), W(bx, by + 1, bz), W(bx + 1, by + 1, bz), 0.5f, 0.5f);
}

__global__ void k_build_lhs(float* __restrict__ adiag, float* __restrict__ t0, float* __restrict__ t1,
                            float* __restrict__ t2, uint8_t* __restrict__ cls, FineP F, int l, int d0, int d1,
                            int d2, float dt, float rho, float dxl0, float dxl1, float dxl2) {
  size_t total = (size_t)d0 * d1 * d2;
  const int dims[3] = {d0, d1, d2};
  const float dxl[3] = {dxl0, dxl1, dxl2};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    int c[3];
    c[2] = (int)(i % d2);
    c[1] = (int)((i / d2) % d1);
    c[0] = (int)(i / ((size_t)d1 * d2));
    float cphi = approx_phi(F, c[0], c[1], c[2], l);
    cls[i] = (cphi > 0.0f) ? 0 : ((cphi < 0.0f) ? 2 : 1);
    // +a face terms (independent of this cell's class: the neighbour's row uses them)
    float tp[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      tp[a] = 0.f;
      if (c[a] + 1 < dims[a]) {
        int f[3] = {c[0], c[1], c[2]};
        f[a] += 1;
        float wv = approx_w(F, a, f[0], f[1], f[2], l);
        if (wv > 0.0f) tp[a] = wv * dt / (rho * dxl[a] * dxl[a]);
      }
    }
    t0[i] = tp[0];
    t1[i] = tp[1];
    t2[i] = tp[2];
    float diag = 0.0f;
    if (!(cphi > 0.0f)) {
      // right, left, top, bottom, front, back (PressureSolver.cpp:557-661)
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const int a = k >> 1;
        const int s = (k & 1) ? -1 : 1;
        if (s > 0 ? !(c[a] + 1 < dims[a]) : !(c[a] - 1 >= 0)) continue;
        int f[3] = {c[0], c[1], c[2]};
        if (s > 0) f[a] += 1;
        float wv = approx_w(F, a, f[0], f[1], f[2], l);
        if (!(wv > 0.0f)) continue;
        float term = wv * dt / (rho * dxl[a] * dxl[a]);
        int nb[3] = {c[0], c[1], c[2]};
        nb[a] += s;
        float ph = approx_phi(F, nb[0], nb[1], nb[2], l);
        if (ph < 0.0f) {
          diag += term;
        } else {
          float frac = max_tenth(fraction_in_region(cphi, ph));
          diag += term / frac;
        }
      }
    }
    adiag[i] = diag;
  }
}

// BuildRHS, PressureSolver.cpp:682-698
__global__ void k_build_rhs(float* __restrict__ b, const float* __restrict__ phi, const float* __restrict__ vx,
                            const float* __restrict__ vy, const float* __restrict__ vz, const float* __restrict__ lwx,
                            const float* __restrict__ lwy, const float* __restrict__ lwz, GridP g) {
  const int N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
  size_t total = (size_t)N0 * N1 * N2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    float r = 0.0f;
    if (!(phi[i] > 0.0f)) {
      int zi = (int)(i % N2), yi = (int)((i / N2) % N1), xi = (int)(i / ((size_t)N1 * N2));
      size_t fx0 = ((size_t)xi * N1 + yi) * N2 + zi, fx1 = fx0 + (size_t)N1 * N2;
      size_t fy0 = ((size_t)xi * (N1 + 1) + yi) * N2 + zi, fy1 = fy0 + N2;
      size_t fz0 = ((size_t)xi * N1 + yi) * (N2 + 1) + zi, fz1 = fz0 + 1;
      r -= lwx[fx1] * vx[fx1] / g.dx[0];
      r += lwx[fx0] * vx[fx0] / g.dx[0];
      r -= lwy[fy1] * vy[fy1] / g.dx[1];
      r += lwy[fy0] * vy[fy0] / g.dx[1];
      r -= lwz[fz1] * vz[fz1] / g.dx[2];
      r += lwz[fz0] * vz[fz0] / g.dx[2];
    }
    b[i] = r;
  }
}

// ---------------------------------------------------------------- V-cycle
// Smooth, PressureSolver.cpp:821-837: mode 0 is (1. - w) * z + w * term with the
// reference's double/float promotion (quirk Q4); mode 1 is z + w * term.
// one colour pass over the active rows; kLevel0 only gives the finest level its own
// symbol in kernel traces.  Thread -> (row, z of the pass's colour in that row); the rows in
// per-XCD runs (xcd_run).
template <bool kLevel0>
__global__ __launch_bounds__(kBlock) void k_smooth(float* __restrict__ z, const float* __restrict__ r, LvP L,
                                                   const int* __restrict__ rows, const int* __restrict__ nrows,
                                                   int phase, float w, int mode) {
  const int d1 = L.d[1], d2 = L.d[2];
  const int H = (d2 + 1) >> 1;  // cells of one colour per row (at most)
  const int R = H >= kBlock ? 1 : kBlock / H;
  const int sub = H >= kBlock ? 0 : (int)threadIdx.x / H;
  if (sub >= R) return;
  const int h0 = (int)threadIdx.x - sub * (H >= kBlock ? 0 : H), hs = H >= kBlock ? kBlock : H;
  const int kcap = L.d[0] * d1 - 1;
  const XcdRun run = xcd_run(*nrows, R, sub);
  const int kstep = run.step;
  int k = run.k;
  int row_next = rows[min(k, kcap)];
  for (; k < run.end; k += kstep) {
    const int row = row_next;
    row_next = rows[min(k + kstep, kcap)];
    const int xi = row / d1, yi = row - xi * d1;
    const int par = (xi + yi + phase) & 1;
    for (int h = h0; h < H; h += hs) {
      const int zi = 2 * h + par;
      if (zi >= d2) continue;
      const size_t i = (size_t)row * d2 + zi;
      // finest level: most cells of a visited row are air (adiag 0), skipped after one load
      if (kLevel0 && is_small(L.adiag[i])) continue;
      const AxAt g = get_ax_full(L, z, r, xi, yi, zi, i);
      const float ad = g.ad, rv = g.v, zv = g.zc;
      if (is_small(ad)) continue;
      float term = (rv - g.ax) / ad;
      if (mode == 0) {
        float wt = w * term;
        z[i] = (float)((1.0 - (double)w) * (double)zv + (double)wt);
      } else {
        z[i] = zv + w * term;
      }
    }
  }
}

// The finest level's colour pass over its segment list (k_smooth<true> restricted to the z-segments
// that hold a liquid cell: a segment without one has adiag 0 everywhere, and k_smooth skips such
// cells): eight segments of kSegZ cells per 256-thread block, a thread one cell of the pass's colour
// per segment; the segments in per-XCD runs of the list.
__global__ __launch_bounds__(kBlock) void k_smooth_seg(float* __restrict__ z, const float* __restrict__ r, LvP L,
                                                       const int* __restrict__ segs, const int* __restrict__ nsegs,
                                                       int nseg, int phase, float w, int mode) {
  static_assert(kBlock == 8 * (kSegZ / 2), "eight segments per block");
  const int d1 = L.d[1], d2 = L.d[2];
  const int sub = (int)threadIdx.x >> 5, h = (int)threadIdx.x & 31;
  const int kcap = L.d[0] * d1 * nseg - 1;
  const XcdRun run = xcd_run(*nsegs, 8, sub);
  const int kstep = run.step;
  int k = run.k;
  int seg_next = segs[min(k, kcap)];
  for (; k < run.end; k += kstep) {
    const int id = seg_next;
    seg_next = segs[min(k + kstep, kcap)];
    const int row = id / nseg, sg = id - row * nseg;
    const int xi = row / d1, yi = row - xi * d1;
    const int zi = sg * kSegZ + 2 * h + ((xi + yi + phase) & 1);
    const size_t i = (size_t)row * d2 + zi;
    if (is_small(L.adiag[i])) continue;
    const AxAt g = get_ax_full(L, z, r, xi, yi, zi, i);
    const float ad = g.ad, rv = g.v, zv = g.zc;
    if (is_small(ad)) continue;
    float term = (rv - g.ax) / ad;
    if (mode == 0) {
      float wt = w * term;
      z[i] = (float)((1.0 - (double)w) * (double)zv + (double)wt);
    } else {
      z[i] = zv + w * term;
    }
  }
}

// One red-black sweep (both colours) in one kernel: Smooth(l, first) followed by
// Smooth(l, 1 - first).  The first colour uses the other colour's old values and
// the second colour uses the first colour's new values, exactly as two separate
// colour passes.  A workgroup owns an 8 x 8 x 64 tile: it stages the old z of
// the tile plus a 2-cell halo in LDS, updates the first colour on the tile plus
// a 1-cell halo (the halo cells are recomputed by the neighbouring tiles too,
// with identical results), then the second colour on the tile, and writes the
// tile.  z is read from zin and written to zout (ping-pong) so tiles never see
// each other's writes; every cell of the tile is written.  Each array element is
// fetched from HBM about once per sweep instead of once per colour.
constexpr int kSwX = 8, kSwY = 8, kSwZ = 64;
constexpr int kSwHX = kSwX + 4, kSwHY = kSwY + 4, kSwHZ = kSwZ + 4;

template <int MODE>
__device__ __forceinline__ float relax(float zc, float r, float ax, float ad, float w) {
  float term = (r - ax) / ad;
  if (MODE == 0) {
    float wt = w * term;
    return (float)((1.0 - (double)w) * (double)zc + (double)wt);
  }
  return zc + w * term;
}

// GetAx with the 7 z values supplied (same order and conditions as get_ax)
__device__ __forceinline__ float get_ax7(const LvP& L, size_t c, int x, int y, int w, float zc, float zxp, float zxm,
                                        float zyp, float zym, float zzp, float zzm) {
  float res = zc * L.adiag[c];
  if (L.cls[c] == 0) return res;
  const size_t sx = (size_t)L.d[1] * L.d[2], sy = (size_t)L.d[2];
  if (x + 1 < L.d[0] && L.cls[c + sx] == 2) res += zxp * (-L.t0[c]);
  if (x - 1 >= 0 && L.cls[c - sx] == 2) res += zxm * (-L.t0[c - sx]);
  if (y + 1 < L.d[1] && L.cls[c + sy] == 2) res += zyp * (-L.t1[c]);
  if (y - 1 >= 0 && L.cls[c - sy] == 2) res += zym * (-L.t1[c - sy]);
  if (w + 1 < L.d[2] && L.cls[c + 1] == 2) res += zzp * (-L.t2[c]);
  if (w - 1 >= 0 && L.cls[c - 1] == 2) res += zzm * (-L.t2[c - 1]);
  return res;
}

// The operands of one cell's update, loaded without any test (out-of-range neighbours read
// the cell itself): z and class of the cell and its six neighbours, its diagonal, its three
// +face terms and the three -face terms of its minus neighbours, and r.
// a cell's stencil for the row sweep; T is the type of z (float, or double for the coarsest level
// in FP64), the coefficients stay FP32 and are promoted
template <class T>
struct Sten {
  T zc, zxp, zxm, zyp, zym, zzp, zzm;
  float ad, rv, txp, txm, typ, tym, tzp, tzm;
  int k, kxp, kxm, kyp, kym, kzp, kzm;
  bool hxp, hxm, hyp, hym, hzp, hzm;
};
template <class T>
__device__ __forceinline__ Sten<T> sten_load(const LvP& L, const T* __restrict__ z, const float* __restrict__ r, int x,
                                             int y, int w, size_t c) {
  Sten<T> s;
  const size_t sx = (size_t)L.d[1] * L.d[2], sy = (size_t)L.d[2];
  s.hxp = x + 1 < L.d[0];
  s.hxm = x >= 1;
  s.hyp = y + 1 < L.d[1];
  s.hym = y >= 1;
  s.hzp = w + 1 < L.d[2];
  s.hzm = w >= 1;
  const size_t cxp = s.hxp ? c + sx : c, cxm = s.hxm ? c - sx : c, cyp = s.hyp ? c + sy : c,
               cym = s.hym ? c - sy : c, czp = s.hzp ? c + 1 : c, czm = s.hzm ? c - 1 : c;
  s.zc = z[c];
  s.ad = L.adiag[c];
  s.rv = r[c];
  // the classes the stencil tests, from the cell's coupling mask: 0 where cls == 0, 2 where a
  // neighbour exists with cls 2 (the only values it compares against)
  const int m = L.cmask[c];
  s.k = (m & 64) ? 0 : 1;
  s.kxp = (m & 1) ? 2 : 0;
  s.kxm = (m & 2) ? 2 : 0;
  s.kyp = (m & 4) ? 2 : 0;
  s.kym = (m & 8) ? 2 : 0;
  s.kzp = (m & 16) ? 2 : 0;
  s.kzm = (m & 32) ? 2 : 0;
  s.zxp = z[cxp];
  s.zxm = z[cxm];
  s.zyp = z[cyp];
  s.zym = z[cym];
  s.zzp = z[czp];
  s.zzm = z[czm];
  s.txp = L.t0[c];
  s.txm = L.t0[cxm];
  s.typ = L.t1[c];
  s.tym = L.t1[cym];
  s.tzp = L.t2[c];
  s.tzm = L.t2[czm];
  return s;
}
// the loaded values are complete here (see get_ax_full)
template <class T>
__device__ __forceinline__ void sten_pin(Sten<T>& s) {
  asm volatile("" : "+v"(s.zc), "+v"(s.ad), "+v"(s.rv), "+v"(s.k), "+v"(s.kxp), "+v"(s.kxm), "+v"(s.kyp),
               "+v"(s.kym), "+v"(s.kzp), "+v"(s.kzm));
  asm volatile("" : "+v"(s.zxp), "+v"(s.zxm), "+v"(s.zyp), "+v"(s.zym), "+v"(s.zzp), "+v"(s.zzm), "+v"(s.txp),
               "+v"(s.txm), "+v"(s.typ), "+v"(s.tym), "+v"(s.tzp), "+v"(s.tzm));
}
// GetAx of the cell with the given neighbour z values (order and conditions of get_ax)
template <class T>
__device__ __forceinline__ T sten_ax(const Sten<T>& s, T zc, T zxp, T zxm, T zyp, T zym, T zzp, T zzm) {
  const T res0 = zc * (T)s.ad;
  T res = res0;
  res = (s.hxp && s.kxp == 2) ? res + zxp * (T)(-s.txp) : res;
  res = (s.hxm && s.kxm == 2) ? res + zxm * (T)(-s.txm) : res;
  res = (s.hyp && s.kyp == 2) ? res + zyp * (T)(-s.typ) : res;
  res = (s.hym && s.kym == 2) ? res + zym * (T)(-s.tym) : res;
  res = (s.hzp && s.kzp == 2) ? res + zzp * (T)(-s.tzp) : res;
  res = (s.hzm && s.kzm == 2) ? res + zzm * (T)(-s.tzm) : res;
  return s.k == 0 ? res0 : res;
}
template <int MODE, class T>
__device__ __forceinline__ T sten_relax(const Sten<T>& s, T ax, T zc, float w) {
  if (is_small(s.ad)) return zc;
  const T term = ((T)s.rv - ax) / (T)s.ad;
  if (MODE == 0) {
    const T wt = (T)w * term;
    return (T)((1.0 - (double)w) * (double)zc + (double)wt);
  }
  return zc + (T)w * term;
}

// One red-black sweep on the listed 4 x 4 x 64 tiles of a level (tiles holding a restriction
// row: a row of nonzero r or of the z-support), with the algorithm of k_rbgs_sweep: the tile and
// a 2-cell halo of zin in LDS, the first colour updated on the tile plus a 1-cell halo, then the
// second colour on the tile, and every cell of the tile written to zout.  Rows outside the listed
// tiles are read by no later stage except through a zero diagonal or a masked neighbour, and
// hold zero in both buffers of a solve.  Loops run over one colour's cells only, and each cell's
// coefficients are loaded at once.  On the 128^3 level of the 256^3 collision it measured 19 us
// per sweep against 2 x 6.5 us for two colour passes (each thread updates ~5 halo cells of the
// first colour in sequence), so it is the coarseSweep "tiles" option, not the default.
constexpr int kTsX = 4, kTsY = 4, kTsZ = 64;
constexpr int kTsHX = kTsX + 4, kTsHY = kTsY + 4, kTsHZ = kTsZ + 4;

struct Coef {
  float ad, rv, txp, txm, typ, tym, tzp, tzm;
  int k, kxp, kxm, kyp, kym, kzp, kzm;
  bool hxp, hxm, hyp, hym, hzp, hzm;
};
__device__ __forceinline__ Coef coef_load(const LvP& L, const float* __restrict__ r, int x, int y, int w, size_t c) {
  Coef s;
  const size_t sx = (size_t)L.d[1] * L.d[2], sy = (size_t)L.d[2];
  s.hxp = x + 1 < L.d[0];
  s.hxm = x >= 1;
  s.hyp = y + 1 < L.d[1];
  s.hym = y >= 1;
  s.hzp = w + 1 < L.d[2];
  s.hzm = w >= 1;
  const size_t cxp = s.hxp ? c + sx : c, cxm = s.hxm ? c - sx : c, cyp = s.hyp ? c + sy : c,
               cym = s.hym ? c - sy : c, czp = s.hzp ? c + 1 : c, czm = s.hzm ? c - 1 : c;
  s.ad = L.adiag[c];
  s.rv = r[c];
  // the classes the stencil tests, from the cell's coupling mask: 0 where cls == 0, 2 where a
  // neighbour exists with cls 2 (the only values it compares against)
  const int m = L.cmask[c];
  s.k = (m & 64) ? 0 : 1;
  s.kxp = (m & 1) ? 2 : 0;
  s.kxm = (m & 2) ? 2 : 0;
  s.kyp = (m & 4) ? 2 : 0;
  s.kym = (m & 8) ? 2 : 0;
  s.kzp = (m & 16) ? 2 : 0;
  s.kzm = (m & 32) ? 2 : 0;
  s.txp = L.t0[c];
  s.txm = L.t0[cxm];
  s.typ = L.t1[c];
  s.tym = L.t1[cym];
  s.tzp = L.t2[c];
  s.tzm = L.t2[czm];
  return s;
}
// the relaxed value of a cell from its coefficients and the 7 z values (GetAx order, then Smooth)
template <int MODE>
__device__ __forceinline__ float coef_relax(const Coef& s, float zc, float zxp, float zxm, float zyp, float zym,
                                            float zzp, float zzm, float w) {
  if (is_small(s.ad)) return zc;
  const float res0 = zc * s.ad;
  float res = res0;
  res = (s.hxp && s.kxp == 2) ? res + zxp * (-s.txp) : res;
  res = (s.hxm && s.kxm == 2) ? res + zxm * (-s.txm) : res;
  res = (s.hyp && s.kyp == 2) ? res + zyp * (-s.typ) : res;
  res = (s.hym && s.kym == 2) ? res + zym * (-s.tym) : res;
  res = (s.hzp && s.kzp == 2) ? res + zzp * (-s.tzp) : res;
  res = (s.hzm && s.kzm == 2) ? res + zzm * (-s.tzm) : res;
  const float ax = s.k == 0 ? res0 : res;
  return relax<MODE>(zc, s.rv, ax, s.ad, w);
}

template <int MODE>
__global__ void __launch_bounds__(256) k_tile_sweep(float* __restrict__ zout, const float* __restrict__ zin,
                                                     const float* __restrict__ r, LvP L, const int* __restrict__ tiles,
                                                     const int* __restrict__ ntiles, int tcap, int first, float w) {
  __shared__ float sz[kTsHX * kTsHY * kTsHZ];
  const int d0 = L.d[0], d1 = L.d[1], d2 = L.d[2];
  const int nty = (d1 + kTsY - 1) / kTsY, nzt = (d2 + kTsZ - 1) / kTsZ;
  const int nt = *ntiles;
  const size_t sx = (size_t)d1 * d2, sy = (size_t)d2;
  const int tid = threadIdx.x;
  constexpr int PX = kTsHY * kTsHZ;  // LDS x-stride
  for (int kk = blockIdx.x; kk < nt * nzt; kk += gridDim.x) {
    const int t = tiles[min(kk / nzt, tcap - 1)];
    const int bz = kk - (kk / nzt) * nzt;
    const int bx = t / nty, by = t - bx * nty;
    const int x0 = bx * kTsX, y0 = by * kTsY, z0 = bz * kTsZ;
    for (int i = tid; i < kTsHX * kTsHY * kTsHZ; i += 256) {
      const int xx = i / PX, rem = i - xx * PX, yy = rem / kTsHZ, zz = rem - yy * kTsHZ;
      const int x = x0 - 2 + xx, y = y0 - 2 + yy, zc = z0 - 2 + zz;
      float v = 0.f;
      if (x >= 0 && x < d0 && y >= 0 && y < d1 && zc >= 0 && zc < d2) v = zin[(size_t)x * sx + (size_t)y * sy + zc];
      sz[i] = v;
    }
    __syncthreads();
    // first colour on the tile plus a 1-cell halo (it reads second-colour values only)
    constexpr int RX = kTsX + 2, RY = kTsY + 2, RZH = (kTsZ + 2) / 2;
    for (int i = tid; i < RX * RY * RZH; i += 256) {
      const int xx = i / (RY * RZH), rem = i - xx * (RY * RZH), yy = rem / RZH, j = rem - yy * RZH;
      const int x = x0 - 1 + xx, y = y0 - 1 + yy;
      const int zs = z0 - 1, zc = zs + (((first - x - y - zs) % 2 + 2) % 2) + 2 * j;
      if (x < 0 || x >= d0 || y < 0 || y >= d1 || zc < 0 || zc >= d2 || zc > z0 + kTsZ) continue;
      const size_t c = (size_t)x * sx + (size_t)y * sy + zc;
      const Coef cf = coef_load(L, r, x, y, zc, c);
      const int k = (xx + 1) * PX + (yy + 1) * kTsHZ + (zc - z0 + 2);
      sz[k] = coef_relax<MODE>(cf, sz[k], sz[k + PX], sz[k - PX], sz[k + kTsHZ], sz[k - kTsHZ], sz[k + 1], sz[k - 1], w);
    }
    __syncthreads();
    // second colour on the tile, then every cell of the tile out
    constexpr int TZH = kTsZ / 2;
    for (int i = tid; i < kTsX * kTsY * TZH; i += 256) {
      const int xx = i / (kTsY * TZH), rem = i - xx * (kTsY * TZH), yy = rem / TZH, j = rem - yy * TZH;
      const int x = x0 + xx, y = y0 + yy;
      const int zc = z0 + (((1 - first - x - y - z0) % 2 + 2) % 2) + 2 * j;
      if (x >= d0 || y >= d1 || zc >= d2) continue;
      const size_t c = (size_t)x * sx + (size_t)y * sy + zc;
      const Coef cf = coef_load(L, r, x, y, zc, c);
      const int k = (xx + 2) * PX + (yy + 2) * kTsHZ + (zc - z0 + 2);
      zout[c] = coef_relax<MODE>(cf, sz[k], sz[k + PX], sz[k - PX], sz[k + kTsHZ], sz[k - kTsHZ], sz[k + 1], sz[k - 1], w);
    }
    for (int i = tid; i < kTsX * kTsY * TZH; i += 256) {
      const int xx = i / (kTsY * TZH), rem = i - xx * (kTsY * TZH), yy = rem / TZH, j = rem - yy * TZH;
      const int x = x0 + xx, y = y0 + yy;
      const int zc = z0 + (((first - x - y - z0) % 2 + 2) % 2) + 2 * j;
      if (x >= d0 || y >= d1 || zc >= d2) continue;
      zout[(size_t)x * sx + (size_t)y * sy + zc] = sz[(xx + 2) * PX + (yy + 2) * kTsHZ + (zc - z0 + 2)];
    }
    __syncthreads();
  }
}

// flag[tile] = 1 when one of the 4 x 4 rows of the (x, y) tile is flagged in rowflag
__global__ void k_tile_flags(const int* __restrict__ rowflag, int* __restrict__ flag, int d0, int d1) {
  const int ntx = (d0 + kTsX - 1) / kTsX, nty = (d1 + kTsY - 1) / kTsY;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntx * nty) return;
  const int bx = t / nty, by = t - bx * nty;
  int f = 0;
  for (int a = 0; a < kTsX; ++a)
    for (int b = 0; b < kTsY; ++b) {
      const int x = bx * kTsX + a, y = by * kTsY + b;
      if (x < d0 && y < d1) f |= rowflag[x * d1 + y];
    }
  flag[t] = f ? 1 : 0;
}

// One red-black sweep (Smooth(l, first) then Smooth(l, 1 - first)) on the listed rows of a
// coarse level in one launch.  A block of five waves owns a 64-cell segment of a row, lane = z.
// A cell of the second colour needs the first-colour results of its six neighbours: the two in
// its row come from the neighbouring lanes of wave 0 (or, at a segment end, are recomputed by
// wave 0 from zin), the four in the rows x +- 1, y +- 1 are recomputed from zin by waves 1-4, one
// row each, and handed to wave 0 through LDS (a first-colour update depends on zin only).  zin is
// read, zout written (every cell of the segment), so no block sees another's writes; the values
// are those of two colour passes, operation for operation.  Five short waves instead of one that
// evaluates six stencils: every launch starts with cold L2s, so the cost is the latency of one
// wave's chain of loads and dependent FP64/FP32 arithmetic, and one wave per SIMD hid none of it.
// The rows are the level's z-support (liquid rows and the rows a coarser correction was prolonged
// into, whose cells are copied): every value read later from z lies there; other rows are only
// ever multiplied by a zero diagonal.
constexpr int kSweepWaves = 5;
template <int MODE, class T = float>
__global__ __launch_bounds__(64 * kSweepWaves) void k_sweep_rows(T* __restrict__ zout, const T* __restrict__ zin,
                                                                  const float* __restrict__ r, LvP L,
                                                                  const int* __restrict__ rows,
                                                                  const int* __restrict__ nrows, int first, float w) {
  __shared__ T s_nb[4][64];  // first-colour updates of the rows x + 1, x - 1, y + 1, y - 1
  const int d0 = L.d[0], d1 = L.d[1], d2 = L.d[2];
  const int nseg = (d2 + 63) >> 6;
  const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
  const int nr = *nrows;
  const size_t sx = (size_t)d1 * d2, sy = (size_t)d2;
  // workgroups go to the 8 XCDs in turn: XCD q takes the q-th eighth of the list's segments, so
  // the rows x +- 1 of a row (a plane of the list away) are mostly read through the same L2
  const int nwork = nr * nseg, chunk = (nwork + 7) >> 3;
  const int q = (int)blockIdx.x & 7, per_xcd = ((int)gridDim.x - q + 7) >> 3;
  for (int j = (int)blockIdx.x >> 3; j < chunk; j += per_xcd) {
    const int k = q * chunk + j;  // (list index, segment) = (k / nseg, k % nseg)
    if (k >= nwork) break;
    const int row = rows[k / nseg];
    const int seg = k - (k / nseg) * nseg;
    const int x = row / d1, y = row - x * d1;
    const int zr0 = seg * 64 + lane;
    const int wz = min(zr0, d2 - 1);
    const size_t c = (size_t)row * d2 + wz;
    const bool first_colour = ((x + y + wz) & 1) == first;
    Sten<T> own;
    T after1 = 0, zl = 0, zr = 0;
    if (wv > 0) {
      // wave a + 1: the first-colour update of the neighbour in row x + 1, x - 1, y + 1 or y - 1
      // (second-colour lanes whose neighbour exists; the others load their own cell, unused)
      const int a = wv - 1;
      const int nx = a == 0 ? x + 1 : (a == 1 ? x - 1 : x), ny = a == 2 ? y + 1 : (a == 3 ? y - 1 : y);
      const bool has = !first_colour && nx >= 0 && nx < d0 && ny >= 0 && ny < d1;
      const size_t cn = has ? (a == 0 ? c + sx : (a == 1 ? c - sx : (a == 2 ? c + sy : c - sy))) : c;
      Sten<T> sn = sten_load<T>(L, zin, r, has ? nx : x, has ? ny : y, wz, cn);
      sten_pin(sn);
      s_nb[a][lane] = sten_relax<MODE>(sn, sten_ax(sn, sn.zc, sn.zxp, sn.zxm, sn.zyp, sn.zym, sn.zzp, sn.zzm), sn.zc, w);
    } else {
      own = sten_load<T>(L, zin, r, x, y, wz, c);
      sten_pin(own);
      // the z neighbour in the next / previous segment (second-colour lanes at a segment end; with
      // one segment per row no lane reads it)
      T vb = own.zc;
      if (nseg > 1) {
        const bool nzp = !first_colour && lane == 63 && wz + 1 < d2;
        const bool nzm = !first_colour && lane == 0 && wz >= 1;
        Sten<T> sb = sten_load<T>(L, zin, r, x, y, nzp ? wz + 1 : (nzm ? wz - 1 : wz), nzp ? c + 1 : (nzm ? c - 1 : c));
        sten_pin(sb);
        vb = sten_relax<MODE>(sb, sten_ax(sb, sb.zc, sb.zxp, sb.zxm, sb.zyp, sb.zym, sb.zzp, sb.zzm), sb.zc, w);
      }
      // first colour: own update
      const T v_own = sten_relax<MODE>(own, sten_ax(own, own.zc, own.zxp, own.zxm, own.zyp, own.zym, own.zzp,
                                                    own.zzm), own.zc, w);
      // the segment after the first colour: first-colour lanes hold their update, the others zin
      after1 = first_colour ? v_own : own.zc;
      // the shuffles run with every lane of the wave active (a lane outside a shuffle supplies
      // no value to the others); segment-end lanes then take their recomputed neighbour
      const T sl = __shfl(after1, lane - 1 < 0 ? 0 : lane - 1);
      const T sr = __shfl(after1, lane + 1 > 63 ? 63 : lane + 1);
      zl = lane == 0 ? vb : sl;
      zr = lane == 63 ? vb : sr;
    }
    __syncthreads();
    if (wv == 0) {
      T out = after1;
      if (!first_colour)
        out = sten_relax<MODE>(own, sten_ax(own, own.zc, s_nb[0][lane], s_nb[1][lane], s_nb[2][lane], s_nb[3][lane],
                                            zr, zl), own.zc, w);
      if (zr0 < d2) zout[c] = out;
    }
    __syncthreads();  // s_nb is rewritten by the next segment
  }
}

// One Chebyshev step of the coarsest level's solve at a cell (the oracle's coarsest_chebyshev):
// d = a d + b (r - A z) / ad (the solve's first step: d = b (r - A z) / ad), z + d; a cell with a
// negligible diagonal keeps z (its d, read by no one, is 0)
template <class T>
__device__ __forceinline__ void cheb_cell(const Sten<T>& s, T ax, T zc, T dprev, double a, double b, bool first,
                                          T& zn, T& dn) {
  if (is_small(s.ad)) {
    zn = zc;
    dn = (T)0;
    return;
  }
  const T t = ((T)s.rv - ax) / (T)s.ad;
  dn = first ? (T)b * t : (T)a * dprev + (T)b * t;
  zn = zc + dn;
}

struct ChebSteps {
  double a0, b0, a1, b1;  // coefficients of the launch's first and second step
  int first;              // the first step is the solve's step 0 (no previous direction)
  int two;                // two steps in this launch
};

// Chebyshev steps on the coarsest level's rows (its z-support), one or two per launch, with the
// layout of k_sweep_rows: a block of five waves per 64-cell row segment.  A step updates every
// cell from the previous iterate, so the second step of a cell needs the first step's values at
// its six neighbours: the two in its row come from wave 0's neighbouring lanes (at a segment end
// recomputed by wave 0), the four in the rows x +- 1, y +- 1 are recomputed by waves 1-4 from
// zin / din and handed over through LDS.  zin / din are read and zout / dout written, so no block
// sees another's writes, and every value is that of two separate steps, operation for operation.
// With zf (the solve's last launch) it also writes z rounded to float for the prolongation, over
// the same cells the conversion k_z64_to_f32 covers (the z-support rows).
template <class T>
__global__ __launch_bounds__(64 * kSweepWaves) void k_cheb_rows(T* __restrict__ zout, T* __restrict__ dout,
                                                                 const T* __restrict__ zin, const T* __restrict__ din,
                                                                 const float* __restrict__ r, LvP L,
                                                                 const int* __restrict__ rows,
                                                                 const int* __restrict__ nrows, ChebSteps q,
                                                                 float* __restrict__ zf) {
  __shared__ T s_nb[4][64];  // first-step z of the rows x + 1, x - 1, y + 1, y - 1
  const int d0 = L.d[0], d1 = L.d[1], d2 = L.d[2];
  const int nseg = (d2 + 63) >> 6;
  const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
  const int nr = *nrows;
  const size_t sx = (size_t)d1 * d2, sy = (size_t)d2;
  const bool first = q.first != 0, two = q.two != 0;
  // XCD q takes the q-th eighth of the list's segments (as k_sweep_rows)
  const int nwork = nr * nseg, chunk = (nwork + 7) >> 3;
  const int xq = (int)blockIdx.x & 7, per_xcd = ((int)gridDim.x - xq + 7) >> 3;
  for (int j = (int)blockIdx.x >> 3; j < chunk; j += per_xcd) {
    const int k = xq * chunk + j;  // (list index, segment) = (k / nseg, k % nseg)
    if (k >= nwork) break;
    const int row = rows[k / nseg];
    const int seg = k - (k / nseg) * nseg;
    const int x = row / d1, y = row - x * d1;
    const int zr0 = seg * 64 + lane;
    const int wz = min(zr0, d2 - 1);
    const size_t c = (size_t)row * d2 + wz;
    Sten<T> own;
    T zA = 0, dA = 0, zl = 0, zr = 0;
    if (wv > 0) {
      if (two) {
        // wave a + 1: the first step at the neighbour in row x + 1, x - 1, y + 1 or y - 1 (lanes
        // whose neighbour lies outside the grid load their own cell, unused)
        const int a = wv - 1;
        const int nx = a == 0 ? x + 1 : (a == 1 ? x - 1 : x), ny = a == 2 ? y + 1 : (a == 3 ? y - 1 : y);
        const bool has = nx >= 0 && nx < d0 && ny >= 0 && ny < d1;
        const size_t cn = has ? (a == 0 ? c + sx : (a == 1 ? c - sx : (a == 2 ? c + sy : c - sy))) : c;
        Sten<T> sn = sten_load<T>(L, zin, r, has ? nx : x, has ? ny : y, wz, cn);
        const T dp = first ? (T)0 : din[cn];
        sten_pin(sn);
        T zn, dn;
        cheb_cell<T>(sn, sten_ax(sn, sn.zc, sn.zxp, sn.zxm, sn.zyp, sn.zym, sn.zzp, sn.zzm), sn.zc, dp, q.a0, q.b0,
                     first, zn, dn);
        s_nb[a][lane] = zn;
      }
    } else {
      own = sten_load<T>(L, zin, r, x, y, wz, c);
      const T dp = first ? (T)0 : din[c];
      sten_pin(own);
      cheb_cell<T>(own, sten_ax(own, own.zc, own.zxp, own.zxm, own.zyp, own.zym, own.zzp, own.zzm), own.zc, dp,
                   q.a0, q.b0, first, zA, dA);
      if (two) {
        // the z neighbour in the next / previous segment (lane 63 / lane 0; with one segment per
        // row no lane reads it)
        T vb = zA;
        if (nseg > 1) {
          const bool nzp = lane == 63 && wz + 1 < d2, nzm = lane == 0 && wz >= 1;
          const int wb = nzp ? wz + 1 : (nzm ? wz - 1 : wz);
          const size_t cb = nzp ? c + 1 : (nzm ? c - 1 : c);
          Sten<T> sb = sten_load<T>(L, zin, r, x, y, wb, cb);
          const T db = first ? (T)0 : din[cb];
          sten_pin(sb);
          T dn;
          cheb_cell<T>(sb, sten_ax(sb, sb.zc, sb.zxp, sb.zxm, sb.zyp, sb.zym, sb.zzp, sb.zzm), sb.zc, db, q.a0, q.b0,
                       first, vb, dn);
        }
        // the shuffles run with every lane of the wave active
        const T sl = __shfl(zA, lane - 1 < 0 ? 0 : lane - 1);
        const T sr = __shfl(zA, lane + 1 > 63 ? 63 : lane + 1);
        zl = lane == 0 ? vb : sl;
        zr = lane == 63 ? vb : sr;
      }
    }
    __syncthreads();
    if (wv == 0) {
      T zo = zA, dO = dA;
      if (two)
        cheb_cell<T>(own, sten_ax(own, zA, s_nb[0][lane], s_nb[1][lane], s_nb[2][lane], s_nb[3][lane], zr, zl), zA, dA,
                     q.a1, q.b1, false, zo, dO);
      if (zr0 < d2) {
        zout[c] = zo;
        dout[c] = dO;
        if (zf) zf[c] = (float)zo;  // the solve's last launch: z for the prolongation
      }
    }
    __syncthreads();  // s_nb is rewritten by the next segment
  }
}

template <int MODE>
__global__ void __launch_bounds__(256) k_rbgs_sweep(float* __restrict__ zout, const float* __restrict__ zin,
                                                     const float* __restrict__ r, LvP L, int first, float w) {
  __shared__ float sz[kSwHX * kSwHY * kSwHZ];
  const int d0 = L.d[0], d1 = L.d[1], d2 = L.d[2];
  const int nby = (d1 + kSwY - 1) / kSwY, nbz = (d2 + kSwZ - 1) / kSwZ;
  const int bz = blockIdx.x % nbz, by = (blockIdx.x / nbz) % nby, bx = blockIdx.x / (nbz * nby);
  const int x0 = bx * kSwX, y0 = by * kSwY, z0 = bz * kSwZ;
  const size_t sx = (size_t)d1 * d2, sy = (size_t)d2;
  const int tid = threadIdx.x;
  constexpr int PX = kSwHY * kSwHZ;  // LDS x-stride
  for (int i = tid; i < kSwHX * kSwHY * kSwHZ; i += 256) {
    int xx = i / PX, rem = i - xx * PX, yy = rem / kSwHZ, zz = rem - yy * kSwHZ;
    int x = x0 - 2 + xx, y = y0 - 2 + yy, zc = z0 - 2 + zz;
    float v = 0.f;
    if (x >= 0 && x < d0 && y >= 0 && y < d1 && zc >= 0 && zc < d2) v = zin[(size_t)x * sx + (size_t)y * sy + zc];
    sz[i] = v;
  }
  __syncthreads();
  // first colour on the tile plus a 1-cell halo (reads only second-colour neighbours)
  constexpr int RX = kSwX + 2, RY = kSwY + 2, RZ = kSwZ + 2;
  for (int i = tid; i < RX * RY * RZ; i += 256) {
    int xx = i / (RY * RZ), rem = i - xx * (RY * RZ), yy = rem / RZ, zz = rem - yy * RZ;
    int x = x0 - 1 + xx, y = y0 - 1 + yy, zc = z0 - 1 + zz;
    if (x < 0 || x >= d0 || y < 0 || y >= d1 || zc < 0 || zc >= d2) continue;
    if (((x + y + zc) & 1) != first) continue;
    size_t c = (size_t)x * sx + (size_t)y * sy + zc;
    float ad = L.adiag[c];
    if (is_small(ad)) continue;
    int k = (xx + 1) * PX + (yy + 1) * kSwHZ + (zz + 1);
    float ax = get_ax7(L, c, x, y, zc, sz[k], sz[k + PX], sz[k - PX], sz[k + kSwHZ], sz[k - kSwHZ], sz[k + 1],
                       sz[k - 1]);
    sz[k] = relax<MODE>(sz[k], r[c], ax, ad, w);
  }
  __syncthreads();
  // second colour on the tile; write the whole tile
  for (int i = tid; i < kSwX * kSwY * kSwZ; i += 256) {
    int xx = i / (kSwY * kSwZ), rem = i - xx * (kSwY * kSwZ), yy = rem / kSwZ, zz = rem - yy * kSwZ;
    int x = x0 + xx, y = y0 + yy, zc = z0 + zz;
    if (x >= d0 || y >= d1 || zc >= d2) continue;
    size_t c = (size_t)x * sx + (size_t)y * sy + zc;
    int k = (xx + 2) * PX + (yy + 2) * kSwHZ + (zz + 2);
    float zv = sz[k];
    if (((x + y + zc) & 1) != first) {
      float ad = L.adiag[c];
      if (!is_small(ad)) {
        float ax = get_ax7(L, c, x, y, zc, zv, sz[k + PX], sz[k - PX], sz[k + kSwHZ], sz[k - kSwHZ], sz[k + 1],
                           sz[k - 1]);
        zv = relax<MODE>(zv, r[c], ax, ad, w);
      }
    }
    zout[c] = zv;
  }
}

// Restrict, PressureSolver.cpp:796-819, with one thread per fine cell: a block takes 32 coarse
// cells of one restriction row, i.e. 4 fine rows x 64 fine z (one wave per fine row segment,
// coalesced), each thread the residual of its fine cell with all stencil loads issued together,
// and the coarse cell then sums its 8 children from LDS in the reference's order.  Eight times
// the threads of a one-coarse-cell-per-thread form keep the loads in flight (finest level
// 32.8 -> 21.6 us, level 1 12.4 -> 5.2 us per launch at 256^3).
// The coarse level's first pre-smoothing colour pass (colour 0: cells with x + y + z even) from z = 0,
// folded into the kernel that zeroes z: with every z zero, (A z)[c] is exactly +0 (0 * adiag and the
// coupled terms 0 * -term are +0 or -0, and +0 + -0 = +0), so the pass leaves z = relax(0, r / adiag)
// on the colour's cells of nonzero diagonal and 0 elsewhere, the operations of k_smooth with zv = 0.
__device__ __forceinline__ float first_pass_from_zero(float rv, float ad, float w, int mode) {
  if (is_small(ad)) return 0.0f;
  const float zv = 0.0f;
  const float term = (rv - 0.0f) / ad;
  if (mode == 0) {
    const float wt = w * term;
    return (float)((1.0 - (double)w) * (double)zv + (double)wt);
  }
  return zv + w * term;
}

// adc / first: the coarse level's adiag and, when first != 0, the first pre-smoothing pass (mode,
// w) is folded in (first_pass_from_zero)
__global__ __launch_bounds__(kBlock) void k_restrict(float* __restrict__ rc, float* __restrict__ zc,
                                                          const float* __restrict__ rf, const float* __restrict__ zf,
                                                          LvP Lf, int c1, int c2, const int* __restrict__ rows,
                                                          const int* __restrict__ nrows, int nrow_cap,
                                                          double* __restrict__ zc64, const float* __restrict__ adc,
                                                          int first, float w, int mode) {
  static_assert(kBlock == 256, "4 fine rows x 64 fine z per block");
  __shared__ float s_d[8][32];
  const int nch = (c2 + 31) >> 5;
  const int nr = *nrows;
  const int tid = (int)threadIdx.x, qx = tid >> 7, qy = (tid >> 6) & 1, wl = tid & 63;
  const int q = (qx << 2) | (qy << 1) | (wl & 1);
  for (int k = blockIdx.x; k < nr * nch; k += gridDim.x) {
    const int row = rows[min(k / nch, nrow_cap - 1)];
    const int ch = k - (k / nch) * nch;
    const int xi = row / c1, yi = row - xi * c1;
    const int zi = ch * 32 + (wl >> 1);
    const bool out = tid < 32 && ch * 32 + tid < c2;  // the threads that write the coarse cells
    const size_t i = (size_t)row * c2 + ch * 32 + tid;
    const float av = (first && out) ? adc[i] : 0.0f;  // issued before the fine-level work
    if (zi < c2) {
      const int x = 2 * xi + qx, y = 2 * yi + qy, w = 2 * zi + (wl & 1);
      const size_t f = ((size_t)x * Lf.d[1] + y) * Lf.d[2] + w;
      const AxAt g = get_ax_full(Lf, zf, rf, x, y, w, f);
      s_d[q][wl >> 1] = g.v - g.ax;
    }
    __syncthreads();
    if (out) {
      float acc = 0.0f;
#pragma unroll
      for (int j = 0; j < 8; ++j) acc = (float)((double)acc + 0.125 * (double)s_d[j][tid]);
      rc[i] = acc;
      zc[i] = (first && ((xi + yi + ch * 32 + tid) & 1) == 0) ? first_pass_from_zero(acc, av, w, mode) : 0.0f;
      if (zc64) zc64[i] = 0.0;  // the coarsest level's FP64 z
    }
    __syncthreads();
  }
}

// Prolong, PressureSolver.cpp:774-794, from the coarse level's z-support rows (its z is zero
// on every other row: nothing but restriction, which clears it, writes there)
__global__ void k_prolong(float* __restrict__ zf, const float* __restrict__ zc, int f1, int f2, int c1, int c2,
                          const int* __restrict__ rows, const int* __restrict__ nrows) {
  const RowIter it = row_iter(c2);
  const int nr = *nrows;
  if (!it.on) return;
  for (int k = blockIdx.x * it.R + it.sub, kst = gridDim.x * it.R, rn = k < nr ? rows[k] : 0; k < nr;
       k += kst) {  // the next row id loads with this row's operands
    const int row = rn;
    rn = k + kst < nr ? rows[k + kst] : 0;
    const int xi = row / c1, yi = row - xi * c1;
    for (int zi = it.z0; zi < c2; zi += it.zs) {
      const float v = zc[(size_t)row * c2 + zi];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        int x = 2 * xi + ((q >> 2) & 1), y = 2 * yi + ((q >> 1) & 1), w = 2 * zi + (q & 1);
        zf[((size_t)x * f1 + y) * f2 + w] += v;
      }
    }
  }
}

// z = (float) z64 on the rows of a level's row list (the FP64 coarsest level's result, for the
// prolongation)
__global__ void k_z64_to_f32(float* __restrict__ z, const double* __restrict__ z64, const int* __restrict__ rows,
                             const int* __restrict__ nrows, int d2) {
  const RowIter it = row_iter(d2);
  const int nr = *nrows;
  if (!it.on) return;
  for (int k = blockIdx.x * it.R + it.sub; k < nr; k += gridDim.x * it.R) {
    const size_t base = (size_t)rows[k] * d2;
    for (int zi = it.z0; zi < d2; zi += it.zs) z[base + zi] = (float)z64[base + zi];
  }
}

// z = 0 on the rows of a level's row list
__global__ void k_zero_rows(float* __restrict__ z, const int* __restrict__ rows, const int* __restrict__ nrows,
                            int d2) {
  const RowIter it = row_iter(d2);
  const int nr = *nrows;
  if (!it.on) return;
  for (int k = blockIdx.x * it.R + it.sub, kst = gridDim.x * it.R, rn = k < nr ? rows[k] : 0; k < nr;
       k += kst) {  // the next row id loads with this row's operands
    const size_t base = (size_t)rn * d2;
    rn = k + kst < nr ? rows[k + kst] : 0;
    for (int zi = it.z0; zi < d2; zi += it.zs) z[base + zi] = 0.0f;
  }
}

// PrecondType::None / Jacobi, PressureSolver.cpp:721-743
__global__ void k_precond_simple(float* __restrict__ z, const float* __restrict__ r, const float* __restrict__ adiag,
                                 const uint8_t* __restrict__ cls, size_t n, int jacobi) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    if (jacobi) {
      float dn = adiag[i];
      if (is_small(dn)) continue;
      z[i] = r[i] / dn;
    } else {
      if (cls[i] == 0) continue;
      z[i] = r[i];
    }
  }
}

// ---------------------------------------------------------------- PCG vectors
// the sum of n per-block FP64 partials, in a fixed order (the same in every block); every
// thread of the block calls it
__device__ double sum_partials_block(const double* __restrict__ part, int n) {
  __shared__ double s[kBlock];
  double a = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) a += part[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int k = kBlock / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
    __syncthreads();
  }
  double t = s[0];
  __syncthreads();
  return t;
}
template <int NV>
// The block's sum in a fixed pairwise order: thread i takes i + 128, then i + 64 (through LDS: the
// four waves), then wave 0 continues with i + 32, 16, 8, 4, 2, 1 by wavefront shuffles, the same
// pairs as a shared-memory tree.
__device__ __forceinline__ void block_partials(double (&v)[NV], double* __restrict__ partial) {
  static_assert(kBlock == 256, "two cross-wave stages, then one wave");
  __shared__ double s[NV][kBlock];
  const int t = (int)threadIdx.x;
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k][t] = v[k];
  __syncthreads();
  if (t < 128) {
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k][t] += s[k][t + 128];
  }
  __syncthreads();
  if (t < 64) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double a = s[k][t] + s[k][t + 64];
#pragma unroll
      for (int st = 32; st > 0; st >>= 1) a += __shfl_down(a, st);  // lanes < st keep the tree's pairs
      if (t == 0) partial[(size_t)k * gridDim.x + blockIdx.x] = a;
    }
  }
}

// r = b on liquid cells (else 0); x = p = z0 = 0 (Init + PressureSolver.cpp:845-853)
__global__ void k_cg_init(float* __restrict__ r, float* __restrict__ x, float* __restrict__ p, float* __restrict__ z,
                          const float* __restrict__ b, const uint8_t* __restrict__ cls, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    r[i] = (cls[i] != 0) ? b[i] : 0.0f;
    x[i] = 0.0f;
    p[i] = 0.0f;
    z[i] = 0.0f;
  }
}

// p = z on liquid cells; partials of r.z and r.r (:856-867; r is zero off the liquid cells)
__global__ void k_cg_first(float* __restrict__ p, const float* __restrict__ z, const float* __restrict__ r,
                           const uint8_t* __restrict__ cls, const int* __restrict__ rows,
                           const int* __restrict__ nrows, int d2, double* __restrict__ partial) {
  double v[2] = {0.0, 0.0};
  const RowIter it = row_iter(d2);
  const int nr = *nrows;
  if (it.on)
    for (int k = blockIdx.x * it.R + it.sub, kst = gridDim.x * it.R, rn = k < nr ? rows[k] : 0; k < nr;
         k += kst) {  // the next row id loads with this row's operands
      const size_t base = (size_t)rn * d2;
      rn = k + kst < nr ? rows[k + kst] : 0;
      for (int zi = it.z0; zi < d2; zi += it.zs) {
        const size_t i = base + zi;
        float zv = z[i], rv = r[i];
        if (cls[i] != 0) p[i] = zv;
        v[0] += (double)rv * (double)zv;
        v[1] += (double)rv * (double)rv;
      }
    }
  block_partials<2>(v, partial);
}

// q = A p and partial p.q over interior liquid cells (:873-882)
__global__ void k_cg_spmv(float* __restrict__ q, const float* __restrict__ p, LvP L, const int* __restrict__ rows,
                          const int* __restrict__ nrows, double* __restrict__ partial) {
  const int d0 = L.d[0], d1 = L.d[1], d2 = L.d[2];
  double v[1] = {0.0};
  const RowIter it = row_iter(d2);
  // the row id is fetched together with the row count (rows holds d0 * d1 entries)
  const int kcap = d0 * d1 - 1, kstep = gridDim.x * it.R;
  int k = blockIdx.x * it.R + it.sub;
  int row_next = rows[min(k, kcap)];
  const int nr = *nrows;
  if (it.on)
    for (; k < nr; k += kstep) {
      const int row = row_next;
      row_next = rows[min(k + kstep, kcap)];
      const int xi = row / d1, yi = row - xi * d1;
      if (xi < 1 || xi > d0 - 2 || yi < 1 || yi > d1 - 2) continue;
      const size_t base = (size_t)row * d2;
      for (int zi = it.z0; zi < d2; zi += it.zs) {
        if (zi < 1 || zi > d2 - 2) continue;
        const size_t i = base + zi;
        const AxAt g = get_ax_full(L, p, p, xi, yi, zi, i);
        if (g.cls == 0) continue;
        q[i] = g.ax;
        v[0] += (double)g.zc * (double)g.ax;
      }
    }
  block_partials<1>(v, partial);
}

// x += alpha p; r -= alpha q; partial r.r over interior liquid cells (:889-902)
// alpha = rho / (p.q) (CGSolver<3>::Solve, :879-883) from the SpMV's partials in every block, then
// x += alpha p, r -= alpha q (and r.r) on the updated cells; with zero_z also z = 0 on every cell
// of the rows, the symmetric preconditioner's first step (z is read nowhere else).  rho is
// sc[6 + slot]; block 0 records sc[5] = p.q, sc[2] = alpha and the NaN flag bit 1.
// zero_z: z = 0 on the liquid rows for the next preconditioner application (symmetric V-cycle);
// with first != 0 the level's first pre-smoothing colour pass is folded in (first_pass_from_zero,
// mode 1, the relaxation w; adiag the level-0 diagonal) and the V-cycle skips that pass
__global__ void k_cg_update_xr(float* __restrict__ x, float* __restrict__ r, const float* __restrict__ p,
                               const float* __restrict__ q, const uint8_t* __restrict__ cls, float* __restrict__ sc,
                               int slot, int* __restrict__ flag, const double* __restrict__ part_pq, int npq, int d0,
                               int d1, int d2, const int* __restrict__ rows, const int* __restrict__ nrows,
                               double* __restrict__ partial, float* __restrict__ z, int zero_z,
                               const float* __restrict__ adiag, int first, float w, int nseg) {
  const float denom = (float)sum_partials_block(part_pq, npq);
  const float alpha = sc[6 + slot] / denom;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[5] = denom;
    sc[2] = alpha;
    if (isnan(alpha) || isinf(alpha)) atomicOr(flag, 1);
  }
  double v[1] = {0.0};
  const RowIter it = list_iter(d2, nseg);
  // the next item id is fetched with the current item's operands; a cell's class and its four
  // operands are loaded together
  const int kcap = d0 * d1 * max(nseg, 1) - 1, kstep = gridDim.x * it.R;
  int k = blockIdx.x * it.R + it.sub;
  int id_next = rows[min(k, kcap)];
  const int nr = *nrows;
  if (it.on)
    for (; k < nr; k += kstep) {
      const WorkItem wi = work_item(id_next, nseg, d2);
      id_next = rows[min(k + kstep, kcap)];
      const int row = wi.row;
      const int xi = row / d1, yi = row - xi * d1;
      const bool inner = xi >= 1 && xi <= d0 - 2 && yi >= 1 && yi <= d1 - 2;
      if (!inner && !zero_z) continue;
      const size_t base = (size_t)row * d2;
      for (int zi = wi.zlo + it.z0; zi < wi.zhi; zi += it.zs) {
        const size_t i = base + zi;
        const bool colour0 = first && ((xi + yi + zi) & 1) == 0;
        float zv = 0.0f;
        if (!inner || zi < 1 || zi > d2 - 2) {  // r stays: the first pass reads it as it is
          if (colour0) zv = first_pass_from_zero(r[i], adiag[i], w, 1);
          if (zero_z) z[i] = zv;
          continue;
        }
        int kc = cls[i];
        float xv = x[i], pv = p[i], rv = r[i], qv = q[i];
        float av = first ? adiag[i] : 0.0f;  // with the other operands: one wait, not two
        asm volatile("" : "+v"(kc), "+v"(xv), "+v"(pv), "+v"(rv), "+v"(qv), "+v"(av));
        if (kc != 0) {
          x[i] = xv + alpha * pv;
          rv = rv - alpha * qv;
          r[i] = rv;
          v[0] += (double)rv * (double)rv;
        }
        if (colour0) zv = first_pass_from_zero(rv, av, w, 1);
        if (zero_z) z[i] = zv;
      }
    }
  block_partials<1>(v, partial);
}

// partial r.z (r is zero off the liquid cells).  Block 0 also forms the residual norm from the
// r.r partials of the x/r update, sc[4] = ||r|| (:903-906), and copies sc[0..7] and the NaN flag
// into pinned host memory: the host's per-iteration read.
__global__ void k_cg_dot_rz(const float* __restrict__ r, const float* __restrict__ z, const int* __restrict__ rows,
                            const int* __restrict__ nrows, int d2, double* __restrict__ partial,
                            const double* __restrict__ part_rr, int npart, float* __restrict__ sc,
                            const int* __restrict__ flag, float* __restrict__ host_sc, int* __restrict__ host_flag,
                            int nseg) {
  if (blockIdx.x == 0) {  // uniform over the block
    const double rr = sum_partials_block(part_rr, npart);
    if (threadIdx.x == 0) {
      sc[4] = sqrtf((float)rr);
#pragma unroll
      for (int k = 0; k < 8; ++k) host_sc[k] = sc[k];
      *host_flag = *flag;
    }
  }
  double v[1] = {0.0};
  const RowIter it = list_iter(d2, nseg);
  const int kstep = gridDim.x * it.R;
  int k = blockIdx.x * it.R + it.sub;
  const int nr = *nrows;
  if (it.on)
    for (int id_next = k < nr ? rows[k] : 0; k < nr; k += kstep) {
      const WorkItem wi = work_item(id_next, nseg, d2);
      const size_t base = (size_t)wi.row * d2;
      id_next = k + kstep < nr ? rows[k + kstep] : 0;
      for (int zi = wi.zlo + it.z0; zi < wi.zhi; zi += it.zs) v[0] += (double)r[base + zi] * (double)z[base + zi];
    }
  block_partials<1>(v, partial);
}

// p = z + beta p (:913-921) fused with the next iteration's q = A p and p.q (:879-881).
// beta = rho_new / rho from the r.z partials in every block; block 0 stores rho_new in the other
// slot (no block reads it here), sc[0] = sc[1] = rho_new (m_oldRZ = m_newRZ, :922), sc[3] = beta
// and the NaN flag bit 2.  The new p goes to pout (pin is read, so no thread sees another's
// write): on every liquid cell of the rows, z + beta pin; a cell's SpMV takes its neighbours' new
// values from z and pin the same way (a cell with cls == 0 keeps pin, which no later stage reads
// with a nonzero weight).  q and the p.q partials are those of k_cg_spmv on the new p, with the
// same thread mapping and order.
__global__ void k_cg_p_spmv(float* __restrict__ pout, float* __restrict__ q, const float* __restrict__ pin,
                            const float* __restrict__ z, LvP L, const int* __restrict__ rows,
                            const int* __restrict__ nrows, float* __restrict__ sc, int slot, int* __restrict__ flag,
                            const double* __restrict__ part_rz, int npart, double* __restrict__ partial, int nseg) {
  const float newRZ = (float)sum_partials_block(part_rz, npart);
  const float beta = newRZ / sc[6 + slot];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[6 + (slot ^ 1)] = newRZ;
    sc[1] = newRZ;
    sc[3] = beta;
    sc[0] = newRZ;
    if (isnan(beta) || isinf(beta)) atomicOr(flag, 2);
  }
  const int d0 = L.d[0], d1 = L.d[1], d2 = L.d[2];
  const size_t sx = (size_t)d1 * d2, sy = (size_t)d2;
  double v[1] = {0.0};
  const RowIter it = list_iter(d2, nseg);
  const int kcap = d0 * d1 * max(nseg, 1) - 1;
  const XcdRun run = xcd_run(*nrows, it.R, it.sub);  // the grid (npartial blocks) is a multiple of 8
  const int kstep = run.step;
  int k = run.k;
  int id_next = rows[min(k, kcap)];
  if (it.on)
    for (; k < run.end; k += kstep) {
      const WorkItem wi = work_item(id_next, nseg, d2);
      id_next = rows[min(k + kstep, kcap)];
      const int row = wi.row;
      const int xi = row / d1, yi = row - xi * d1;
      const bool irow = xi >= 1 && xi <= d0 - 2 && yi >= 1 && yi <= d1 - 2;
      const size_t base = (size_t)row * d2;
      for (int zi = wi.zlo + it.z0; zi < wi.zhi; zi += it.zs) {
        const size_t i = base + zi;
        if (!(irow && zi >= 1 && zi <= d2 - 2)) {  // a cell on the box faces: the p update only
          if (L.cls[i] != 0) pout[i] = z[i] + beta * pin[i];
          continue;
        }
        // interior cell: every neighbour exists
        int m = L.cmask[i];
        float zc = z[i], zxp = z[i + sx], zxm = z[i - sx], zyp = z[i + sy], zym = z[i - sy], zzp = z[i + 1],
              zzm = z[i - 1];
        float pc = pin[i], pxp = pin[i + sx], pxm = pin[i - sx], pyp = pin[i + sy], pym = pin[i - sy],
              pzp = pin[i + 1], pzm = pin[i - 1];
        float ad = L.adiag[i], txp = L.t0[i], txm = L.t0[i - sx], typ = L.t1[i], tym = L.t1[i - sy], tzp = L.t2[i],
              tzm = L.t2[i - 1];
        // the loads complete here, together (see get_ax_full)
        asm volatile("" : "+v"(m), "+v"(zc), "+v"(zxp), "+v"(zxm), "+v"(zyp), "+v"(zym), "+v"(zzp), "+v"(zzm));
        asm volatile("" : "+v"(pc), "+v"(pxp), "+v"(pxm), "+v"(pyp), "+v"(pym), "+v"(pzp), "+v"(pzm), "+v"(ad),
                     "+v"(txp), "+v"(txm), "+v"(typ), "+v"(tym), "+v"(tzp), "+v"(tzm));
        if (m & 64) continue;  // cls == 0
        // a neighbour's new p where its cls != 0 (mask bits 8-13; every neighbour of an interior cell
        // exists), its off-diagonal term where its cls == 2 (bits 0-5)
        const float nc = zc + beta * pc;
        const float nxp = (m & 256) ? zxp + beta * pxp : pxp, nxm = (m & 512) ? zxm + beta * pxm : pxm;
        const float nyp = (m & 1024) ? zyp + beta * pyp : pyp, nym = (m & 2048) ? zym + beta * pym : pym;
        const float nzp = (m & 4096) ? zzp + beta * pzp : pzp, nzm = (m & 8192) ? zzm + beta * pzm : pzm;
        float res = nc * ad;
        res = (m & 1) ? res + nxp * (-txp) : res;
        res = (m & 2) ? res + nxm * (-txm) : res;
        res = (m & 4) ? res + nyp * (-typ) : res;
        res = (m & 8) ? res + nym * (-tym) : res;
        res = (m & 16) ? res + nzp * (-tzp) : res;
        res = (m & 32) ? res + nzm * (-tzm) : res;
        pout[i] = nc;
        q[i] = res;
        v[0] += (double)nc * (double)res;
      }
    }
  block_partials<1>(v, partial);
}


// sc = [oldRZ, newRZ, alpha, beta, residual, denom, rho slot 0, rho slot 1]; k_sc_first seeds it
// sc64 = [rz64, rr64, denom64, res64]
__global__ void k_sc_first(const double* __restrict__ part, int nb, float* __restrict__ sc, double* __restrict__ sc64) {
  double rz = sum_partials_block(part, nb);
  double rr = sum_partials_block(part + nb, nb);
  if (threadIdx.x == 0) {
    sc[0] = (float)rz;
    sc[6] = (float)rz;  // rho of the first iteration (slot 0)
    sc[4] = sqrtf((float)rr);
    sc64[0] = rz;
    sc64[1] = rr;
  }
}

// ComputeDivergence (LiquidModel.cpp:1065-1083) over the level-0 liquid rows: the rows holding a
// cell with cls != 0 are exactly those holding a cell with liquidPhi <= 0 (the cells the reference
// counts; a NaN phi has cls 1 and is counted as the reference's "> 0.0" test lets it through).
// Per block: the FP64 sum, the max |div| and the count, at partial[b], [nb + b], [2 nb + b].
__global__ void k_div_rows(const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz,
                           const float* __restrict__ lwx, const float* __restrict__ lwy,
                           const float* __restrict__ lwz, const float* __restrict__ phi,
                           const int* __restrict__ rows, const int* __restrict__ nrows, GridP g,
                           double* __restrict__ partial) {
  const int N1 = g.N[1], N2 = g.N[2];
  const size_t plane = (size_t)N1 * N2;
  const RowIter it = row_iter(N2);
  const int kstep = gridDim.x * it.R;
  const int nr = *nrows;
  double acc = 0.0;
  float mx = 0.f;
  int cnt = 0;
  if (it.on)
    for (int k = blockIdx.x * it.R + it.sub; k < nr; k += kstep) {
      const int row = rows[k];
      const int xi = row / N1, yi = row - xi * N1;
      const size_t base = (size_t)row * N2;
      const size_t fy0 = ((size_t)xi * (N1 + 1) + yi) * N2, fz0 = base + (size_t)row;  // (x*N1+y)*(N2+1)
      for (int zi = it.z0; zi < N2; zi += it.zs) {
        if (phi[base + zi] > 0.0f) continue;
        const size_t fx0 = base + zi, fx1 = fx0 + plane;
        const size_t y0 = fy0 + zi, y1 = y0 + N2;
        const size_t z0 = fz0 + zi, z1 = z0 + 1;
        float d = lwx[fx1] * vx[fx1] - lwx[fx0] * vx[fx0];
        d += lwy[y1] * vy[y1] - lwy[y0] * vy[y0];
        d += lwz[z1] * vz[z1] - lwz[z0] * vz[z0];
        d /= g.dx[0];
        acc += (double)d;
        const float ad = fabsf(d);
        mx = ad > mx ? ad : mx;
        ++cnt;
      }
    }
  __shared__ double s_acc[kBlock];
  __shared__ float s_mx[kBlock];
  __shared__ int s_cnt[kBlock];
  s_acc[threadIdx.x] = acc;
  s_mx[threadIdx.x] = mx;
  s_cnt[threadIdx.x] = cnt;
  __syncthreads();
  for (int st = kBlock / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      s_acc[threadIdx.x] += s_acc[threadIdx.x + st];
      s_mx[threadIdx.x] = fmaxf(s_mx[threadIdx.x], s_mx[threadIdx.x + st]);
      s_cnt[threadIdx.x] += s_cnt[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = s_acc[0];
    partial[gridDim.x + blockIdx.x] = (double)s_mx[0];
    partial[2 * gridDim.x + blockIdx.x] = (double)s_cnt[0];
  }
}

// sums of the divergence partials in a fixed order: out[0] = sum (as_mean: sum / count, 0 without
// cells), out[1] = max, out[2] = count (out may be pinned host memory read after the step's wait)
__global__ void k_div_final(const double* __restrict__ partial, int nb, double* __restrict__ out, int as_mean) {
  const double sum = sum_partials_block(partial, nb);
  const double cnt = sum_partials_block(partial + 2 * nb, nb);
  __shared__ double s_mx[kBlock];
  double m = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) m = fmax(m, partial[nb + i]);
  s_mx[threadIdx.x] = m;
  __syncthreads();
  for (int st = kBlock / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) s_mx[threadIdx.x] = fmax(s_mx[threadIdx.x], s_mx[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = as_mean ? (cnt > 0.0 ? sum / cnt : 0.0) : sum;
    out[1] = s_mx[0];
    out[2] = cnt;
  }
}

}  // namespace

void launch_divergence_rows(Ctx& c, double* out) {
  const Level& L = c.lv[0];
  const int nb = c.npartial;
  hipLaunchKernelGGL(k_div_rows, dim3(nb), dim3(kBlock), 0, c.s, c.v[0], c.v[1], c.v[2], c.lw[0], c.lw[1], c.lw[2],
                     c.phi, L.rows, L.nrows, c.g, c.partial);
  hipLaunchKernelGGL(k_div_final, dim3(1), dim3(kBlock), 0, c.s, c.partial, nb, out, 0);
}

void launch_div_final_on(hipStream_t s, const double* partial, int nb, double* out, int as_mean) {
  hipLaunchKernelGGL(k_div_final, dim3(1), dim3(kBlock), 0, s, partial, nb, out, as_mean);
}

void launch_build_rhs(Ctx& c) {
  hipLaunchKernelGGL(k_build_rhs, dim3(grid_for(c.ncell)), dim3(kBlock), 0, c.s, c.rhs, c.phi, c.v[0], c.v[1], c.v[2],
                     c.lw[0], c.lw[1], c.lw[2], c.g);
}

void launch_build_lhs(Ctx& c, float rho, float dt) {
  FineP F;
  for (int k = 0; k < 3; ++k) F.N[k] = c.g.N[k];
  F.phi = c.phi;
  F.lw0 = c.lw[0];
  F.lw1 = c.lw[1];
  F.lw2 = c.lw[2];
  for (size_t l = 0; l < c.lv.size(); ++l) {
    Level& L = c.lv[l];
    float f = (float)(1 << l);
    hipLaunchKernelGGL(k_build_lhs, dim3(grid_for(L.n)), dim3(kBlock), 0, c.s, L.adiag, L.t[0], L.t[1], L.t[2], L.cls,
                       F, (int)l, L.d[0], L.d[1], L.d[2], dt, rho, f * c.g.dx[0], f * c.g.dx[1],
                       f * c.g.dx[2]);
    hipLaunchKernelGGL(k_build_cmask, dim3(grid_for(L.n)), dim3(kBlock), 0, c.s, L.cmask, L.cls, L.d[0], L.d[1],
                       L.d[2]);
    const int nrow = L.d[0] * L.d[1];
    hipLaunchKernelGGL(k_row_flags, dim3((unsigned)div_up(nrow, kBlock / 64)), dim3(kBlock), 0, c.s, L.cls, L.aflag,
                       nrow, L.d[2], L.smask);
    hipLaunchKernelGGL(k_row_compact, dim3(1), dim3(1024), 0, c.s, L.aflag, nrow, L.rows, L.nrows);
    if (L.segs)
      hipLaunchKernelGGL(k_seg_compact, dim3(1), dim3(1024), 0, c.s, L.rows, L.nrows, L.smask, L.d[2] / kSegZ, L.segs,
                         L.nsegs);
  }
  // levels >= 1: rows where restriction makes r nonzero, finest first (each needs the finer
  // level's set) ...
  const int L = (int)c.lv.size();
  for (int l = 1; l < L; ++l) {
    Level& Lc = c.lv[l];
    const Level& Lf = c.lv[l - 1];
    const int nrow = Lc.d[0] * Lc.d[1];
    hipLaunchKernelGGL(k_qrow_flags, dim3((unsigned)div_up(nrow, kBlock)), dim3(kBlock), 0, c.s, Lc.aflag, Lf.aflag,
                       l >= 2 ? Lf.qflag : nullptr, Lc.qflag, Lc.d[0], Lc.d[1], Lf.d[1]);
  }
  // ... the z-supports, coarsest first ...
  for (int l = L - 1; l >= 1; --l) {
    Level& Lv = c.lv[l];
    const int nrow = Lv.d[0] * Lv.d[1];
    const bool coarser = l + 1 < L;
    hipLaunchKernelGGL(k_zrow_flags, dim3((unsigned)div_up(nrow, kBlock)), dim3(kBlock), 0, c.s, Lv.aflag,
                       coarser ? c.lv[l + 1].zflag : nullptr, Lv.zflag, Lv.d[0], Lv.d[1],
                       coarser ? c.lv[l + 1].d[0] : 0, coarser ? c.lv[l + 1].d[1] : 0);
  }
  // ... then the restriction rows = r rows or z-support, and the prolongation rows = z-support.
  // The q flags of the next level were built from the r rows alone, as they must be.
  for (int l = 1; l < L; ++l) {
    Level& Lv = c.lv[l];
    const int nrow = Lv.d[0] * Lv.d[1];
    hipLaunchKernelGGL(k_row_compact, dim3(1), dim3(1024), 0, c.s, Lv.zflag, nrow, Lv.zrows, Lv.nzrows);
  }
  for (int l = 1; l < L; ++l) {
    Level& Lv = c.lv[l];
    const int nrow = Lv.d[0] * Lv.d[1];
    hipLaunchKernelGGL(k_or_flags, dim3((unsigned)div_up(nrow, kBlock)), dim3(kBlock), 0, c.s, Lv.qflag, Lv.zflag,
                       nrow);
    hipLaunchKernelGGL(k_row_compact, dim3(1), dim3(1024), 0, c.s, Lv.qflag, nrow, Lv.qrows, Lv.nqrows);
    // tiles of the restriction rows (every row where z or r can be nonzero), for the tiled sweep
    if (c.cfg.coarse_sweep == kCoarseTiles) {
      const int ntxy = (int)(div_up(Lv.d[0], kTsX) * div_up(Lv.d[1], kTsY));
      hipLaunchKernelGGL(k_tile_flags, dim3((unsigned)div_up(ntxy, kBlock)), dim3(kBlock), 0, c.s, Lv.qflag, Lv.tflag,
                         Lv.d[0], Lv.d[1]);
      hipLaunchKernelGGL(k_row_compact, dim3(1), dim3(1024), 0, c.s, Lv.tflag, ntxy, Lv.tiles, Lv.ntiles);
    }
  }
}

namespace {

// blocks for a colour pass over the active rows (enough for all rows, capped; the kernel strides)
unsigned smooth_blocks(const Level& L) {
  const int H = (L.d[2] + 1) / 2;
  const int R = H >= kBlock ? 1 : kBlock / H;
  const size_t need = div_up((size_t)L.d[0] * L.d[1], (size_t)R);
  return (unsigned)(div_up(std::max<size_t>(1, std::min<size_t>(need, 8192)), 8) * 8);  // xcd_run: 8 | grid
}

// blocks for k_smooth_seg: eight segments per block over the list's capacity, capped (the kernel
// strides); a multiple of 8 for xcd_run
unsigned seg_blocks(const Level& L) {
  const size_t need = div_up((size_t)L.d[0] * L.d[1] * (L.d[2] / kSegZ), (size_t)8);
  return (unsigned)(div_up(std::max<size_t>(1, std::min<size_t>(need, 8192)), 8) * 8);
}

// blocks for a kernel over the cells of a level's row list (RowIter mapping; the kernel strides)
unsigned row_blocks(const Level& L) {
  const int R = L.d[2] >= kBlock ? 1 : kBlock / L.d[2];
  const size_t need = div_up((size_t)L.d[0] * L.d[1], (size_t)R);
  return (unsigned)std::max<size_t>(1, std::min<size_t>(need, 8192));
}

// blocks for k_sweep_rows: one per row segment of the level (the list's capacity; blocks past the
// list's length exit at once), striding beyond 65536
unsigned row_sweep_blocks(const Level& L) {
  const size_t nseg = div_up((size_t)L.d[2], (size_t)64);
  return (unsigned)std::max<size_t>(1, std::min<size_t>((size_t)L.d[0] * L.d[1] * nseg, 65536));
}

// one red-black sweep on level l: Smooth(l, first) then Smooth(l, 1 - first).
// whether level l is smoothed by two colour passes per sweep (the last branch of sweep() but one)
bool colour_passes(const Ctx& c, int l) {
  const Level& L = c.lv[l];
  if (l > 0 && c.cfg.coarse_sweep == kCoarseRows && (long long)L.n <= c.cfg.row_sweep_max_cells) return false;
  if (l > 0 && c.cfg.coarse_sweep == kCoarseTiles) return false;
  return (long long)L.n < c.cfg.fused_sweep_min_cells;
}

// Large levels use the fused tile kernel; small ones (launch-bound) two colour passes.  skip_first:
// the first colour pass has been folded into the kernel that zeroed z (colour passes only).
void sweep(Ctx& c, int l, int first, int mode, bool skip_first = false) {
  Level& L = c.lv[l];
  if (l > 0 && c.cfg.coarse_sweep == kCoarseRows && (long long)L.n <= c.cfg.row_sweep_max_cells) {
    const unsigned nb = row_sweep_blocks(L);
    // over the z-support rows (liquid rows plus the rows the coarser correction reaches): the
    // output buffer must hold every value a later prolongation reads
    if (mode == 0)
      hipLaunchKernelGGL(k_sweep_rows<0>, dim3(nb), dim3(64 * kSweepWaves), 0, c.s, L.z2, L.z, L.r, lvp(L), L.zrows,
                         L.nzrows, first, c.cfg.omega);
    else
      hipLaunchKernelGGL(k_sweep_rows<1>, dim3(nb), dim3(64 * kSweepWaves), 0, c.s, L.z2, L.z, L.r, lvp(L), L.zrows,
                         L.nzrows, first, c.cfg.omega);
    std::swap(L.z, L.z2);
    return;
  }
  if (l > 0 && c.cfg.coarse_sweep == kCoarseTiles) {
    const int ntxy = (int)(div_up(L.d[0], kTsX) * div_up(L.d[1], kTsY));
    const unsigned nb = (unsigned)std::max<size_t>(
        1, std::min<size_t>((size_t)ntxy * div_up((size_t)L.d[2], (size_t)kTsZ), 8192));
    if (mode == 0)
      hipLaunchKernelGGL(k_tile_sweep<0>, dim3(nb), dim3(256), 0, c.s, L.z2, L.z, L.r, lvp(L), L.tiles, L.ntiles, ntxy,
                         first, c.cfg.omega);
    else
      hipLaunchKernelGGL(k_tile_sweep<1>, dim3(nb), dim3(256), 0, c.s, L.z2, L.z, L.r, lvp(L), L.tiles, L.ntiles, ntxy,
                         first, c.cfg.omega);
    std::swap(L.z, L.z2);
    return;
  }
  if ((long long)L.n < c.cfg.fused_sweep_min_cells) {
    const unsigned nb = smooth_blocks(L);
    for (int k = skip_first ? 1 : 0; k < 2; ++k) {
      const int phase = k == 0 ? first : 1 - first;
      if (l == 0 && L.segs)
        hipLaunchKernelGGL(k_smooth_seg, dim3(seg_blocks(L)), dim3(kBlock), 0, c.s, L.z, L.r, lvp(L), L.segs, L.nsegs,
                           L.d[2] / kSegZ, phase, c.cfg.omega, mode);
      else if (l == 0)
        hipLaunchKernelGGL(k_smooth<true>, dim3(nb), dim3(kBlock), 0, c.s, L.z, L.r, lvp(L), L.rows, L.nrows, phase,
                           c.cfg.omega, mode);
      else
        hipLaunchKernelGGL(k_smooth<false>, dim3(nb), dim3(kBlock), 0, c.s, L.z, L.r, lvp(L), L.rows, L.nrows, phase,
                           c.cfg.omega, mode);
    }
    return;
  }
  unsigned nb = (unsigned)(div_up(L.d[0], kSwX) * div_up(L.d[1], kSwY) * div_up(L.d[2], kSwZ));
  if (mode == 0)
    hipLaunchKernelGGL(k_rbgs_sweep<0>, dim3(nb), dim3(256), 0, c.s, L.z2, L.z, L.r, lvp(L), first, c.cfg.omega);
  else
    hipLaunchKernelGGL(k_rbgs_sweep<1>, dim3(nb), dim3(256), 0, c.s, L.z2, L.z, L.r, lvp(L), first, c.cfg.omega);
  std::swap(L.z, L.z2);
}

// `fwd` sweeps starting with colour `first`, then `rev` starting with the other colour, on
// level l: consecutive launches, timed as one run when profiling.  skip_first: the first sweep's
// first colour pass has been folded into the kernel that zeroed z (colour_passes(l) only).
void sweeps(Ctx& c, int l, int first, int fwd, int rev, int mode, bool skip_first = false) {
  if (fwd + rev <= 0) return;
  const bool passes = colour_passes(c, l);
  ProfScope ps(c, l == 0 ? LH_PROF_SMOOTH_L0 : (l == 1 ? LH_PROF_SMOOTH_L1 : LH_PROF_SMOOTH_L2),
               (fwd + rev) * (passes ? 2 : 1) - (skip_first ? 1 : 0));
  for (int it = 0; it < fwd; ++it) sweep(c, l, first, mode, skip_first && it == 0);
  for (int it = 0; it < rev; ++it) sweep(c, l, 1 - first, mode, skip_first && fwd == 0 && it == 0);
}

// The coarsest level's sweeps in FP64 (coarsest_fp64): `fwd` row sweeps starting with colour 0,
// then `rev` starting with colour 1, on z64 (zeroed by the restriction) with r promoted, then z =
// (float) z64 on the z-support rows for the prolongation
void coarsest_sweeps64(Ctx& c, int fwd, int rev, int mode) {
  const int l = (int)c.lv.size() - 1;
  Level& L = c.lv[l];
  const unsigned nb = row_sweep_blocks(L);
  {
    ProfScope ps(c, l == 1 ? LH_PROF_SMOOTH_L1 : LH_PROF_SMOOTH_L2, fwd + rev);
    for (int it = 0; it < fwd + rev; ++it) {
      const int first = it < fwd ? 0 : 1;
      if (mode == 0)
        hipLaunchKernelGGL((k_sweep_rows<0, double>), dim3(nb), dim3(64 * kSweepWaves), 0, c.s, L.z64b, L.z64, L.r,
                           lvp(L), L.zrows, L.nzrows, first, c.cfg.omega);
      else
        hipLaunchKernelGGL((k_sweep_rows<1, double>), dim3(nb), dim3(64 * kSweepWaves), 0, c.s, L.z64b, L.z64, L.r,
                           lvp(L), L.zrows, L.nzrows, first, c.cfg.omega);
      std::swap(L.z64, L.z64b);
    }
  }
  hipLaunchKernelGGL(k_z64_to_f32, dim3(row_blocks(L)), dim3(kBlock), 0, c.s, L.z, L.z64, L.zrows, L.nzrows, L.d[2]);
}

// The coarsest level solved by coarsest_cheb Chebyshev steps on its Jacobi-scaled system over the
// interval [cheb_lo, 2] (in FP64 on z64 / d64, from z64 = 0 as the restriction leaves it), two
// steps per launch, then z = (float) z64 on the z-support rows for the prolongation.  The step
// coefficients follow the oracle's recurrence in the same double operations: theta = (2 + lo) / 2,
// delta = (2 - lo) / 2, sigma = theta / delta, rho_0 = 1 / sigma; step 0: b = 1 / theta; step k:
// rho = 1 / (2 sigma - rho_prev), a = rho rho_prev, b = 2 rho / delta.  The last launch also
// writes z = (float) z64 for the prolongation.
void coarsest_chebyshev64(Ctx& c) {
  const int l = (int)c.lv.size() - 1;
  Level& L = c.lv[l];
  const int steps = c.cfg.coarsest_cheb;
  const double lo = (double)c.cfg.cheb_lo, hi = 2.0;
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
  double rho_prev = 1.0 / sigma;
  std::vector<double> A(steps), B(steps);
  for (int k = 0; k < steps; ++k) {
    A[k] = 0.0;
    B[k] = 1.0 / theta;
    if (k > 0) {
      const double rho = 1.0 / (2.0 * sigma - rho_prev);
      A[k] = rho * rho_prev;
      B[k] = 2.0 * rho / delta;
      rho_prev = rho;
    }
  }
  const unsigned nb = row_sweep_blocks(L);
  {
    ProfScope ps(c, l == 1 ? LH_PROF_SMOOTH_L1 : LH_PROF_SMOOTH_L2, (steps + 1) / 2);
    for (int k = 0; k < steps; k += 2) {
      ChebSteps q;
      q.two = k + 1 < steps ? 1 : 0;
      q.first = k == 0 ? 1 : 0;
      q.a0 = A[k];
      q.b0 = B[k];
      q.a1 = q.two ? A[k + 1] : 0.0;
      q.b1 = q.two ? B[k + 1] : 0.0;
      hipLaunchKernelGGL(k_cheb_rows<double>, dim3(nb), dim3(64 * kSweepWaves), 0, c.s, L.z64b, L.d64b, L.z64, L.d64,
                         L.r, lvp(L), L.zrows, L.nzrows, q, k + 2 >= steps ? L.z : (float*)nullptr);
      std::swap(L.z64, L.z64b);
      std::swap(L.d64, L.d64b);
    }
  }
}

// ApplyPreconditioner, PressureSolver.cpp:719-772 (symmetric mode: z[0] = 0 first,
// post-smoothing and the second half of the coarsest sweeps in reverse colour order)
// first0: the caller's zeroing of z has also done level 0's first pre-smoothing colour pass
// (k_cg_update_xr with first, symmetric V-cycle with colour passes on level 0)
void apply_precond(Ctx& c, bool z_zeroed, bool first0 = false) {
  Level& L0 = c.lv[0];
  if (c.cfg.precond != LH_PRECOND_MULTIGRID) {
    hipLaunchKernelGGL(k_precond_simple, dim3(grid_for(L0.n)), dim3(kBlock), 0, c.s, L0.z, L0.r, L0.adiag, L0.cls,
                       L0.n, c.cfg.precond == LH_PRECOND_JACOBI ? 1 : 0);
    return;
  }
  const int L = (int)c.lv.size();
  const int sym = (c.cfg.smoother == LH_SMOOTHER_SYMMETRIC);
  const int mode = sym ? 1 : 0;
  // z[0] = 0 where it is read: on the liquid rows (elsewhere it only meets zero adiag and r),
  // unless the caller has cleared it
  if (sym && !z_zeroed)
    hipLaunchKernelGGL(k_zero_rows, dim3(row_blocks(L0)), dim3(kBlock), 0, c.s, L0.z, L0.rows, L0.nrows, L0.d[2]);
  bool folded = first0;  // level l - 1's first pre-smoothing pass done by the kernel that zeroed its z
  for (int l = 1; l < L; ++l) {
    sweeps(c, l - 1, 0, c.cfg.num_pre, 0, mode, folded);
    Level& Lc = c.lv[l];
    Level& Lf = c.lv[l - 1];
    // the restriction zeroes z of level l; a level that pre-smooths with colour passes gets its
    // first pass there
    folded = l < L - 1 && c.cfg.num_pre > 0 && colour_passes(c, l);
    const size_t nch = div_up((size_t)Lc.d[2], (size_t)32);
    const unsigned nb = (unsigned)std::min<size_t>((size_t)Lc.d[0] * Lc.d[1] * nch, 16384);
    hipLaunchKernelGGL(k_restrict, dim3(nb), dim3(kBlock), 0, c.s, Lc.r, Lc.z, Lf.r, Lf.z, lvp(Lf), Lc.d[1], Lc.d[2],
                       Lc.qrows, Lc.nqrows, Lc.d[0] * Lc.d[1], Lc.z64, Lc.adiag, folded ? 1 : 0, c.cfg.omega,
                       mode);
  }
  if (c.cfg.coarsest_cheb > 0)
    coarsest_chebyshev64(c);
  else if (c.lv[L - 1].z64)
    coarsest_sweeps64(c, c.cfg.num_final, sym ? c.cfg.num_final : 0, mode);
  else
    sweeps(c, L - 1, 0, c.cfg.num_final, sym ? c.cfg.num_final : 0, mode);
  for (int l = L - 2; l >= 0; --l) {
    Level& Lc = c.lv[l + 1];
    Level& Lf = c.lv[l];
    hipLaunchKernelGGL(k_prolong, dim3(row_blocks(Lc)), dim3(kBlock), 0, c.s, Lf.z, Lc.z, Lf.d[1], Lf.d[2], Lc.d[1],
                       Lc.d[2], Lc.zrows, Lc.nzrows);
    sweeps(c, l, sym ? 1 : 0, c.cfg.num_post, 0, mode);
  }
}

}  // namespace

namespace {

// ---------------------------------------------------------------- warm start
// x = x0 on interior liquid cells (the cells the PCG updates, :875-877), 0 elsewhere
__global__ void k_cg_set_x0(float* __restrict__ x, float* __restrict__ p, float* __restrict__ z,
                            const float* __restrict__ x0, const uint8_t* __restrict__ cls, int d0, int d1, int d2) {
  size_t total = (size_t)d0 * d1 * d2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    int zi = (int)(i % d2), yi = (int)((i / d2) % d1), xi = (int)(i / ((size_t)d1 * d2));
    bool in = xi >= 1 && xi <= d0 - 2 && yi >= 1 && yi <= d1 - 2 && zi >= 1 && zi <= d2 - 2 && cls[i] != 0;
    x[i] = in ? x0[i] : 0.0f;
    p[i] = 0.0f;
    z[i] = 0.0f;
  }
}
// r = b - A x on liquid cells (0 elsewhere); partials of |b|^2 and |r|^2 over liquid cells
__global__ void k_cg_warm_residual(float* __restrict__ r, const float* __restrict__ x, const float* __restrict__ b,
                                   LvP L, double* __restrict__ partial) {
  const int d0 = L.d[0], d1 = L.d[1], d2 = L.d[2];
  size_t total = (size_t)d0 * d1 * d2;
  double v[2] = {0.0, 0.0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    float rv = 0.0f;
    if (L.cls[i] != 0) {
      int zi = (int)(i % d2), yi = (int)((i / d2) % d1), xi = (int)(i / ((size_t)d1 * d2));
      rv = b[i] - get_ax(L, x, xi, yi, zi, i);
      v[0] += (double)b[i] * (double)b[i];
      v[1] += (double)rv * (double)rv;
    }
    r[i] = rv;
  }
  block_partials<2>(v, partial);
}
__global__ void k_sc_norms(const double* __restrict__ part, int nb, double* __restrict__ sc64) {
  double bb = sum_partials_block(part, nb);
  double rr = sum_partials_block(part + nb, nb);
  if (threadIdx.x == 0) {
    sc64[4] = bb;
    sc64[5] = rr;
  }
}

// ---------------------------------------------------------------- FP64 refinement
__device__ __forceinline__ double get_ax64(const LvP& L, const double* __restrict__ z, int x, int y, int w, size_t c) {
  double res = z[c] * (double)L.adiag[c];
  if (L.cls[c] == 0) return res;
  const size_t sx = (size_t)L.d[1] * L.d[2], sy = (size_t)L.d[2];
  if (x + 1 < L.d[0] && L.cls[c + sx] == 2) res += z[c + sx] * (double)(-L.t0[c]);
  if (x - 1 >= 0 && L.cls[c - sx] == 2) res += z[c - sx] * (double)(-L.t0[c - sx]);
  if (y + 1 < L.d[1] && L.cls[c + sy] == 2) res += z[c + sy] * (double)(-L.t1[c]);
  if (y - 1 >= 0 && L.cls[c - sy] == 2) res += z[c - sy] * (double)(-L.t1[c - sy]);
  if (w + 1 < L.d[2] && L.cls[c + 1] == 2) res += z[c + 1] * (double)(-L.t2[c]);
  if (w - 1 >= 0 && L.cls[c - 1] == 2) res += z[c - 1] * (double)(-L.t2[c - 1]);
  return res;
}
__device__ __forceinline__ bool interior_liquid(const LvP& L, size_t i, int& xi, int& yi, int& zi) {
  const int d1 = L.d[1], d2 = L.d[2];
  zi = (int)(i % d2);
  yi = (int)((i / d2) % d1);
  xi = (int)(i / ((size_t)d1 * d2));
  return xi >= 1 && xi <= L.d[0] - 2 && yi >= 1 && yi <= d1 - 2 && zi >= 1 && zi <= d2 - 2 && L.cls[i] != 0;
}
// r64 = b - A x64 on the interior liquid cells of the level-0 liquid rows (0 on their other
// cells); partials of |b|^2, |r64|^2 there.  Off those rows r64 is never read: the PCG of the
// correction takes its right-hand side on liquid cells only.
__global__ void k_ir_residual(double* __restrict__ r64, const double* __restrict__ x64, const float* __restrict__ b,
                              LvP L, const int* __restrict__ rows, const int* __restrict__ nrows,
                              double* __restrict__ partial) {
  const int d1 = L.d[1], d2 = L.d[2];
  double v[2] = {0.0, 0.0};
  const RowIter it = row_iter(d2);
  const int nr = *nrows;
  if (it.on)
    for (int k = blockIdx.x * it.R + it.sub, kst = gridDim.x * it.R, rn = k < nr ? rows[k] : 0; k < nr;
         k += kst) {  // the next row id loads with this row's operands
      const size_t base = (size_t)rn * d2;
      rn = k + kst < nr ? rows[k + kst] : 0;
      for (int zi = it.z0; zi < d2; zi += it.zs) {
        const size_t i = base + zi;
        int xi, yi, zz;
        double rv = 0.0;
        if (interior_liquid(L, i, xi, yi, zz)) {
          rv = (double)b[i] - get_ax64(L, x64, xi, yi, zz, i);
          v[0] += (double)b[i] * (double)b[i];
          v[1] += rv * rv;
        }
        r64[i] = rv;
      }
    }
  (void)d1;
  block_partials<2>(v, partial);
}
// FP32 right-hand side of the correction equation, scaled to unit norm (level-0 liquid rows)
__global__ void k_ir_rhs(float* __restrict__ rf, const double* __restrict__ r64, const double* __restrict__ sc64,
                         const int* __restrict__ rows, const int* __restrict__ nrows, int d2) {
  const double inv = (sc64[5] > 0.0) ? 1.0 / sqrt(sc64[5]) : 0.0;
  const RowIter it = row_iter(d2);
  const int nr = *nrows;
  if (!it.on) return;
  for (int k = blockIdx.x * it.R + it.sub, kst = gridDim.x * it.R, rn = k < nr ? rows[k] : 0; k < nr;
       k += kst) {  // the next row id loads with this row's operands
    const size_t base = (size_t)rn * d2;
    rn = k + kst < nr ? rows[k + kst] : 0;
    for (int zi = it.z0; zi < d2; zi += it.zs) rf[base + zi] = (float)(r64[base + zi] * inv);
  }
}
// x64 += ||r|| e over the level-0 liquid rows (e, the correction, is zero off the PCG's cells)
__global__ void k_ir_update(double* __restrict__ x64, const float* __restrict__ e, const double* __restrict__ sc64,
                            const int* __restrict__ rows, const int* __restrict__ nrows, int d2) {
  const double s = sqrt(sc64[5]);
  const RowIter it = row_iter(d2);
  const int nr = *nrows;
  if (!it.on) return;
  for (int k = blockIdx.x * it.R + it.sub, kst = gridDim.x * it.R, rn = k < nr ? rows[k] : 0; k < nr;
       k += kst) {  // the next row id loads with this row's operands
    const size_t base = (size_t)rn * d2;
    rn = k + kst < nr ? rows[k + kst] : 0;
    for (int zi = it.z0; zi < d2; zi += it.zs) x64[base + zi] += s * (double)e[base + zi];
  }
}
__global__ void k_ir_to_f32(float* __restrict__ out, const double* __restrict__ x64, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = (float)x64[i];
}
__global__ void k_f32_to_f64(double* __restrict__ out, const float* __restrict__ in, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = (double)in[i];
}

}  // namespace

// Init (:666-680: r and z of every level zeroed), BuildRHS (:682-698), BuildLHS (:544-664).
// Level 0's r and z are written over the whole grid by every solve's first kernel (k_cg_init,
// or k_cg_set_x0 and k_cg_warm_residual) before anything reads them, so only levels >= 1 are
// cleared here.
void build_system(Ctx& c, float rho, float dt) {
  for (size_t l = 1; l < c.lv.size(); ++l) {
    Level& L = c.lv[l];
    launch_fill_f32(c, L.r, L.n, 0.f);
    launch_fill_f32(c, L.z, L.n, 0.f);
    // the tiled sweep's ping-pong partner: zero outside the tiles for the whole solve, as z
    if (c.cfg.coarse_sweep == kCoarseTiles) launch_fill_f32(c, L.z2, L.n, 0.f);
  }
  launch_build_rhs(c);
  launch_build_lhs(c, rho, dt);
}

// The PCG loop of CGSolver<3>::Solve (PressureSolver.cpp:845-931) on the operator
// already built, right-hand side b (device, Array3 layout), optional initial
// guess x0 (device).  With x0, the PCG starts from x0 restricted to the cells it
// updates and falls back to x = 0 when ||b - A x0|| > ||b||.  The stop rule is
// the configured one; in relative mode it is ||r|| <= tol * ||b||.  Result in c.x.
namespace {

// The level-0 work list of the PCG vector kernels (the x/r update, r.z, the p update + SpMV): the
// z-segments holding liquid when the level keeps them, else the rows holding liquid.  The cells
// they leave out have cls == 0 (r = 0, no x / r / p update, zero weight in every sum and stencil).
struct WorkList {
  const int* list;
  const int* n;
  int nseg;
};
WorkList level0_work(const Level& L0) {
  if (L0.segs) return {L0.segs, L0.nsegs, L0.d[2] / kSegZ};
  return {L0.rows, L0.nrows, 0};
}

// One PCG iteration up to the r.z the host waits for (PressureSolver.cpp:879-906): x/r update,
// preconditioner, r.z.  part_pq / part_rr: the p.q and r.r partials in c.partial.
void pcg_iteration_launches(Ctx& c, int slot, double* part_pq, double* part_rr) {
  Level& L0 = c.lv[0];
  const int nb = c.npartial;
  const int sym = c.cfg.precond == LH_PRECOND_MULTIGRID && c.cfg.smoother == LH_SMOOTHER_SYMMETRIC;
  // the symmetric V-cycle zeroes level-0 z here; with colour passes on level 0 (and more than one
  // level, so that level 0 pre-smooths) its first pass is folded in
  const bool first0 = sym && c.lv.size() > 1 && c.cfg.num_pre > 0 && colour_passes(c, 0);
  const WorkList wl = level0_work(L0);
  hipLaunchKernelGGL(k_cg_update_xr, dim3(nb), dim3(kBlock), 0, c.s, c.x, L0.r, c.p, c.q, L0.cls, c.dsc, slot,
                     c.dflag + 2, part_pq, nb, L0.d[0], L0.d[1], L0.d[2], wl.list, wl.n, part_rr, L0.z, sym,
                     L0.adiag, first0 ? 1 : 0, c.cfg.omega, wl.nseg);
  apply_precond(c, sym != 0, first0);
  hipLaunchKernelGGL(k_cg_dot_rz, dim3(nb), dim3(kBlock), 0, c.s, L0.r, L0.z, wl.list, wl.n, L0.d[2], c.partial,
                     part_rr, nb, c.dsc, c.dflag + 2, c.h_pin_dev, c.h_flag_dev, wl.nseg);
}

// the levels' ping-pong buffers, which the preconditioner's host code swaps as it launches
std::vector<void*> level_buffers(const Ctx& c) {
  std::vector<void*> v;
  for (const Level& L : c.lv) {
    v.push_back(L.z);
    v.push_back(L.z2);
    v.push_back(L.z64);
    v.push_back(L.z64b);
    v.push_back(L.d64);
    v.push_back(L.d64b);
  }
  return v;
}
void set_level_buffers(Ctx& c, const std::vector<void*>& v) {
  size_t k = 0;
  for (Level& L : c.lv) {
    L.z = (float*)v[k++];
    L.z2 = (float*)v[k++];
    L.z64 = (double*)v[k++];
    L.z64b = (double*)v[k++];
    L.d64 = (double*)v[k++];
    L.d64b = (double*)v[k++];
  }
}

// pcg_iteration_launches as an instantiated hipGraph (pcgGraph).  Every argument of those launches
// is fixed for the context's lifetime except the p buffer (p ping-pongs with p2), the rho slot and
// the levels' ping-pong buffers, so a graph is captured once per combination of them and replayed
// in every later solve; the grids are sized from the row lists' capacities and the kernels read
// the lists and their lengths on the device.  The captured launches contain no memset (this
// runtime has given wrong results on replays of memset nodes) and no event (profiling uses the
// direct launches).
bool pcg_iteration_graph(Ctx& c, int slot, double* part_pq, double* part_rr) {
  const std::vector<void*> before = level_buffers(c);
  for (const Ctx::PcgGraph& g : c.pcg_graphs)
    if (g.p == c.p && g.slot == slot && g.before == before) {
      if (!hip_ok(c, hipGraphLaunch(g.exec, c.s), "PCG iteration graph launch")) return false;
      set_level_buffers(c, g.after);
      return true;
    }
  if (!hip_ok(c, hipStreamBeginCapture(c.s, hipStreamCaptureModeThreadLocal), "PCG iteration capture")) return false;
  pcg_iteration_launches(c, slot, part_pq, part_rr);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(c.s, &graph);
  const std::vector<void*> after = level_buffers(c);
  set_level_buffers(c, before);  // the captured launches have not run yet
  if (!hip_ok(c, e, "PCG iteration capture")) return false;
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (!hip_ok(c, ei, "PCG iteration graph instantiation")) return false;
  Ctx::PcgGraph g;
  g.exec = exec;
  g.p = c.p;
  g.slot = slot;
  g.before = before;
  g.after = after;
  c.pcg_graphs.push_back(g);
  if (!hip_ok(c, hipGraphLaunch(exec, c.s), "PCG iteration graph launch")) return false;
  set_level_buffers(c, after);
  return true;
}

}  // namespace

int pcg_solve(Ctx& c, const float* b, const float* x0, float tol, int tol_mode, int max_iter, lh_solve_stats& S) {
  Level& L0 = c.lv[0];
  const size_t n = L0.n;
  const int nb = c.npartial;
  LvP P0 = lvp(L0);
  hipMemsetAsync(c.dflag + 2, 0, sizeof(int), c.s);
  for (size_t l = 1; l < c.lv.size(); ++l) launch_fill_f32(c, c.lv[l].z, c.lv[l].n, 0.f);
  double bnorm = -1.0;
  S.warm_start_used = 0;
  if (x0) {
    hipLaunchKernelGGL(k_cg_set_x0, dim3(grid_for(n)), dim3(kBlock), 0, c.s, c.x, c.p, L0.z, x0, L0.cls, L0.d[0],
                       L0.d[1], L0.d[2]);
    hipLaunchKernelGGL(k_cg_warm_residual, dim3(nb), dim3(kBlock), 0, c.s, L0.r, c.x, b, P0, c.partial);
    hipLaunchKernelGGL(k_sc_norms, dim3(1), dim3(kBlock), 0, c.s, c.partial, nb, c.dsc64);
    hipMemcpyAsync(c.h_pin + 16, c.dsc64 + 4, 2 * sizeof(double), hipMemcpyDeviceToHost, c.s);
    if (!hip_ok(c, hipStreamSynchronize(c.s), "warm start")) return LH_ERR_DEVICE;
    bnorm = (double)sqrtf((float)c.h_pin[16]);
    double r0 = (double)sqrtf((float)c.h_pin[17]);
    S.warm_start_used = (r0 <= bnorm) ? 1 : 0;  // fallback: the zero guess is better
  }
  if (!S.warm_start_used)
    hipLaunchKernelGGL(k_cg_init, dim3(grid_for(n)), dim3(kBlock), 0, c.s, L0.r, c.x, c.p, L0.z, b, L0.cls, n);
  apply_precond(c, true);  // k_cg_init and k_cg_set_x0 clear z
  hipLaunchKernelGGL(k_cg_first, dim3(nb), dim3(kBlock), 0, c.s, c.p, L0.z, L0.r, L0.cls, L0.rows, L0.nrows, L0.d[2],
                     c.partial);
  hipLaunchKernelGGL(k_sc_first, dim3(1), dim3(kBlock), 0, c.s, c.partial, nb, c.dsc, c.dsc64);
  hipMemcpyAsync(c.h_pin, c.dsc, 8 * sizeof(float), hipMemcpyDeviceToHost, c.s);
  // partials: r.z in [0, nb), p.q in [nb, 2 nb), r.r in [2 nb, 3 nb) (no kernel reads the range
  // it writes); rho alternates between sc[6] and sc[7].  An iteration's p update and the next
  // q = A p are queued before the host has read the residual, so the GPU runs them while the host
  // waits: when that residual shows convergence they are discarded (they write only p2, q, the
  // partial sums, sc[0, 1, 3], the other rho slot and the NaN flag; x and r are untouched).
  double* part_pq = c.partial + (size_t)nb;
  double* part_rr = c.partial + 2 * (size_t)nb;
  hipEventRecord(c.ev_iter, c.s);
  {
    ProfScope ps(c, LH_PROF_SPMV);
    hipLaunchKernelGGL(k_cg_spmv, dim3(nb), dim3(kBlock), 0, c.s, c.q, c.p, P0, L0.rows, L0.nrows, part_pq);
  }
  if (!hip_ok(c, hipEventSynchronize(c.ev_iter), "solve init")) return LH_ERR_DEVICE;
  float hsc[8];
  std::memcpy(hsc, c.h_pin, sizeof(hsc));
  float residual = hsc[4];
  const float res0 = residual;
  if (bnorm < 0.0) bnorm = (double)res0;  // cold start: ||r0|| = ||b|| over liquid cells
  auto done = [&](float res) {
    if (tol_mode == LH_TOL_ABSOLUTE) return res < tol;
    return (double)res <= (double)tol * bnorm;
  };
  int it = 0;
  bool conv = false, nan = false;
  for (it = 0; it < max_iter; ++it) {
    if (done(residual)) {
      conv = true;
      break;
    }
    // q = A p of this iteration is already queued
    const int slot = it & 1;
    if (c.cfg.pcg_graph && !c.prof) {
      if (!pcg_iteration_graph(c, slot, part_pq, part_rr)) return LH_ERR_DEVICE;
    } else {
      pcg_iteration_launches(c, slot, part_pq, part_rr);
    }
    hipEventRecord(c.ev_iter, c.s);
    {  // p = z + beta p and the next iteration's q = A p, while the host waits for the event and decides
      ProfScope ps(c, LH_PROF_SPMV);
      const WorkList wl = level0_work(L0);
      hipLaunchKernelGGL(k_cg_p_spmv, dim3(nb), dim3(kBlock), 0, c.s, c.p2, c.q, c.p, L0.z, P0, wl.list, wl.n,
                         c.dsc, slot, c.dflag + 2, c.partial, nb, part_pq, wl.nseg);
    }
    std::swap(c.p, c.p2);
    if (!hip_ok(c, hipEventSynchronize(c.ev_iter), "solve iteration")) return LH_ERR_DEVICE;
    std::memcpy(hsc, c.h_pin, sizeof(hsc));
    residual = hsc[4];
    if (*c.h_flag) {
      nan = true;
      ++it;
      break;
    }
  }
  if (!conv && !nan && it >= max_iter) conv = done(residual);
  S.iterations += it;
  S.converged = conv ? 1 : 0;
  S.initial_residual = res0;
  S.final_residual = residual;
  S.rhs_norm = bnorm;
  if (nan) {
    c.err = "PCG produced a NaN/Inf alpha or beta";
    return LH_ERR_NAN;
  }
  if (!conv) {
    c.err = "PCG did not reach the tolerance within maxIterations";
    return LH_ERR_NOT_CONVERGED;
  }
  return LH_OK;
}

namespace {

bool ensure_fp64(Ctx& c) {
  if (c.x64) return true;
  if (hipMalloc((void**)&c.x64, c.ncell * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&c.r64, c.ncell * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&c.rf, c.ncell * sizeof(float)) != hipSuccess) {
    c.err = "hipMalloc failed (FP64 refinement buffers)";
    return false;
  }
  return true;
}

// Mixed-precision iterative refinement: the true residual r = b - A x is formed
// in FP64 (FP64 arithmetic on the FP32 operator) over the cells the PCG updates;
// each correction A e = r / ||r|| is solved by the FP32 MGPCG to the configured
// tolerance and accumulated into the FP64 solution, until ||r|| <= tol64 ||b||.
int refine_solve(Ctx& c, const float* b, const float* x0, lh_solve_stats& S) {
  if (!ensure_fp64(c)) return LH_ERR_OOM;
  Level& L0 = c.lv[0];
  const size_t n = L0.n;
  const int nb = c.npartial;
  LvP P0 = lvp(L0);
  if (x0) {
    hipLaunchKernelGGL(k_cg_set_x0, dim3(grid_for(n)), dim3(kBlock), 0, c.s, c.x, c.p, L0.z, x0, L0.cls, L0.d[0],
                       L0.d[1], L0.d[2]);
    hipLaunchKernelGGL(k_f32_to_f64, dim3(grid_for(n)), dim3(kBlock), 0, c.s, c.x64, c.x, n);
  } else {
    hipMemsetAsync(c.x64, 0, n * sizeof(double), c.s);
  }
  int rc = LH_OK;
  for (int outer = 0;; ++outer) {
    hipLaunchKernelGGL(k_ir_residual, dim3(nb), dim3(kBlock), 0, c.s, c.r64, c.x64, b, P0, L0.rows, L0.nrows,
                       c.partial);
    hipLaunchKernelGGL(k_sc_norms, dim3(1), dim3(kBlock), 0, c.s, c.partial, nb, c.dsc64);
    hipMemcpyAsync(c.h_pin + 16, c.dsc64 + 4, 2 * sizeof(double), hipMemcpyDeviceToHost, c.s);
    if (!hip_ok(c, hipStreamSynchronize(c.s), "refinement")) return LH_ERR_DEVICE;
    const double bn = std::sqrt(c.h_pin[16]), rn = std::sqrt(c.h_pin[17]);
    S.true_residual = (bn > 0.0) ? rn / bn : 0.0;
    S.outer_iterations = outer;
    if (outer == 0 && x0 && rn > bn) {
      // the warm start is worse than the zero guess: restart from x = 0
      hipMemsetAsync(c.x64, 0, n * sizeof(double), c.s);
      x0 = nullptr;
      continue;
    }
    if (rn <= c.cfg.refined_tol * bn || bn == 0.0) break;
    if (outer >= c.cfg.max_outer) {
      c.err = "FP64 refinement did not reach refinedTolerance";
      rc = LH_ERR_NOT_CONVERGED;
      break;
    }
    hipLaunchKernelGGL(k_ir_rhs, dim3(row_blocks(L0)), dim3(kBlock), 0, c.s, c.rf, c.r64, c.dsc64, L0.rows, L0.nrows,
                       L0.d[2]);
    // inner tolerance: the correction only has to bring ||r|| under tol64 ||b||, so it is
    // solved to max(tol, 0.25 tol64 ||b|| / ||r||) relative (at most 0.1), never tighter than
    // the configured FP32 tolerance
    double inner = std::max((double)c.cfg.tol, 0.25 * c.cfg.refined_tol * bn / rn);
    inner = std::min(inner, 0.1);
    int irc = pcg_solve(c, c.rf, nullptr, (float)inner, LH_TOL_RELATIVE, c.cfg.max_iter, S);
    if (irc == LH_ERR_DEVICE || irc == LH_ERR_NAN) return irc;
    hipLaunchKernelGGL(k_ir_update, dim3(row_blocks(L0)), dim3(kBlock), 0, c.s, c.x64, c.x, c.dsc64, L0.rows,
                       L0.nrows, L0.d[2]);
  }
  S.warm_start_used = x0 ? 1 : 0;
  S.converged = (rc == LH_OK) ? 1 : 0;
  S.rhs_norm = std::sqrt(c.h_pin[16]);
  hipLaunchKernelGGL(k_ir_to_f32, dim3(grid_for(n)), dim3(kBlock), 0, c.s, c.x, c.x64, n);
  return rc;
}

}  // namespace

// solve A x = b on the built operator in the configured precision; x -> pressure
int solve_built(Ctx& c, const float* b, const float* x0, lh_solve_stats* st, bool record_start) {
  lh_solve_stats S;
  std::memset(&S, 0, sizeof(S));
  if (record_start) hipEventRecord(c.ev0, c.s);
  int rc;
  if (c.cfg.refined) rc = refine_solve(c, b, x0, S);
  else rc = pcg_solve(c, b, x0, c.cfg.tol, c.cfg.tol_mode, c.cfg.max_iter, S);
  if (rc == LH_ERR_DEVICE) return rc;
  // m_grid.pressure.CopyFrom(m_x), :933
  hipMemcpyAsync(c.pressure, c.x, c.ncell * sizeof(float), hipMemcpyDeviceToDevice, c.s);
  hipEventRecord(c.ev1, c.s);
  if (!hip_ok(c, hipEventSynchronize(c.ev1), "solve")) return LH_ERR_DEVICE;
  float ms = 0.f;
  hipEventElapsedTime(&ms, c.ev0, c.ev1);
  S.solve_ms = ms;
  if (st) *st = S;
  return rc;
}

namespace {
// partials of (x . b, x . A x) over the level-0 liquid cells (x is zero outside the PCG set)
__global__ void k_ritz_partials(const float* __restrict__ x, const float* __restrict__ b, LvP L,
                                double* __restrict__ partial) {
  const int d0 = L.d[0], d1 = L.d[1], d2 = L.d[2];
  size_t total = (size_t)d0 * d1 * d2;
  double v[2] = {0.0, 0.0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    if (L.cls[i] == 0 || x[i] == 0.0f) continue;
    int zi = (int)(i % d2), yi = (int)((i / d2) % d1), xi = (int)(i / ((size_t)d1 * d2));
    const double ax = (double)get_ax(L, x, xi, yi, zi, i);
    v[0] += (double)x[i] * (double)b[i];
    v[1] += (double)x[i] * ax;
  }
  block_partials<2>(v, partial);
}
__global__ void k_ritz_final(const double* __restrict__ part, int nb, double* __restrict__ sc64) {
  double xb = sum_partials_block(part, nb);
  double xax = sum_partials_block(part + nb, nb);
  if (threadIdx.x == 0) sc64[6] = (xax > 0.0) ? xb / xax : 0.0;
}
__global__ void k_ritz_scale(float* __restrict__ x, const double* __restrict__ sc64, size_t n) {
  const float a = (float)sc64[6];
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] *= a;
}
}  // namespace

namespace {
// ay = A y on liquid cells; partials of (y . b, y . A y, xs . b) over liquid cells
__global__ void k_energy_partials(const float* __restrict__ y, const float* __restrict__ b,
                                  const float* __restrict__ xs, LvP L, float* __restrict__ ay,
                                  double* __restrict__ partial) {
  const int d0 = L.d[0], d1 = L.d[1], d2 = L.d[2];
  size_t total = (size_t)d0 * d1 * d2;
  double v[3] = {0.0, 0.0, 0.0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    float a = 0.0f;
    if (L.cls[i] != 0) {
      int zi = (int)(i % d2), yi = (int)((i / d2) % d1), xi = (int)(i / ((size_t)d1 * d2));
      a = get_ax(L, y, xi, yi, zi, i);
      v[0] += (double)y[i] * (double)b[i];
      v[1] += (double)y[i] * (double)a;
      v[2] += (double)xs[i] * (double)b[i];
    }
    ay[i] = a;
  }
  block_partials<3>(v, partial);
}
__global__ void k_energy_final(const double* __restrict__ part, int nb, double* __restrict__ out) {
  double s0 = sum_partials_block(part, nb);
  double s1 = sum_partials_block(part + nb, nb);
  double s2 = sum_partials_block(part + 2 * nb, nb);
  if (threadIdx.x == 0) {
    out[0] = s0;
    out[1] = s1;
    out[2] = s2;
  }
}
}  // namespace

// ay = A y and out[0..2] = (y . b, y . A y, xs . b) on the built level-0 operator (device)
void launch_energy_terms(Ctx& c, const float* y, const float* b, const float* xs, float* ay, double* out) {
  Level& L0 = c.lv[0];
  const int nb = c.npartial;
  hipLaunchKernelGGL(k_energy_partials, dim3(nb), dim3(kBlock), 0, c.s, y, b, xs, lvp(L0), ay, c.partial);
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kBlock), 0, c.s, c.partial, nb, out);
}

// x0 <- alpha x0, alpha = (x0 . b) / (x0 . A x0): the multiple of the guess closest to
// the solution in the A-norm, so ||x* - alpha x0||_A <= ||x*||_A (the zero guess's error)
void launch_warm_ritz(Ctx& c, float* x0, const float* b) {
  Level& L0 = c.lv[0];
  const int nb = c.npartial;
  hipLaunchKernelGGL(k_ritz_partials, dim3(nb), dim3(kBlock), 0, c.s, x0, b, lvp(L0), c.partial);
  hipLaunchKernelGGL(k_ritz_final, dim3(1), dim3(kBlock), 0, c.s, c.partial, nb, c.dsc64);
  hipLaunchKernelGGL(k_ritz_scale, dim3(grid_for(L0.n)), dim3(kBlock), 0, c.s, x0, c.dsc64, L0.n);
}

// CGSolver<3>::Solve, PressureSolver.cpp:839-934: Init, BuildRHS, BuildLHS, PCG
// from x = 0 (or the installed / learned initial guess), pressure = x.
int run_solve(Ctx& c, float rho, float dt, lh_solve_stats* st) {
  hipEventRecord(c.ev0, c.s);
  build_system(c, rho, dt);
  c.last_dt = dt;
  const float* x0 = c.warm_x0;
  if (c.warm_learned && c.net) {
    float* g = nullptr;
    int rc = nn_predict(c, c.rhs, &g);
    if (rc) return rc;
    x0 = g;
  }
  return solve_built(c, c.rhs, x0, st, false);
}

}  // namespace lh
