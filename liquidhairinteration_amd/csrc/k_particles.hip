// k_particles.hip -- particle side of the liquid step on gfx950.
//
//  * binning: particles are counting-sorted into half-cell bins (2N per axis,
//    x-major, z fastest) and physically permuted into bin order every step;
//    inside a bin they are ordered by particle id, so the order is a pure
//    function of positions and ids (deterministic).  Replaces the spatial hash
//    of NeighborSearch.h:59-131.
//  * ComputeLiquidPhi (LiquidModel.cpp:539-560) as a per-cell gather: a cell
//    takes the min over particles whose cell lies within +-ceil(r/dx) of it.
//    min is order-free, so the result equals the reference's bit for bit.
//  * Particle2Grid (LiquidModel.cpp:613-721) as a per-face gather with no float
//    atomics.  A face of component a is reached by particles with |gridPos -
//    faceCentre| < 1.5 on every axis, i.e. by exactly 6 half-cells per axis; the
//    6 half-cells along z are contiguous in bin order, so a face reads 36
//    contiguous particle runs.  Normalisation, gravity (LiquidModel.cpp:898-910)
//    and the optional tank acceleration are fused into the epilogue.
//  * Grid2Particle + AdvectParticle + GetCFL's max|v| (LiquidModel.cpp:723-935,
//    LiquidModel.h:49-55) fused into one per-particle kernel.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "lh_ctx.h"
#include "lh_math.h"

namespace lh {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(size_t n) {
  size_t b = div_up(n, kBlock);
  if (b > 65536 * 8) b = 65536 * 8;
  return (unsigned)(b ? b : 1);
}

// ------------------------------------------------------------------ binning
struct BinP {
  int H[3];
  float mn[3], dx[3];
  float inv[3];  // 1 / dx per axis when every dx is a power of two (then x / dx == x * inv exactly)
};

// The particle records of a buffer (one particle: R0 at r0[i], R1..R3 at r0[cap + 3 i .. + 2]); the id is
// R0.w, read 16 bytes apart by the kernels that need it beside or instead of the position.
__device__ __forceinline__ int rec_id(const float4* __restrict__ r0, size_t i) { return __float_as_int(r0[i].w); }

// key[i] = the particle's half-cell bin, arrival[i] = its arrival rank among the bin's particles
// (the atomic's order: any order serves, the rank kernel orders each bin by id), id[i] = its id
// (a dense copy of R0.w for the rank kernel's comparisons)
template <bool kPow2>
__global__ void k_bin_key(const float4* __restrict__ r0, int64_t P, BinP b, int* __restrict__ key,
                          int* __restrict__ arrival, int* __restrict__ id, int* __restrict__ count,
                          int* __restrict__ bflag, int* __restrict__ err) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < P; i += (int64_t)gridDim.x * blockDim.x) {
    const float4 R = r0[i];
    const float x[3] = {R.x, R.y, R.z};
    int h[3];
    bool bad = false;
    for (int a = 0; a < 3; ++a) {
      float g = kPow2 ? (x[a] - b.mn[a]) * b.inv[a] : (x[a] - b.mn[a]) / b.dx[a];
      float g2 = 2.0f * g;
      if (!(g2 >= 0.0f) || !(g2 < (float)b.H[a])) {
        bad = true;
        g2 = (g2 >= (float)b.H[a]) ? (float)(b.H[a] - 1) : 0.0f;
      }
      h[a] = (int)g2;
    }
    if (bad) atomicOr(err, 1);
    int k = (h[0] * b.H[1] + h[1]) * b.H[2] + h[2];
    key[i] = k;
    id[i] = __float_as_int(R.w);
    arrival[i] = atomicAdd(&count[k], 1);
    bflag[k / kScanBins] = 1;
  }
}

// slot_owner[start[key] + arrival] = i; every particle of a bin clears its count for the next
// binning (the scan has read the counts)
__global__ void k_bin_scatter(const int* __restrict__ key, const int* __restrict__ arrival, int64_t P,
                              const int* __restrict__ start, int* __restrict__ count, int* __restrict__ slot_owner) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < P; i += (int64_t)gridDim.x * blockDim.x) {
    const int k = key[i];
    slot_owner[start[k] + arrival[i]] = (int)i;
    count[k] = 0;
  }
}

// place particle i at start[key] + (number of particles in its bin with a smaller id)
__global__ void k_bin_rank(const int* __restrict__ id, int64_t P, const int* __restrict__ key,
                           const int* __restrict__ start, const int* __restrict__ slot_owner, int* __restrict__ order) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < P; i += (int64_t)gridDim.x * blockDim.x) {
    const int k = key[i];
    const int s0 = start[k], s1 = start[k + 1];
    int rank = 0;
    if (s1 - s0 > 1) {  // a bin of one particle needs no comparison
      const int myid = id[i];
      for (int s = s0; s < s1; ++s) {
        const int j = slot_owner[s];
        const int idj = id[j];
        rank += (idj < myid) ? 1 : 0;
      }
    }
    order[s0 + rank] = (int)i;
  }
}

// the records in bin order: destination d takes the record of source order[d], four 16-byte loads
// from two contiguous runs (R0; R1..R3) and four coalesced 16-byte stores
__global__ void k_bin_gather(const float4* __restrict__ src, float4* __restrict__ dst, int64_t cap, int64_t P,
                             const int* __restrict__ order) {
  const float4* __restrict__ ss = src + cap;
  float4* __restrict__ ds = dst + cap;
  for (int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; d < P; d += (int64_t)gridDim.x * blockDim.x) {
    // particle ids are 0..P-1, so the ranks of a bin are distinct and every slot is written; the
    // guard only keeps a corrupt order inside the buffer
    size_t i = (size_t)(unsigned)order[d];
    if (i >= (size_t)P) i = (size_t)d;
    const float4 a = src[i], b = ss[3 * i], c = ss[3 * i + 1], e = ss[3 * i + 2];
    dst[d] = a;
    ds[3 * (size_t)d] = b;
    ds[3 * (size_t)d + 1] = c;
    ds[3 * (size_t)d + 2] = e;
  }
}

// ------------------------------------------------------------------ bin scan
// bin_start = exclusive sum of bin_count over nbins + 1 entries.  The counts are zero outside
// the bins that hold particles, so the scan works on blocks of kScanBins bins (256 threads x 8)
// and reads the counts of the blocks k_bin_key flagged only: (1) k_scan_bsum sums each flagged
// block (an unflagged block sums to 0), (2) an exclusive scan of the block sums gives the block
// offsets, (3) k_scan_write writes every start: a flagged block's from its counts, an unflagged
// block's all equal to its offset; it clears the flags for the next binning.
static_assert(kScanBins == 256 * 8, "bin scan block layout");

__device__ __forceinline__ void scan_load8(const int* __restrict__ count, size_t base, size_t n, int v[8]) {
  if (base + 8 <= n) {
    const int4 a = *reinterpret_cast<const int4*>(count + base);
    const int4 b = *reinterpret_cast<const int4*>(count + base + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = base + j < n ? count[base + j] : 0;
  }
}
__device__ __forceinline__ void scan_store8(int* __restrict__ start, size_t base, size_t n, const int v[8]) {
  if (base + 8 <= n) {
    *reinterpret_cast<int4*>(start + base) = make_int4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<int4*>(start + base + 4) = make_int4(v[4], v[5], v[6], v[7]);
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (base + j < n) start[base + j] = v[j];
  }
}

__global__ __launch_bounds__(256) void k_scan_bsum(const int* __restrict__ count, const int* __restrict__ bflag,
                                                   int* __restrict__ bsum, size_t n) {
  __shared__ int s_w[4];
  const size_t b = blockIdx.x;
  if (bflag[b] == 0) {  // uniform over the block
    if (threadIdx.x == 0) bsum[b] = 0;
    return;
  }
  int v[8];
  scan_load8(count, b * kScanBins + threadIdx.x * 8, n, v);
  int t = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
  for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) bsum[b] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

__global__ __launch_bounds__(256) void k_scan_write(const int* __restrict__ count, int* __restrict__ bflag,
                                                    const int* __restrict__ boff, int* __restrict__ start, size_t n) {
  __shared__ int s_w[4];
  const size_t b = blockIdx.x;
  const size_t base = b * kScanBins + threadIdx.x * 8;
  const int off = boff[b];
  int v[8];
  if (bflag[b] == 0) {  // uniform over the block: no particle in it
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = off;
    scan_store8(start, base, n, v);
    return;
  }
  scan_load8(count, base, n, v);
  int t = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int c = v[j];
    v[j] = t;
    t += c;
  }
  // inclusive scan of the thread totals over the wave, then over the 4 waves
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int inc = t;
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(inc, o);
    if (lane >= o) inc += u;
  }
  if (lane == 63) s_w[wave] = inc;
  __syncthreads();  // every thread has read bflag[b] before this point
  int pre = off + inc - t;
  for (int w = 0; w < wave; ++w) pre += s_w[w];
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] += pre;
  scan_store8(start, base, n, v);
  if (threadIdx.x == 0) bflag[b] = 0;
}

// ------------------------------------------------------------------ P2G
struct P2GP {
  GridP g;
  int H[3];
  float gdt;         // gravity * dt, added to interior vy faces
  int tank_axis;
  float adt;         // tank acceleration * dt on interior faces of tank_axis
  float inv[3];      // 1 / dx per axis when every dx is a power of two (then x / dx == x * inv exactly)
};

// Tiles of 1 x 4 x 64 faces (or cells) per 256-thread workgroup: each
// wavefront owns 64 consecutive z of one (x, y) row, so its lanes walk the same
// bin rows and read overlapping, nearly contiguous particle windows.  A tile
// that no particle can reach writes its closed-form value and leaves; the test
// reads the two end offsets of every (hx, hy) bin row of the tile's reach.
constexpr int kTX = 1, kTY = 4, kTZ = 64;
// P2G tiles: MX x kPTY x kPTZ faces, one (y, z) row of MX x-faces per thread
constexpr int kPTY = 4, kPTZ = 64;

__device__ bool tile_has_particles(const int* __restrict__ start, const int lo[3], const int hi[3], int H1, int H2) {
  const int ny = hi[1] - lo[1] + 1;
  const int pairs = (hi[0] - lo[0] + 1) * ny;
  int cnt = 0;
  for (int pr = threadIdx.x; pr < pairs; pr += blockDim.x) {
    int hx = lo[0] + pr / ny, hy = lo[1] + pr % ny;
    size_t k0 = ((size_t)hx * H1 + hy) * H2;
    cnt += start[k0 + hi[2] + 1] - start[k0 + lo[2]];
  }
  return __syncthreads_or(cnt != 0) != 0;
}

// Blocks are dispatched round-robin over the 8 XCDs (block b on XCD b % 8), each XCD with
// its own L2.  Tiles adjacent in x read most of the same particles, so tiles are numbered x
// fastest and dealt to the XCDs in chunks of kXChunk consecutive x-tiles, which keeps a
// chunk's shared particles in one L2.  Each group of 8 consecutive chunks is spread over the
// 8 XCDs with a rotation hashed from the group: a fixed chunk-to-XCD map repeats with the
// rows whenever a row is a multiple of 8 chunks long, and then the XCDs whose x stripes
// hold the liquid get most of the work.  The grid is a multiple of 8 kXChunk blocks
// (tile_grid); blocks past the last tile return false.
constexpr int kXcds = 8, kXChunk = 4;
inline unsigned tile_grid(size_t ntiles) {
  return (unsigned)(div_up(ntiles, (size_t)kXcds * kXChunk) * kXcds * kXChunk);
}

template <int TX = kTX, int TY = kTY, int TZ = kTZ>
__device__ __forceinline__ bool tile_coords(int d0, int d1, int d2, int c[3], int c0[3]) {
  static_assert(TY * TZ <= kBlock, "one (y, z) per thread");
  const int ntx = (d0 + TX - 1) / TX, nty = (d1 + TY - 1) / TY, ntz = (d2 + TZ - 1) / TZ;
  const int x = (int)(blockIdx.x % kXcds), j = (int)(blockIdx.x / kXcds);
  const unsigned grp = (unsigned)(j / kXChunk);
  const int rot = (int)((grp * 2654435761u) >> 29);  // 0..7
  const int t = ((int)grp * kXcds + ((x + rot) & (kXcds - 1))) * kXChunk + j % kXChunk;
  if (t >= ntx * nty * ntz) return false;
  const int bx = t % ntx, by = (t / ntx) % nty, bz = t / (ntx * nty);
  c0[0] = bx * TX;
  c0[1] = by * TY;
  c0[2] = bz * TZ;
  c[0] = c0[0];
  c[1] = c0[1] + (int)threadIdx.x / TZ;
  c[2] = c0[2] + (int)threadIdx.x % TZ;
  return true;
}

// ------------------------------------------------------------------ x-marching P2G, three components
// The particle records, in bin order after the permutation: R0 = (x, y, z, id) in one array (the
// positions liquid phi reads, 16 bytes apart) and R1 = (C.col(0), v_y), R2 = (C.col(1), v_z),
// R3 = (C.col(2), v_x) in a second (48 bytes per particle).
//
// A 256-thread workgroup owns MX x positions x 4 y x 64 z of all three face grids: each thread
// one (y, z) column with its x-, y- and z-faces at x = c0 .. c0 + MX - 1 (those that exist: the
// x-face grid is (N0 + 1) x N1 x N2, and so on).  A face of component a is reached by the
// particles of the half-cells [2c - 3, 2c + 2] on axis a and [2c - 2, 2c + 3] on the other axes
// (every particle outside has weight exactly zero), so the union [2c - 3, 2c + 3] per axis holds
// the particles of all three faces of a column: the thread marches over the x half-slabs of the
// tile's reach in ascending order, reads each slab's 7 x 7 union run once, and adds every record
// to the (at most 3) consecutive faces of each component whose reach holds the slab, kept in
// sliding windows of registers (the x-faces' window runs half a cell ahead of the y- and
// z-faces').  The components share the record, its grid coordinates, the y and z weights of
// the two face offsets (cell centre and cell face) and the x weights of the y- and z-faces.  A
// face receives its nonzero contributions in ascending slabs and, within a slab, its y columns
// and z bins in order, and zero weights add nothing: the bin order the oracle sums in, so the
// result is bit-identical to one component per launch.  A thread reads 2 MX + 6 slabs for its MX
// x positions, so a longer row re-reads each slab fewer times but leaves fewer busy waves: 4 x
// positions suit liquid that fills a few percent of the box (the 256^3 collision, 0.4 particles
// per cell: 1.54 ms against 1.60 / 1.64 ms with 8 / 12), 12 a grid whose liquid fills its tiles
// (the 256^3 sloshing tank, 2.9 particles per cell: 5.62 ms against 6.87 ms with 4).
constexpr int kMXSparse = 4, kMXDense = 12;

// record index of the t-th particle of a thread's run of up to 7 bin columns: column j holds
// the t in [cum[j-1], cum[j]), at records t + off[j]
__device__ __forceinline__ int p2g_rec(int t, const int off[7], const int cum[7]) {
  int o = off[6];
#pragma unroll
  for (int j = 5; j >= 0; --j) o = (t < cum[j]) ? off[j] : o;
  return t + o;
}

// a thread's column: grid minimum, dx (1 / dx when kPow2), the face coordinates of its (y, z)
// on the cell-face (i) and cell-centre (h) offsets, in grid units and in world units
struct P2GCol {
  float mn0, mn1, mn2, dx0, dx1, dx2;
  float ccy_i, ccy_h, ccz_i, ccz_h;
  float fy_i, fy_h, fz_i, fz_h;
};

// one particle record into the weight and momentum sums of the first NX x-faces and the first
// NYZ y- and z-faces of the windows (LiquidModel.cpp:613-721): weight quad(gx - ccx) quad(gy -
// ccy) quad(gz - ccz) with grid coordinates (x - min) / dx (an exact power-of-two multiply when
// kPow2), momentum weight * (v_a + C.col(a) . (x - face)).  A record outside a face's reach has
// weight exactly +0 and adds exactly nothing: the weight sum stays as it is, and the momentum
// term is +-0 (velocities and affine terms are finite), which leaves a sum that started at +0
// unchanged (such a sum is never -0: x + -x and +0 + -0 are +0 in round-to-nearest).
template <bool kPow2, int NX, int NYZ>
__device__ __forceinline__ void p2g_acc(float4 R0, float4 R1, float4 R2, float4 R3, const P2GCol& k,
                                        const float ccx0[3], const float fx0[3], const float ccxh[3],
                                        const float fxh[3], float ws[3][3], float vs[3][3]) {
  const float gx = kPow2 ? (R0.x - k.mn0) * k.dx0 : (R0.x - k.mn0) / k.dx0;
  const float gy = kPow2 ? (R0.y - k.mn1) * k.dx1 : (R0.y - k.mn1) / k.dx1;
  const float gz = kPow2 ? (R0.z - k.mn2) * k.dx2 : (R0.z - k.mn2) / k.dx2;
  const float wyi = quad_w(gy - k.ccy_i), wyh = quad_w(gy - k.ccy_h);
  const float wzi = quad_w(gz - k.ccz_i), wzh = quad_w(gz - k.ccz_h);
  // the y and z part of each component's affine term (sum3(a, b, c) == a + (b + c))
  const float yz0 = R1.y * (R0.y - k.fy_h) + R1.z * (R0.z - k.fz_h);
  const float yz1 = R2.y * (R0.y - k.fy_i) + R2.z * (R0.z - k.fz_h);
  const float yz2 = R3.y * (R0.y - k.fy_h) + R3.z * (R0.z - k.fz_i);
#pragma unroll
  for (int f = 0; f < NX; ++f) {
    const float wgt = quad_w(gx - ccx0[f]) * wyh * wzh;
    const float dot = R1.x * (R0.x - fx0[f]) + yz0;
    ws[0][f] = ws[0][f] + wgt;
    vs[0][f] = vs[0][f] + wgt * (R3.w + dot);
  }
#pragma unroll
  for (int f = 0; f < NYZ; ++f) {
    const float wx = quad_w(gx - ccxh[f]);
    const float w1 = wx * wyi * wzh, w2 = wx * wyh * wzi;
    const float dxf = R0.x - fxh[f];
    const float d1 = R2.x * dxf + yz1, d2 = R3.x * dxf + yz2;
    ws[1][f] = ws[1][f] + w1;
    vs[1][f] = vs[1][f] + w1 * (R1.w + d1);
    ws[2][f] = ws[2][f] + w2;
    vs[2][f] = vs[2][f] + w2 * (R2.w + d2);
  }
}

// the records t in [0, total) of a thread's run of bin columns in one slab, two in flight
template <bool kPow2, int NX, int NYZ>
__device__ __forceinline__ void p2g_slab(const float4* __restrict__ pos, const float4* __restrict__ rest,
                                         const int off[7], const int cum[7], int total, const P2GCol& k,
                                         const float ccx0[3], const float fx0[3], const float ccxh[3],
                                         const float fxh[3], float ws[3][3], float vs[3][3]) {
  int t = 0;
  for (; t + 1 < total; t += 2) {
    const size_t i0 = (size_t)p2g_rec(t, off, cum), i1 = (size_t)p2g_rec(t + 1, off, cum);
    const float4 A0 = pos[i0], B0 = rest[3 * i0], C0 = rest[3 * i0 + 1], D0 = rest[3 * i0 + 2];
    const float4 A1 = pos[i1], B1 = rest[3 * i1], C1 = rest[3 * i1 + 1], D1 = rest[3 * i1 + 2];
    p2g_acc<kPow2, NX, NYZ>(A0, B0, C0, D0, k, ccx0, fx0, ccxh, fxh, ws, vs);
    p2g_acc<kPow2, NX, NYZ>(A1, B1, C1, D1, k, ccx0, fx0, ccxh, fxh, ws, vs);
  }
  if (t < total) {
    const size_t i0 = (size_t)p2g_rec(t, off, cum);
    p2g_acc<kPow2, NX, NYZ>(pos[i0], rest[3 * i0], rest[3 * i0 + 1], rest[3 * i0 + 2], k, ccx0, fx0, ccxh, fxh, ws,
                            vs);
  }
}

// Gather form of Particle2Grid for the three velocity components, each face summing the
// particles of the half-cell bins around it in bin order.  Within a slab the (up to 7) y
// columns of a thread's union window are concatenated into one run, so a wave's lanes wait for
// the longest run of 7 columns instead of the longest of each column.
template <bool kPow2, int MX>
__global__ void __launch_bounds__(256) k_p2g(float* __restrict__ vx, float* __restrict__ vy, float* __restrict__ vz,
                                              const float4* __restrict__ pos, const float4* __restrict__ rest,
                                              const int* __restrict__ start, P2GP q) {
  const int N0 = q.g.N[0], N1 = q.g.N[1], N2 = q.g.N[2];
  int c[3], c0[3];
  // tiles over the union of the three face grids, (N0 + 1) x (N1 + 1) x (N2 + 1)
  if (!tile_coords<MX, kPTY, kPTZ>(N0 + 1, N1 + 1, N2 + 1, c, c0)) return;
  const int T[3] = {MX, kPTY, kPTZ};
  int blo[3], bhi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    blo[k] = max(2 * c0[k] - 3, 0);
    bhi[k] = min(2 * (c0[k] + T[k] - 1) + 3, q.H[k] - 1);
  }
  const bool any = tile_has_particles(start, blo, bhi, q.H[1], q.H[2]);
  // the 64 lanes of a wave share their y (kPTZ == 64): y and everything derived from it is
  // wave-uniform, and lives in scalar registers
  static_assert(kPTZ == 64, "one y per wave");
  c[1] = __builtin_amdgcn_readfirstlane(c[1]);
  if (c[1] > N1 || c[2] > N2) return;
  // the faces of this column: x-faces (x <= N0, y < N1, z < N2), y-faces (x < N0, y <= N1,
  // z < N2), z-faces (x < N0, y < N1, z <= N2)
  const bool has[3] = {c[1] < N1 && c[2] < N2, c[1] <= N1 && c[2] < N2, c[1] < N1 && c[2] <= N2};
  const int xend0 = min(c0[0] + MX, N0 + 1), xend12 = min(c0[0] + MX, N0);
  float* const vout[3] = {vx, vy, vz};
  auto finish = [&](int a, int x, float ws, float vs) {
    if (!has[a]) return;
    float v = !is_small(ws) ? vs / ws : 0.0f;
    // interior ranges of ApplyGravity (a == 1) and of the tank acceleration
    if (a == 1 || a == q.tank_axis) {
      const int cc[3] = {x, c[1], c[2]};
      bool in = true;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int h = (k == a) ? q.g.N[k] - 1 : q.g.N[k] - 2;
        if (cc[k] < 1 || cc[k] > h) in = false;
      }
      if (in) {
        if (a == 1) v += q.gdt;
        if (a == q.tank_axis) v += q.adt;
      }
    }
    const int d1 = N1 + (a == 1), d2 = N2 + (a == 2);
    vout[a][((size_t)x * d1 + c[1]) * d2 + c[2]] = v;
  };
  if (!any) {
    for (int x = c0[0]; x < xend0; ++x) finish(0, x, 0.f, 0.f);
    for (int x = c0[0]; x < xend12; ++x) {
      finish(1, x, 0.f, 0.f);
      finish(2, x, 0.f, 0.f);
    }
    return;
  }
  P2GCol k;
  k.mn0 = q.g.mn[0];
  k.mn1 = q.g.mn[1];
  k.mn2 = q.g.mn[2];
  k.dx0 = kPow2 ? q.inv[0] : q.g.dx[0];
  k.dx1 = kPow2 ? q.inv[1] : q.g.dx[1];
  k.dx2 = kPow2 ? q.inv[2] : q.g.dx[2];
  k.ccy_i = (float)c[1];
  k.ccy_h = (float)c[1] + 0.5f;
  k.ccz_i = (float)c[2];
  k.ccz_h = (float)c[2] + 0.5f;
  k.fy_i = k.ccy_i * q.g.dx[1] + k.mn1;
  k.fy_h = k.ccy_h * q.g.dx[1] + k.mn1;
  k.fz_i = k.ccz_i * q.g.dx[2] + k.mn2;
  k.fz_h = k.ccz_h * q.g.dx[2] + k.mn2;
  // the union window in y and z
  const int ylo = max(2 * c[1] - 3, 0), yhi = min(2 * c[1] + 3, q.H[1] - 1);
  const int zlo = max(2 * c[2] - 3, 0), zhi = min(2 * c[2] + 3, q.H[2] - 1);
  // x reach of an x-face (0) and of a y- or z-face (1)
  auto xlo = [&](int m, int x) { return max(m == 0 ? 2 * x - 3 : 2 * x - 2, 0); };
  auto xhi = [&](int m, int x) { return min((m == 0 ? 2 * x - 3 : 2 * x - 2) + 5, q.H[0] - 1); };
  float ws[3][3], vs[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int f = 0; f < 3; ++f) ws[a][f] = vs[a][f] = 0.f;
  int cw0 = c0[0], cw12 = c0[0];  // first face of each window
  for (int hx = blo[0]; hx <= bhi[0]; ++hx) {
    // the faces whose reach holds slab hx: a prefix of each window (the reaches of consecutive
    // faces are ordered and at most 3 overlap)
    float ccx0[3], fx0[3], ccxh[3], fxh[3];
    int nx = 0, nyz = 0;
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const int x0 = cw0 + f, x1 = cw12 + f;
      nx += (x0 < xend0 && xlo(0, x0) <= hx && hx <= xhi(0, x0)) ? 1 : 0;
      nyz += (x1 < xend12 && xlo(1, x1) <= hx && hx <= xhi(1, x1)) ? 1 : 0;
      ccx0[f] = (float)x0;
      fx0[f] = ccx0[f] * q.g.dx[0] + k.mn0;
      ccxh[f] = (float)x1 + 0.5f;
      fxh[f] = ccxh[f] * q.g.dx[0] + k.mn0;
    }
    if (nx + nyz > 0) {
      int off[7], cum[7], run = 0;
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        const int hy = ylo + j;
        int s0 = 0, n = 0;
        if (hy <= yhi) {
          const size_t k0 = ((size_t)hx * q.H[1] + hy) * q.H[2];
          s0 = start[k0 + zlo];
          n = start[k0 + zhi + 1] - s0;
        }
        off[j] = s0 - run;
        run += n;
        cum[j] = run;
      }
#define LH_P2G_CASE(A, B) \
  case A * 4 + B:         \
    p2g_slab<kPow2, A, B>(pos, rest, off, cum, run, k, ccx0, fx0, ccxh, fxh, ws, vs); \
    break;
      switch (nx * 4 + nyz) {
        LH_P2G_CASE(0, 1) LH_P2G_CASE(0, 2) LH_P2G_CASE(0, 3)
        LH_P2G_CASE(1, 0) LH_P2G_CASE(1, 1) LH_P2G_CASE(1, 2) LH_P2G_CASE(1, 3)
        LH_P2G_CASE(2, 0) LH_P2G_CASE(2, 1) LH_P2G_CASE(2, 2) LH_P2G_CASE(2, 3)
        LH_P2G_CASE(3, 0) LH_P2G_CASE(3, 1) LH_P2G_CASE(3, 2) LH_P2G_CASE(3, 3)
        default: break;
      }
#undef LH_P2G_CASE
    }
    // faces whose last slab this was
    while (cw0 < xend0 && xhi(0, cw0) <= hx) {
      finish(0, cw0, ws[0][0], vs[0][0]);
      ws[0][0] = ws[0][1];
      vs[0][0] = vs[0][1];
      ws[0][1] = ws[0][2];
      vs[0][1] = vs[0][2];
      ws[0][2] = 0.f;
      vs[0][2] = 0.f;
      ++cw0;
    }
    while (cw12 < xend12 && xhi(1, cw12) <= hx) {
      finish(1, cw12, ws[1][0], vs[1][0]);
      finish(2, cw12, ws[2][0], vs[2][0]);
#pragma unroll
      for (int a = 1; a < 3; ++a) {
        ws[a][0] = ws[a][1];
        vs[a][0] = vs[a][1];
        ws[a][1] = ws[a][2];
        vs[a][1] = vs[a][2];
        ws[a][2] = 0.f;
        vs[a][2] = 0.f;
      }
      ++cw12;
    }
  }
}

// ------------------------------------------------------------------ P2G with shared records
// The gather above has every lane evaluate every record of its union window, so the x and y
// parts of one record's weights and momenta are evaluated by each of the ~3.5 lanes (z) of a
// wave that reach it, although they depend only on the record and the wave's y row and x
// positions.  Here a wave (one y row of 64 z columns at the tile's x positions) evaluates them
// once per record: per x slab and y column it takes the records of the whole wave's z window
// (half-cell bins [2 z0 - 3, 2 (z0 + 63) + 3]) up to kShareCap at a time, one or more per lane,
// and leaves per record in LDS 24 floats: per component and active x face the product w_x w_y,
// the y parts of the affine terms, the velocities, x, z and the x and z rows of C (the gather's
// operations, in its order).  Then each lane runs the gather's loop over its own window's records
// in bin order, reading them from LDS and finishing each with its x parts, its z weight and its z
// part: the same operations on the same values in the same order, so the result is bit-identical
// to k_p2g.  A lane's window holds ~7 of the ~134 records of a column (one particle per
// half-cell), and a chunk holding most of a column keeps the lanes busy.  Storing the x parts
// instead (28 floats) costs occupancy: 10 waves per CU instead of 12, 1 261 against 1 220 us on
// the 256^3 collision.
constexpr int kShareCap = 128;  // records per chunk (two rounds of the wave's lanes)
constexpr int kShareF = 24;     // floats per record in LDS
struct P2GShared {
  // per record: R0.z, C.col(0).z, C.col(1).z, C.col(2).z | Y0, Y1, Y2, v_x | v_y, v_z, R0.x, C.col(0).x |
  // C.col(1).x, C.col(2).x, A0[0..1] | A0[2], A1[0..2] | A2[0..2], -
  float r[kShareCap][kShareF];
};

// LDS hand-over between the lanes of one wave: the wave's LDS operations complete in order, so
// keeping the compiler from moving them across this point is enough
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// a record's terms for the wave's faces into LDS slot `slot`: A = w_x w_y per component and active x
// face, Y = C.col(a).y (y - face y), the velocities, x, z and the x and z rows of C
template <bool kPow2, int NX, int NYZ>
__device__ __forceinline__ void p2g_share_record(float4 R0, float4 R1, float4 R2, float4 R3, const P2GCol& k,
                                                 const float ccx0[3], const float fx0[3], const float ccxh[3],
                                                 const float fxh[3], P2GShared& sh, int slot) {
  const float gx = kPow2 ? (R0.x - k.mn0) * k.dx0 : (R0.x - k.mn0) / k.dx0;
  const float gy = kPow2 ? (R0.y - k.mn1) * k.dx1 : (R0.y - k.mn1) / k.dx1;
  const float wyi = quad_w(gy - k.ccy_i), wyh = quad_w(gy - k.ccy_h);
  float o[kShareF];
#pragma unroll
  for (int f = 0; f < kShareF; ++f) o[f] = 0.f;
  o[0] = R0.z;
  o[1] = R1.z;
  o[2] = R2.z;
  o[3] = R3.z;
  o[4] = R1.y * (R0.y - k.fy_h);
  o[5] = R2.y * (R0.y - k.fy_i);
  o[6] = R3.y * (R0.y - k.fy_h);
  o[7] = R3.w;
  o[8] = R1.w;
  o[9] = R2.w;
  o[10] = R0.x;
  o[11] = R1.x;
  o[12] = R2.x;
  o[13] = R3.x;
#pragma unroll
  for (int f = 0; f < NX; ++f) o[14 + f] = quad_w(gx - ccx0[f]) * wyh;
#pragma unroll
  for (int f = 0; f < NYZ; ++f) {
    const float wx = quad_w(gx - ccxh[f]);
    o[17 + f] = wx * wyi;
    o[20 + f] = wx * wyh;
  }
  float4* d = reinterpret_cast<float4*>(sh.r[slot]);
#pragma unroll
  for (int q = 0; q < kShareF / 4; ++q) d[q] = make_float4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
}

__device__ __forceinline__ void p2g_share_load(const P2GShared& sh, int s, float4 v[kShareF / 4]) {
  const float4* p = reinterpret_cast<const float4*>(sh.r[s]);
#pragma unroll
  for (int q = 0; q < kShareF / 4; ++q) v[q] = p[q];
}

// one LDS slot's terms into this lane's faces (p2g_acc with the x and y terms read back)
template <bool kPow2, int NX, int NYZ>
__device__ __forceinline__ void p2g_share_add(const float4 v[kShareF / 4], const P2GCol& k, const float fx0[3],
                                              const float fxh[3], float ws[3][3], float vs[3][3]) {
  float o[kShareF];
#pragma unroll
  for (int q = 0; q < kShareF / 4; ++q) {
    o[4 * q] = v[q].x;
    o[4 * q + 1] = v[q].y;
    o[4 * q + 2] = v[q].z;
    o[4 * q + 3] = v[q].w;
  }
  // per component and active x face: the product w_x w_y and the x part of the affine term, formed
  // here from the record's x and C.col(a).x (the same operations as in p2g_acc)
  float A0[3], E0[3], A1[3], E1[3], A2[3], E2[3];
#pragma unroll
  for (int f = 0; f < NX; ++f) {
    A0[f] = o[14 + f];
    E0[f] = o[11] * (o[10] - fx0[f]);
  }
#pragma unroll
  for (int f = 0; f < NYZ; ++f) {
    A1[f] = o[17 + f];
    A2[f] = o[20 + f];
    const float dxf = o[10] - fxh[f];
    E1[f] = o[12] * dxf;
    E2[f] = o[13] * dxf;
  }
  const float z = o[0];
  const float gz = kPow2 ? (z - k.mn2) * k.dx2 : (z - k.mn2) / k.dx2;
  const float wzh = quad_w(gz - k.ccz_h);
  const float th = z - k.fz_h;
  const float yz0 = o[4] + o[1] * th, yz1 = o[5] + o[2] * th;
#pragma unroll
  for (int f = 0; f < NX; ++f) {
    const float w = A0[f] * wzh;
    ws[0][f] = ws[0][f] + w;
    vs[0][f] = vs[0][f] + w * (o[7] + (E0[f] + yz0));
  }
  if (NYZ > 0) {
    const float wzi = quad_w(gz - k.ccz_i);
    const float yz2 = o[6] + o[3] * (z - k.fz_i);
#pragma unroll
    for (int f = 0; f < NYZ; ++f) {
      const float w1 = A1[f] * wzh, w2 = A2[f] * wzi;
      ws[1][f] = ws[1][f] + w1;
      vs[1][f] = vs[1][f] + w1 * (o[8] + (E1[f] + yz1));
      ws[2][f] = ws[2][f] + w2;
      vs[2][f] = vs[2][f] + w2 * (o[9] + (E2[f] + yz2));
    }
  }
}

// one y column of a wave's slab: the records [s0, s0 + n) of the wave's window, of which this
// lane's window is [s0 + a, s0 + b), kShareCap at a time
template <bool kPow2, int NX, int NYZ>
__device__ __forceinline__ void p2g_share_column(const float4* __restrict__ pos, const float4* __restrict__ rest,
                                                 int s0, int n, int a, int b, const P2GCol& k, const float ccx0[3],
                                                 const float fx0[3], const float ccxh[3], const float fxh[3],
                                                 P2GShared& sh, int lane, float ws[3][3], float vs[3][3]) {
  for (int cb = 0; cb < n; cb += kShareCap) {
    const int cn = min(n - cb, kShareCap);
#pragma unroll
    for (int r = 0; r < (kShareCap + 63) / 64; ++r) {
      const int t = r * 64 + lane;
      if (t < cn) {
        const size_t i = (size_t)(s0 + cb + t);
        p2g_share_record<kPow2, NX, NYZ>(pos[i], rest[3 * i], rest[3 * i + 1], rest[3 * i + 2], k, ccx0, fx0, ccxh,
                                         fxh, sh, t);
      }
    }
    wave_lds_sync();
    const int lo = max(a, cb) - cb, hi = min(b, cb + cn) - cb;
    for (int t = lo; t < hi; ++t) {
      float4 v[kShareF / 4];
      p2g_share_load(sh, t, v);
      p2g_share_add<kPow2, NX, NYZ>(v, k, fx0, fxh, ws, vs);
    }
    wave_lds_sync();
  }
}

// x positions per tile, as for the gather: a longer row evaluates each record for fewer tiles
// but leaves fewer busy waves (same-box sweeps with 2-wave workgroups: the 256^3 collision, 0.4
// particles per cell, 3 / 4 / 6 / 8 positions 1288 / 1262 / 1300 / 1517 us; the 256^3 sloshing
// tank, 2.9 per cell, 8 / 10 / 12 / 16 positions 4846 / 4875 / 4780 / 4712 us, 12 the fastest
// step).  The macros let a build sweep them again.
#ifndef LH_P2G_SHARED_SPARSE_MX
#define LH_P2G_SHARED_SPARSE_MX 4
#endif
#ifndef LH_P2G_SHARED_DENSE_MX
#define LH_P2G_SHARED_DENSE_MX 12
#endif
constexpr int kMXSharedSparse = LH_P2G_SHARED_SPARSE_MX, kMXSharedDense = LH_P2G_SHARED_DENSE_MX;

// y rows (waves) per workgroup of the shared-record P2G.  Each wave holds 12 KB of LDS, so the
// workgroup size sets how many fit a CU: with 2 waves 6 workgroups, 12 waves (the VGPR limit at
// ~130 registers); with 4 waves 8.  Same box: 4 / 2 / 1 waves 1435 / 1271 / 1283 us on the
// collision, 5592 / 4870 / 4799 us on the sloshing tank (with 28-float records).
#ifndef LH_P2G_SHARED_TY
#define LH_P2G_SHARED_TY 2
#endif
constexpr int kSTY = LH_P2G_SHARED_TY;

template <bool kPow2, int MX>
__global__ void __launch_bounds__(64 * kSTY) k_p2g_shared(float* __restrict__ vx, float* __restrict__ vy,
                                                     float* __restrict__ vz, const float4* __restrict__ pos,
                                                     const float4* __restrict__ rest, const int* __restrict__ start,
                                                     P2GP q) {
  __shared__ P2GShared shm[kSTY];
  const int N0 = q.g.N[0], N1 = q.g.N[1], N2 = q.g.N[2];
  int c[3], c0[3];
  if (!tile_coords<MX, kSTY, kPTZ>(N0 + 1, N1 + 1, N2 + 1, c, c0)) return;
  const int T[3] = {MX, kSTY, kPTZ};
  int blo[3], bhi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    blo[k] = max(2 * c0[k] - 3, 0);
    bhi[k] = min(2 * (c0[k] + T[k] - 1) + 3, q.H[k] - 1);
  }
  const bool any = tile_has_particles(start, blo, bhi, q.H[1], q.H[2]);
  static_assert(kPTZ == 64, "one y per wave");
  c[1] = __builtin_amdgcn_readfirstlane(c[1]);
  // a wave past the last y row leaves; a lane past the last z stays (it evaluates records for
  // the others) and owns no face
  if (c[1] > N1) return;
  const bool zok = c[2] <= N2;
  const bool has[3] = {c[1] < N1 && c[2] < N2, c[1] <= N1 && c[2] < N2, c[1] < N1 && c[2] <= N2};
  const int xend0 = min(c0[0] + MX, N0 + 1), xend12 = min(c0[0] + MX, N0);
  float* const vout[3] = {vx, vy, vz};
  auto finish = [&](int a, int x, float ws, float vs) {
    if (!has[a]) return;
    float v = !is_small(ws) ? vs / ws : 0.0f;
    if (a == 1 || a == q.tank_axis) {
      const int cc[3] = {x, c[1], c[2]};
      bool in = true;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int h = (k == a) ? q.g.N[k] - 1 : q.g.N[k] - 2;
        if (cc[k] < 1 || cc[k] > h) in = false;
      }
      if (in) {
        if (a == 1) v += q.gdt;
        if (a == q.tank_axis) v += q.adt;
      }
    }
    const int d1 = N1 + (a == 1), d2 = N2 + (a == 2);
    vout[a][((size_t)x * d1 + c[1]) * d2 + c[2]] = v;
  };
  if (!any) {
    for (int x = c0[0]; x < xend0; ++x) finish(0, x, 0.f, 0.f);
    for (int x = c0[0]; x < xend12; ++x) {
      finish(1, x, 0.f, 0.f);
      finish(2, x, 0.f, 0.f);
    }
    return;
  }
  const int lane = (int)threadIdx.x & 63;
  P2GShared& sh = shm[threadIdx.x >> 6];
  P2GCol k;
  k.mn0 = q.g.mn[0];
  k.mn1 = q.g.mn[1];
  k.mn2 = q.g.mn[2];
  k.dx0 = kPow2 ? q.inv[0] : q.g.dx[0];
  k.dx1 = kPow2 ? q.inv[1] : q.g.dx[1];
  k.dx2 = kPow2 ? q.inv[2] : q.g.dx[2];
  k.ccy_i = (float)c[1];
  k.ccy_h = (float)c[1] + 0.5f;
  k.ccz_i = (float)c[2];
  k.ccz_h = (float)c[2] + 0.5f;
  k.fy_i = k.ccy_i * q.g.dx[1] + k.mn1;
  k.fy_h = k.ccy_h * q.g.dx[1] + k.mn1;
  k.fz_i = k.ccz_i * q.g.dx[2] + k.mn2;
  k.fz_h = k.ccz_h * q.g.dx[2] + k.mn2;
  const int H2 = q.H[2];
  const int ylo = max(2 * c[1] - 3, 0), yhi = min(2 * c[1] + 3, q.H[1] - 1);
  // the wave's union window in z, and this lane's (inside it: a lane past the last face has none)
  const int wzlo = max(2 * c0[2] - 3, 0), wzhi = min(2 * (c0[2] + kPTZ - 1) + 3, H2 - 1);
  const int zlo = max(2 * c[2] - 3, 0), zhi = min(2 * c[2] + 3, H2 - 1);
  auto xlo = [&](int m, int x) { return max(m == 0 ? 2 * x - 3 : 2 * x - 2, 0); };
  auto xhi = [&](int m, int x) { return min((m == 0 ? 2 * x - 3 : 2 * x - 2) + 5, q.H[0] - 1); };
  float ws[3][3], vs[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int f = 0; f < 3; ++f) ws[a][f] = vs[a][f] = 0.f;
  int cw0 = c0[0], cw12 = c0[0];
  for (int hx = blo[0]; hx <= bhi[0]; ++hx) {
    float ccx0[3], fx0[3], ccxh[3], fxh[3];
    int nx = 0, nyz = 0;
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const int x0 = cw0 + f, x1 = cw12 + f;
      nx += (x0 < xend0 && xlo(0, x0) <= hx && hx <= xhi(0, x0)) ? 1 : 0;
      nyz += (x1 < xend12 && xlo(1, x1) <= hx && hx <= xhi(1, x1)) ? 1 : 0;
      ccx0[f] = (float)x0;
      fx0[f] = ccx0[f] * q.g.dx[0] + k.mn0;
      ccxh[f] = (float)x1 + 0.5f;
      fxh[f] = ccxh[f] * q.g.dx[0] + k.mn0;
    }
    if (nx + nyz > 0) {
      for (int hy = ylo; hy <= yhi; ++hy) {
        const size_t k0 = ((size_t)hx * q.H[1] + hy) * H2;
        const int s0 = start[k0 + wzlo];
        const int n = start[k0 + wzhi + 1] - s0;
        if (n == 0) continue;
        int a = 0, b = 0;
        if (zok) {
          a = start[k0 + zlo] - s0;
          b = start[k0 + zhi + 1] - s0;
        }
#define LH_P2G_CASE(A, B)                                                                                \
  case A * 4 + B:                                                                                        \
    p2g_share_column<kPow2, A, B>(pos, rest, s0, n, a, b, k, ccx0, fx0, ccxh, fxh, sh, lane, ws, vs); \
    break;
        switch (nx * 4 + nyz) {
          LH_P2G_CASE(0, 1) LH_P2G_CASE(0, 2) LH_P2G_CASE(0, 3)
          LH_P2G_CASE(1, 0) LH_P2G_CASE(1, 1) LH_P2G_CASE(1, 2) LH_P2G_CASE(1, 3)
          LH_P2G_CASE(2, 0) LH_P2G_CASE(2, 1) LH_P2G_CASE(2, 2) LH_P2G_CASE(2, 3)
          LH_P2G_CASE(3, 0) LH_P2G_CASE(3, 1) LH_P2G_CASE(3, 2) LH_P2G_CASE(3, 3)
          default: break;
        }
      }
#undef LH_P2G_CASE
    }
    while (cw0 < xend0 && xhi(0, cw0) <= hx) {
      finish(0, cw0, ws[0][0], vs[0][0]);
      ws[0][0] = ws[0][1];
      vs[0][0] = vs[0][1];
      ws[0][1] = ws[0][2];
      vs[0][1] = vs[0][2];
      ws[0][2] = 0.f;
      vs[0][2] = 0.f;
      ++cw0;
    }
    while (cw12 < xend12 && xhi(1, cw12) <= hx) {
      finish(1, cw12, ws[1][0], vs[1][0]);
      finish(2, cw12, ws[2][0], vs[2][0]);
#pragma unroll
      for (int a = 1; a < 3; ++a) {
        ws[a][0] = ws[a][1];
        vs[a][0] = vs[a][1];
        ws[a][1] = ws[a][2];
        vs[a][1] = vs[a][2];
        ws[a][2] = 0.f;
        vs[a][2] = 0.f;
      }
      ++cw12;
    }
  }
}

// ------------------------------------------------------------------ liquid phi
// min over particles in cells c +- dW of |centre - x_p| - r.  The min is
// order-free, so bin rows are visited nearest first (the cell's own 2 x 2 rows)
// and a row is skipped when its distance lower bound already exceeds the
// current min by a margin far above the float error of the distance: skipped
// particles cannot change the result.
__global__ void __launch_bounds__(256) k_liquid_phi(float* __restrict__ phi, const float4* __restrict__ rec,
                                                     int64_t cap, const int* __restrict__ start, GridP g, int H0,
                                                     int H1, int H2, int dW0, int dW1, int dW2, float radius,
                                                     float init) {
  const int N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
  int c[3], c0[3];
  if (!tile_coords(N0, N1, N2, c, c0)) return;
  const int T[3] = {kTX, kTY, kTZ};
  const int dW[3] = {dW0, dW1, dW2};
  const int H[3] = {H0, H1, H2};
  int blo[3], bhi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    blo[k] = max(0, 2 * (c0[k] - dW[k]));
    bhi[k] = min(H[k] - 1, 2 * (c0[k] + T[k] - 1 + dW[k]) + 1);
  }
  const bool any = tile_has_particles(start, blo, bhi, H1, H2);
  const int xi = c[0], yi = c[1], zi = c[2];
  if (xi >= N0 || yi >= N1 || zi >= N2) return;
  float best = init;
  if (any) {
    const float cpx = ((float)xi + 0.5f) * g.dx[0] + g.mn[0];
    const float cpy = ((float)yi + 0.5f) * g.dx[1] + g.mn[1];
    const float cpz = ((float)zi + 0.5f) * g.dx[2] + g.mn[2];
    const int hx0 = max(0, 2 * (xi - dW0)), hx1 = min(H0 - 1, 2 * (xi + dW0) + 1);
    const int hy0 = max(0, 2 * (yi - dW1)), hy1 = min(H1 - 1, 2 * (yi + dW1) + 1);
    const int hz0 = max(0, 2 * (zi - dW2)), hz1 = min(H2 - 1, 2 * (zi + dW2) + 1);
    // bound slack: 1e-3 dx plus the float error of world coordinates
    const float margin = 1e-3f * g.dx[0] + 1e-6f * (fabsf(cpx) + fabsf(cpy) + fabsf(cpz));
    auto phi_of = [&](int s) {
      const float4 A = rec[s];  // x y z of the bin-ordered particle
      return norm3(cpx - A.x, cpy - A.y, cpz - A.z) - radius;
    };
    auto run = [&](int s, int s1) {
      for (; s + 3 < s1; s += 4) {  // four loads in flight; the min is exact in any order
        const float p0 = phi_of(s), p1 = phi_of(s + 1), p2 = phi_of(s + 2), p3 = phi_of(s + 3);
        const float m01 = (p0 < p1) ? p0 : p1, m23 = (p2 < p3) ? p2 : p3;
        const float m = (m01 < m23) ? m01 : m23;
        best = (best < m) ? best : m;
      }
      for (; s < s1; ++s) {
        const float ph = phi_of(s);
        best = (best < ph) ? best : ph;
      }
    };
    // the cell's own 8 half-cells first (a tight bound early), then the rest of its own 2 x 2
    // half-cell rows, cut in z as below; starts loaded together
    {
      int S0[4], S1[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const size_t k0 = ((size_t)(2 * xi + (j >> 1)) * H1 + 2 * yi + (j & 1)) * H2;
        S0[j] = start[k0 + 2 * zi];
        S1[j] = start[k0 + 2 * zi + 2];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) run(S0[j], S1[j]);
      const float lim = best + margin + radius;
      const int w2 = (int)floorf(2.f * lim / g.dx[2]);
      const int lo = max(hz0, 2 * zi - w2), hi = min(hz1, 2 * zi + 1 + w2);
      int A0[4], A1[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const size_t k0 = ((size_t)(2 * xi + (j >> 1)) * H1 + 2 * yi + (j & 1)) * H2;
        A0[j] = start[k0 + lo];
        S0[j] = start[k0 + 2 * zi];
        A1[j] = start[k0 + 2 * zi + 2];
        S1[j] = start[k0 + hi + 1];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        run(A0[j], S0[j]);
        run(A1[j], S1[j]);
      }
    }
    // the other rows, pruned by the distance bound against the running minimum (it only
    // decreases, so a row pruned once stays prunable): a particle lowers the minimum only within
    // lim of the cell centre, so half-slabs and rows farther than lim in x, (x, y) are skipped and
    // a row's z half-cells are cut to those within sqrt(lim^2 - gx^2 - gy^2) in z.  Per x
    // half-cell, the bounds of up to 10 rows are formed first, and their starts loaded together
    // only when one of them is in reach (for a cell inside the liquid none is).  The half-slabs
    // are visited from the cell's own two outwards, so the nearest particles tighten the bound
    // first.
    constexpr int kRowBatch = 10;
    for (int m = 0; m < 2 * (hx1 - hx0 + 1); ++m) {
      // centre out: 2xi, 2xi + 1, 2xi - 1, 2xi + 2, ...
      const int hx = (m & 1) ? 2 * xi + 1 + (m >> 1) : 2 * xi - (m >> 1);
      if (hx < hx0 || hx > hx1) continue;
      const float gx = fmaxf(0.f, fmaxf(0.5f * hx - (xi + 0.5f), (xi + 0.5f) - 0.5f * (hx + 1))) * g.dx[0];
      const float lim = best + margin + radius;
      if (gx > lim) continue;
      for (int hyb = hy0; hyb <= hy1; hyb += kRowBatch) {
        float rem[kRowBatch];
        bool use[kRowBatch];
        bool any_row = false;
#pragma unroll
        for (int j = 0; j < kRowBatch; ++j) {
          const int hy = hyb + j;
          const float gy = fmaxf(0.f, fmaxf(0.5f * hy - (yi + 0.5f), (yi + 0.5f) - 0.5f * (hy + 1))) * g.dx[1];
          const bool own = (hx >> 1) == xi && (hy >> 1) == yi;
          rem[j] = lim * lim - gx * gx - gy * gy;
          use[j] = hy <= hy1 && !own && rem[j] >= 0.f;
          any_row = any_row || use[j];
        }
        if (!any_row) continue;
        int S0[kRowBatch], S1[kRowBatch];
#pragma unroll
        for (int j = 0; j < kRowBatch; ++j) {
          const int hy = hyb + j;
          const int w2 = (int)floorf(2.f * sqrtf(fmaxf(rem[j], 0.f)) / g.dx[2]);
          const int lo = max(hz0, 2 * zi - w2), hi = min(hz1, 2 * zi + 1 + w2);
          const size_t k0 = ((size_t)hx * H1 + min(hy, hy1)) * H2;
          S0[j] = use[j] ? start[k0 + lo] : 0;
          S1[j] = use[j] ? start[k0 + hi + 1] : 0;
        }
#pragma unroll
        for (int j = 0; j < kRowBatch; ++j) run(S0[j], S1[j]);
      }
    }
  }
  phi[((size_t)xi * N1 + yi) * N2 + zi] = best;
}

// ------------------------------------------------------------------ G2P + advect
struct GP {
  GridP g;
  float dt;
  float inv[3];  // 1 / dx per axis when every dx is a power of two (then x / dx == x * inv exactly)
};

// grid coordinate (x - min) / dx on axis k: with kPow2 every dx is a power of two and the
// division is the exact multiply by 1 / dx
template <bool kPow2>
__device__ __forceinline__ float to_grid(const GridP& g, const float* inv, int k, float x) {
  return kPow2 ? (x - g.mn[k]) * inv[k] : (x - g.mn[k]) / g.dx[k];
}

template <bool kPow2>
__device__ __forceinline__ float solid_phi_in(const float* __restrict__ sp, const GridP& g, const float* inv, float x,
                                              float y, float z) {
  float lx = to_grid<kPow2>(g, inv, 0, x);
  float ly = to_grid<kPow2>(g, inv, 1, y);
  float lz = to_grid<kPow2>(g, inv, 2, z);
  int xi, yi, zi;
  float fx, fy, fz;
  barycentric(lx, 0, g.N[0] + 1, xi, fx);
  barycentric(ly, 0, g.N[1] + 1, yi, fy);
  barycentric(lz, 0, g.N[2] + 1, zi, fz);
  const int n1 = g.N[1] + 1, n2 = g.N[2] + 1;
  auto S = [&](int a, int b, int c) { return sp[((size_t)a * n1 + b) * n2 + c]; };
  return trilerp(S(xi, yi, zi), S(xi + 1, yi, zi), S(xi, yi + 1, zi), S(xi + 1, yi + 1, zi), S(xi, yi, zi + 1),
                 S(xi + 1, yi, zi + 1), S(xi, yi + 1, zi + 1), S(xi + 1, yi + 1, zi + 1), fx, fy, fz);
}
__device__ __forceinline__ bool in_grid(const GridP& g, const float mx[3], float x, float y, float z) {
  return g.mn[0] <= x && g.mn[1] <= y && g.mn[2] <= z && x <= mx[0] && y <= mx[1] && z <= mx[2];
}
// central-difference normalised gradient at a point whose +-eps samples are all
// inside the grid (Grid.cpp:183-191)
__device__ void grad_from_samples(float left, float right, float bottom, float top, float front, float back,
                                  float out[3]) {
  const float eps = 1e-4f;
  const float two_eps = (float)(2.0 * (double)eps);
  float a = (right - left) / two_eps, b = (top - bottom) / two_eps, c = (front - back) / two_eps;
  float n2 = a * a + (b * b + c * c);
  if (n2 > 0.f) {
    float n = sqrtf(n2);
    a = a / n;
    b = b / n;
    c = c / n;
  }
  out[0] = a;
  out[1] = b;
  out[2] = c;
}
template <bool kPow2>
__device__ void grad_inner(const float* sp, const GridP& g, const float* inv, float x, float y, float z,
                           float out[3]) {
  const float eps = 1e-4f;
  grad_from_samples(solid_phi_in<kPow2>(sp, g, inv, x - eps, y, z), solid_phi_in<kPow2>(sp, g, inv, x + eps, y, z),
                    solid_phi_in<kPow2>(sp, g, inv, x, y - eps, z), solid_phi_in<kPow2>(sp, g, inv, x, y + eps, z),
                    solid_phi_in<kPow2>(sp, g, inv, x, y, z + eps), solid_phi_in<kPow2>(sp, g, inv, x, y, z - eps),
                    out);
}
// Grid<3>::GetSolidPhi (Grid.cpp:128-155) with the out-of-grid extrapolation
template <bool kPow2 = false>
__device__ float solid_phi_any(const float* sp, const GridP& g, const float mx[3], float x, float y, float z,
                               const float* inv = nullptr) {
  if (in_grid(g, mx, x, y, z)) return solid_phi_in<kPow2>(sp, g, inv, x, y, z);
  float vx = clampf(x, g.mn[0] + g.dx[0], mx[0] - g.dx[0]);
  float vy = clampf(y, g.mn[1] + g.dx[1], mx[1] - g.dx[1]);
  float vz = clampf(z, g.mn[2] + g.dx[2], mx[2] - g.dx[2]);
  float gr[3];
  grad_inner<kPow2>(sp, g, inv, vx, vy, vz, gr);
  return solid_phi_in<kPow2>(sp, g, inv, vx, vy, vz) + gr[0] * (x - vx) + gr[1] * (y - vy) + gr[2] * (z - vz);
}
// Grid<3>::GetSolidGradient (Grid.cpp:175-192)
template <bool kPow2 = false>
__device__ void solid_grad_any(const float* sp, const GridP& g, const float mx[3], float x, float y, float z,
                               float out[3], const float* inv = nullptr) {
  if (!in_grid(g, mx, x, y, z)) {
    float vx = clampf(x, g.mn[0] + g.dx[0], mx[0] - g.dx[0]);
    float vy = clampf(y, g.mn[1] + g.dx[1], mx[1] - g.dx[1]);
    float vz = clampf(z, g.mn[2] + g.dx[2], mx[2] - g.dx[2]);
    grad_inner<kPow2>(sp, g, inv, vx, vy, vz, out);
    return;
  }
  const float eps = 1e-4f;
  grad_from_samples(solid_phi_any<kPow2>(sp, g, mx, x - eps, y, z, inv),
                    solid_phi_any<kPow2>(sp, g, mx, x + eps, y, z, inv),
                    solid_phi_any<kPow2>(sp, g, mx, x, y - eps, z, inv),
                    solid_phi_any<kPow2>(sp, g, mx, x, y + eps, z, inv),
                    solid_phi_any<kPow2>(sp, g, mx, x, y, z + eps, inv),
                    solid_phi_any<kPow2>(sp, g, mx, x, y, z - eps, inv), out);
}

// The G2P taps of one axis: the nodes j in [0, dim) whose centre c (j on the component's own
// axis, j + 0.5 on the other two) has a nonzero B-spline weight, |g - c| < 1.5.  They are
// consecutive and at most 3.  With b = floor(g) - 1 they lie among b .. b + 3 (integer centres;
// b + 3 only when b has weight zero) or b .. b + 2 (half-integer centres); the reference scans
// ceil(radius / dx) >= 2 nodes either side of (int)g, a window that holds every one of them.
// Out: the first tap j0 and, for taps j0 .. j0 + 2, the weight and unsigned gradient computed
// from d = g - c exactly as the reference does, zero for a tap outside [0, dim) or of weight
// zero.  A zero weight adds exactly nothing to the sums (velocities are finite and a running
// sum that starts at +0 is never -0), so the sums over the 27 taps equal the reference's over
// the nonzero taps.
template <bool kHalf>
__device__ __forceinline__ void g2p_axis(float g, int dim, int& j0, float w[3], float gw[3]) {
  constexpr int NC = kHalf ? 3 : 4;
  const float fb = floorf(g) - 1.0f;
  const int b = (int)fb;
  float d[NC], wc[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    const float cc = kHalf ? (fb + (float)m) + 0.5f : fb + (float)m;  // (float)j (+ 0.5f), exactly
    d[m] = g - cc;
    wc[m] = quad_w(d[m]);
  }
  // integer centres: the taps start at b + 1 when b has weight zero
  const bool sh = !kHalf && wc[0] == 0.f;
  j0 = sh ? b + 1 : b;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    float wm = wc[m], dm = d[m];
    if (!kHalf) {
      wm = sh ? wc[m + 1] : wm;
      dm = sh ? d[m + 1] : dm;
    }
    const bool in = (unsigned)(j0 + m) < (unsigned)dim;
    w[m] = in ? wm : 0.f;
    gw[m] = in ? quad_gw(dm) : 0.f;
  }
}

__device__ __forceinline__ float ld_at(const float* __restrict__ base, unsigned byte_off) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + byte_off);
}

// Grid2Particle (LiquidModel.cpp:724-822) fused with AdvectParticle and the max |v| of GetCFL.
// Each axis's taps of both centre kinds are found once and shared by the three components.
template <bool kPow2>
__global__ void __launch_bounds__(256) k_g2p_advect(float4* __restrict__ r0, int64_t cap, int64_t P,
                                                     const float* __restrict__ vx, const float* __restrict__ vy,
                                                     const float* __restrict__ vz, const float* __restrict__ sp, GP q,
                                                     float mx0, float mx1, float mx2, int do_g2p, int do_advect,
                                                     unsigned* __restrict__ maxbits) {
  const float mx[3] = {mx0, mx1, mx2};
  float4* __restrict__ rs = r0 + cap;
  float local_max = 0.f;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < P; i += (int64_t)gridDim.x * blockDim.x) {
    const float4 R0 = r0[i];
    float pos[3] = {R0.x, R0.y, R0.z};
    float vel[3];
    if (do_g2p) {
      // per axis k: taps of the integer centres (the face axis, N_k + 1 nodes) and of the
      // half-integer centres (N_k nodes)
      int ji[3], jh[3];
      float wi[3][3], gi[3][3], wh[3][3], gh[3][3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float gk = to_grid<kPow2>(q.g, q.inv, k, pos[k]);
        g2p_axis<false>(gk, q.g.N[k] + 1, ji[k], wi[k], gi[k]);
        g2p_axis<true>(gk, q.g.N[k], jh[k], wh[k], gh[k]);
      }
      float Cm[9];
      const float* V[3] = {vx, vy, vz};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int d1 = q.g.N[1] + (a == 1), d2 = q.g.N[2] + (a == 2);
        const int dims[3] = {q.g.N[0] + (a == 0), d1, d2};
        // this component's taps: integer centres on axis a, half-integer on the others
        int j0[3];
        const float* w[3];
        const float* gw[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          j0[k] = (k == a) ? ji[k] : jh[k];
          w[k] = (k == a) ? wi[k] : wh[k];
          gw[k] = (k == a) ? gi[k] : gh[k];
        }
        // byte offsets of the 27 taps; a tap outside the grid (weight zero) reads the nearest node
        unsigned ox[3], oy[3], oz[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          ox[m] = (unsigned)min(max(j0[0] + m, 0), dims[0] - 1) * ((unsigned)d1 * (unsigned)d2 * 4u);
          oy[m] = (unsigned)min(max(j0[1] + m, 0), dims[1] - 1) * ((unsigned)d2 * 4u);
          oz[m] = (unsigned)min(max(j0[2] + m, 0), dims[2] - 1) * 4u;
        }
        // all 27 loads in flight, then the sums in the reference's tap order (x, y, z ascending)
        const float* Va = V[a];
        float fv[3][3][3];
#pragma unroll
        for (int ix = 0; ix < 3; ++ix)
#pragma unroll
          for (int iy = 0; iy < 3; ++iy) {
            const unsigned oxy = ox[ix] + oy[iy];
#pragma unroll
            for (int iz = 0; iz < 3; ++iz) fv[ix][iy][iz] = ld_at(Va, oxy + oz[iz]);
          }
        float acc = 0.f, c0 = 0.f, c1 = 0.f, c2 = 0.f;
#pragma unroll
        for (int ix = 0; ix < 3; ++ix)
#pragma unroll
          for (int iy = 0; iy < 3; ++iy)
#pragma unroll
            for (int iz = 0; iz < 3; ++iz) {
              const float wX = w[0][ix], wY = w[1][iy], wZ = w[2][iz];
              const float gX = gw[0][ix], gY = gw[1][iy], gZ = gw[2][iz];
              const float weight = wX * wY * wZ;
              const float g0 = gX * wY * wZ, g1 = wX * gY * wZ, g2 = wX * wY * gZ;
              const float f = fv[ix][iy][iz];
              acc = acc + f * weight;
              c0 = c0 + f * g0;
              c1 = c1 + f * g1;
              c2 = c2 + f * g2;
            }
        vel[a] = acc;
        Cm[3 * a + 0] = c0;
        Cm[3 * a + 1] = c1;
        Cm[3 * a + 2] = c2;
      }
      rs[3 * i] = make_float4(Cm[0], Cm[1], Cm[2], vel[1]);
      rs[3 * i + 1] = make_float4(Cm[3], Cm[4], Cm[5], vel[2]);
      rs[3 * i + 2] = make_float4(Cm[6], Cm[7], Cm[8], vel[0]);
    } else {
      vel[0] = rs[3 * i + 2].w;
      vel[1] = rs[3 * i].w;
      vel[2] = rs[3 * i + 1].w;
    }
    float m3 = max3(fabsf(vel[0]), fabsf(vel[1]), fabsf(vel[2]));
    local_max = (local_max > m3) ? local_max : m3;
    if (do_advect) {
      float p0 = pos[0] + vel[0] * q.dt, p1 = pos[1] + vel[1] * q.dt, p2 = pos[2] + vel[2] * q.dt;
      float s = solid_phi_any<kPow2>(sp, q.g, mx, p0, p1, p2, q.inv);
      if (s < 0.0f) {
        float gr[3];
        solid_grad_any<kPow2>(sp, q.g, mx, p0, p1, p2, gr, q.inv);
        float sh = s + kEps;
        p0 -= sh * gr[0];
        p1 -= sh * gr[1];
        p2 -= sh * gr[2];
      }
      pos[0] = p0;
      pos[1] = p1;
      pos[2] = p2;
    }
    if (do_advect) r0[i] = make_float4(pos[0], pos[1], pos[2], R0.w);
  }
  // block max, then one atomic per block (max of non-negative floats is order-free)
  __shared__ float sm[kBlock];
  sm[threadIdx.x] = local_max;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] = fmaxf(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(maxbits, __float_as_uint(sm[0]));
}

// ------------------------------------------------------------------ CorrectVolume
// LiquidModel<3>::CorrectVolume, LiquidModel.cpp:1085-1132, for the particles with
// id % correctCycle == correctStep.  The reference's hash search (NeighborSearch.h:22-57,
// 131-136) files each particle under the hash slots of the 3x3x3 cells around its own cell
// (clamped to the grid), 2 NumParticles slots, and scans the slots of the 3x3x3 cells
// around the querying particle, so a neighbour closer than the radius is visited
//   visits = #{(C1, C2) in R(pi) x R(npi) : slot(C1) == slot(C2)}
// times: once per cell the two regions share, plus once per pair of distinct cells whose
// slots collide.  Without collisions among the cells involved this is the overlap product
//   prod_axes |[max(c1-1,0), min(c1+1,N-1)] & [max(c2-1,0), min(c2+1,N-1)]|.
// A neighbour within the radius (2 dx) lies within 2 cells per axis, so C1 and C2 are at
// most 4 cells apart: a per-cell flag (slot shared with another cell within 4) marks the
// regions that need the exact count; all others take the product.  Neighbours come from the
// half-cell bins (cells within 2, plus one half-cell of slack) in bin order, and the random
// kick of coincident particles (the reference's random_device-seeded mt19937) comes from a
// counter generator.
struct CVP {
  GridP g;
  int H[3];
  int cycle, step;
  float dt, wf, radius;
  unsigned long long seed, count;
  float mx[3];
  int nslot;
};

__device__ __forceinline__ uint64_t cv_splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// uniform [0, 1) for component k of the v-th visit of neighbour npid by particle pid in
// the count-th CorrectVolume call
__device__ __forceinline__ float cv_rand(uint64_t seed, uint64_t count, int64_t pid, int64_t npid, int v, int k) {
  uint64_t h = cv_splitmix(seed ^ 0x436F7272566F6Cull);
  h = cv_splitmix(h ^ count);
  h = cv_splitmix(h ^ (uint64_t)pid);
  h = cv_splitmix(h ^ ((uint64_t)npid * 128u + (uint64_t)v * 4u + (uint64_t)k));
  return (float)(h >> 40) * (1.0f / 16777216.0f);
}
__device__ __forceinline__ int cv_cell(float p, float mn, float dx) { return (int)((p - mn) / dx); }
// hash slot of cell (x, y, z) among m, NeighborSearch.h:45 with the constants of :134-136: the
// products wrap at 32 bits, % truncates toward zero, then the result is made >= 0
__device__ __forceinline__ int ns_slot(int x, int y, int z, int m) {
  const uint32_t h = (73856093u * (uint32_t)x) ^ (19349663u * (uint32_t)y) ^ (83492791u * (uint32_t)z);
  const int hs = (int)h;
  return ((hs % m) + m) % m;
}

// flag[cell] = 1 when another cell within 4 per axis (inside the grid) has its slot
__global__ void __launch_bounds__(256) k_cv_flags(uint8_t* __restrict__ flag, int N0, int N1, int N2, int m) {
  const int64_t n = (int64_t)N0 * N1 * N2;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(i % N2), y = (int)((i / N2) % N1), x = (int)(i / ((int64_t)N1 * N2));
    const int s = ns_slot(x, y, z, m);
    int hit = 0;
    for (int x2 = max(x - 4, 0); x2 <= min(x + 4, N0 - 1) && !hit; ++x2)
      for (int y2 = max(y - 4, 0); y2 <= min(y + 4, N1 - 1) && !hit; ++y2)
        for (int z2 = max(z - 4, 0); z2 <= min(z + 4, N2 - 1); ++z2)
          if (ns_slot(x2, y2, z2, m) == s && (x2 != x || y2 != y || z2 != z)) {
            hit = 1;
            break;
          }
    flag[i] = (uint8_t)hit;
  }
}

// new positions of the selected particles into out (3 x cap), bin-ordered particles
__global__ void __launch_bounds__(256) k_correct_volume(const float4* __restrict__ r0, int64_t cap, int64_t P,
                                                        const int* __restrict__ start,
                                                        const float* __restrict__ solid,
                                                        const uint8_t* __restrict__ flag, CVP q,
                                                        float* __restrict__ out) {
  const int* N = q.g.N;
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < P; s += (int64_t)gridDim.x * blockDim.x) {
    const int pid = rec_id(r0, (size_t)s);
    if (pid % q.cycle != q.step) continue;
    const float4 Rs = r0[s];
    const float p[3] = {Rs.x, Rs.y, Rs.z};
    const float r = q.radius, r2 = r * r;
    int lo[3], hi[3], hlo[3], hhi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int cc = cv_cell(p[a], q.g.mn[a], q.g.dx[a]);
      lo[a] = min(max(cc - 1, 0), N[a] - 1);
      hi[a] = min(max(cc + 1, 0), N[a] - 1);
      hlo[a] = max(0, 2 * (cc - 2) - 1);
      hhi[a] = min(q.H[a] - 1, 2 * (cc + 2) + 2);
    }
    bool exact = false;  // a cell of the region shares its slot with a cell it can meet
    for (int x1 = lo[0]; x1 <= hi[0]; ++x1)
      for (int y1 = lo[1]; y1 <= hi[1]; ++y1)
        for (int z1 = lo[2]; z1 <= hi[2]; ++z1) exact |= flag[((size_t)x1 * N[1] + y1) * N[2] + z1] != 0;
    int reg[27], nreg = 0;  // slots of the region's cells (exact count only)
    if (exact)
      for (int x1 = lo[0]; x1 <= hi[0]; ++x1)
        for (int y1 = lo[1]; y1 <= hi[1]; ++y1)
          for (int z1 = lo[2]; z1 <= hi[2]; ++z1) reg[nreg++] = ns_slot(x1, y1, z1, q.nslot);
    float rep[3] = {0.f, 0.f, 0.f};
    for (int hx = hlo[0]; hx <= hhi[0]; ++hx)
      for (int hy = hlo[1]; hy <= hhi[1]; ++hy) {
        const size_t k0 = ((size_t)hx * q.H[1] + hy) * q.H[2];
        const int s1 = start[k0 + hhi[2] + 1];
        for (int t = start[k0 + hlo[2]]; t < s1; ++t) {
          const int npid = rec_id(r0, (size_t)t);
          if (npid == pid) continue;
          const float4 Rt = r0[t];
          const float qq[3] = {Rt.x, Rt.y, Rt.z};
          const float d0 = p[0] - qq[0], d1 = p[1] - qq[1], d2 = p[2] - qq[2];
          const float sq = d0 * d0 + (d1 * d1 + d2 * d2);
          if (sq >= r2) continue;
          int l2[3], h2[3];
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            const int c2 = cv_cell(qq[a], q.g.mn[a], q.g.dx[a]);
            l2[a] = min(max(c2 - 1, 0), N[a] - 1);
            h2[a] = min(max(c2 + 1, 0), N[a] - 1);
          }
          int mult = 0;
          if (exact) {
            for (int x2 = l2[0]; x2 <= h2[0]; ++x2)
              for (int y2 = l2[1]; y2 <= h2[1]; ++y2)
                for (int z2 = l2[2]; z2 <= h2[2]; ++z2) {
                  const int s2 = ns_slot(x2, y2, z2, q.nslot);
                  for (int k = 0; k < nreg; ++k) mult += reg[k] == s2 ? 1 : 0;
                }
          } else {
            mult = 1;
#pragma unroll
            for (int a = 0; a < 3; ++a) mult *= max(0, min(hi[a], h2[a]) - max(lo[a], l2[a]) + 1);
          }
          const float dist = sqrtf(sq);
          const float w = q.wf * quad_w(dist / r);
          for (int v = 0; v < mult; ++v) {
            if (dist > 1e-4f * r) {
              const float wr = w * r;
              rep[0] += wr * (d0 / dist);
              rep[1] += wr * (d1 / dist);
              rep[2] += wr * (d2 / dist);
            } else {
#pragma unroll
              for (int k = 0; k < 3; ++k) rep[k] += r / q.dt * cv_rand(q.seed, q.count, pid, npid, v, k);
            }
          }
        }
      }
    float np[3] = {p[0] + rep[0] * q.dt, p[1] + rep[1] * q.dt, p[2] + rep[2] * q.dt};
    const float sphi = solid_phi_any(solid, q.g, q.mx, np[0], np[1], np[2]);
    if (sphi < 0.0f) {
      float gr[3];
      solid_grad_any(solid, q.g, q.mx, np[0], np[1], np[2], gr);
#pragma unroll
      for (int k = 0; k < 3; ++k) np[k] -= sphi * gr[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) out[(size_t)k * cap + s] = np[k];
  }
}
// positions of the selected particles <- out
__global__ void k_correct_commit(float4* __restrict__ r0, int64_t cap, int64_t P, const float* __restrict__ out,
                                 int cycle, int step) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < P; s += (int64_t)gridDim.x * blockDim.x) {
    const int pid = rec_id(r0, (size_t)s);
    if (pid % cycle != step) continue;
    const float4 R = r0[s];
    r0[s] = make_float4(out[s], out[(size_t)cap + s], out[2 * (size_t)cap + s], R.w);
  }
}

}  // namespace

// ------------------------------------------------------------------ bin scan (host)
// hipCUB scans the block sums: nbins / kScanBins + 1 entries
size_t bin_scan_bytes(size_t nblocks) {
  size_t bytes = 0;
  hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const int*)nullptr, (int*)nullptr, (int)nblocks);
  return bytes;
}
int bin_scan(Ctx& c) {
  const size_t n = c.nbins + 1;
  hipLaunchKernelGGL(k_scan_bsum, dim3((unsigned)c.nbblk), dim3(256), 0, c.s, c.bin_count, c.bin_bflag, c.bin_bsum, n);
  size_t bytes = c.scan_bytes;
  if (!hip_ok(c, hipcub::DeviceScan::ExclusiveSum(c.scan_tmp, bytes, c.bin_bsum, c.bin_boff, (int)c.nbblk, c.s),
              "bin scan"))
    return LH_ERR_DEVICE;
  hipLaunchKernelGGL(k_scan_write, dim3((unsigned)c.nbblk), dim3(256), 0, c.s, c.bin_count, c.bin_bflag, c.bin_boff,
                     c.bin_start, n);
  return LH_OK;
}

// dx = 2^-k on every axis: (x - min) / dx == (x - min) * 2^k exactly; inv receives 2^k
static bool pow2_inverse(const GridP& g, float inv[3]) {
  bool pow2 = true;
  for (int k = 0; k < 3; ++k) {
    int e = 0;
    const float m = std::frexp(g.dx[k], &e);
    inv[k] = std::ldexp(1.0f, 1 - e);
    pow2 = pow2 && m == 0.5f && std::isnormal(inv[k]) && g.dx[k] * inv[k] == 1.0f;
  }
  return pow2;
}

int run_binning(Ctx& c) {
  c.bins_valid = false;
  if (c.P == 0) return LH_OK;
  BinP b;
  for (int a = 0; a < 3; ++a) {
    b.H[a] = c.H[a];
    b.mn[a] = c.g.mn[a];
    b.dx[a] = c.g.dx[a];
  }
  const bool pow2 = pow2_inverse(c.g, b.inv);
  hipMemsetAsync(c.dflag + 1, 0, sizeof(int), c.s);
  unsigned gp = grid_for((size_t)c.P);
  const float4* rec = (const float4*)c.prec[c.cur];
  if (pow2)
    hipLaunchKernelGGL(k_bin_key<true>, dim3(gp), dim3(kBlock), 0, c.s, rec, c.P, b, c.pkey, c.porder, c.pid,
                       c.bin_count, c.bin_bflag, c.dflag + 1);
  else
    hipLaunchKernelGGL(k_bin_key<false>, dim3(gp), dim3(kBlock), 0, c.s, rec, c.P, b, c.pkey, c.porder, c.pid,
                       c.bin_count, c.bin_bflag, c.dflag + 1);
  int rc = bin_scan(c);
  if (rc) return rc;
  // porder holds the arrival ranks until k_bin_rank writes the order over them
  hipLaunchKernelGGL(k_bin_scatter, dim3(gp), dim3(kBlock), 0, c.s, c.pkey, c.porder, c.P, c.bin_start, c.bin_count,
                     c.pslot);
  hipLaunchKernelGGL(k_bin_rank, dim3(gp), dim3(kBlock), 0, c.s, c.pid, c.P, c.pkey, c.bin_start, c.pslot,
                     c.porder);
  hipLaunchKernelGGL(k_bin_gather, dim3(gp), dim3(kBlock), 0, c.s, rec, (float4*)c.prec[1 - c.cur], c.cap, c.P,
                     c.porder);
  c.cur = 1 - c.cur;
  c.bins_valid = true;
  return LH_OK;
}

void launch_liquid_phi(Ctx& c) {
  if (c.P == 0 || !c.bins_valid) return;
  ProfScope ps(c, LH_PROF_LIQUID_PHI);
  unsigned nt = tile_grid(div_up(c.g.N[0], kTX) * div_up(c.g.N[1], kTY) * div_up(c.g.N[2], kTZ));
  hipLaunchKernelGGL(k_liquid_phi, dim3(nt), dim3(kBlock), 0, c.s, c.phi, (const float4*)c.prec[c.cur], c.cap, c.bin_start, c.g,
                     c.H[0], c.H[1], c.H[2], c.dW[0], c.dW[1], c.dW[2], c.radius, 3 * c.g.dx[0]);
}


void launch_p2g(Ctx& c, bool fuse_body_forces, float dt) {
  if (c.P == 0 || !c.bins_valid) return;
  P2GP q;
  q.g = c.g;
  for (int k = 0; k < 3; ++k) q.H[k] = c.H[k];
  q.gdt = 0.f;
  q.adt = 0.f;
  q.tank_axis = -1;
  if (fuse_body_forces) {
    q.gdt = c.cfg.gravity * dt;
    if (c.cfg.tank_axis >= 0 && c.cfg.tank_axis <= 2 && c.cfg.tank_amp != 0.f) {
      double acc = (double)c.cfg.tank_amp * std::sin(2.0 * M_PI * (double)c.cfg.tank_freq * c.time);
      q.adt = (float)acc * dt;
      q.tank_axis = c.cfg.tank_axis;
    }
  }
  const bool pow2 = pow2_inverse(c.g, q.inv);
  if (!fuse_body_forces) q.gdt = 0.f;
  ProfScope ps(c, LH_PROF_P2G);
  // at least one particle per cell: the liquid fills its tiles, and longer rows pay
  const bool dense = (size_t)c.P >= c.ncell;
  const int mx = dense ? kMXDense : kMXSparse;
  const unsigned nt = tile_grid(div_up(c.g.N[0] + 1, mx) * div_up(c.g.N[1] + 1, kPTY) * div_up(c.g.N[2] + 1, kPTZ));
  const float4* pos = (const float4*)c.prec[c.cur];
  const float4* rest = pos + c.cap;
  if (!c.cfg.p2g_gather) {
    const int ms = dense ? kMXSharedDense : kMXSharedSparse;
    const unsigned ns = tile_grid(div_up(c.g.N[0] + 1, ms) * div_up(c.g.N[1] + 1, kSTY) * div_up(c.g.N[2] + 1, kPTZ));
    if (pow2 && dense)
      hipLaunchKernelGGL((k_p2g_shared<true, kMXSharedDense>), dim3(ns), dim3(64 * kSTY), 0, c.s, c.v[0], c.v[1], c.v[2],
                         pos, rest, c.bin_start, q);
    else if (pow2)
      hipLaunchKernelGGL((k_p2g_shared<true, kMXSharedSparse>), dim3(ns), dim3(64 * kSTY), 0, c.s, c.v[0], c.v[1], c.v[2],
                         pos, rest, c.bin_start, q);
    else if (dense)
      hipLaunchKernelGGL((k_p2g_shared<false, kMXSharedDense>), dim3(ns), dim3(64 * kSTY), 0, c.s, c.v[0], c.v[1], c.v[2],
                         pos, rest, c.bin_start, q);
    else
      hipLaunchKernelGGL((k_p2g_shared<false, kMXSharedSparse>), dim3(ns), dim3(64 * kSTY), 0, c.s, c.v[0], c.v[1], c.v[2],
                         pos, rest, c.bin_start, q);
    return;
  }
  if (pow2 && dense)
    hipLaunchKernelGGL((k_p2g<true, kMXDense>), dim3(nt), dim3(kBlock), 0, c.s, c.v[0], c.v[1], c.v[2], pos, rest,
                       c.bin_start, q);
  else if (pow2)
    hipLaunchKernelGGL((k_p2g<true, kMXSparse>), dim3(nt), dim3(kBlock), 0, c.s, c.v[0], c.v[1], c.v[2], pos, rest,
                       c.bin_start, q);
  else if (dense)
    hipLaunchKernelGGL((k_p2g<false, kMXDense>), dim3(nt), dim3(kBlock), 0, c.s, c.v[0], c.v[1], c.v[2], pos, rest,
                       c.bin_start, q);
  else
    hipLaunchKernelGGL((k_p2g<false, kMXSparse>), dim3(nt), dim3(kBlock), 0, c.s, c.v[0], c.v[1], c.v[2], pos, rest,
                       c.bin_start, q);
}

int run_correct_volume(Ctx& c, float dt) {
  if (c.P == 0) return LH_OK;
  int rc = run_binning(c);  // neighbours of the advected positions
  if (rc) return rc;
  CVP q;
  q.g = c.g;
  for (int a = 0; a < 3; ++a) {
    q.H[a] = c.H[a];
    q.mx[a] = c.cfg.mx[a];
  }
  q.cycle = std::max(1, c.cfg.correct_cycle);
  q.step = c.correct_step;
  q.dt = dt;
  q.wf = sqrtf(c.g.dx[0] * c.g.dx[1] * c.g.dx[2]);  // std::sqrt(dx.x() * dx.y() * dx.z())
  q.radius = c.radius;
  q.seed = c.cfg.seed;
  q.count = c.correct_count;
  if (c.P > (int64_t)(INT32_MAX / 2)) {
    c.err = "CorrectVolume: the hash table of 2 x NumParticles slots exceeds a 32-bit index";
    return LH_ERR_DOMAIN;
  }
  q.nslot = (int)(2 * c.P);  // HashNeighborSearch: scale factor 2 x NumParticles
  const size_t ncell = (size_t)c.g.N[0] * c.g.N[1] * c.g.N[2];
  if (!c.cv_flag) {
    if (hipMalloc((void**)&c.cv_flag, ncell) != hipSuccess) {
      c.cv_flag = nullptr;
      c.err = "hipMalloc failed (CorrectVolume cell flags)";
      return LH_ERR_DEVICE;
    }
    c.cv_nslot = 0;
  }
  if (c.cv_nslot != q.nslot) {
    hipLaunchKernelGGL(k_cv_flags, dim3(grid_for(ncell)), dim3(kBlock), 0, c.s, c.cv_flag, c.g.N[0], c.g.N[1],
                       c.g.N[2], q.nslot);
    c.cv_nslot = q.nslot;
  }
  float* out = c.prec[1 - c.cur];  // free between permutations: 3 x cap floats of it
  const unsigned gp = grid_for((size_t)c.P);
  hipLaunchKernelGGL(k_correct_volume, dim3(gp), dim3(kBlock), 0, c.s, (const float4*)c.prec[c.cur], c.cap, c.P,
                     c.bin_start, c.solid_phi, c.cv_flag, q, out);
  hipLaunchKernelGGL(k_correct_commit, dim3(gp), dim3(kBlock), 0, c.s, (float4*)c.prec[c.cur], c.cap, c.P, out,
                     q.cycle, q.step);
  c.bins_valid = false;
  c.correct_step = (q.step + 1) % q.cycle;
  c.correct_count++;
  return LH_OK;
}

void launch_g2p_advect(Ctx& c, bool do_g2p, bool do_advect, float dt) {
  hipMemsetAsync(c.dmaxv, 0, sizeof(unsigned), c.s);
  if (c.P == 0) return;
  GP q;
  q.g = c.g;
  q.dt = dt;
  const bool pow2 = pow2_inverse(c.g, q.inv);
  unsigned gp = grid_for((size_t)c.P);
  ProfScope ps(c, LH_PROF_G2P);
  // the face arrays' byte offsets fit 32 bits: grids are under 2^31 half-cell bins (N < 645 per axis)
  if (pow2)
    hipLaunchKernelGGL(k_g2p_advect<true>, dim3(gp), dim3(kBlock), 0, c.s, (float4*)c.prec[c.cur], c.cap, c.P, c.v[0], c.v[1],
                       c.v[2], c.solid_phi, q, c.cfg.mx[0], c.cfg.mx[1], c.cfg.mx[2], do_g2p ? 1 : 0,
                       do_advect ? 1 : 0, c.dmaxv);
  else
    hipLaunchKernelGGL(k_g2p_advect<false>, dim3(gp), dim3(kBlock), 0, c.s, (float4*)c.prec[c.cur], c.cap, c.P, c.v[0], c.v[1],
                       c.v[2], c.solid_phi, q, c.cfg.mx[0], c.cfg.mx[1], c.cfg.mx[2], do_g2p ? 1 : 0,
                       do_advect ? 1 : 0, c.dmaxv);
}

}  // namespace lh
