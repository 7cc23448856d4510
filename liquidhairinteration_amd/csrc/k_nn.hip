// k_nn.hip -- the learned initial guess (lh_nn.h): BF16 implicit-GEMM 3x3x3
// convolutions on MFMA (forward, backward data, backward weights), pooling,
// upsample + concatenation, the masked MSE loss and Adam.
//
// Layout: channels-last tensors [x][y][z][c] (z fastest, the Array3 order of the
// grid), BF16 (round to nearest even), FP32 accumulation.  A convolution is the GEMM
//   out[v][co] = sum_k W[co][k] * im2col[v][k],  k = tap * Cin + ci,
// mapped on v_mfma_f32_16x16x32_bf16 with the weights as the 16 x 32 A operand and
// 16 voxels x 32 K of the implicit im2col as the B operand: lane l holds
// A[l % 16][8 (l / 16) .. +7] and B[8 (l / 16) .. +7][l % 16] -- 8 consecutive K,
// i.e. 8 consecutive input channels of one tap, one 16-byte load -- and receives
// C[4 (l / 16) + i][l % 16], four consecutive output channels of one voxel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "lh_ctx.h"
#include "lh_nn.h"

namespace lh {

const int kNnCin[kNnLayers] = {2, 16, 32, 64, 160, 96, 48, 16};
const int kNnCout[kNnLayers] = {16, 32, 64, 96, 64, 32, 16, 1};
const int kNnLevel[kNnLayers] = {0, 1, 2, 3, 2, 1, 0, 0};

namespace {

typedef uint16_t bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

constexpr int kBlk = 256;

__host__ __device__ inline float bf2f(bf16_t h) { return __builtin_bit_cast(float, (unsigned)h << 16); }
// FP32 -> BF16, round to nearest even (NaN stays NaN)
__host__ __device__ inline bf16_t f2bf(float f) {
  unsigned u = __builtin_bit_cast(unsigned, f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ inline bool bf_pos(bf16_t h) { return h != 0 && !(h & 0x8000u); }
constexpr int kpad32(int k) { return (k + 31) / 32 * 32; }

__device__ inline bf16x8 ld8(const bf16_t* p) { return __builtin_bit_cast(bf16x8, *(const u32x4*)p); }
__device__ inline bf16x8 zero8() { return __builtin_bit_cast(bf16x8, u32x4{0u, 0u, 0u, 0u}); }
__device__ inline void tap_off(int tap, int& ox, int& oy, int& oz) {
  ox = tap / 9 - 1;
  oy = (tap / 3) % 3 - 1;
  oz = tap % 3 - 1;
}

// ------------------------------------------------------------ convolution (MFMA)
// EPI 0: out = bf16(relu(acc + bias)); 1: out = bf16(act > 0 ? acc : 0) (backward data through
// the ReLU that produced act); 2: out = bf16(acc).  One wave: all COUT x (16 NVB voxels).
template <int CIN, int COUT, int NVB, int EPI>
__global__ __launch_bounds__(kBlk) void k_conv(const bf16_t* __restrict__ in, const bf16_t* __restrict__ W,
                                               const float* __restrict__ bias, const bf16_t* __restrict__ act,
                                               bf16_t* __restrict__ out, int d0, int d1, int d2) {
  static_assert(COUT % 16 == 0 && (CIN % 8 == 0 || CIN == 2), "channel counts");
  constexpr int MB = COUT / 16, KP = kpad32(27 * CIN), KS = KP / 32;
  const int lane = threadIdx.x & 63, row = lane & 15, kq = lane >> 4;
  const int V = d0 * d1 * d2;
  const int vb = (blockIdx.x * (kBlk / 64) + (threadIdx.x >> 6)) * (16 * NVB);
  if (vb >= V) return;
  int px[NVB], py[NVB], pz[NVB];
  bool pv[NVB];
#pragma unroll
  for (int nb = 0; nb < NVB; ++nb) {
    const int v = vb + nb * 16 + row;
    pv[nb] = v < V;
    pz[nb] = v % d2;
    py[nb] = (v / d2) % d1;
    px[nb] = v / (d1 * d2);
  }
  f32x4 acc[MB][NVB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NVB; ++nb) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16_t* wrow = W + (size_t)row * KP + kq * 8;
  for (int ks = 0; ks < KS; ++ks) {
    const int k0 = ks * 32 + kq * 8;
    bf16x8 a[MB], b[NVB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) a[mb] = ld8(wrow + (size_t)mb * 16 * KP + ks * 32);
    if constexpr (CIN % 8 == 0) {
      const int tap = k0 / CIN, ci = k0 - tap * CIN;
      int ox, oy, oz;
      tap_off(tap, ox, oy, oz);
#pragma unroll
      for (int nb = 0; nb < NVB; ++nb) {
        const int x = px[nb] + ox, y = py[nb] + oy, z = pz[nb] + oz;
        const bool ok = pv[nb] && tap < 27 && (unsigned)x < (unsigned)d0 && (unsigned)y < (unsigned)d1 &&
                        (unsigned)z < (unsigned)d2;
        b[nb] = ok ? ld8(in + (size_t)((x * d1 + y) * d2 + z) * CIN + ci) : zero8();
      }
    } else {  // CIN == 2: a lane's 8 K values are 4 taps x 2 channels
#pragma unroll
      for (int nb = 0; nb < NVB; ++nb) {
        u32x4 q;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int tap = k0 / 2 + j;
          int ox, oy, oz;
          tap_off(tap, ox, oy, oz);
          const int x = px[nb] + ox, y = py[nb] + oy, z = pz[nb] + oz;
          const bool ok = pv[nb] && tap < 27 && (unsigned)x < (unsigned)d0 && (unsigned)y < (unsigned)d1 &&
                          (unsigned)z < (unsigned)d2;
          q[j] = ok ? *(const unsigned*)(in + (size_t)((x * d1 + y) * d2 + z) * 2) : 0u;
        }
        b[nb] = __builtin_bit_cast(bf16x8, q);
      }
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
      for (int nb = 0; nb < NVB; ++nb)
        acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mb], b[nb], acc[mb][nb], 0, 0, 0);
  }
#pragma unroll
  for (int nb = 0; nb < NVB; ++nb) {
    if (!pv[nb]) continue;
    const size_t v = (size_t)(vb + nb * 16 + row);
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      const int co = mb * 16 + kq * 4;
      bf16_t o[4];
      if constexpr (EPI == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] = f2bf(fmaxf(acc[mb][nb][i] + bias[co + i], 0.f));
      } else if constexpr (EPI == 1) {
        const u32x2 am = *(const u32x2*)(act + v * COUT + co);
        const bf16_t h[4] = {(bf16_t)(am[0] & 0xffffu), (bf16_t)(am[0] >> 16), (bf16_t)(am[1] & 0xffffu),
                             (bf16_t)(am[1] >> 16)};
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] = bf_pos(h[i]) ? f2bf(acc[mb][nb][i]) : (bf16_t)0;
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] = f2bf(acc[mb][nb][i]);
      }
      *(u32x2*)(out + v * COUT + co) = u32x2{(unsigned)o[0] | ((unsigned)o[1] << 16), (unsigned)o[2] | ((unsigned)o[3] << 16)};
    }
  }
}

// The same convolution over LDS tiles: a workgroup (4 waves) computes TX x TY x 16 output voxels.
// The input over the tile and a one-voxel halo, (TX + 2) (TY + 2) 18 voxels x CIN channels, is
// staged in LDS once (voxel stride CIN + 8 BF16: an odd multiple of 16 bytes, so the 16-byte reads
// of 16 consecutive voxels hit distinct banks), and every tap reads its B fragments from there
// instead of gathering them from global memory 27 times.  The weights pass through LDS in chunks of
// KC columns of K, double-buffered: each thread fetches its share of chunk c + 1 into registers
// before the MFMAs of chunk c and stores it after them, so the global latency hides behind the
// chunk's arithmetic.  Each wave owns TX TY / 4 z-runs of 16 voxels (NVB B blocks sharing each A
// fragment).  The MFMA mapping, the K order (k = tap * CIN + ci) and the epilogues are those of
// k_conv, so the output is identical.  Tiles are numbered z fastest and dealt to the 8 XCDs in
// contiguous runs, so tiles that share halo voxels share an L2.
template <int CIN, int COUT, int TX, int TY, int KC, int EPI>
__global__ __launch_bounds__(256) void k_conv_tile(const bf16_t* __restrict__ in, const bf16_t* __restrict__ W,
                                                   const float* __restrict__ bias, const bf16_t* __restrict__ act,
                                                   bf16_t* __restrict__ out, int d0, int d1, int d2) {
  static_assert(COUT % 16 == 0 && CIN % 8 == 0 && (TX * TY) % 4 == 0 && KC % 32 == 0, "channel counts / tile");
  constexpr int TZ = 16, NVB = TX * TY / 4, MB = COUT / 16, KP = kpad32(27 * CIN), KS = KP / 32;
  constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2, NH = HX * HY * HZ, VS = CIN + 8;
  constexpr int WS = KC + 8;                         // LDS row stride of a weight chunk (odd x 16 bytes)
  constexpr int NCH = (KP + KC - 1) / KC;            // weight chunks
  constexpr int WP = COUT * KC / 8;                  // 16-byte pieces of a chunk
  constexpr int WPT = (WP + 255) / 256;              // pieces per thread
  __shared__ __attribute__((aligned(16))) bf16_t s_in[NH * VS];
  __shared__ __attribute__((aligned(16))) bf16_t s_w[2][COUT * WS];
  // per K step and lane group: the LDS offset of the B fragment from the lane's halo-corner voxel
  // (tap (ox, oy, oz) -> ((ox + 1) HY + oy + 1) HZ + oz + 1 voxels, plus the channel); past 27 taps
  // the weights are zero and the offset 0 reads a finite in-tile value
  __shared__ uint16_t s_off[KS * 4];
  const int ntz = (d2 + TZ - 1) / TZ, nty = (d1 + TY - 1) / TY, ntx = (d0 + TX - 1) / TX;
  const int nt = ntx * nty * ntz;
  const int run = (nt + 7) >> 3;
  const int t = ((int)blockIdx.x & 7) * run + ((int)blockIdx.x >> 3);
  if ((int)blockIdx.x >= 8 * run || t >= nt) return;  // uniform over the block
  const int tz = t % ntz, ty = (t / ntz) % nty, tx = t / (ntz * nty);
  const int X0 = tx * TX, Y0 = ty * TY, Z0 = tz * TZ;
  // weight chunk c: rows co, columns [c KC, c KC + KC) of W (zero past KP)
  auto fetch_w = [&](int c, u32x4 (&q)[WPT]) {
#pragma unroll
    for (int j = 0; j < WPT; ++j) {
      const int p = (int)threadIdx.x + 256 * j;
      const int co = p / (KC / 8), k8 = p - co * (KC / 8);
      const int k = c * KC + k8 * 8;
      q[j] = (p < WP && k < KP) ? *(const u32x4*)(W + (size_t)co * KP + k) : u32x4{0u, 0u, 0u, 0u};
    }
  };
  auto store_w = [&](int buf, const u32x4 (&q)[WPT]) {
#pragma unroll
    for (int j = 0; j < WPT; ++j) {
      const int p = (int)threadIdx.x + 256 * j;
      const int co = p / (KC / 8), k8 = p - co * (KC / 8);
      if (p < WP) *(u32x4*)(&s_w[buf][co * WS + k8 * 8]) = q[j];
    }
  };
  {
    u32x4 q[WPT];
    fetch_w(0, q);
    store_w(0, q);
  }
  for (int i = threadIdx.x; i < KS * 4; i += 256) {
    const int k0 = (i >> 2) * 32 + (i & 3) * 8;
    const int tap = k0 / CIN, ci = k0 - tap * CIN;
    int ox, oy, oz;
    tap_off(tap, ox, oy, oz);
    s_off[i] = tap < 27 ? (uint16_t)((((ox + 1) * HY + (oy + 1)) * HZ + (oz + 1)) * VS + ci) : (uint16_t)0;
  }
  // stage the input tile with its halo (zero outside the grid), 8 channels per 16-byte load
  constexpr int C8 = CIN / 8;
  for (int i = threadIdx.x; i < NH * C8; i += 256) {
    const int hv = i / C8, c8 = i - hv * C8;
    const int hz = hv % HZ, hy = (hv / HZ) % HY, hx = hv / (HZ * HY);
    const int x = X0 - 1 + hx, y = Y0 - 1 + hy, z = Z0 - 1 + hz;
    u32x4 q = u32x4{0u, 0u, 0u, 0u};
    if ((unsigned)x < (unsigned)d0 && (unsigned)y < (unsigned)d1 && (unsigned)z < (unsigned)d2)
      q = *(const u32x4*)(in + (size_t)((x * d1 + y) * d2 + z) * CIN + c8 * 8);
    *(u32x4*)(s_in + hv * VS + c8 * 8) = q;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, row = lane & 15, kq = lane >> 4, wv = threadIdx.x >> 6;
  int rx[NVB], ry[NVB], base[NVB];
#pragma unroll
  for (int nb = 0; nb < NVB; ++nb) {
    const int rid = wv * NVB + nb;
    rx[nb] = rid / TY;
    ry[nb] = rid - rx[nb] * TY;
    base[nb] = ((rx[nb] * HY + ry[nb]) * HZ + row) * VS;  // the halo-corner voxel of the lane's output voxel
  }
  f32x4 acc[MB][NVB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NVB; ++nb) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int c = 0; c < NCH; ++c) {
    u32x4 q[WPT];
    if (c + 1 < NCH) fetch_w(c + 1, q);
    const bf16_t* wl = &s_w[c & 1][row * WS + kq * 8];
#pragma unroll
    for (int kk = 0; kk < KC / 32; ++kk) {
      const int ks = c * (KC / 32) + kk;
      if (ks < KS) {  // uniform
        const int o = s_off[ks * 4 + kq];
        bf16x8 a[MB], b[NVB];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) a[mb] = __builtin_bit_cast(bf16x8, *(const u32x4*)(wl + mb * 16 * WS + kk * 32));
#pragma unroll
        for (int nb = 0; nb < NVB; ++nb) b[nb] = __builtin_bit_cast(bf16x8, *(const u32x4*)(s_in + base[nb] + o));
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
          for (int nb = 0; nb < NVB; ++nb)
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mb], b[nb], acc[mb][nb], 0, 0, 0);
      }
    }
    if (c + 1 < NCH) {
      store_w((c + 1) & 1, q);  // the other buffer: its last readers passed the previous barrier
      __syncthreads();
    }
  }
#pragma unroll
  for (int nb = 0; nb < NVB; ++nb) {
    const int x = X0 + rx[nb], y = Y0 + ry[nb], z = Z0 + row;
    if (x >= d0 || y >= d1 || z >= d2) continue;
    const size_t v = (size_t)((x * d1 + y) * d2 + z);
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      const int co = mb * 16 + kq * 4;
      bf16_t o[4];
      if constexpr (EPI == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] = f2bf(fmaxf(acc[mb][nb][i] + bias[co + i], 0.f));
      } else if constexpr (EPI == 1) {
        const u32x2 am = *(const u32x2*)(act + v * COUT + co);
        const bf16_t h[4] = {(bf16_t)(am[0] & 0xffffu), (bf16_t)(am[0] >> 16), (bf16_t)(am[1] & 0xffffu),
                             (bf16_t)(am[1] >> 16)};
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] = bf_pos(h[i]) ? f2bf(acc[mb][nb][i]) : (bf16_t)0;
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] = f2bf(acc[mb][nb][i]);
      }
      *(u32x2*)(out + v * COUT + co) = u32x2{(unsigned)o[0] | ((unsigned)o[1] << 16), (unsigned)o[2] | ((unsigned)o[3] << 16)};
    }
  }
}

// dW[co][k] += sum_v g[v][co] * in[v + off(tap(k))][ci(k)], k < 27 CIN.  MFMA with the
// voxels as the reduction: A = g^T (16 co x 32 voxels), B = im2col (32 voxels x 16 k).
// Grid: x voxel chunks, y groups of NBW 16-column blocks of k, z groups of MBW 16-row blocks of co.
template <int CIN, int COUT, int MBW, int NBW>
__global__ __launch_bounds__(64) void k_conv_wgrad(const bf16_t* __restrict__ g, const bf16_t* __restrict__ in,
                                                   float* __restrict__ dW, int d0, int d1, int d2, int chunk) {
  constexpr int KR = 27 * CIN, KP = kpad32(KR), NBT = (KR + 15) / 16;
  const int lane = threadIdx.x, row = lane & 15, kq = lane >> 4;
  const int V = d0 * d1 * d2;
  const int v_begin = blockIdx.x * chunk;
  const int v_end = min(V, v_begin + chunk);
  const int nb0 = blockIdx.y * NBW, mb0 = blockIdx.z * MBW;
  int ox[NBW], oy[NBW], oz[NBW], ci[NBW];
  bool kok[NBW];
#pragma unroll
  for (int nb = 0; nb < NBW; ++nb) {
    const int k = (nb0 + nb) * 16 + row;
    kok[nb] = (nb0 + nb) < NBT && k < KR;
    const int tap = kok[nb] ? k / CIN : 13;
    ci[nb] = kok[nb] ? k - tap * CIN : 0;
    tap_off(tap, ox[nb], oy[nb], oz[nb]);
  }
  f32x4 acc[MBW][NBW];
#pragma unroll
  for (int mb = 0; mb < MBW; ++mb)
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int vs = v_begin; vs < v_end; vs += 32) {
    bf16_t ua[MBW][8], ub[NBW][8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int v = vs + kq * 8 + i;
      const bool vok = v < v_end;
      const int z = v % d2, y = (v / d2) % d1, x = v / (d1 * d2);
#pragma unroll
      for (int mb = 0; mb < MBW; ++mb) ua[mb][i] = vok ? g[(size_t)v * COUT + (mb0 + mb) * 16 + row] : (bf16_t)0;
#pragma unroll
      for (int nb = 0; nb < NBW; ++nb) {
        const int xx = x + ox[nb], yy = y + oy[nb], zz = z + oz[nb];
        const bool ok = vok && kok[nb] && (unsigned)xx < (unsigned)d0 && (unsigned)yy < (unsigned)d1 &&
                        (unsigned)zz < (unsigned)d2;
        ub[nb][i] = ok ? in[(size_t)((xx * d1 + yy) * d2 + zz) * CIN + ci[nb]] : (bf16_t)0;
      }
    }
    bf16x8 a[MBW], b[NBW];
#pragma unroll
    for (int mb = 0; mb < MBW; ++mb) {
      u32x4 q;
#pragma unroll
      for (int j = 0; j < 4; ++j) q[j] = (unsigned)ua[mb][2 * j] | ((unsigned)ua[mb][2 * j + 1] << 16);
      a[mb] = __builtin_bit_cast(bf16x8, q);
    }
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb) {
      u32x4 q;
#pragma unroll
      for (int j = 0; j < 4; ++j) q[j] = (unsigned)ub[nb][2 * j] | ((unsigned)ub[nb][2 * j + 1] << 16);
      b[nb] = __builtin_bit_cast(bf16x8, q);
    }
#pragma unroll
    for (int mb = 0; mb < MBW; ++mb)
#pragma unroll
      for (int nb = 0; nb < NBW; ++nb)
        acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mb], b[nb], acc[mb][nb], 0, 0, 0);
  }
#pragma unroll
  for (int nb = 0; nb < NBW; ++nb) {
    if (!kok[nb]) continue;
    const int k = (nb0 + nb) * 16 + row;
#pragma unroll
    for (int mb = 0; mb < MBW; ++mb)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        unsafeAtomicAdd(dW + (size_t)((mb0 + mb) * 16 + kq * 4 + i) * KP + k, acc[mb][nb][i]);
  }
}

// db[co] += sum_v g[v][co] over a voxel chunk per block
template <int COUT>
__global__ __launch_bounds__(kBlk) void k_bias_grad(const bf16_t* __restrict__ g, float* __restrict__ db, int V,
                                                    int chunk) {
  constexpr int R = kBlk / COUT;
  __shared__ float sh[kBlk];
  const int t = threadIdx.x, c = t % COUT, r = t / COUT;
  float s = 0.f;
  if (r < R) {
    const int e = min(V, (int)(blockIdx.x + 1) * chunk);
    for (int v = blockIdx.x * chunk + r; v < e; v += R) s += bf2f(g[(size_t)v * COUT + c]);
  }
  sh[t] = s;
  __syncthreads();
  if (t < COUT) {
    float tot = 0.f;
    for (int rr = 0; rr < R; ++rr) tot += sh[rr * COUT + t];
    unsafeAtomicAdd(db + t, tot);
  }
}

// ------------------------------------------------------------ output layer (16 -> 1)
// y[v] = bias + sum_tap sum_ci W[tap * 16 + ci] d[v + off(tap)][ci], FP32, BF16 weights
__global__ __launch_bounds__(kBlk) void k_out_fwd(const bf16_t* __restrict__ d, const bf16_t* __restrict__ w16,
                                                  const float* __restrict__ bias, float* __restrict__ y, int d0,
                                                  int d1, int d2) {
  __shared__ float sw[432];
  for (int t = threadIdx.x; t < 432; t += kBlk) sw[t] = bf2f(w16[t]);
  __syncthreads();
  const int V = d0 * d1 * d2;
  const int v = blockIdx.x * kBlk + threadIdx.x;
  if (v >= V) return;
  const int z = v % d2, yy = (v / d2) % d1, x = v / (d1 * d2);
  float acc = bias[0];
  for (int tap = 0; tap < 27; ++tap) {
    int ox, oy, oz;
    tap_off(tap, ox, oy, oz);
    const int xx = x + ox, y2 = yy + oy, zz = z + oz;
    if ((unsigned)xx >= (unsigned)d0 || (unsigned)y2 >= (unsigned)d1 || (unsigned)zz >= (unsigned)d2) continue;
    const bf16_t* p = d + (size_t)((xx * d1 + y2) * d2 + zz) * 16;
    const u32x4 q0 = *(const u32x4*)p, q1 = *(const u32x4*)(p + 8);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc += sw[tap * 16 + 2 * j] * bf2f((bf16_t)(q0[j] & 0xffffu));
      acc += sw[tap * 16 + 2 * j + 1] * bf2f((bf16_t)(q0[j] >> 16));
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc += sw[tap * 16 + 8 + 2 * j] * bf2f((bf16_t)(q1[j] & 0xffffu));
      acc += sw[tap * 16 + 8 + 2 * j + 1] * bf2f((bf16_t)(q1[j] >> 16));
    }
  }
  y[v] = acc;
}

// g_d[u][ci] = d[u][ci] > 0 ? sum_tap W[tap][ci] gy[u - off(tap)] : 0
__global__ __launch_bounds__(kBlk) void k_out_bwd_data(const float* __restrict__ gy, const bf16_t* __restrict__ w16,
                                                       const bf16_t* __restrict__ d, bf16_t* __restrict__ gd, int d0,
                                                       int d1, int d2) {
  __shared__ float sw[432];
  for (int t = threadIdx.x; t < 432; t += kBlk) sw[t] = bf2f(w16[t]);
  __syncthreads();
  const int V = d0 * d1 * d2;
  const int u = blockIdx.x * kBlk + threadIdx.x;
  if (u >= V) return;
  const int z = u % d2, yy = (u / d2) % d1, x = u / (d1 * d2);
  float s[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) s[c] = 0.f;
  for (int tap = 0; tap < 27; ++tap) {
    int ox, oy, oz;
    tap_off(tap, ox, oy, oz);
    const int xx = x - ox, y2 = yy - oy, zz = z - oz;
    if ((unsigned)xx >= (unsigned)d0 || (unsigned)y2 >= (unsigned)d1 || (unsigned)zz >= (unsigned)d2) continue;
    const float gv = gy[(xx * d1 + y2) * d2 + zz];
#pragma unroll
    for (int c = 0; c < 16; ++c) s[c] += sw[tap * 16 + c] * gv;
  }
  const bf16_t* dp = d + (size_t)u * 16;
  bf16_t* gp = gd + (size_t)u * 16;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const u32x4 q = *(const u32x4*)(dp + 8 * h);
    u32x4 o;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bf16_t lo = bf_pos((bf16_t)(q[j] & 0xffffu)) ? f2bf(s[8 * h + 2 * j]) : (bf16_t)0;
      const bf16_t hi = bf_pos((bf16_t)(q[j] >> 16)) ? f2bf(s[8 * h + 2 * j + 1]) : (bf16_t)0;
      o[j] = (unsigned)lo | ((unsigned)hi << 16);
    }
    *(u32x4*)(gp + 8 * h) = o;
  }
}

// dW[tap * 16 + ci] += sum_v gy[v] d[v + off(tap)][ci]; thread 432 accumulates db = sum gy
__global__ __launch_bounds__(448) void k_out_wgrad(const float* __restrict__ gy, const bf16_t* __restrict__ d,
                                                   float* __restrict__ dW, float* __restrict__ db, int d0, int d1,
                                                   int d2, int chunk) {
  const int t = threadIdx.x;
  if (t > 432) return;
  const int V = d0 * d1 * d2;
  const int vb = blockIdx.x * chunk, ve = min(V, vb + chunk);
  float s = 0.f;
  if (t == 432) {
    for (int v = vb; v < ve; ++v) s += gy[v];
    unsafeAtomicAdd(db, s);
    return;
  }
  const int tap = t / 16, ci = t % 16;
  int ox, oy, oz;
  tap_off(tap, ox, oy, oz);
  for (int v = vb; v < ve; ++v) {
    const float gv = gy[v];
    if (gv == 0.f) continue;
    const int z = v % d2, y = (v / d2) % d1, x = v / (d1 * d2);
    const int xx = x + ox, yy = y + oy, zz = z + oz;
    if ((unsigned)xx >= (unsigned)d0 || (unsigned)yy >= (unsigned)d1 || (unsigned)zz >= (unsigned)d2) continue;
    s += gv * bf2f(d[(size_t)((xx * d1 + yy) * d2 + zz) * 16 + ci]);
  }
  unsafeAtomicAdd(dW + t, s);
}

// ------------------------------------------------------------ pyramid plumbing
// out[vc][c] = bf16(0.125 * sum of the 8 children), coarse dims c0 x c1 x c2
template <int C>
__global__ __launch_bounds__(kBlk) void k_pool(const bf16_t* __restrict__ in, bf16_t* __restrict__ out, int c0,
                                               int c1, int c2) {
  constexpr int G = C / 8;
  const size_t t = (size_t)blockIdx.x * kBlk + threadIdx.x;
  const size_t Vc = (size_t)c0 * c1 * c2;
  if (t >= Vc * G) return;
  const int g = (int)(t % G);
  const int vc = (int)(t / G);
  const int Z = vc % c2, Y = (vc / c2) % c1, X = vc / (c1 * c2);
  const int f1 = 2 * c1, f2 = 2 * c2;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int ch = 0; ch < 8; ++ch) {
    const int x = 2 * X + (ch >> 2), y = 2 * Y + ((ch >> 1) & 1), z = 2 * Z + (ch & 1);
    const u32x4 q = *(const u32x4*)(in + (size_t)((x * f1 + y) * f2 + z) * C + g * 8);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      s[2 * j] += bf2f((bf16_t)(q[j] & 0xffffu));
      s[2 * j + 1] += bf2f((bf16_t)(q[j] >> 16));
    }
  }
  u32x4 o;
#pragma unroll
  for (int j = 0; j < 4; ++j) o[j] = (unsigned)f2bf(s[2 * j] * 0.125f) | ((unsigned)f2bf(s[2 * j + 1] * 0.125f) << 16);
  *(u32x4*)(out + (size_t)vc * C + g * 8) = o;
}

// out[v] = [coarse[parent(v)] (CU channels), skip[v] (CS channels)], fine dims
template <int CU, int CS>
__global__ __launch_bounds__(kBlk) void k_upcat(const bf16_t* __restrict__ coarse, const bf16_t* __restrict__ skip,
                                                bf16_t* __restrict__ out, int d0, int d1, int d2) {
  constexpr int CT = CU + CS, G = CT / 8;
  const size_t t = (size_t)blockIdx.x * kBlk + threadIdx.x;
  const size_t V = (size_t)d0 * d1 * d2;
  if (t >= V * G) return;
  const int g = (int)(t % G);
  const int v = (int)(t / G);
  const int z = v % d2, y = (v / d2) % d1, x = v / (d1 * d2);
  const bf16_t* src;
  if (g * 8 < CU) src = coarse + (size_t)(((x >> 1) * (d1 >> 1) + (y >> 1)) * (d2 >> 1) + (z >> 1)) * CU + g * 8;
  else src = skip + (size_t)v * CS + (g * 8 - CU);
  *(u32x4*)(out + (size_t)v * CT + g * 8) = *(const u32x4*)src;
}

// gc[vc][c] = bf16(sum over the 8 children of gcat[child][c]), c < CU; coarse dims
template <int CU, int CT>
__global__ __launch_bounds__(kBlk) void k_upcat_bwd(const bf16_t* __restrict__ gcat, bf16_t* __restrict__ gc, int c0,
                                                    int c1, int c2) {
  constexpr int G = CU / 8;
  const size_t t = (size_t)blockIdx.x * kBlk + threadIdx.x;
  const size_t Vc = (size_t)c0 * c1 * c2;
  if (t >= Vc * G) return;
  const int g = (int)(t % G);
  const int vc = (int)(t / G);
  const int Z = vc % c2, Y = (vc / c2) % c1, X = vc / (c1 * c2);
  const int f1 = 2 * c1, f2 = 2 * c2;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int ch = 0; ch < 8; ++ch) {
    const int x = 2 * X + (ch >> 2), y = 2 * Y + ((ch >> 1) & 1), z = 2 * Z + (ch & 1);
    const u32x4 q = *(const u32x4*)(gcat + (size_t)((x * f1 + y) * f2 + z) * CT + g * 8);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      s[2 * j] += bf2f((bf16_t)(q[j] & 0xffffu));
      s[2 * j + 1] += bf2f((bf16_t)(q[j] >> 16));
    }
  }
  u32x4 o;
#pragma unroll
  for (int j = 0; j < 4; ++j) o[j] = (unsigned)f2bf(s[2 * j]) | ((unsigned)f2bf(s[2 * j + 1]) << 16);
  *(u32x4*)(gc + (size_t)vc * CU + g * 8) = o;
}

// gradient of an encoder activation a (C channels) that feeds both the pool below and a
// skip connection: ga = gcat[v][CU + c] + (a > 0 ? gp[parent][c] / 8 : 0); fine dims
template <int C, int CU, int CT>
__global__ __launch_bounds__(kBlk) void k_skip_combine(const bf16_t* __restrict__ gcat, const bf16_t* __restrict__ a,
                                                       const bf16_t* __restrict__ gp, bf16_t* __restrict__ ga, int d0,
                                                       int d1, int d2) {
  constexpr int G = C / 8;
  const size_t t = (size_t)blockIdx.x * kBlk + threadIdx.x;
  const size_t V = (size_t)d0 * d1 * d2;
  if (t >= V * G) return;
  const int g = (int)(t % G);
  const int v = (int)(t / G);
  const int z = v % d2, y = (v / d2) % d1, x = v / (d1 * d2);
  const size_t par = (size_t)(((x >> 1) * (d1 >> 1) + (y >> 1)) * (d2 >> 1) + (z >> 1));
  const u32x4 qs = *(const u32x4*)(gcat + (size_t)v * CT + CU + g * 8);
  const u32x4 qa = *(const u32x4*)(a + (size_t)v * C + g * 8);
  const u32x4 qp = *(const u32x4*)(gp + par * C + g * 8);
  u32x4 o;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    bf16_t r[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int sh = 16 * h;
      float s = bf2f((bf16_t)((qs[j] >> sh) & 0xffffu));
      if (bf_pos((bf16_t)((qa[j] >> sh) & 0xffffu))) s += bf2f((bf16_t)((qp[j] >> sh) & 0xffffu)) * 0.125f;
      r[h] = f2bf(s);
    }
    o[j] = (unsigned)r[0] | ((unsigned)r[1] << 16);
  }
  *(u32x4*)(ga + (size_t)v * C + g * 8) = o;
}

// ------------------------------------------------------------ inputs, loss, optimiser
template <int K>
__device__ inline void block_sum(const double (&v)[K], double* __restrict__ part) {
  __shared__ double sh[K][kBlk];
  for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = kBlk / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < K; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < K; ++k) part[k * gridDim.x + blockIdx.x] = sh[k][0];
}

__device__ inline bool interior(int x, int y, int z, int d0, int d1, int d2) {
  return x >= 1 && x <= d0 - 2 && y >= 1 && y <= d1 - 2 && z >= 1 && z <= d2 - 2;
}

// partials of (sum b^2 over liquid cells, number of liquid cells, number of interior liquid
// cells, sum p^2 over interior liquid cells); p may be null
__global__ __launch_bounds__(kBlk) void k_nn_stats(const float* __restrict__ b, const uint8_t* __restrict__ cls,
                                                   const float* __restrict__ p, double* __restrict__ part, int d0,
                                                   int d1, int d2) {
  const int V = d0 * d1 * d2;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int v = blockIdx.x * kBlk + threadIdx.x; v < V; v += gridDim.x * kBlk) {
    if (!cls[v]) continue;
    const int z = v % d2, y = (v / d2) % d1, x = v / (d1 * d2);
    s[0] += (double)b[v] * (double)b[v];
    s[1] += 1.0;
    if (interior(x, y, z, d0, d1, d2)) {
      s[2] += 1.0;
      if (p) s[3] += (double)p[v] * (double)p[v];
    }
  }
  block_sum<4>(s, part);
}
// sc[0] = 1 / rms(b over liquid), sc[1] = n liquid, sc[2] = n interior liquid (the loss mask),
// sc[4] = 1 / rms(p over the mask) / target_scale
__global__ __launch_bounds__(kBlk) void k_nn_stats_final(const double* __restrict__ part, int nb, float target_scale,
                                                         double* __restrict__ sc) {
  __shared__ double sh[4][kBlk];
  for (int k = 0; k < 4; ++k) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += kBlk) s += part[k * nb + i];
    sh[k][threadIdx.x] = s;
  }
  __syncthreads();
  for (int s = kBlk / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double bb = sh[0][0], n = sh[1][0], nm = sh[2][0], pp = sh[3][0];
    sc[0] = (n > 0.0 && bb > 0.0) ? 1.0 / sqrt(bb / n) : 0.0;
    sc[1] = n;
    sc[2] = nm;
    sc[4] = (nm > 0.0 && pp > 0.0) ? 1.0 / (sqrt(pp / nm) * (double)target_scale) : 0.0;
  }
}

// network input (b / rms(b), cell type), loss mask, and (with p) the training target
// p / rms(p over the mask) / target_scale.  Cell type: +1 liquid, -1 all six faces closed
// (liquidW = 0), 0 otherwise.
__global__ __launch_bounds__(kBlk) void k_nn_input(const float* __restrict__ b, const uint8_t* __restrict__ cls,
                                                   const float* __restrict__ w0, const float* __restrict__ w1,
                                                   const float* __restrict__ w2, const double* __restrict__ sc,
                                                   const float* __restrict__ p, bf16_t* __restrict__ xin,
                                                   float* __restrict__ m, float* __restrict__ t, int d0, int d1,
                                                   int d2) {
  const int V = d0 * d1 * d2;
  const int v = blockIdx.x * kBlk + threadIdx.x;
  if (v >= V) return;
  const int z = v % d2, y = (v / d2) % d1, x = v / (d1 * d2);
  const float inv_s = (float)sc[0];
  float ty, bv = 0.f;
  if (cls[v]) {
    ty = 1.f;
    bv = b[v] * inv_s;
  } else {
    const bool closed = w0[(x * d1 + y) * d2 + z] == 0.f && w0[((x + 1) * d1 + y) * d2 + z] == 0.f &&
                        w1[(x * (d1 + 1) + y) * d2 + z] == 0.f && w1[(x * (d1 + 1) + y + 1) * d2 + z] == 0.f &&
                        w2[(x * d1 + y) * (d2 + 1) + z] == 0.f && w2[(x * d1 + y) * (d2 + 1) + z + 1] == 0.f;
    ty = closed ? -1.f : 0.f;
  }
  const bool mk = cls[v] && interior(x, y, z, d0, d1, d2);
  *(unsigned*)(xin + (size_t)v * 2) = (unsigned)f2bf(bv) | ((unsigned)f2bf(ty) << 16);
  m[v] = mk ? 1.f : 0.f;
  if (p) t[v] = mk ? p[v] * (float)sc[4] : 0.f;
}

__global__ __launch_bounds__(kBlk) void k_f32_to_bf16(const float* __restrict__ in, bf16_t* __restrict__ out,
                                                      size_t n) {
  for (size_t i = blockIdx.x * (size_t)kBlk + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlk)
    out[i] = f2bf(in[i]);
}
__global__ __launch_bounds__(kBlk) void k_bf16_to_f32(const bf16_t* __restrict__ in, float* __restrict__ out,
                                                      size_t n) {
  for (size_t i = blockIdx.x * (size_t)kBlk + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlk)
    out[i] = bf2f(in[i]);
}

// masked MSE: loss = sum m (y - t)^2 / n_mask, gy = 2 m (y - t) / n_mask; partial sums of m (y-t)^2
__global__ __launch_bounds__(kBlk) void k_nn_loss(const float* __restrict__ y, const float* __restrict__ t,
                                                  const float* __restrict__ m, const double* __restrict__ sc,
                                                  float* __restrict__ gy, double* __restrict__ part, int V) {
  const float inv_n = sc[2] > 0.0 ? (float)(1.0 / sc[2]) : 0.f;
  double s[1] = {0.0};
  for (int v = blockIdx.x * kBlk + threadIdx.x; v < V; v += gridDim.x * kBlk) {
    const float d = (y[v] - t[v]) * m[v];
    gy[v] = 2.f * d * inv_n;
    s[0] += (double)d * (double)d;
  }
  block_sum<1>(s, part);
}
__global__ __launch_bounds__(kBlk) void k_nn_loss_final(const double* __restrict__ part, int nb,
                                                        double* __restrict__ sc) {
  __shared__ double sh[kBlk];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += kBlk) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = kBlk / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) sc[3] = sc[2] > 0.0 ? sh[0] / sc[2] : 0.0;
}

// energy loss L = 1 - (y.b)^2 / ((y.Ay)(x*.b)) of y = m * out (e = {y.b, y.Ay, x*.b}):
// gy = m * (-2 (y.b) / ((y.Ay)(x*.b)) b + 2 (y.b)^2 / ((y.Ay)^2 (x*.b)) Ay); sc[3] = L
__global__ __launch_bounds__(kBlk) void k_nn_energy_grad(const float* __restrict__ ay, const float* __restrict__ b,
                                                         const float* __restrict__ m, const double* __restrict__ e,
                                                         double* __restrict__ sc, float* __restrict__ gy, int V) {
  const double yb = e[0], yay = e[1], xb = e[2];
  const bool ok = yay > 0.0 && xb > 0.0;
  const float c1 = ok ? (float)(-2.0 * yb / (yay * xb)) : 0.f;
  const float c2 = ok ? (float)(2.0 * yb * yb / (yay * yay * xb)) : 0.f;
  for (int v = blockIdx.x * kBlk + threadIdx.x; v < V; v += gridDim.x * kBlk) gy[v] = m[v] * (c1 * b[v] + c2 * ay[v]);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[3] = ok ? 1.0 - yb * yb / (yay * xb) : 1.0;
}

__global__ __launch_bounds__(kBlk) void k_mask_mul(const float* __restrict__ y, const float* __restrict__ m,
                                                   float* __restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)kBlk + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlk) out[i] = y[i] * m[i];
}

// Adam (FP32 master weights), then the BF16 copy
__global__ __launch_bounds__(kBlk) void k_adam(float* __restrict__ w, const float* __restrict__ g,
                                               float* __restrict__ m, float* __restrict__ v, bf16_t* __restrict__ w16,
                                               size_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (size_t i = blockIdx.x * (size_t)kBlk + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlk) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    w16[i] = f2bf(wi);
  }
}

// Wt[ci][tap' COUT + co] = W[co][(26 - tap') CIN + ci]: backward data is the forward
// convolution of the output gradient with transposed, spatially flipped weights
__global__ __launch_bounds__(kBlk) void k_wt(const bf16_t* __restrict__ w, bf16_t* __restrict__ wt, int cin, int cout,
                                             int kp, int kpt) {
  const int n = cin * 27 * cout;
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  const int co = i % cout, tap = (i / cout) % 27, ci = i / (27 * cout);
  wt[(size_t)ci * kpt + tap * cout + co] = w[(size_t)co * kp + (26 - tap) * cin + ci];
}

// ------------------------------------------------------------ host helpers
template <typename T>
bool nalloc(Ctx& c, T*& p, size_t n) {
  if (hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
    p = nullptr;
    c.err = "hipMalloc failed (warm-start network)";
    return false;
  }
  return true;
}
template <typename T>
void nfree(T*& p) {
  if (p) hipFree(p);
  p = nullptr;
}
inline unsigned nblocks(size_t n) { return (unsigned)div_up(n, kBlk); }

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double gauss(uint64_t& s) {
  const double u1 = ((splitmix(s) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
  const double u2 = (splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

void free_acts(Net& n) {
  nfree(n.x0); nfree(n.a0); nfree(n.p0); nfree(n.a1); nfree(n.p1); nfree(n.a2); nfree(n.p2); nfree(n.a3);
  nfree(n.c2); nfree(n.d2); nfree(n.c1); nfree(n.d1); nfree(n.c0); nfree(n.d0);
  nfree(n.y); nfree(n.t); nfree(n.m); nfree(n.gy); nfree(n.warm); nfree(n.stage);
  nfree(n.g_d0); nfree(n.g_c0); nfree(n.g_d1); nfree(n.g_c1); nfree(n.g_d2); nfree(n.g_c2); nfree(n.g_a3);
  nfree(n.g_p2); nfree(n.g_a2); nfree(n.g_p1); nfree(n.g_a1); nfree(n.g_p0); nfree(n.g_a0);
  n.dims[0] = n.dims[1] = n.dims[2] = 0;
  n.train_bufs = false;
}

int ensure_acts(Ctx& c, Net& n, const int d[3], bool train) {
  for (int a = 0; a < 3; ++a)
    if (d[a] < 8 || d[a] % 8 != 0) {
      c.err = "warm-start network: grid dimensions must be multiples of 8";
      return LH_ERR_ARG;
    }
  if (n.dims[0] == d[0] && n.dims[1] == d[1] && n.dims[2] == d[2] && (!train || n.train_bufs)) return LH_OK;
  free_acts(n);
  const size_t V0 = (size_t)d[0] * d[1] * d[2], V1 = V0 / 8, V2 = V1 / 8, V3 = V2 / 8;
  bool ok = nalloc(c, n.x0, V0 * 2) && nalloc(c, n.a0, V0 * 16) && nalloc(c, n.p0, V1 * 16) &&
            nalloc(c, n.a1, V1 * 32) && nalloc(c, n.p1, V2 * 32) && nalloc(c, n.a2, V2 * 64) &&
            nalloc(c, n.p2, V3 * 64) && nalloc(c, n.a3, V3 * 96) && nalloc(c, n.c2, V2 * 160) &&
            nalloc(c, n.d2, V2 * 64) && nalloc(c, n.c1, V1 * 96) && nalloc(c, n.d1, V1 * 32) &&
            nalloc(c, n.c0, V0 * 48) && nalloc(c, n.d0, V0 * 16) && nalloc(c, n.y, V0) && nalloc(c, n.t, V0) &&
            nalloc(c, n.m, V0) && nalloc(c, n.warm, V0) && nalloc(c, n.stage, 2 * V0);
  if (ok && train)
    ok = nalloc(c, n.gy, V0) && nalloc(c, n.g_d0, V0 * 16) && nalloc(c, n.g_c0, V0 * 48) &&
         nalloc(c, n.g_d1, V1 * 32) && nalloc(c, n.g_c1, V1 * 96) && nalloc(c, n.g_d2, V2 * 64) &&
         nalloc(c, n.g_c2, V2 * 160) && nalloc(c, n.g_a3, V3 * 96) && nalloc(c, n.g_p2, V3 * 64) &&
         nalloc(c, n.g_a2, V2 * 64) && nalloc(c, n.g_p1, V2 * 32) && nalloc(c, n.g_a1, V1 * 32) &&
         nalloc(c, n.g_p0, V1 * 16) && nalloc(c, n.g_a0, V0 * 16);
  if (!ok) {
    free_acts(n);
    return LH_ERR_OOM;
  }
  for (int a = 0; a < 3; ++a) n.dims[a] = d[a];
  n.train_bufs = train;
  return LH_OK;
}

// rebuild the transposed weights of layers 1..6 from w16
void rebuild_wt(Ctx& c, Net& n) {
  for (int l = 1; l <= 6; ++l) {
    const int cnt = kNnCin[l] * 27 * kNnCout[l];
    hipLaunchKernelGGL(k_wt, dim3(nblocks(cnt)), dim3(kBlk), 0, c.s, n.w16 + n.off[l], n.wt16 + n.wtoff[l], kNnCin[l],
                       kNnCout[l], n.kp[l], n.kpt[l]);
  }
}

// upload FP32 master weights from a host buffer in the padded layout; BF16 copy,
// transposed weights; Adam state reset
int upload_params(Ctx& c, Net& n, const std::vector<float>& hw) {
  std::vector<bf16_t> h16(n.npad);
  for (size_t i = 0; i < n.npad; ++i) h16[i] = f2bf(hw[i]);
  hipMemcpyAsync(n.w32, hw.data(), n.npad * sizeof(float), hipMemcpyHostToDevice, c.s);
  hipMemcpyAsync(n.w16, h16.data(), n.npad * sizeof(bf16_t), hipMemcpyHostToDevice, c.s);
  hipMemsetAsync(n.am, 0, n.npad * sizeof(float), c.s);
  hipMemsetAsync(n.av, 0, n.npad * sizeof(float), c.s);
  rebuild_wt(c, n);
  n.adam_t = 0;
  return hip_ok(c, hipStreamSynchronize(c.s), "warm-start parameters") ? LH_OK : LH_ERR_DEVICE;
}

// logical layout (per layer W[co][tap][ci] then bias[co]) <-> padded layout
void logical_to_padded(const Net& n, const float* lg, std::vector<float>& pd) {
  pd.assign(n.npad, 0.f);
  size_t q = 0;
  for (int l = 0; l < kNnLayers; ++l) {
    const int kr = 27 * kNnCin[l];
    for (int co = 0; co < kNnCout[l]; ++co)
      for (int k = 0; k < kr; ++k) pd[n.off[l] + (size_t)co * n.kp[l] + k] = lg[q++];
    for (int co = 0; co < kNnCout[l]; ++co) pd[n.boff[l] + co] = lg[q++];
  }
}
void padded_to_logical(const Net& n, const std::vector<float>& pd, float* lg) {
  size_t q = 0;
  for (int l = 0; l < kNnLayers; ++l) {
    const int kr = 27 * kNnCin[l];
    for (int co = 0; co < kNnCout[l]; ++co)
      for (int k = 0; k < kr; ++k) lg[q++] = pd[n.off[l] + (size_t)co * n.kp[l] + k];
    for (int co = 0; co < kNnCout[l]; ++co) lg[q++] = pd[n.boff[l] + co];
  }
}

#define CONV(CI, CO, NVB, EPI, IN, WP, BIAS, ACT, OUT, D)                                                          \
  hipLaunchKernelGGL((k_conv<CI, CO, NVB, EPI>), dim3((unsigned)div_up((size_t)(D)[0] * (D)[1] * (D)[2], 64 * (NVB))), \
                     dim3(kBlk), 0, c.s, IN, WP, BIAS, ACT, OUT, (D)[0], (D)[1], (D)[2])
// the LDS-tiled form: TX x TY x 16 voxels per workgroup, the grid 8 x the per-XCD run of tiles
#define CONVT(CI, CO, TX, TY, KC, EPI, IN, WP, BIAS, ACT, OUT, D)                                                \
  hipLaunchKernelGGL((k_conv_tile<CI, CO, TX, TY, KC, EPI>),                                                     \
                     dim3((unsigned)(8 * div_up(div_up((D)[0], TX) * div_up((D)[1], TY) * div_up((D)[2], 16), 8))), \
                     dim3(256), 0, c.s, IN, WP, BIAS, ACT, OUT, (D)[0], (D)[1], (D)[2])

void forward(Ctx& c, Net& n) {
  ProfScope ps(c, LH_PROF_NN_FWD);
  int D[4][3];
  for (int a = 0; a < 3; ++a)
    for (int l = 0; l < 4; ++l) D[l][a] = n.dims[a] >> l;
  const float* w = n.w32;
  const bf16_t* h = n.w16;
  CONV(2, 16, 4, 0, n.x0, h + n.off[0], w + n.boff[0], nullptr, n.a0, D[0]);
  hipLaunchKernelGGL(k_pool<16>, dim3(nblocks((size_t)D[1][0] * D[1][1] * D[1][2] * 2)), dim3(kBlk), 0, c.s, n.a0, n.p0,
                     D[1][0], D[1][1], D[1][2]);
  CONVT(16, 32, 4, 4, 128, 0, n.p0, h + n.off[1], w + n.boff[1], nullptr, n.a1, D[1]);
  hipLaunchKernelGGL(k_pool<32>, dim3(nblocks((size_t)D[2][0] * D[2][1] * D[2][2] * 4)), dim3(kBlk), 0, c.s, n.a1, n.p1,
                     D[2][0], D[2][1], D[2][2]);
  CONVT(32, 64, 4, 4, 64, 0, n.p1, h + n.off[2], w + n.boff[2], nullptr, n.a2, D[2]);
  hipLaunchKernelGGL(k_pool<64>, dim3(nblocks((size_t)D[3][0] * D[3][1] * D[3][2] * 8)), dim3(kBlk), 0, c.s, n.a2, n.p2,
                     D[3][0], D[3][1], D[3][2]);
  CONVT(64, 96, 2, 2, 64, 0, n.p2, h + n.off[3], w + n.boff[3], nullptr, n.a3, D[3]);
  hipLaunchKernelGGL((k_upcat<96, 64>), dim3(nblocks((size_t)D[2][0] * D[2][1] * D[2][2] * 20)), dim3(kBlk), 0, c.s,
                     n.a3, n.a2, n.c2, D[2][0], D[2][1], D[2][2]);
  CONVT(160, 64, 2, 2, 128, 0, n.c2, h + n.off[4], w + n.boff[4], nullptr, n.d2, D[2]);
  hipLaunchKernelGGL((k_upcat<64, 32>), dim3(nblocks((size_t)D[1][0] * D[1][1] * D[1][2] * 12)), dim3(kBlk), 0, c.s,
                     n.d2, n.a1, n.c1, D[1][0], D[1][1], D[1][2]);
  CONVT(96, 32, 2, 2, 128, 0, n.c1, h + n.off[5], w + n.boff[5], nullptr, n.d1, D[1]);
  hipLaunchKernelGGL((k_upcat<32, 16>), dim3(nblocks((size_t)D[0][0] * D[0][1] * D[0][2] * 6)), dim3(kBlk), 0, c.s,
                     n.d1, n.a0, n.c0, D[0][0], D[0][1], D[0][2]);
  {
    ProfScope pd(c, LH_PROF_CONV_D0);
    CONVT(48, 16, 4, 4, 128, 0, n.c0, h + n.off[6], w + n.boff[6], nullptr, n.d0, D[0]);
  }
  hipLaunchKernelGGL(k_out_fwd, dim3(nblocks((size_t)D[0][0] * D[0][1] * D[0][2])), dim3(kBlk), 0, c.s, n.d0,
                     h + n.off[7], w + n.boff[7], n.y, D[0][0], D[0][1], D[0][2]);
}

inline int wchunk(size_t V) { return (int)std::max<size_t>(4096, div_up(V, 256) / 32 * 32 + 32); }

#define WGRAD(CI, CO, MBW, NBW, G, IN, L, D)                                                                       \
  do {                                                                                                            \
    const size_t Vw = (size_t)(D)[0] * (D)[1] * (D)[2];                                                          \
    const int ch = wchunk(Vw);                                                                                    \
    constexpr int nbt = (27 * (CI) + 15) / 16;                                                                    \
    hipLaunchKernelGGL((k_conv_wgrad<CI, CO, MBW, NBW>), dim3((unsigned)div_up(Vw, ch), (nbt + (NBW)-1) / (NBW), \
                                                              (CO) / 16 / (MBW)),                                 \
                       dim3(64), 0, c.s, G, IN, n.g32 + n.off[L], (D)[0], (D)[1], (D)[2], ch);                    \
    hipLaunchKernelGGL(k_bias_grad<CO>, dim3((unsigned)div_up(Vw, 4096)), dim3(kBlk), 0, c.s, G, n.g32 + n.boff[L], \
                       (int)Vw, 4096);                                                                            \
  } while (0)

// gradient of the loss already formed in n.gy; accumulates into n.g32 (zeroed here)
void backward(Ctx& c, Net& n) {
  ProfScope ps(c, LH_PROF_NN_BWD);
  int D[4][3];
  for (int a = 0; a < 3; ++a)
    for (int l = 0; l < 4; ++l) D[l][a] = n.dims[a] >> l;
  size_t Vl[4];
  for (int l = 0; l < 4; ++l) Vl[l] = (size_t)D[l][0] * D[l][1] * D[l][2];
  const bf16_t* wt = n.wt16;
  hipMemsetAsync(n.g32, 0, n.npad * sizeof(float), c.s);
  // OUT (16 -> 1)
  {
    const int ch = 1024;
    hipLaunchKernelGGL(k_out_wgrad, dim3((unsigned)div_up(Vl[0], ch)), dim3(448), 0, c.s, n.gy, n.d0,
                       n.g32 + n.off[7], n.g32 + n.boff[7], D[0][0], D[0][1], D[0][2], ch);
    hipLaunchKernelGGL(k_out_bwd_data, dim3(nblocks(Vl[0])), dim3(kBlk), 0, c.s, n.gy, n.w16 + n.off[7], n.d0, n.g_d0,
                       D[0][0], D[0][1], D[0][2]);
  }
  // D0 (48 -> 16) at level 0
  WGRAD(48, 16, 1, 4, n.g_d0, n.c0, 6, D[0]);
  CONVT(16, 48, 4, 4, 128, 1, n.g_d0, wt + n.wtoff[6], nullptr, n.c0, n.g_c0, D[0]);
  hipLaunchKernelGGL((k_upcat_bwd<32, 48>), dim3(nblocks(Vl[1] * 4)), dim3(kBlk), 0, c.s, n.g_c0, n.g_d1, D[1][0],
                     D[1][1], D[1][2]);
  // D1 (96 -> 32) at level 1
  WGRAD(96, 32, 2, 4, n.g_d1, n.c1, 5, D[1]);
  CONVT(32, 96, 4, 4, 64, 1, n.g_d1, wt + n.wtoff[5], nullptr, n.c1, n.g_c1, D[1]);
  hipLaunchKernelGGL((k_upcat_bwd<64, 96>), dim3(nblocks(Vl[2] * 8)), dim3(kBlk), 0, c.s, n.g_c1, n.g_d2, D[2][0],
                     D[2][1], D[2][2]);
  // D2 (160 -> 64) at level 2
  WGRAD(160, 64, 2, 4, n.g_d2, n.c2, 4, D[2]);
  CONVT(64, 160, 2, 2, 64, 1, n.g_d2, wt + n.wtoff[4], nullptr, n.c2, n.g_c2, D[2]);
  hipLaunchKernelGGL((k_upcat_bwd<96, 160>), dim3(nblocks(Vl[3] * 12)), dim3(kBlk), 0, c.s, n.g_c2, n.g_a3, D[3][0],
                     D[3][1], D[3][2]);
  // E3 (64 -> 96) at level 3
  WGRAD(64, 96, 2, 4, n.g_a3, n.p2, 3, D[3]);
  CONVT(96, 64, 2, 2, 64, 2, n.g_a3, wt + n.wtoff[3], nullptr, nullptr, n.g_p2, D[3]);
  hipLaunchKernelGGL((k_skip_combine<64, 96, 160>), dim3(nblocks(Vl[2] * 8)), dim3(kBlk), 0, c.s, n.g_c2, n.a2, n.g_p2,
                     n.g_a2, D[2][0], D[2][1], D[2][2]);
  // E2 (32 -> 64) at level 2
  WGRAD(32, 64, 2, 4, n.g_a2, n.p1, 2, D[2]);
  CONVT(64, 32, 2, 4, 64, 2, n.g_a2, wt + n.wtoff[2], nullptr, nullptr, n.g_p1, D[2]);
  hipLaunchKernelGGL((k_skip_combine<32, 64, 96>), dim3(nblocks(Vl[1] * 4)), dim3(kBlk), 0, c.s, n.g_c1, n.a1, n.g_p1,
                     n.g_a1, D[1][0], D[1][1], D[1][2]);
  // E1 (16 -> 32) at level 1
  WGRAD(16, 32, 2, 4, n.g_a1, n.p0, 1, D[1]);
  CONVT(32, 16, 4, 4, 128, 2, n.g_a1, wt + n.wtoff[1], nullptr, nullptr, n.g_p0, D[1]);
  hipLaunchKernelGGL((k_skip_combine<16, 32, 48>), dim3(nblocks(Vl[0] * 2)), dim3(kBlk), 0, c.s, n.g_c0, n.a0, n.g_p0,
                     n.g_a0, D[0][0], D[0][1], D[0][2]);
  // E0 (2 -> 16) at level 0
  WGRAD(2, 16, 1, 4, n.g_a0, n.x0, 0, D[0]);
}

void loss(Ctx& c, Net& n) {
  const size_t V = (size_t)n.dims[0] * n.dims[1] * n.dims[2];
  const int nb = (int)std::min<size_t>(1024, nblocks(V));
  hipLaunchKernelGGL(k_nn_loss, dim3(nb), dim3(kBlk), 0, c.s, n.y, n.t, n.m, n.sc, n.gy, n.part, (int)V);
  hipLaunchKernelGGL(k_nn_loss_final, dim3(1), dim3(kBlk), 0, c.s, n.part, nb, n.sc);
}

void adam(Ctx& c, Net& n, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  n.adam_t++;
  const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, n.adam_t)));
  const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, n.adam_t)));
  hipLaunchKernelGGL(k_adam, dim3((unsigned)std::min<size_t>(4096, nblocks(n.npad))), dim3(kBlk), 0, c.s, n.w32, n.g32,
                     n.am, n.av, n.w16, n.npad, lr, b1, b2, eps, c1, c2);
  rebuild_wt(c, n);
}

// network input, loss mask and (with_target) training target from the built operator of c
// and the right-hand side b (device, level-0 layout)
void prep_from_ctx(Ctx& c, Net& n, const float* b, bool with_target) {
  const int* d = c.g.N;
  const size_t V = c.ncell;
  const int nb = (int)std::min<size_t>(1024, nblocks(V));
  const uint8_t* cls = c.lv[0].cls;
  const float* p = with_target ? c.pressure : nullptr;
  hipLaunchKernelGGL(k_nn_stats, dim3(nb), dim3(kBlk), 0, c.s, b, cls, p, n.part, d[0], d[1], d[2]);
  hipLaunchKernelGGL(k_nn_stats_final, dim3(1), dim3(kBlk), 0, c.s, n.part, nb, n.target_scale, n.sc);
  hipLaunchKernelGGL(k_nn_input, dim3(nblocks(V)), dim3(kBlk), 0, c.s, b, cls, c.lw[0], c.lw[1], c.lw[2], n.sc, p,
                     n.x0, n.m, n.t, d[0], d[1], d[2]);
}

int check_net(Ctx& c) {
  if (!c.net) {
    c.err = "no warm-start network: call lh_warmstart_init or lh_warmstart_load first";
    return LH_ERR_ARG;
  }
  return LH_OK;
}

constexpr char kMagic[8] = {'L', 'H', 'W', 'S', 'N', 'E', 'T', '1'};

}  // namespace

void nn_free(Net* n) {
  if (!n) return;
  free_acts(*n);
  nfree(n->w32); nfree(n->g32); nfree(n->am); nfree(n->av); nfree(n->w16); nfree(n->wt16);
  nfree(n->part); nfree(n->sc);
  delete n;
}

int nn_create(Ctx& c, uint64_t seed) {
  nn_free(c.net);
  c.net = nullptr;
  Net* n = new Net();
  size_t o = 0, ot = 0, nl = 0;
  for (int l = 0; l < kNnLayers; ++l) {
    n->kp[l] = kpad32(27 * kNnCin[l]);
    n->kpt[l] = kpad32(27 * kNnCout[l]);
    n->off[l] = o;
    n->boff[l] = o + (size_t)kNnCout[l] * n->kp[l];
    o = n->boff[l] + (size_t)(kNnCout[l] + 31) / 32 * 32;
    n->wtoff[l] = ot;
    ot += (size_t)kNnCin[l] * n->kpt[l];
    nl += (size_t)kNnCout[l] * 27 * kNnCin[l] + kNnCout[l];
  }
  n->npad = o;
  n->nwt = ot;
  n->nlog = nl;
  bool ok = nalloc(c, n->w32, o) && nalloc(c, n->g32, o) && nalloc(c, n->am, o) && nalloc(c, n->av, o) &&
            nalloc(c, n->w16, o) && nalloc(c, n->wt16, ot) && nalloc(c, n->part, 4 * 1024) && nalloc(c, n->sc, 16);
  if (!ok) {
    nn_free(n);
    return LH_ERR_OOM;
  }
  hipMemsetAsync(n->wt16, 0, ot * sizeof(bf16_t), c.s);
  hipMemsetAsync(n->sc, 0, 16 * sizeof(double), c.s);
  // He-normal weights, zero biases
  std::vector<float> hw(o, 0.f);
  uint64_t s = seed * 0x2545F4914F6CDD1Dull + 0x1234567ull;
  for (int l = 0; l < kNnLayers; ++l) {
    const double std = std::sqrt((l == kNnLayers - 1 ? 1.0 : 2.0) / (27.0 * kNnCin[l]));
    for (int co = 0; co < kNnCout[l]; ++co)
      for (int k = 0; k < 27 * kNnCin[l]; ++k) hw[n->off[l] + (size_t)co * n->kp[l] + k] = (float)(std * gauss(s));
  }
  c.net = n;
  return upload_params(c, *n, hw);
}

int nn_get_params(Ctx& c, float* host, int64_t count, bool grads) {
  if (int rc = check_net(c)) return rc;
  Net& n = *c.net;
  if (!host || count != (int64_t)n.nlog) {
    c.err = "warm-start parameters: bad buffer size";
    return LH_ERR_ARG;
  }
  std::vector<float> pd(n.npad);
  hipMemcpyAsync(pd.data(), grads ? n.g32 : n.w32, n.npad * sizeof(float), hipMemcpyDeviceToHost, c.s);
  if (!hip_ok(c, hipStreamSynchronize(c.s), "warm-start parameters")) return LH_ERR_DEVICE;
  padded_to_logical(n, pd, host);
  return LH_OK;
}

int nn_set_params(Ctx& c, const float* host, int64_t count) {
  if (int rc = check_net(c)) return rc;
  Net& n = *c.net;
  if (!host || count != (int64_t)n.nlog) {
    c.err = "warm-start parameters: bad buffer size";
    return LH_ERR_ARG;
  }
  std::vector<float> pd;
  logical_to_padded(n, host, pd);
  return upload_params(c, n, pd);
}

int nn_save(Ctx& c, const char* path) {
  if (int rc = check_net(c)) return rc;
  Net& n = *c.net;
  std::vector<float> lg(n.nlog);
  if (int rc = nn_get_params(c, lg.data(), (int64_t)n.nlog, false)) return rc;
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    c.err = std::string("cannot open ") + path;
    return LH_ERR_ARG;
  }
  const int32_t nl = kNnLayers;
  const int64_t np = (int64_t)n.nlog;
  bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(&nl, 4, 1, f) == 1 &&
            std::fwrite(kNnCin, 4, kNnLayers, f) == (size_t)kNnLayers &&
            std::fwrite(kNnCout, 4, kNnLayers, f) == (size_t)kNnLayers &&
            std::fwrite(&n.target_scale, 4, 1, f) == 1 && std::fwrite(&np, 8, 1, f) == 1 &&
            std::fwrite(lg.data(), 4, lg.size(), f) == lg.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    c.err = std::string("write failed: ") + path;
    return LH_ERR_ARG;
  }
  return LH_OK;
}

int nn_load(Ctx& c, const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    c.err = std::string("cannot open ") + path;
    return LH_ERR_ARG;
  }
  char mg[8];
  int32_t nl = 0, ci[kNnLayers], co[kNnLayers];
  float ts = 1.f;
  int64_t np = 0;
  bool ok = std::fread(mg, 1, 8, f) == 8 && std::memcmp(mg, kMagic, 8) == 0 && std::fread(&nl, 4, 1, f) == 1 &&
            nl == kNnLayers && std::fread(ci, 4, kNnLayers, f) == (size_t)kNnLayers &&
            std::fread(co, 4, kNnLayers, f) == (size_t)kNnLayers && std::fread(&ts, 4, 1, f) == 1 &&
            std::fread(&np, 8, 1, f) == 1;
  for (int l = 0; ok && l < kNnLayers; ++l) ok = ci[l] == kNnCin[l] && co[l] == kNnCout[l];
  std::vector<float> lg;
  if (ok && np > 0 && np < (1 << 26)) {
    lg.resize((size_t)np);
    ok = std::fread(lg.data(), 4, lg.size(), f) == lg.size();
  } else {
    ok = false;
  }
  std::fclose(f);
  if (!ok) {
    c.err = std::string("not a warm-start network file of this architecture: ") + path;
    return LH_ERR_ARG;
  }
  if (int rc = nn_create(c, 0)) return rc;
  if (np != (int64_t)c.net->nlog) {
    c.err = "warm-start network file: parameter count mismatch";
    return LH_ERR_ARG;
  }
  c.net->target_scale = ts;
  return nn_set_params(c, lg.data(), np);
}

namespace {

// host (V x 2 floats, channels last) -> x0 (BF16)
void upload_input(Ctx& c, Net& n, const float* in, size_t V) {
  hipMemcpyAsync(n.stage, in, V * 2 * sizeof(float), hipMemcpyHostToDevice, c.s);
  hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)std::min<size_t>(4096, nblocks(2 * V))), dim3(kBlk), 0, c.s,
                     n.stage, n.x0, 2 * V);
}


}  // namespace

int nn_forward_host(Ctx& c, const int dims[3], const float* in, float* y) {
  if (int rc = check_net(c)) return rc;
  Net& n = *c.net;
  if (!in || !y) {
    c.err = "warm-start forward: null buffer";
    return LH_ERR_ARG;
  }
  if (int rc = ensure_acts(c, n, dims, false)) return rc;
  const size_t V = (size_t)dims[0] * dims[1] * dims[2];
  upload_input(c, n, in, V);
  forward(c, n);
  hipMemcpyAsync(y, n.y, V * sizeof(float), hipMemcpyDeviceToHost, c.s);
  return hip_ok(c, hipStreamSynchronize(c.s), "warm-start forward") ? LH_OK : LH_ERR_DEVICE;
}

int nn_step_host(Ctx& c, const int dims[3], const float* in, const float* target, const float* mask, float lr,
                 float* grad, float* loss_out) {
  if (int rc = check_net(c)) return rc;
  Net& n = *c.net;
  if (!in || !target || !mask) {
    c.err = "warm-start gradient: null buffer";
    return LH_ERR_ARG;
  }
  if (int rc = ensure_acts(c, n, dims, true)) return rc;
  const size_t V = (size_t)dims[0] * dims[1] * dims[2];
  double nm = 0.0;
  for (size_t i = 0; i < V; ++i) nm += mask[i];
  upload_input(c, n, in, V);
  hipMemcpyAsync(n.t, target, V * sizeof(float), hipMemcpyHostToDevice, c.s);
  hipMemcpyAsync(n.m, mask, V * sizeof(float), hipMemcpyHostToDevice, c.s);
  hipMemcpyAsync(n.sc + 2, &nm, sizeof(double), hipMemcpyHostToDevice, c.s);
  forward(c, n);
  loss(c, n);
  backward(c, n);
  if (lr > 0.f) adam(c, n, lr);
  double l = 0.0;
  hipMemcpyAsync(&l, n.sc + 3, sizeof(double), hipMemcpyDeviceToHost, c.s);
  if (!hip_ok(c, hipStreamSynchronize(c.s), "warm-start gradient")) return LH_ERR_DEVICE;
  if (loss_out) *loss_out = (float)l;
  if (grad) return nn_get_params(c, grad, (int64_t)n.nlog, true);
  return LH_OK;
}

int nn_train_last_solve(Ctx& c, float lr, float* loss_out) {
  if (int rc = check_net(c)) return rc;
  Net& n = *c.net;
  if (c.last_dt <= 0.f || c.lv.empty()) {
    c.err = "warm-start training: no pressure system has been built";
    return LH_ERR_ARG;
  }
  if (int rc = ensure_acts(c, n, c.g.N, true)) return rc;
  prep_from_ctx(c, n, c.rhs, true);
  forward(c, n);
  if (n.loss_mode == 1) {
    const size_t V = c.ncell;
    const unsigned nbk = (unsigned)std::min<size_t>(4096, nblocks(V));
    hipLaunchKernelGGL(k_mask_mul, dim3(nbk), dim3(kBlk), 0, c.s, n.y, n.m, n.warm, V);
    launch_energy_terms(c, n.warm, c.rhs, c.pressure, n.t, n.sc + 5);
    hipLaunchKernelGGL(k_nn_energy_grad, dim3(nbk), dim3(kBlk), 0, c.s, n.t, c.rhs, n.m, n.sc + 5, n.sc, n.gy, (int)V);
  } else {
    loss(c, n);
  }
  backward(c, n);
  if (lr > 0.f) adam(c, n, lr);
  double l = 0.0;
  hipMemcpyAsync(&l, n.sc + 3, sizeof(double), hipMemcpyDeviceToHost, c.s);
  if (!hip_ok(c, hipStreamSynchronize(c.s), "warm-start training")) return LH_ERR_DEVICE;
  if (loss_out) *loss_out = (float)l;
  return LH_OK;
}

int nn_sample_last_solve(Ctx& c, float* in, float* target, float* mask) {
  if (int rc = check_net(c)) return rc;
  Net& n = *c.net;
  if (c.last_dt <= 0.f || c.lv.empty()) {
    c.err = "warm-start sample: no pressure system has been built";
    return LH_ERR_ARG;
  }
  if (int rc = ensure_acts(c, n, c.g.N, false)) return rc;
  const size_t V = c.ncell;
  prep_from_ctx(c, n, c.rhs, true);
  hipLaunchKernelGGL(k_bf16_to_f32, dim3((unsigned)std::min<size_t>(4096, nblocks(2 * V))), dim3(kBlk), 0, c.s, n.x0,
                     n.stage, 2 * V);
  if (in) hipMemcpyAsync(in, n.stage, 2 * V * sizeof(float), hipMemcpyDeviceToHost, c.s);
  if (target) hipMemcpyAsync(target, n.t, V * sizeof(float), hipMemcpyDeviceToHost, c.s);
  if (mask) hipMemcpyAsync(mask, n.m, V * sizeof(float), hipMemcpyDeviceToHost, c.s);
  return hip_ok(c, hipStreamSynchronize(c.s), "warm-start sample") ? LH_OK : LH_ERR_DEVICE;
}

int nn_predict(Ctx& c, const float* b, float** x0) {
  *x0 = nullptr;
  if (int rc = check_net(c)) return rc;
  Net& n = *c.net;
  if (int rc = ensure_acts(c, n, c.g.N, n.train_bufs)) return rc;
  const size_t V = c.ncell;
  prep_from_ctx(c, n, b, false);
  forward(c, n);
  hipLaunchKernelGGL(k_mask_mul, dim3((unsigned)std::min<size_t>(4096, nblocks(V))), dim3(kBlk), 0, c.s, n.y, n.m,
                     n.warm, V);
  launch_warm_ritz(c, n.warm, b);
  *x0 = n.warm;
  return LH_OK;
}

}  // namespace lh
