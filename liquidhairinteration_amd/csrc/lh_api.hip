// lh_api.hip -- the C ABI of include/lh_api.h: scene construction from the
// reference's JSON config (Scene.cpp:13-20, SceneParameter.h:9-152), device
// allocation (Grid.cpp:86-98, PressureSolver.cpp:422-447), initial particles
// (LiquidModel.cpp:506-512, ParticleSystem.cpp:44-64) and the step driver
// (LiquidModel.cpp:514-537, main.cpp:47-53).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "lh_ctx.h"
#include "lh_json.h"
#include "lh_math.h"
#include "lh_nn.h"

using lh::Ctx;

struct lh_ctx {
  Ctx c;
};

namespace lh {
namespace {

// ------------------------------------------------------------------ config
void parse_vec3(const Json& j, float out[3]) {
  for (int a = 0; a < 3; ++a) out[a] = j.at((size_t)a).as_float();
}

void emit_sdf(const Json& n, std::vector<float>& prog) {
  const std::string& t = n.at("type").as_string();
  if (t == "Rect" || t == "Sphere") {
    float c[3] = {0, 0, 0}, h[3] = {0, 0, 0};
    if (n.has("center")) parse_vec3(n.at("center"), c);
    if (n.has("halfExtent")) parse_vec3(n.at("halfExtent"), h);
    float rad = n.has("radius") ? n.at("radius").as_float() : 0.f;
    float ins = (n.has("useInside") && n.at("useInside").as_bool()) ? 1.f : 0.f;
    float node[12] = {t == "Rect" ? 0.f : 1.f, c[0], c[1], c[2], h[0], h[1], h[2], rad, ins, 0, 0, 0};
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

Config parse_config(const std::string& text) {
  Json j = Json::parse(text);
  Config c;
  // SolverParameter (SceneParameter.h:20-44)
  const Json& s = j.at("SolverParameter");
  c.max_iter = s.at("maxIterations").as_int();
  c.tol = s.at("tolerance").as_float();
  c.num_level = s.at("numLevel").as_int();
  c.num_pre = s.at("numPreSmooth").as_int();
  c.num_post = s.at("numPostSmooth").as_int();
  c.num_final = s.at("numFinalSmooth").as_int();
  c.omega = s.at("smoothingFactor").as_float();
  const std::string& pt = s.at("preconditionerType").as_string();
  if (pt == "None") c.precond = LH_PRECOND_NONE;
  else if (pt == "Jacobi") c.precond = LH_PRECOND_JACOBI;
  else if (pt == "Multigrid") c.precond = LH_PRECOND_MULTIGRID;
  else throw std::runtime_error("Unknown preconditioner type");
  std::string tm = s.get("toleranceMode", "absolute");
  if (tm == "absolute") c.tol_mode = LH_TOL_ABSOLUTE;
  else if (tm == "relative") c.tol_mode = LH_TOL_RELATIVE;
  else throw std::runtime_error("Unknown toleranceMode " + tm);
  std::string sm = s.get("smoother", "reference");
  if (sm == "reference") c.smoother = LH_SMOOTHER_REFERENCE;
  else if (sm == "symmetric") c.smoother = LH_SMOOTHER_SYMMETRIC;
  else throw std::runtime_error("Unknown smoother " + sm);
  c.fused_sweep_min_cells = (long long)s.get("fusedSweepMinCells", (double)(1LL << 62));
  std::string cs = s.get("coarseSweep", "rows");
  if (cs == "rows") c.coarse_sweep = kCoarseRows;
  else if (cs == "tiles") c.coarse_sweep = kCoarseTiles;
  else if (cs == "passes") c.coarse_sweep = kCoarsePasses;
  else throw std::runtime_error("Unknown coarseSweep " + cs);
  c.row_sweep_max_cells = (long long)s.get("rowSweepMaxCells", (double)(1LL << 18));
  // LiquidParameter (SceneParameter.h:55-63)
  const Json& lq = j.at("LiquidParameter");
  c.density = lq.at("density").as_float();
  c.viscosity = lq.at("viscosity").as_float();
  c.num_particles = lq.at("numParticles").as_int();
  c.surface_tension = lq.at("surfaceTension").as_float();
  c.correct_cycle = lq.at("correctCycle").as_int();
  // SceneParameter (SceneParameter.h:139-151)
  c.cfg_dt = j.at("dt").as_float();
  c.gravity = j.at("gravity").as_float();
  const Json& gd = j.at("gridDimension");
  const Json& mn = j.at("minCoord");
  const Json& mx = j.at("maxCoord");
  for (int a = 0; a < 3; ++a) {
    c.N[a] = gd.at((size_t)a).as_int();
    c.mn[a] = mn.at((size_t)a).as_float();
    c.mx[a] = mx.at((size_t)a).as_float();
  }
  // build extensions (all optional; defaults reproduce the reference)
  if (j.has("Boundary")) emit_sdf(j.at("Boundary"), c.sdf);
  else {
    // Scene.cpp:18: OperatorBoundary(None) over RectBoundary(0, 0.9, useInside)
    float rect[12] = {0, 0, 0, 0, 0.9f, 0.9f, 0.9f, 0, 1.f, 0, 0, 0};
    float none[12] = {13, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    c.sdf.assign(rect, rect + 12);
    c.sdf.insert(c.sdf.end(), none, none + 12);
  }
  if (c.sdf.size() / 12 > (size_t)kMaxSdfNodes) throw std::runtime_error("Boundary tree too large");
  if (j.has("FluidBlocks")) {
    const Json& bl = j.at("FluidBlocks");
    for (size_t b = 0; b < bl.size(); ++b) {
      const Json& B = bl.at(b);
      float lo[3], hi[3], v[3] = {0, 0, 0};
      parse_vec3(B.at("min"), lo);
      parse_vec3(B.at("max"), hi);
      if (B.has("velocity")) parse_vec3(B.at("velocity"), v);
      float sig = B.has("velocityNoise") ? B.at("velocityNoise").as_float() : 0.f;
      float rec[10] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], v[0], v[1], v[2], sig};
      c.blocks.insert(c.blocks.end(), rec, rec + 10);
    }
  }
  c.k_ppc = (int)j.get("particlesPerCellAxis", 2.0);
  c.seed = (uint64_t)j.get("seed", 0.0);
  if (j.has("TankAcceleration")) {
    const Json& t = j.at("TankAcceleration");
    c.tank_amp = (float)t.get("amplitude", 0.0);
    c.tank_freq = (float)t.get("frequency", 1.0);
    c.tank_axis = (int)t.get("axis", 0.0);
  }
  std::string prec = j.get("precision", "fp32");
  if (prec == "refined") c.refined = 1;
  else if (prec != "fp32") throw std::runtime_error("Unknown precision " + prec);
  c.refined_tol = j.get("refinedTolerance", 1e-9);
  std::string cp = s.get("coarsestPrecision", c.refined ? "fp64" : "fp32");
  if (cp == "fp64") c.coarsest_fp64 = 1;
  else if (cp != "fp32") throw std::runtime_error("Unknown coarsestPrecision " + cp);
  std::string csol = s.get("coarsestSolver", "sweeps");
  if (csol == "chebyshev") {
    const double steps = s.get("coarsestSteps", 12.0), lo = s.get("chebyshevLower", 0.01);
    if (!(steps >= 1 && steps <= 1000) || steps != (double)(int)steps)
      throw std::runtime_error("coarsestSteps must be an integer in [1, 1000]");
    if (!(lo > 0.0 && lo < 2.0)) throw std::runtime_error("chebyshevLower must lie in (0, 2)");
    if (!c.coarsest_fp64) throw std::runtime_error("coarsestSolver chebyshev needs coarsestPrecision fp64");
    c.coarsest_cheb = (int)steps;
    c.cheb_lo = (float)lo;
  } else if (csol != "sweeps") {
    throw std::runtime_error("Unknown coarsestSolver " + csol);
  }
  c.pcg_graph = (s.has("pcgGraph") && s.at("pcgGraph").as_bool()) ? 1 : 0;
  std::string pf = j.get("p2gForm", "shared");
  if (pf == "gather") c.p2g_gather = 1;
  else if (pf != "shared") throw std::runtime_error("Unknown p2gForm " + pf);
  c.first_dt = (float)j.get("firstDt", 0.01);
  c.correct_volume = (j.has("correctVolume") && j.at("correctVolume").as_bool()) ? 1 : 0;
  c.steps = (int)j.get("steps", 0.0);
  for (int a = 0; a < 3; ++a)
    if (c.N[a] < 3) throw std::runtime_error("gridDimension must be >= 3 on every axis");
  return c;
}

// ------------------------------------------------------------------ memory
template <class T>
bool dalloc(Ctx& c, T*& p, size_t n) {
  p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)&p, n * sizeof(T));
  if (e != hipSuccess) {
    c.err = std::string("hipMalloc failed: ") + hipGetErrorString(e);
    p = nullptr;
    return false;
  }
  return true;
}

int alloc_ctx(Ctx& c) {
  GridP& g = c.g;
  for (int a = 0; a < 3; ++a) {
    g.N[a] = c.cfg.N[a];
    g.mn[a] = c.cfg.mn[a];
    float interval = c.cfg.mx[a] - c.cfg.mn[a];
    g.dx[a] = (float)(interval / (float)c.cfg.N[a]);  // Grid.cpp:94-97
  }
  // ParticleSystem(numParticles, dx.x): m_defaultRadius = 2 * dx.x
  c.radius = 2.0f * g.dx[0];
  for (int a = 0; a < 3; ++a) {
    c.dW[a] = (int)std::ceil(c.radius / g.dx[a]);
    if (c.dW[a] < 2) {
      c.err = "cells longer than dx.x on some axis are not supported (ceil(2dx.x/dx_a) < 2)";
      return LH_ERR_CONFIG;
    }
  }
  const int N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
  c.ncell = (size_t)N0 * N1 * N2;
  c.nface[0] = (size_t)(N0 + 1) * N1 * N2;
  c.nface[1] = (size_t)N0 * (N1 + 1) * N2;
  c.nface[2] = (size_t)N0 * N1 * (N2 + 1);
  c.nnode = (size_t)(N0 + 1) * (N1 + 1) * (N2 + 1);
  size_t fmax = std::max(c.nface[0], std::max(c.nface[1], c.nface[2]));
  bool ok = true;
  for (int a = 0; a < 3; ++a) {
    ok = ok && dalloc(c, c.v[a], c.nface[a]) && dalloc(c, c.tv[a], c.nface[a]) && dalloc(c, c.lw[a], c.nface[a]);
  }
  ok = ok && dalloc(c, c.pressure, c.ncell) && dalloc(c, c.phi, c.ncell) && dalloc(c, c.rhs, c.ncell) &&
       dalloc(c, c.solid_phi, c.nnode) && dalloc(c, c.is_valid, c.nnode) && dalloc(c, c.xstate, fmax);
  c.nsdf = (int)(c.cfg.sdf.size() / 12);
  ok = ok && dalloc(c, c.d_sdf, (size_t)12 * kMaxSdfNodes);  // room for any tree lh_set_boundary accepts
  for (int a = 0; a < 3; ++a) c.H[a] = 2 * c.cfg.N[a];
  c.nbins = (size_t)c.H[0] * c.H[1] * c.H[2];
  if (c.nbins >= (size_t)0x7fffffff) {
    c.err = "grid too large for 32-bit bin offsets";
    return LH_ERR_CONFIG;
  }
  c.nbblk = div_up(c.nbins + 1, (size_t)kScanBins);
  c.scan_bytes = bin_scan_bytes(c.nbblk);
  ok = ok && dalloc(c, c.bin_count, c.nbins + 1) && dalloc(c, c.bin_start, c.nbins + 1) &&
       dalloc(c, c.bin_bflag, c.nbblk) && dalloc(c, c.bin_bsum, c.nbblk) && dalloc(c, c.bin_boff, c.nbblk) &&
       dalloc(c, c.scan_tmp, c.scan_bytes);
  // multigrid levels
  int L = std::max(1, c.cfg.num_level);
  if (c.cfg.precond != LH_PRECOND_MULTIGRID) L = 1;
  c.lv.resize(L);
  for (int l = 0; l < L; ++l) {
    Level& q = c.lv[l];
    for (int a = 0; a < 3; ++a) q.d[a] = c.cfg.N[a] / (1 << l);
    if (q.d[0] < 1 || q.d[1] < 1 || q.d[2] < 1) {
      c.err = "numLevel too deep for gridDimension";
      return LH_ERR_CONFIG;
    }
    q.n = (size_t)q.d[0] * q.d[1] * q.d[2];
    ok = ok && dalloc(c, q.adiag, q.n) && dalloc(c, q.t[0], q.n) && dalloc(c, q.t[1], q.n) &&
         dalloc(c, q.t[2], q.n) && dalloc(c, q.cls, q.n) && dalloc(c, q.cmask, q.n) && dalloc(c, q.r, q.n) &&
         dalloc(c, q.z, q.n) &&
         dalloc(c, q.z2, q.n) && dalloc(c, q.rows, (size_t)q.d[0] * q.d[1]) && dalloc(c, q.nrows, 1) &&
         dalloc(c, q.aflag, (size_t)q.d[0] * q.d[1]) && dalloc(c, q.qrows, (size_t)q.d[0] * q.d[1]) &&
         dalloc(c, q.nqrows, 1) && dalloc(c, q.qflag, (size_t)q.d[0] * q.d[1]) &&
         dalloc(c, q.zrows, (size_t)q.d[0] * q.d[1]) && dalloc(c, q.nzrows, 1) &&
         dalloc(c, q.zflag, (size_t)q.d[0] * q.d[1]) && dalloc(c, q.tflag, div_up(q.d[0], 4) * div_up(q.d[1], 4)) &&
         dalloc(c, q.tiles, div_up(q.d[0], 4) * div_up(q.d[1], 4)) && dalloc(c, q.ntiles, 1);
    if (l == 0 && q.d[2] % kSegZ == 0 && q.d[2] / kSegZ <= 31)
      ok = ok && dalloc(c, q.smask, (size_t)q.d[0] * q.d[1]) &&
           dalloc(c, q.segs, (size_t)q.d[0] * q.d[1] * (q.d[2] / kSegZ)) && dalloc(c, q.nsegs, 1);
  }
  if (c.cfg.coarsest_fp64 && c.cfg.precond == LH_PRECOND_MULTIGRID) {
    // the FP64 coarsest sweeps are row sweeps: the coarsest level must be a coarse level swept by rows
    Level& q = c.lv[L - 1];
    if (L < 2 || c.cfg.coarse_sweep != kCoarseRows || (long long)q.n > c.cfg.row_sweep_max_cells) {
      c.err = "coarsestPrecision fp64 needs numLevel >= 2, coarseSweep rows and a coarsest level of at most "
              "rowSweepMaxCells cells";
      return LH_ERR_CONFIG;
    }
    ok = ok && dalloc(c, q.z64, q.n) && dalloc(c, q.z64b, q.n);
    if (c.cfg.coarsest_cheb > 0) ok = ok && dalloc(c, q.d64, q.n) && dalloc(c, q.d64b, q.n);
    if (ok) {
      hipMemsetAsync(q.z64, 0, q.n * sizeof(double), c.s);
      hipMemsetAsync(q.z64b, 0, q.n * sizeof(double), c.s);
      if (q.d64) {
        hipMemsetAsync(q.d64, 0, q.n * sizeof(double), c.s);
        hipMemsetAsync(q.d64b, 0, q.n * sizeof(double), c.s);
      }
    }
  }
  c.npartial = 1024;  // also the PCG vector kernels' grid: a multiple of 8 (per-XCD row runs)
  ok = ok && dalloc(c, c.p, c.ncell) && dalloc(c, c.p2, c.ncell) && dalloc(c, c.x, c.ncell) && dalloc(c, c.q, c.ncell) &&
       dalloc(c, c.partial, 4 * (size_t)c.npartial) && dalloc(c, c.dsc, 16) && dalloc(c, c.dsc64, 16) &&
       dalloc(c, c.dmaxv, 4) && dalloc(c, c.dflag, 16);
  if (!ok) return LH_ERR_OOM;
  if (hipHostMalloc((void**)&c.h_pin, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c.h_flag, 16 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
    c.err = "hipHostMalloc failed";
    return LH_ERR_OOM;
  }
  if (hipHostGetDevicePointer((void**)&c.h_pin_dev, c.h_pin, 0) != hipSuccess ||
      hipHostGetDevicePointer((void**)&c.h_flag_dev, c.h_flag, 0) != hipSuccess) {
    c.err = "hipHostGetDevicePointer failed";
    return LH_ERR_DEVICE;
  }
  hipMemcpyAsync(c.d_sdf, c.cfg.sdf.data(), c.cfg.sdf.size() * sizeof(float), hipMemcpyHostToDevice, c.s);
  hipMemsetAsync(c.bin_count, 0, (c.nbins + 1) * sizeof(int), c.s);  // [nbins] stays 0 (the scan total)
  hipMemsetAsync(c.bin_bflag, 0, c.nbblk * sizeof(int), c.s);
  launch_fill_f32(c, c.p2, c.ncell, 0.f);
  hipMemsetAsync(c.is_valid, 0, c.nnode, c.s);
  hipMemsetAsync(c.dflag, 0, 16 * sizeof(int), c.s);
  launch_fill_f32(c, c.solid_phi, c.nnode, 0.f);
  for (int a = 0; a < 3; ++a) {
    launch_fill_f32(c, c.lw[a], c.nface[a], 0.f);
    launch_fill_f32(c, c.v[a], c.nface[a], 0.f);
    launch_fill_f32(c, c.tv[a], c.nface[a], 0.f);
  }
  launch_fill_f32(c, c.pressure, c.ncell, 0.f);
  launch_fill_f32(c, c.phi, c.ncell, 3 * g.dx[0]);
  launch_fill_f32(c, c.rhs, c.ncell, 0.f);
  for (auto& q : c.lv) {
    hipMemsetAsync(q.nrows, 0, sizeof(int), c.s);
    hipMemsetAsync(q.nqrows, 0, sizeof(int), c.s);
    hipMemsetAsync(q.nzrows, 0, sizeof(int), c.s);
    launch_fill_f32(c, q.adiag, q.n, 0.f);
    for (int a = 0; a < 3; ++a) launch_fill_f32(c, q.t[a], q.n, 0.f);
    launch_fill_u8(c, q.cls, q.n, 0);
    launch_fill_f32(c, q.r, q.n, 0.f);
    launch_fill_f32(c, q.z, q.n, 0.f);
    launch_fill_f32(c, q.z2, q.n, 0.f);
  }
  if (!hip_ok(c, hipStreamSynchronize(c.s), "init")) return LH_ERR_DEVICE;
  return LH_OK;
}

void free_ctx(Ctx& c) {
  auto F = [](void* p) {
    if (p) hipFree(p);
  };
  for (int a = 0; a < 3; ++a) {
    F(c.v[a]);
    F(c.tv[a]);
    F(c.lw[a]);
  }
  F(c.pressure); F(c.phi); F(c.rhs); F(c.solid_phi); F(c.is_valid); F(c.xstate); F(c.d_sdf);
  F(c.bin_count); F(c.bin_start); F(c.bin_bflag); F(c.bin_bsum); F(c.bin_boff); F(c.scan_tmp); F(c.pkey); F(c.pslot); F(c.porder); F(c.pid);
  F(c.prec[0]); F(c.prec[1]); F(c.cv_flag);
  for (auto& q : c.lv) {
    F(q.adiag); F(q.t[0]); F(q.t[1]); F(q.t[2]); F(q.cls); F(q.cmask); F(q.r); F(q.z); F(q.z2); F(q.z64); F(q.z64b); F(q.d64); F(q.d64b);
    F(q.rows); F(q.nrows);
    F(q.smask); F(q.segs); F(q.nsegs);
    F(q.aflag); F(q.qrows); F(q.nqrows); F(q.qflag); F(q.zrows); F(q.nzrows); F(q.zflag); F(q.tflag); F(q.tiles); F(q.ntiles);
  }
  F(c.p); F(c.p2); F(c.x); F(c.q); F(c.partial); F(c.dsc); F(c.dsc64); F(c.dmaxv); F(c.dflag);
  F(c.x64); F(c.r64); F(c.rf); F(c.b_user); F(c.x0_user);
  F(c.xv_snap); F(c.xv_final); F(c.xcount); F(c.xlist[0]); F(c.xlist[1]); F(c.xlist[2]);
  nn_free(c.net);
  c.net = nullptr;
  if (c.h_pin) hipHostFree(c.h_pin);
  if (c.h_flag) hipHostFree(c.h_flag);
  if (c.ev0) hipEventDestroy(c.ev0);
  if (c.ev1) hipEventDestroy(c.ev1);
  if (c.ev_iter) hipEventDestroy(c.ev_iter);
  for (auto& g : c.pcg_graphs)
    if (g.exec) hipGraphExecDestroy(g.exec);
  c.pcg_graphs.clear();
  for (int k = 0; k < LH_PROF_COUNT; ++k)
    for (auto e : c.prof_ev[k]) hipEventDestroy(e);
  if (c.s) hipStreamDestroy(c.s);
}

int ensure_particle_capacity(Ctx& c, int64_t n) {
  if (n <= c.cap && c.prec[0]) return LH_OK;
  if (c.prec[0]) hipFree(c.prec[0]);
  if (c.prec[1]) hipFree(c.prec[1]);
  if (c.pkey) hipFree(c.pkey);
  if (c.pslot) hipFree(c.pslot);
  if (c.porder) hipFree(c.porder);
  if (c.pid) hipFree(c.pid);
  c.prec[0] = c.prec[1] = nullptr;
  c.pkey = c.pslot = c.porder = c.pid = nullptr;
  int64_t cap = std::max<int64_t>(n, 1);
  bool ok = dalloc(c, c.prec[0], (size_t)kRecFloats * cap) && dalloc(c, c.prec[1], (size_t)kRecFloats * cap) &&
            dalloc(c, c.pkey, cap) && dalloc(c, c.pslot, cap) && dalloc(c, c.porder, cap) && dalloc(c, c.pid, cap);
  if (!ok) {
    c.cap = 0;
    return LH_ERR_OOM;
  }
  c.cap = cap;
  c.cur = 0;
  return LH_OK;
}

// upload particles in AoS (pos[3n], vel[3n], C[9n]) order; ids 0..n-1
int upload_particles(Ctx& c, const float* pos, const float* vel, const float* C, int64_t n) {
  int rc = ensure_particle_capacity(c, n);
  if (rc) return rc;
  c.P = n;
  c.bins_valid = false;
  if (n == 0) return LH_OK;
  // the records: R0 for every particle, then R1..R3 for every particle
  std::vector<float> r0((size_t)4 * n), rs((size_t)12 * n);
  for (int64_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) r0[4 * i + a] = pos[3 * i + a];
    const int id = (int)i;
    std::memcpy(&r0[4 * i + 3], &id, 4);
    for (int a = 0; a < 3; ++a) {
      // v_x in R3.w, v_y in R1.w, v_z in R2.w
      rs[12 * i + 4 * ((a + 2) % 3) + 3] = vel ? vel[3 * i + a] : 0.f;
      for (int k = 0; k < 3; ++k) rs[12 * i + 4 * a + k] = C ? C[9 * i + 3 * a + k] : 0.f;
    }
  }
  hipMemcpyAsync(c.prec[c.cur], r0.data(), r0.size() * sizeof(float), hipMemcpyHostToDevice, c.s);
  hipMemcpyAsync(c.prec[c.cur] + (size_t)4 * c.cap, rs.data(), rs.size() * sizeof(float), hipMemcpyHostToDevice,
                 c.s);
  if (!hip_ok(c, hipStreamSynchronize(c.s), "upload particles")) return LH_ERR_DEVICE;
  return LH_OK;
}

// ParticleSystem<3>::SpawnParticles, ParticleSystem.cpp:44-64
void spawn_lattice(Ctx& c, std::vector<float>& pos) {
  const float* mn = c.cfg.mn;
  const float* mx = c.cfg.mx;
  float lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    float off = 0.3f * (mx[a] - mn[a]);  // Scalar(0.3) * (max - min), LiquidModel.cpp:509
    lo[a] = mn[a] + off;
    hi[a] = mx[a] - off;
  }
  float interval = (float)(0.5 * (double)c.radius);
  int d[3];
  for (int a = 0; a < 3; ++a) d[a] = (int)((hi[a] - lo[a]) / interval);
  int64_t total = (int64_t)d[0] * d[1] * d[2];
  int64_t n = std::min<int64_t>(c.cfg.num_particles, total);
  pos.assign(3 * n, 0.f);
  for (int xi = 0; xi < d[0]; ++xi)
    for (int yi = 0; yi < d[1]; ++yi)
      for (int zi = 0; zi < d[2]; ++zi) {
        int64_t idx = (int64_t)xi * d[1] * d[2] + (int64_t)yi * d[2] + zi;
        if (idx >= n) break;
        pos[3 * idx] = (float)xi * interval + lo[0];
        pos[3 * idx + 1] = (float)yi * interval + lo[1];
        pos[3 * idx + 2] = (float)zi * interval + lo[2];
      }
}

// Jittered block seeding (build extension): k^3 particles per cell whose centre
// lies inside the block, one per sub-cell at min + ((k*i + s) + u) * (dx/k),
// u from a SplitMix64 hash of (seed, block, cell, sub-cell, axis); velocity
// v0 + sigma * N(0,1) (Box-Muller in FP64); particles inside solids dropped.
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
  return splitmix(h ^ (sub * 16 + stream));
}
void seed_blocks(Ctx& c, std::vector<float>& pos, std::vector<float>& vel) {
  const int k = std::max(1, c.cfg.k_ppc);
  const GridP& g = c.g;
  const float* prog = c.cfg.sdf.data();
  const int nprog = (int)(c.cfg.sdf.size() / 12);
  size_t nb = c.cfg.blocks.size() / 10;
  for (size_t b = 0; b < nb; ++b) {
    const float* B = &c.cfg.blocks[10 * b];
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = g.N[a];
      hi[a] = -1;
      for (int i = 0; i < g.N[a]; ++i) {
        float cc = ((float)i + 0.5f) * g.dx[a] + g.mn[a];
        if (cc >= B[a] && cc <= B[3 + a]) {
          lo[a] = std::min(lo[a], i);
          hi[a] = i;
        }
      }
    }
    for (int xi = lo[0]; xi <= hi[0]; ++xi)
      for (int yi = lo[1]; yi <= hi[1]; ++yi)
        for (int zi = lo[2]; zi <= hi[2]; ++zi) {
          uint64_t cell = ((uint64_t)xi * g.N[1] + yi) * g.N[2] + zi;
          for (int s = 0; s < k * k * k; ++s) {
            int ss[3] = {s / (k * k), (s / k) % k, s % k};
            int cc[3] = {xi, yi, zi};
            float p[3];
            for (int a = 0; a < 3; ++a) {
              float u = (float)(rng_u64(c.cfg.seed, b, cell, s, a) >> 40) * (1.0f / 16777216.0f);
              p[a] = g.mn[a] + ((float)(k * cc[a] + ss[a]) + u) * (g.dx[a] / (float)k);
            }
            if (!(sdf_eval(prog, nprog, p[0], p[1], p[2]) > 0.0f)) continue;
            float v[3];
            for (int a = 0; a < 3; ++a) {
              v[a] = B[6 + a];
              if (B[9] > 0.f) {
                uint64_t h1 = rng_u64(c.cfg.seed, b, cell, s, 3 + 2 * a);
                uint64_t h2 = rng_u64(c.cfg.seed, b, cell, s, 4 + 2 * a);
                double u1 = ((double)(h1 >> 11) + 1.0) * (1.0 / 9007199254740992.0);
                double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
                double nrm = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
                v[a] = (float)((double)B[6 + a] + (double)B[9] * nrm);
              }
            }
            for (int a = 0; a < 3; ++a) {
              pos.push_back(p[a]);
              vel.push_back(v[a]);
            }
          }
        }
  }
}

size_t field_count(const Ctx& c, int f, int shape[3]) {
  const int* N = c.g.N;
  int s[3] = {N[0], N[1], N[2]};
  switch (f) {
    case LH_FIELD_VX: case LH_FIELD_TEMPVX: case LH_FIELD_LIQUID_WX: s[0] += 1; break;
    case LH_FIELD_VY: case LH_FIELD_TEMPVY: case LH_FIELD_LIQUID_WY: s[1] += 1; break;
    case LH_FIELD_VZ: case LH_FIELD_TEMPVZ: case LH_FIELD_LIQUID_WZ: s[2] += 1; break;
    case LH_FIELD_SOLID_PHI: case LH_FIELD_IS_VALID: s[0] += 1; s[1] += 1; s[2] += 1; break;
    case LH_FIELD_PRESSURE: case LH_FIELD_LIQUID_PHI: case LH_FIELD_RHS: break;
    default: return 0;
  }
  if (shape) {
    shape[0] = s[0]; shape[1] = s[1]; shape[2] = s[2];
  }
  return (size_t)s[0] * s[1] * s[2];
}

void* field_dev(Ctx& c, int f) {
  switch (f) {
    case LH_FIELD_VX: return c.v[0];
    case LH_FIELD_VY: return c.v[1];
    case LH_FIELD_VZ: return c.v[2];
    case LH_FIELD_TEMPVX: return c.tv[0];
    case LH_FIELD_TEMPVY: return c.tv[1];
    case LH_FIELD_TEMPVZ: return c.tv[2];
    case LH_FIELD_LIQUID_WX: return c.lw[0];
    case LH_FIELD_LIQUID_WY: return c.lw[1];
    case LH_FIELD_LIQUID_WZ: return c.lw[2];
    case LH_FIELD_PRESSURE: return c.pressure;
    case LH_FIELD_LIQUID_PHI: return c.phi;
    case LH_FIELD_SOLID_PHI: return c.solid_phi;
    case LH_FIELD_IS_VALID: return c.is_valid;
    case LH_FIELD_RHS: return c.rhs;
    default: return nullptr;
  }
}

int check_domain(Ctx& c) {
  hipMemcpyAsync(c.h_flag + 1, c.dflag + 1, sizeof(int), hipMemcpyDeviceToHost, c.s);
  if (!hip_ok(c, hipStreamSynchronize(c.s), "binning")) return LH_ERR_DEVICE;
  if (c.h_flag[1]) {
    c.err = "a particle left the grid box [minCoord, maxCoord)";
    return LH_ERR_DOMAIN;
  }
  return LH_OK;
}

int read_maxv(Ctx& c) {
  hipMemcpyAsync(c.h_flag + 4, c.dmaxv, sizeof(unsigned), hipMemcpyDeviceToHost, c.s);
  if (!hip_ok(c, hipStreamSynchronize(c.s), "max velocity")) return LH_ERR_DEVICE;
  std::memcpy(&c.last_maxv, c.h_flag + 4, sizeof(float));
  return LH_OK;
}

// ---- step pieces
int do_pre_step(Ctx& c, bool in_step = false) {
  launch_clear(c, in_step);  // Grid<3>::Clear
  int rc = run_binning(c);
  if (rc) return rc;
  launch_liquid_phi(c);  // ComputeLiquidPhi
  return LH_OK;
}

// one Step; with div != nullptr also PostStep's ComputeDivergence (LiquidModel.cpp:530-537,
// 1065-1083) inside the step's one wait: div[0] = mean, div[1] = L-inf, div[2] = count
int do_step(Ctx& c, float dt, lh_solve_stats* st, double* div = nullptr) {
  lh_solve_stats local;
  std::memset(&local, 0, sizeof(local));
  if (!c.bins_valid && c.P > 0) {
    int rc = run_binning(c);
    if (rc) return rc;
  }
  launch_p2g(c, true, dt);  // Particle2Grid + ApplyGravity (+ tank acceleration)
  if (c.P == 0) launch_gravity(c, dt);
  int solve_rc = LH_OK;
  if (c.P > 0) {  // SolvePressure returns early without particles (LiquidModel.cpp:828)
    // SolvePressure passes the scene's density (LiquidModel.cpp:830-831)
    solve_rc = run_solve(c, c.cfg.density, dt, &local);
    if (solve_rc == LH_ERR_DEVICE || solve_rc == LH_ERR_NAN) {
      if (st) *st = local;
      return solve_rc;
    }
    launch_velocity_update(c, dt);
  }
  if (run_extrapolate(c) < 0) return LH_ERR_DEVICE;
  launch_g2p_advect(c, true, true, dt);  // Grid2Particle + AdvectParticle (+ max|v|)
  c.bins_valid = false;
  if (c.cfg.correct_volume) {  // CorrectVolume(dt), LiquidModel.cpp:527 (commented out in the reference)
    int crc = run_correct_volume(c, dt);
    if (crc) return crc;
  }
  c.time += (double)dt;
  c.last_solve = local;
  if (div) {
    double* out = reinterpret_cast<double*>(c.h_pin_dev) + kPinDiv;
    // with particles the solve has just built the level-0 liquid rows from this liquidPhi
    if (c.P > 0) launch_divergence_rows(c, out);
    else launch_divergence_dense(c, out);
  }
  // the domain flag and max |v| with one wait (and the divergence's host-mapped results)
  hipMemcpyAsync(c.h_flag + 1, c.dflag + 1, sizeof(int), hipMemcpyDeviceToHost, c.s);
  hipMemcpyAsync(c.h_flag + 4, c.dmaxv, sizeof(unsigned), hipMemcpyDeviceToHost, c.s);
  if (!hip_ok(c, hipStreamSynchronize(c.s), "step end")) return LH_ERR_DEVICE;
  if (st) *st = local;
  if (c.h_flag[1]) {
    c.err = "a particle left the grid box [minCoord, maxCoord)";
    return LH_ERR_DOMAIN;
  }
  std::memcpy(&c.last_maxv, c.h_flag + 4, sizeof(float));
  if (div) {
    const double cnt = c.h_pin[kPinDiv + 2];
    div[0] = cnt > 0 ? c.h_pin[kPinDiv] / cnt : 0.0;
    div[1] = c.h_pin[kPinDiv + 1];
    div[2] = cnt;
  }
  return LH_OK;
}

float cfl_from(const Ctx& c) {
  // GetCFL: clamp(dx / maxVel, 0.001, 0.1), LiquidModel.h:49-55
  return clampf(c.g.dx[0] / c.last_maxv, 0.001f, 0.1f);
}

template <class F>
int guarded(lh_ctx* ctx, F&& f) {
  if (!ctx) return LH_ERR_ARG;
  try {
    return f(ctx->c);
  } catch (const json_error& e) {
    ctx->c.err = e.what();
    return LH_ERR_CONFIG;
  } catch (const std::exception& e) {
    ctx->c.err = e.what();
    return LH_ERR_CONFIG;
  }
}

int create_common(const std::string& text, lh_ctx** out) {
  if (!out) return LH_ERR_ARG;
  *out = nullptr;
  lh_ctx* ctx = new lh_ctx();
  Ctx& c = ctx->c;
  try {
    c.cfg = parse_config(text);
  } catch (const std::exception& e) {
    c.err = e.what();
    *out = ctx;  // returned so the caller can read lh_last_error, then lh_destroy
    return LH_ERR_CONFIG;
  }
  if (!hip_ok(c, hipStreamCreateWithFlags(&c.s, hipStreamNonBlocking), "stream") ||
      !hip_ok(c, hipEventCreate(&c.ev0), "event") || !hip_ok(c, hipEventCreate(&c.ev1), "event") ||
      !hip_ok(c, hipEventCreateWithFlags(&c.ev_iter, hipEventDisableTiming), "event")) {
    *out = ctx;
    return LH_ERR_DEVICE;
  }
  int rc = alloc_ctx(c);
  *out = ctx;
  return rc;
}

}  // namespace
}  // namespace lh

using namespace lh;

extern "C" {

int lh_create(const char* json_path, lh_ctx** out) {
  if (!json_path) return LH_ERR_ARG;
  std::ifstream f(json_path);
  if (!f.is_open()) {
    if (out) {
      *out = new lh_ctx();
      (*out)->c.err = std::string("cannot open config ") + json_path;
    }
    return LH_ERR_CONFIG;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  return create_common(ss.str(), out);
}

int lh_create_from_string(const char* json_text, lh_ctx** out) {
  if (!json_text) return LH_ERR_ARG;
  return create_common(json_text, out);
}

void lh_destroy(lh_ctx* ctx) {
  if (!ctx) return;
  if (ctx->c.s) hipStreamSynchronize(ctx->c.s);
  free_ctx(ctx->c);
  delete ctx;
}

const char* lh_last_error(const lh_ctx* ctx) { return ctx ? ctx->c.err.c_str() : noalloc_last_error(); }

int lh_grid_info(const lh_ctx* ctx, int dims[3], float min_coord[3], float dx[3]) {
  if (!ctx) return LH_ERR_ARG;
  for (int a = 0; a < 3; ++a) {
    if (dims) dims[a] = ctx->c.g.N[a];
    if (min_coord) min_coord[a] = ctx->c.g.mn[a];
    if (dx) dx[a] = ctx->c.g.dx[a];
  }
  return LH_OK;
}

int lh_field_shape(const lh_ctx* ctx, int field, int shape[3]) {
  if (!ctx) return LH_ERR_ARG;
  return field_count(ctx->c, field, shape) ? LH_OK : LH_ERR_ARG;
}

int lh_set_boundary(lh_ctx* ctx, const char* json) {
  return guarded(ctx, [json](Ctx& c) {
    if (!json) return LH_ERR_ARG;
    std::vector<float> prog;
    emit_sdf(Json::parse(json), prog);
    if (prog.empty() || prog.size() / 12 > (size_t)kMaxSdfNodes) {
      c.err = "Boundary tree too large";
      return LH_ERR_CONFIG;
    }
    c.cfg.sdf = prog;
    c.nsdf = (int)(prog.size() / 12);
    hipMemcpyAsync(c.d_sdf, c.cfg.sdf.data(), c.cfg.sdf.size() * sizeof(float), hipMemcpyHostToDevice, c.s);
    return hip_ok(c, hipStreamSynchronize(c.s), "boundary") ? LH_OK : LH_ERR_DEVICE;
  });
}

int lh_compute_solid_phi(lh_ctx* ctx) {
  return guarded(ctx, [](Ctx& c) {
    launch_solid_phi(c);
    c.xcache_ok = false;  // liquidW changed: the Extrapolate schedule must be re-recorded
    c.solid_ready = true;
    return hip_ok(c, hipStreamSynchronize(c.s), "solid phi") ? LH_OK : LH_ERR_DEVICE;
  });
}

int lh_init(lh_ctx* ctx) {
  return guarded(ctx, [](Ctx& c) {
    // LiquidModel<3>::Init: ComputeSolidPhi, then spawn (LiquidModel.cpp:506-512)
    launch_solid_phi(c);
    c.xcache_ok = false;  // liquidW changed: the Extrapolate schedule must be re-recorded
    c.solid_ready = true;
    std::vector<float> pos, vel;
    if (!c.cfg.blocks.empty()) seed_blocks(c, pos, vel);
    else spawn_lattice(c, pos);
    int64_t n = (int64_t)(pos.size() / 3);
    c.time = 0.0;
    return upload_particles(c, pos.data(), vel.empty() ? nullptr : vel.data(), nullptr, n);
  });
}

int lh_clear_grid(lh_ctx* ctx) {
  return guarded(ctx, [](Ctx& c) {
    launch_clear(c);
    return hip_ok(c, hipStreamSynchronize(c.s), "clear") ? LH_OK : LH_ERR_DEVICE;
  });
}

int lh_compute_liquid_phi(lh_ctx* ctx) {
  return guarded(ctx, [](Ctx& c) {
    int rc = run_binning(c);
    if (rc) return rc;
    launch_liquid_phi(c);
    return check_domain(c);
  });
}

int lh_pre_step(lh_ctx* ctx, float dt) {
  (void)dt;
  return guarded(ctx, [](Ctx& c) {
    int rc = do_pre_step(c);
    if (rc) return rc;
    return check_domain(c);
  });
}

int lh_p2g(lh_ctx* ctx) {
  return guarded(ctx, [](Ctx& c) {
    if (!c.bins_valid) {
      int rc = run_binning(c);
      if (rc) return rc;
    }
    launch_p2g(c, false, 0.f);
    return check_domain(c);
  });
}

int lh_apply_gravity(lh_ctx* ctx, float dt) {
  return guarded(ctx, [dt](Ctx& c) {
    launch_gravity(c, dt);
    return hip_ok(c, hipStreamSynchronize(c.s), "gravity") ? LH_OK : LH_ERR_DEVICE;
  });
}

int lh_solve_pressure(lh_ctx* ctx, float rho, float dt, lh_solve_stats* stats) {
  return guarded(ctx, [rho, dt, stats](Ctx& c) {
    if (!(rho > 0.f) || !std::isfinite(rho)) {
      c.err = "lh_solve_pressure: liquidDensity must be positive and finite";
      return LH_ERR_ARG;
    }
    if (c.P == 0) {
      if (stats) std::memset(stats, 0, sizeof(*stats));
      return LH_OK;
    }
    return run_solve(c, rho, dt, stats);
  });
}

int lh_velocity_update(lh_ctx* ctx, float dt) {
  return guarded(ctx, [dt](Ctx& c) {
    if (c.P > 0) launch_velocity_update(c, dt);
    return hip_ok(c, hipStreamSynchronize(c.s), "velocity update") ? LH_OK : LH_ERR_DEVICE;
  });
}

int lh_extrapolate(lh_ctx* ctx) {
  return guarded(ctx, [](Ctx& c) {
    if (run_extrapolate(c) < 0) return LH_ERR_DEVICE;
    return hip_ok(c, hipStreamSynchronize(c.s), "extrapolate") ? LH_OK : LH_ERR_DEVICE;
  });
}

int lh_g2p(lh_ctx* ctx) {
  return guarded(ctx, [](Ctx& c) {
    launch_g2p_advect(c, true, false, 0.f);
    c.bins_valid = false;  // the bin-ordered copies of v and C are stale: the next P2G re-bins
    return read_maxv(c);
  });
}

int lh_advect(lh_ctx* ctx, float dt) {
  return guarded(ctx, [dt](Ctx& c) {
    launch_g2p_advect(c, false, true, dt);
    c.bins_valid = false;
    return read_maxv(c);
  });
}

int lh_correct_volume(lh_ctx* ctx, float dt) {
  return guarded(ctx, [dt](Ctx& c) {
    int rc = run_correct_volume(c, dt);
    if (rc) return rc;
    return hip_ok(c, hipStreamSynchronize(c.s), "correct volume") ? LH_OK : LH_ERR_DEVICE;
  });
}

int lh_divergence(lh_ctx* ctx, double* mean, double* linf, int* count) {
  return guarded(ctx, [=](Ctx& c) {
    launch_divergence(c, mean, linf, count);
    return hip_ok(c, hipGetLastError(), "divergence") ? LH_OK : LH_ERR_DEVICE;
  });
}

int lh_step(lh_ctx* ctx, float dt) {
  return guarded(ctx, [dt](Ctx& c) { return do_step(c, dt, nullptr); });
}

int lh_post_step(lh_ctx* ctx, float dt, lh_step_stats* stats) {
  (void)dt;
  return guarded(ctx, [stats](Ctx& c) {
    if (stats) {
      stats->solve = c.last_solve;
      launch_divergence(c, &stats->divergence_mean, &stats->divergence_linf, &stats->fluid_cells);
      stats->max_velocity = c.last_maxv;
      stats->cfl_dt = cfl_from(c);
      stats->sim_time = c.time;
    }
    return LH_OK;
  });
}

int lh_get_cfl(lh_ctx* ctx, float* dt) {
  return guarded(ctx, [dt](Ctx& c) {
    if (!dt) return LH_ERR_ARG;
    *dt = cfl_from(c);
    return LH_OK;
  });
}

int lh_run_steps(lh_ctx* ctx, int nsteps, float* dt, lh_step_stats* last) {
  return guarded(ctx, [=](Ctx& c) {
    if (!dt) return LH_ERR_ARG;
    for (int k = 0; k < nsteps; ++k) {
      lh_step_stats st;
      std::memset(&st, 0, sizeof(st));
      hipEvent_t e0, e1;
      hipEventCreate(&e0);
      hipEventCreate(&e1);
      hipEventRecord(e0, c.s);
      double div[3] = {0.0, 0.0, 0.0};
      int rc = do_pre_step(c, true);
      if (rc == LH_OK) rc = do_step(c, *dt, &st.solve, div);  // PostStep's divergence every step
      hipEventRecord(e1, c.s);
      hipEventSynchronize(e1);
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      hipEventDestroy(e0);
      hipEventDestroy(e1);
      if (rc != LH_OK && rc != LH_ERR_NOT_CONVERGED) return rc;
      st.step_ms = ms;
      st.divergence_mean = div[0];
      st.divergence_linf = div[1];
      st.fluid_cells = (int)div[2];
      st.max_velocity = c.last_maxv;
      st.cfl_dt = cfl_from(c);
      st.sim_time = c.time;
      *dt = st.cfl_dt;  // dt = liquidModel.GetCFL(), main.cpp:53
      if (last) *last = st;
    }
    return LH_OK;
  });
}

int lh_num_particles(const lh_ctx* ctx, int64_t* n) {
  if (!ctx || !n) return LH_ERR_ARG;
  *n = ctx->c.P;
  return LH_OK;
}

int lh_set_particles(lh_ctx* ctx, const float* pos, const float* vel, const float* C, int64_t n) {
  return guarded(ctx, [=](Ctx& c) {
    if (n < 0 || (n > 0 && !pos)) return LH_ERR_ARG;
    if (n >= (int64_t)0x7fffffff) {
      c.err = "particle count exceeds 32-bit bin offsets";
      return LH_ERR_ARG;
    }
    return upload_particles(c, pos, vel, C, n);
  });
}

int lh_get_particles(lh_ctx* ctx, float* pos, float* vel, float* C, int64_t n) {
  return guarded(ctx, [=](Ctx& c) {
    if (n != c.P) {
      c.err = "lh_get_particles: n does not match the particle count";
      return LH_ERR_ARG;
    }
    if (n == 0) return LH_OK;
    std::vector<float> r0((size_t)4 * n), rs((size_t)12 * n);
    hipMemcpyAsync(r0.data(), c.prec[c.cur], r0.size() * sizeof(float), hipMemcpyDeviceToHost, c.s);
    hipMemcpyAsync(rs.data(), c.prec[c.cur] + (size_t)4 * c.cap, rs.size() * sizeof(float), hipMemcpyDeviceToHost,
                   c.s);
    if (!hip_ok(c, hipStreamSynchronize(c.s), "download particles")) return LH_ERR_DEVICE;
    for (int64_t i = 0; i < n; ++i) {
      int id;
      std::memcpy(&id, &r0[4 * i + 3], 4);
      if (id < 0 || id >= n) {
        c.err = "corrupt particle id";
        return LH_ERR_DEVICE;
      }
      for (int a = 0; a < 3; ++a) {
        if (pos) pos[3 * (size_t)id + a] = r0[4 * i + a];
        if (vel) vel[3 * (size_t)id + a] = rs[12 * i + 4 * ((a + 2) % 3) + 3];
        if (C)
          for (int k = 0; k < 3; ++k) C[9 * (size_t)id + 3 * a + k] = rs[12 * i + 4 * a + k];
      }
    }
    return LH_OK;
  });
}

int lh_resize_particles(lh_ctx* ctx, int64_t n) {
  return guarded(ctx, [=](Ctx& c) {
    if (n < 0 || n >= (int64_t)0x7fffffff) {
      c.err = "lh_resize_particles: count out of range";
      return LH_ERR_ARG;
    }
    // ParticleSystem<3>::Resize (ParticleSystem.cpp:10-21): the first min(P, n) particles keep
    // their state (in id order), new ones start at zero
    const int64_t keep = std::min<int64_t>(c.P, n);
    std::vector<float> pos(3 * (size_t)n, 0.f), vel(3 * (size_t)n, 0.f), Cm(9 * (size_t)n, 0.f);
    if (keep > 0) {
      std::vector<float> p0(3 * (size_t)c.P), v0(3 * (size_t)c.P), c0(9 * (size_t)c.P);
      int rc = lh_get_particles(ctx, p0.data(), v0.data(), c0.data(), c.P);
      if (rc) return rc;
      std::copy(p0.begin(), p0.begin() + 3 * keep, pos.begin());
      std::copy(v0.begin(), v0.begin() + 3 * keep, vel.begin());
      std::copy(c0.begin(), c0.begin() + 9 * keep, Cm.begin());
    }
    return upload_particles(c, pos.data(), vel.data(), Cm.data(), n);
  });
}

int lh_liquid_density(const lh_ctx* ctx, float* rho) {
  if (!ctx || !rho) return LH_ERR_ARG;
  *rho = ctx->c.cfg.density;
  return LH_OK;
}

int lh_get_field(lh_ctx* ctx, int field, void* host, size_t count) {
  return guarded(ctx, [=](Ctx& c) {
    size_t n = field_count(c, field, nullptr);
    if (!n || n != count || !host) {
      c.err = "lh_get_field: bad field id or element count";
      return LH_ERR_ARG;
    }
    size_t bytes = n * (field == LH_FIELD_IS_VALID ? 1 : 4);
    hipMemcpyAsync(host, field_dev(c, field), bytes, hipMemcpyDeviceToHost, c.s);
    return hip_ok(c, hipStreamSynchronize(c.s), "get field") ? LH_OK : LH_ERR_DEVICE;
  });
}

int lh_set_field(lh_ctx* ctx, int field, const void* host, size_t count) {
  return guarded(ctx, [=](Ctx& c) {
    size_t n = field_count(c, field, nullptr);
    if (!n || n != count || !host) {
      c.err = "lh_set_field: bad field id or element count";
      return LH_ERR_ARG;
    }
    size_t bytes = n * (field == LH_FIELD_IS_VALID ? 1 : 4);
    hipMemcpyAsync(field_dev(c, field), host, bytes, hipMemcpyHostToDevice, c.s);
    if (field >= LH_FIELD_LIQUID_WX && field <= LH_FIELD_LIQUID_WZ) c.xcache_ok = false;
    return hip_ok(c, hipStreamSynchronize(c.s), "set field") ? LH_OK : LH_ERR_DEVICE;
  });
}

int lh_profile_enable(lh_ctx* ctx, int on) {
  if (!ctx) return LH_ERR_ARG;
  Ctx& c = ctx->c;
  hipStreamSynchronize(c.s);
  c.prof = (on != 0);
  for (int k = 0; k < LH_PROF_COUNT; ++k) c.prof_used[k] = c.prof_launch[k] = 0;
  return LH_OK;
}

int lh_profile_read(lh_ctx* ctx, int kernel, int* launches, double* total_ms) {
  if (!ctx || kernel < 0 || kernel >= LH_PROF_COUNT) return LH_ERR_ARG;
  Ctx& c = ctx->c;
  if (!hip_ok(c, hipStreamSynchronize(c.s), "profile")) return LH_ERR_DEVICE;
  double tot = 0.0;
  for (int u = 0; u < c.prof_used[kernel]; ++u) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c.prof_ev[kernel][2 * u], c.prof_ev[kernel][2 * u + 1]) == hipSuccess) tot += ms;
  }
  if (launches) *launches = c.prof_launch[kernel];
  if (total_ms) *total_ms = tot;
  return LH_OK;
}

int lh_build_system(lh_ctx* ctx, float rho, float dt) {
  return guarded(ctx, [rho, dt](Ctx& c) {
    if (!(rho > 0.f) || !std::isfinite(rho)) {
      c.err = "lh_build_system: liquidDensity must be positive and finite";
      return LH_ERR_ARG;
    }
    build_system(c, rho, dt);
    c.last_dt = dt;
    return hip_ok(c, hipStreamSynchronize(c.s), "build system") ? LH_OK : LH_ERR_DEVICE;
  });
}

namespace {
int upload_cells(Ctx& c, float*& dst, const float* host) {
  if (!dst && hipMalloc((void**)&dst, c.ncell * sizeof(float)) != hipSuccess) {
    c.err = "hipMalloc failed";
    return LH_ERR_OOM;
  }
  hipMemcpyAsync(dst, host, c.ncell * sizeof(float), hipMemcpyHostToDevice, c.s);
  return LH_OK;
}
}  // namespace

int lh_solve_rhs(lh_ctx* ctx, const float* b, const float* x0, lh_solve_stats* stats) {
  return guarded(ctx, [=](Ctx& c) {
    const float* db = c.rhs;
    const float* dx0 = nullptr;
    if (b) {
      int rc = upload_cells(c, c.b_user, b);
      if (rc) return rc;
      db = c.b_user;
    }
    if (x0) {
      int rc = upload_cells(c, c.x0_user, x0);
      if (rc) return rc;
      dx0 = c.x0_user;
    } else if (c.warm_learned && c.net) {
      // the learned guess: inference, scaling and r0 are inside the timed solve
      hipEventRecord(c.ev0, c.s);
      float* g = nullptr;
      int rc = nn_predict(c, db, &g);
      if (rc) return rc;
      return solve_built(c, db, g, stats, false);
    }
    return solve_built(c, db, dx0, stats);
  });
}

int lh_get_solution64(lh_ctx* ctx, double* host, size_t count) {
  return guarded(ctx, [=](Ctx& c) {
    if (!host || count != c.ncell) {
      c.err = "lh_get_solution64: bad buffer";
      return LH_ERR_ARG;
    }
    if (c.cfg.refined && c.x64) {
      hipMemcpyAsync(host, c.x64, count * sizeof(double), hipMemcpyDeviceToHost, c.s);
      return hip_ok(c, hipStreamSynchronize(c.s), "get solution") ? LH_OK : LH_ERR_DEVICE;
    }
    std::vector<float> tmp(count);
    hipMemcpyAsync(tmp.data(), c.pressure, count * sizeof(float), hipMemcpyDeviceToHost, c.s);
    if (!hip_ok(c, hipStreamSynchronize(c.s), "get solution")) return LH_ERR_DEVICE;
    for (size_t i = 0; i < count; ++i) host[i] = tmp[i];
    return LH_OK;
  });
}

int lh_solver_level_info(lh_ctx* ctx, int level, int dims[3], int64_t* active_rows) {
  return guarded(ctx, [=](Ctx& c) {
    if (level < 0 || level >= (int)c.lv.size()) {
      c.err = "lh_solver_level_info: no such level";
      return LH_ERR_ARG;
    }
    const Level& L = c.lv[level];
    if (dims)
      for (int a = 0; a < 3; ++a) dims[a] = L.d[a];
    if (active_rows) {
      int n = 0;
      hipMemcpyAsync(&n, L.nrows, sizeof(int), hipMemcpyDeviceToHost, c.s);
      if (!hip_ok(c, hipStreamSynchronize(c.s), "level info")) return LH_ERR_DEVICE;
      *active_rows = n;
    }
    return LH_OK;
  });
}

int lh_solver_level_segments(lh_ctx* ctx, int level, int* segment_cells, int64_t* segments) {
  return guarded(ctx, [=](Ctx& c) {
    if (level < 0 || level >= (int)c.lv.size()) {
      c.err = "lh_solver_level_segments: no such level";
      return LH_ERR_ARG;
    }
    const Level& L = c.lv[level];
    if (segment_cells) *segment_cells = L.segs ? kSegZ : 0;
    if (segments) {
      int n = 0;
      if (L.segs) {
        hipMemcpyAsync(&n, L.nsegs, sizeof(int), hipMemcpyDeviceToHost, c.s);
        if (!hip_ok(c, hipStreamSynchronize(c.s), "level segments")) return LH_ERR_DEVICE;
      }
      *segments = n;
    }
    return LH_OK;
  });
}

int lh_set_warm_start(lh_ctx* ctx, const float* x0, size_t count) {
  return guarded(ctx, [=](Ctx& c) {
    if (!x0) {
      c.warm_x0 = nullptr;
      return LH_OK;
    }
    if (count != c.ncell) {
      c.err = "lh_set_warm_start: bad element count";
      return LH_ERR_ARG;
    }
    int rc = upload_cells(c, c.x0_user, x0);
    if (rc) return rc;
    c.warm_x0 = c.x0_user;
    return hip_ok(c, hipStreamSynchronize(c.s), "warm start") ? LH_OK : LH_ERR_DEVICE;
  });
}

int lh_warmstart_init(lh_ctx* ctx, uint64_t seed) {
  return guarded(ctx, [=](Ctx& c) { return nn_create(c, seed); });
}
int lh_warmstart_load(lh_ctx* ctx, const char* path) {
  if (!path) return LH_ERR_ARG;
  return guarded(ctx, [=](Ctx& c) { return nn_load(c, path); });
}
int lh_warmstart_save(lh_ctx* ctx, const char* path) {
  if (!path) return LH_ERR_ARG;
  return guarded(ctx, [=](Ctx& c) { return nn_save(c, path); });
}
int lh_warmstart_enable(lh_ctx* ctx, int on) {
  return guarded(ctx, [=](Ctx& c) {
    if (on && !c.net) {
      c.err = "no warm-start network: call lh_warmstart_init or lh_warmstart_load first";
      return LH_ERR_ARG;
    }
    c.warm_learned = (on != 0);
    return LH_OK;
  });
}
int lh_warmstart_num_params(lh_ctx* ctx, int64_t* n) {
  return guarded(ctx, [=](Ctx& c) {
    if (!c.net || !n) {
      c.err = "no warm-start network";
      return LH_ERR_ARG;
    }
    *n = (int64_t)c.net->nlog;
    return LH_OK;
  });
}
int lh_warmstart_get_params(lh_ctx* ctx, float* w, int64_t n) {
  return guarded(ctx, [=](Ctx& c) { return nn_get_params(c, w, n, false); });
}
int lh_warmstart_set_params(lh_ctx* ctx, const float* w, int64_t n) {
  return guarded(ctx, [=](Ctx& c) { return nn_set_params(c, w, n); });
}
int lh_warmstart_set_target_scale(lh_ctx* ctx, float scale) {
  return guarded(ctx, [=](Ctx& c) {
    if (!c.net || !(scale > 0.f)) {
      c.err = "lh_warmstart_set_target_scale: no network or scale <= 0";
      return LH_ERR_ARG;
    }
    c.net->target_scale = scale;
    return LH_OK;
  });
}
int lh_warmstart_forward(lh_ctx* ctx, const int dims[3], const float* in, float* y) {
  if (!dims) return LH_ERR_ARG;
  return guarded(ctx, [=](Ctx& c) { return nn_forward_host(c, dims, in, y); });
}
int lh_warmstart_step(lh_ctx* ctx, const int dims[3], const float* in, const float* target, const float* mask,
                      float lr, float* grad, float* loss) {
  if (!dims) return LH_ERR_ARG;
  return guarded(ctx, [=](Ctx& c) { return nn_step_host(c, dims, in, target, mask, lr, grad, loss); });
}
int lh_warmstart_set_loss(lh_ctx* ctx, int mode) {
  return guarded(ctx, [=](Ctx& c) {
    if (!c.net || (mode != 0 && mode != 1)) {
      c.err = "lh_warmstart_set_loss: no network or mode not 0 / 1";
      return LH_ERR_ARG;
    }
    c.net->loss_mode = mode;
    return LH_OK;
  });
}
int lh_warmstart_train(lh_ctx* ctx, float lr, float* grad, float* loss) {
  return guarded(ctx, [=](Ctx& c) {
    int rc = nn_train_last_solve(c, lr, loss);
    if (rc == LH_OK && grad) rc = nn_get_params(c, grad, (int64_t)c.net->nlog, true);
    return rc;
  });
}
int lh_warmstart_sample(lh_ctx* ctx, float* in, float* target, float* mask) {
  return guarded(ctx, [=](Ctx& c) { return nn_sample_last_solve(c, in, target, mask); });
}

int lh_write_positions(lh_ctx* ctx, const char* path, int append) {
  if (!ctx || !path) return LH_ERR_ARG;
  const int64_t n = ctx->c.P;
  std::vector<float> pos((size_t)3 * n);
  int rc = lh_get_particles(ctx, pos.data(), nullptr, nullptr, n);
  if (rc) return rc;
  std::ofstream w(path, append ? (std::ios::out | std::ios::app) : (std::ios::out | std::ios::trunc));
  if (!w.is_open()) {
    ctx->c.err = std::string("Failed to open file: ") + path;  // ParticleSystem.cpp:72-74
    return LH_ERR_ARG;
  }
  w << n << "\n";
  for (size_t i = 0; i < (size_t)3 * n; ++i) w << pos[i] << ",";
  w << "\n";
  return w.good() ? LH_OK : LH_ERR_ARG;
}

int lh_load_positions_line(lh_ctx* ctx, const char* line) {
  if (!ctx || !line) return LH_ERR_ARG;
  const int64_t n = ctx->c.P;
  std::vector<float> pos((size_t)3 * n), vel((size_t)3 * n), C((size_t)9 * n);
  int rc = lh_get_particles(ctx, pos.data(), vel.data(), C.data(), n);
  if (rc) return rc;
  // ParticleSystem<DIM>::LoadFromLine: tokens separated by ',', the first one
  // skipped, then x,y,z per particle; a missing token leaves the value unchanged
  std::istringstream ss(line);
  std::string tok;
  std::getline(ss, tok, ',');
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      if (!std::getline(ss, tok, ',')) continue;
      try {
        pos[3 * i + a] = std::stof(tok);
      } catch (const std::exception&) {
        ctx->c.err = "Failed to recognize line";
        return LH_ERR_ARG;
      }
    }
  return lh_set_particles(ctx, pos.data(), vel.data(), C.data(), n);
}

namespace {
struct SnapHeader {
  char magic[8];
  int32_t n[3];
  int32_t correct_step;
  int64_t particles;
  double time;
  uint64_t correct_count;
  float last_maxv;
  int32_t reserved;
};
size_t valid_count(const Ctx& c) { return (size_t)(c.g.N[0] + 1) * (c.g.N[1] + 1) * (c.g.N[2] + 1); }
}  // namespace

int lh_save(lh_ctx* ctx, const char* path) {
  if (!ctx || !path) return LH_ERR_ARG;
  Ctx& c = ctx->c;
  const int64_t P = c.P;
  std::vector<float> pos(3 * (size_t)P), vel(3 * (size_t)P), C(9 * (size_t)P);
  std::vector<uint8_t> valid(valid_count(c));
  int rc = lh_get_particles(ctx, pos.data(), vel.data(), C.data(), P);
  if (rc) return rc;
  rc = lh_get_field(ctx, LH_FIELD_IS_VALID, valid.data(), valid.size());
  if (rc) return rc;
  SnapHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, "LHSNAP01", 8);
  for (int a = 0; a < 3; ++a) h.n[a] = c.g.N[a];
  h.correct_step = c.correct_step;
  h.particles = P;
  h.time = c.time;
  h.correct_count = c.correct_count;
  h.last_maxv = c.last_maxv;
  std::ofstream f(path, std::ios::binary);
  f.write((const char*)&h, sizeof(h));
  f.write((const char*)pos.data(), pos.size() * sizeof(float));
  f.write((const char*)vel.data(), vel.size() * sizeof(float));
  f.write((const char*)C.data(), C.size() * sizeof(float));
  f.write((const char*)valid.data(), valid.size());
  if (!f) {
    c.err = std::string("lh_save: cannot write ") + path;
    return LH_ERR_ARG;
  }
  return LH_OK;
}

int lh_load(lh_ctx* ctx, const char* path) {
  if (!ctx || !path) return LH_ERR_ARG;
  Ctx& c = ctx->c;
  std::ifstream f(path, std::ios::binary);
  SnapHeader h;
  if (!f || !f.read((char*)&h, sizeof(h)) || std::memcmp(h.magic, "LHSNAP01", 8) != 0) {
    c.err = std::string("lh_load: not a snapshot file: ") + path;
    return LH_ERR_ARG;
  }
  if (h.n[0] != c.g.N[0] || h.n[1] != c.g.N[1] || h.n[2] != c.g.N[2] || h.particles < 0) {
    c.err = "lh_load: the snapshot's grid does not match this context";
    return LH_ERR_ARG;
  }
  const int64_t P = h.particles;
  std::vector<float> pos(3 * (size_t)P), vel(3 * (size_t)P), C(9 * (size_t)P);
  std::vector<uint8_t> valid(valid_count(c));
  f.read((char*)pos.data(), pos.size() * sizeof(float));
  f.read((char*)vel.data(), vel.size() * sizeof(float));
  f.read((char*)C.data(), C.size() * sizeof(float));
  f.read((char*)valid.data(), valid.size());
  if (!f) {
    c.err = std::string("lh_load: truncated snapshot ") + path;
    return LH_ERR_ARG;
  }
  int rc = lh_set_particles(ctx, pos.data(), vel.data(), C.data(), P);
  if (rc) return rc;
  rc = lh_set_field(ctx, LH_FIELD_IS_VALID, valid.data(), valid.size());
  if (rc) return rc;
  c.time = h.time;
  c.correct_step = h.correct_step;
  c.correct_count = h.correct_count;
  c.last_maxv = h.last_maxv;
  return LH_OK;
}

int lh_set_time(lh_ctx* ctx, double t) {
  if (!ctx) return LH_ERR_ARG;
  ctx->c.time = t;
  return LH_OK;
}

int lh_get_time(const lh_ctx* ctx, double* t) {
  if (!ctx || !t) return LH_ERR_ARG;
  *t = ctx->c.time;
  return LH_OK;
}

}  // extern "C"
