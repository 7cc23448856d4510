// lh_ctx.h -- device-side state of one liquid scene and the launchers of the
// HIP kernels (k_grid.hip, k_particles.hip, k_solver.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/lh_api.h"

namespace lh {

struct Net;  // learned initial guess (lh_nn.h)

// Scene configuration: the reference's four parameter structs
// (SceneParameter.h:9-152) plus the build's optional keys.
struct Config {
  int N[3] = {0, 0, 0};
  float mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
  float gravity = -9.81f, density = 1000.f, cfg_dt = 0.01f, viscosity = 0.f, surface_tension = 0.f;
  int correct_cycle = 8;
  int num_particles = 0;
  // SolverParameter
  int max_iter = 100;
  float tol = 1e-5f;
  int tol_mode = LH_TOL_ABSOLUTE;
  int precond = LH_PRECOND_MULTIGRID;
  int num_level = 1, num_pre = 5, num_post = 5, num_final = 10;
  float omega = 0.5f;
  int smoother = LH_SMOOTHER_REFERENCE;
  // levels with at least this many cells use the fused one-kernel red-black sweep
  long long fused_sweep_min_cells = 1LL << 62;  // off: two colour passes measured faster at 256^3

  // red-black sweeps on the coarse levels (>= 1), all with the values of two colour passes:
  //   kCoarseRows (default): levels of at most row_sweep_max_cells cells in one launch over the
  //     z-support rows (k_sweep_rows; 64^3: 5.8 us per sweep against 2 x 3.9 us), larger ones
  //     as two colour passes (the row sweep's four recomputed neighbour updates per cell make it
  //     slow there: 30 us against 2 x 5.6 us at 128^3);
  //   kCoarseTiles: every coarse level in one launch over 4 x 4 x 64 tiles (k_tile_sweep);
  //   kCoarsePasses: two colour passes (k_smooth) on every coarse level
  int coarse_sweep = 1;
  long long row_sweep_max_cells = 1LL << 18;
  // mixed-precision refinement (FP32 inner MGPCG under an FP64 residual loop)
  int refined = 0;
  // the multigrid coarsest level's red-black sweeps in FP64 (SolverParameter.coarsestPrecision;
  // default: on with refined precision): r and z promoted, z rounded back for the prolongation
  int coarsest_fp64 = 0;
  // coarsestSolver "chebyshev": the coarsest level solved by coarsest_cheb Chebyshev steps on
  // [cheb_lo, 2] in FP64 instead of numFinalSmooth sweeps (0: sweeps)
  int coarsest_cheb = 0;
  float cheb_lo = 0.01f;
  // SolverParameter.pcgGraph: each PCG iteration's x/r update, preconditioner and r.z launched as
  // one instantiated hipGraph instead of one launch at a time
  int pcg_graph = 0;
  // "p2gForm": "shared" (default: a wave's records evaluated once per wave and handed to the face
  // lanes through LDS) or "gather" (every lane evaluates every record of its window)
  int p2g_gather = 0;
  double refined_tol = 1e-9;
  int max_outer = 50;
  // boundary SDF program (postfix, 12 floats per node)
  std::vector<float> sdf;
  // jittered fluid blocks: min[3], max[3], vel[3], sigma
  std::vector<float> blocks;
  int k_ppc = 2;
  uint64_t seed = 0;
  // tank acceleration a(t) = amp * sin(2 pi freq t) on axis
  float tank_amp = 0.f, tank_freq = 0.f;
  int tank_axis = -1;
  float first_dt = 0.01f;
  int steps = 0;
  int correct_volume = 0;  // CorrectVolume after AdvectParticle (commented out of the reference's 3D step)
};

struct GridP {
  int N[3];
  float mn[3], dx[3];
};

enum { kCoarsePasses = 0, kCoarseRows = 1, kCoarseTiles = 2 };
constexpr int kPinDiv = 24;  // first of the three pinned host doubles the divergence reduction writes

constexpr int kSegZ = 64;  // cells per z-segment of the finest level's work list

struct Level {
  int d[3] = {0, 0, 0};
  size_t n = 0;
  float* adiag = nullptr;
  float* t[3] = {nullptr, nullptr, nullptr};  // term of the +axis face of each cell
  uint8_t* cls = nullptr;                     // 0: phi > 0, 1: phi == 0 (or NaN), 2: phi < 0
  // the classes a cell's stencil needs, in one load: bits 0-5 the +x, -x, +y, -y, +z, -z neighbour
  // exists and has cls 2 (the off-diagonal terms), bit 6 the cell's own cls is 0, bits 8-13 the
  // neighbour exists and has cls != 0 (built with cls by launch_build_lhs)
  uint16_t* cmask = nullptr;
  float* r = nullptr;
  float* z = nullptr;
  float* z2 = nullptr;  // ping-pong partner of z for the fused red-black sweep
  double* z64 = nullptr;   // the coarsest level's z in FP64 (coarsest_fp64) ...
  double* z64b = nullptr;  // ... and its ping-pong partner
  double* d64 = nullptr;   // the coarsest level's Chebyshev direction (coarsest_cheb) ...
  double* d64b = nullptr;  // ... and its ping-pong partner
  // z-rows (x, y) holding at least one liquid cell (cls != 0), ascending; rebuilt with the operator
  int* rows = nullptr;   // capacity d[0] * d[1]
  int* nrows = nullptr;  // device count
  int* aflag = nullptr;  // per-row flag of that set
  // level 0 when d[2] is a multiple of kSegZ: the kSegZ-cell z-segments of those rows that hold a
  // liquid cell, ascending as row * (d[2] / kSegZ) + segment (the finest colour passes' work list)
  int* smask = nullptr;  // per row, bit s: segment s holds a cell with cls != 0
  int* segs = nullptr;   // capacity d[0] * d[1] * d[2] / kSegZ
  int* nsegs = nullptr;  // device count
  // levels >= 1, rows of the restriction: where it can produce a nonzero r (parents of the
  // finer level's liquid and restriction rows) or must clear z (the z-support below)
  int* qrows = nullptr;
  int* nqrows = nullptr;
  int* qflag = nullptr;
  // levels >= 1, z-support: rows where z can be nonzero -- this level's liquid rows and the
  // children of the coarser level's z-support (prolongation adds there); prolongation rows
  int* zrows = nullptr;
  int* nzrows = nullptr;
  int* zflag = nullptr;
  // levels >= 1: the 4 x 4-row tiles holding a restriction row, for the tiled red-black sweep
  int* tflag = nullptr;
  int* tiles = nullptr;
  int* ntiles = nullptr;
};

// One particle is one 64-byte record of four float4: R0 = (x, y, z, id bits) at float4 index i of a
// record buffer, and R1 = (C.col(0), v_y), R2 = (C.col(1), v_z), R3 = (C.col(2), v_x) at float4
// index cap + 3 i .. + 2 (C column-major: C.col(a) holds the gradient of velocity component a).
// kRecFloats floats per particle per buffer.
constexpr int kRecFloats = 16;

struct Ctx {
  Config cfg;
  GridP g;
  int dW[3] = {2, 2, 2};  // ceil(radius / dx_a), radius = 2 dx_x (ParticleSystem.cpp:6-8)
  float radius = 0.f;
  hipStream_t s = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t ev_iter = nullptr;  // end of a PCG iteration's last kernel the host waits for
  // grid (Grid<3>, Grid.h:137-170)
  float *v[3] = {}, *tv[3] = {}, *lw[3] = {};
  float *pressure = nullptr, *phi = nullptr, *solid_phi = nullptr, *rhs = nullptr;
  uint8_t* is_valid = nullptr;  // (N+1)^3
  uint8_t* xstate = nullptr;    // extrapolation sweep state, sized for the largest face array
  size_t nface[3] = {0, 0, 0}, ncell = 0, nnode = 0;
  float* d_sdf = nullptr;
  int nsdf = 0;
  // particles: record buffers (the layout above), double buffered for the per-step bin-order
  // permutation; prec[cur] holds the state, in bin order after each binning
  int64_t P = 0, cap = 0;
  float* prec[2] = {nullptr, nullptr};
  int cur = 0;
  // half-cell bins: 2N per axis
  int H[3] = {0, 0, 0};
  size_t nbins = 0;
  int* bin_count = nullptr;  // kept at zero between steps (the scatter clears the counts it used)
  int* bin_start = nullptr;  // nbins + 1
  int* pkey = nullptr;       // per particle bin key
  int* pslot = nullptr;      // per bin-order slot, the particle the scatter placed there (arrival order)
  int* porder = nullptr;     // per bin-order slot, the particle (source index) that takes it
  int* pid = nullptr;        // per particle, its id (R0.w), written by the key kernel for the rank kernel
  // the bin scan runs over blocks of kScanBins bins: a flag per block (set by the key kernel
  // for every block that holds a particle, cleared by the scan), the block sums and offsets
  int* bin_bflag = nullptr;
  int* bin_bsum = nullptr;
  int* bin_boff = nullptr;
  size_t nbblk = 0;
  uint8_t* scan_tmp = nullptr;  // temporary storage of the scan of the block sums
  size_t scan_bytes = 0;
  bool bins_valid = false;
  // solver (CGSolver<3>)
  std::vector<Level> lv;
  float *p = nullptr, *x = nullptr, *q = nullptr;
  float* p2 = nullptr;  // ping-pong partner of p: the fused p update + SpMV writes the new p here
  double *x64 = nullptr, *r64 = nullptr;  // FP64 refinement solution and true residual
  float* rf = nullptr;                     // FP32 right-hand side of a refinement correction
  float* b_user = nullptr;                 // right-hand side handed in through lh_solve_rhs
  float* x0_user = nullptr;                // initial guess handed in through lh_solve_rhs / lh_set_warm_start
  float* warm_x0 = nullptr;                // installed initial guess for the next pressure solve (or null)
  Net* net = nullptr;                      // learned initial guess network (or null)
  bool warm_learned = false;               // use the network's guess in every pressure solve
  float last_dt = 0.f;                     // dt of the last built system
  lh_solve_stats last_solve = {};          // statistics of the last step's pressure solve
  double* partial = nullptr;  // per-block FP64 partial sums
  int npartial = 0;
  float* dsc = nullptr;       // device scalars: [0] oldRZ [1] newRZ [2] alpha [3] beta [4] residual [5] p.q [6..7] rho
  double* dsc64 = nullptr;    // device FP64 scalars
  unsigned* dmaxv = nullptr;  // max |v| bits
  int* dflag = nullptr;       // misc device flags (extrapolation progress, domain errors)
  double* h_pin = nullptr;    // pinned host scalars (64; [kPinDiv, kPinDiv + 3): divergence sum, L-inf, count)
  int* h_flag = nullptr;
  // device-side addresses of h_pin and h_flag: the PCG's last scalar kernel of an iteration
  // stores the scalars and the NaN flag there itself instead of two copies per iteration
  float* h_pin_dev = nullptr;
  int* h_flag_dev = nullptr;
  double time = 0.0;
  float last_maxv = 0.f;
  int correct_step = 0;              // m_correctStep (LiquidModel.h:73)
  unsigned long long correct_count = 0;  // CorrectVolume calls so far (random-stream counter)
  // per cell: 1 when its hash slot (2 P slots) equals that of another cell within 4 cells,
  // valid for cv_nslot slots (0: not built)
  uint8_t* cv_flag = nullptr;
  int cv_nslot = 0;
  bool solid_ready = false;
  // Extrapolate schedule cache: which faces each sweep validates, and from which
  // neighbours, is a function of isValid at the start of Extrapolate and of
  // liquidW only; once recorded it is replayed (values only) whenever isValid
  // at the start is bit-identical to the recorded snapshot.
  uint8_t* xv_snap = nullptr;   // isValid when the schedule was recorded
  uint8_t* xv_final = nullptr;  // isValid at the end of that Extrapolate
  uint32_t* xlist[3] = {nullptr, nullptr, nullptr};  // (face, 27-bit neighbour mask) per validated face
  size_t xlist_cap[3] = {0, 0, 0};
  int* xcount = nullptr;        // device counters, component * 5 + sweep
  int xcnt[15] = {};
  bool xcache_ok = false;
  // pcgGraph: instantiated graphs of one PCG iteration, keyed by the p buffer and rho slot they
  // read and the levels' ping-pong buffers at capture; `after` is the buffer state the captured
  // host code left (the preconditioner swaps ping-pong pointers as it launches)
  struct PcgGraph {
    hipGraphExec_t exec = nullptr;
    const float* p = nullptr;
    int slot = 0;
    std::vector<void*> before, after;
  };
  std::vector<PcgGraph> pcg_graphs;
  // live kernel timing: event pairs per instrumented kernel
  bool prof = false;
  std::vector<hipEvent_t> prof_ev[LH_PROF_COUNT];
  int prof_used[LH_PROF_COUNT] = {};    // event pairs recorded
  int prof_launch[LH_PROF_COUNT] = {};  // kernel launches inside them
  std::string err;
};

// brackets n consecutive launches of one kernel (nothing else in between) with one
// hipEventRecord pair when c.prof is set: an event pair costs microseconds of queue time,
// so runs of short launches are timed as a whole and divided by their count
struct ProfScope {
  Ctx& c;
  int k, n;
  hipEvent_t e1 = nullptr;
  ProfScope(Ctx& cc, int kk, int nn = 1) : c(cc), k(kk), n(nn) {
    if (!c.prof || k < 0 || k >= LH_PROF_COUNT || n <= 0) return;
    auto& v = c.prof_ev[k];
    int u = c.prof_used[k];
    while ((int)v.size() < 2 * (u + 1)) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      v.push_back(e);
    }
    hipEventRecord(v[2 * u], c.s);
    e1 = v[2 * u + 1];
  }
  ~ProfScope() {
    if (!e1) return;
    hipEventRecord(e1, c.s);
    c.prof_used[k]++;
    c.prof_launch[k] += n;
  }
};

// ---- launchers (k_grid.hip)
void launch_fill_f32(Ctx& c, float* a, size_t n, float v);
void launch_fill_u8(Ctx& c, uint8_t* a, size_t n, uint8_t v);
void launch_solid_phi(Ctx& c);
// Grid<3>::Clear; within a step that computes liquid phi, P2G and the solve on particles
// (in_step and P > 0) the fields those overwrite entirely (liquidPhi, v, pressure) are skipped
void launch_clear(Ctx& c, bool in_step = false);
void launch_velocity_update(Ctx& c, float dt);
// returns number of propagation rounds used; <0 on device error
int run_extrapolate(Ctx& c);
// ComputeDivergence (LiquidModel.cpp:1065-1083) over every cell; waits and returns the mean, the
// L-inf norm and the count of cells with liquidPhi <= 0
void launch_divergence(Ctx& c, double* h_mean, double* h_linf, int* h_count);
// its launches alone (no wait): out[0] = FP64 sum, out[1] = L-inf, out[2] = count
void launch_divergence_dense(Ctx& c, double* out);
// the same on any arrays of the grid g, with nb blocks whose partials (3 nb doubles) go to
// partial; out[0] is the mean instead of the sum when as_mean
void launch_divergence_arrays(hipStream_t s, const float* const v[3], const float* const lw[3], const float* phi,
                              const GridP& g, double* partial, int nb, double* out, int as_mean);
void launch_gravity(Ctx& c, float dt);
// ---- launchers (k_particles.hip)
constexpr int kScanBins = 2048;        // bins per block of the bin scan
size_t bin_scan_bytes(size_t nblocks);  // temporary storage of the scan of nblocks block sums
int run_binning(Ctx& c);  // LH_OK or LH_ERR_DOMAIN
void launch_liquid_phi(Ctx& c);
void launch_p2g(Ctx& c, bool fuse_body_forces, float dt);
void launch_g2p_advect(Ctx& c, bool do_g2p, bool do_advect, float dt);
int run_correct_volume(Ctx& c, float dt);  // LiquidModel<3>::CorrectVolume; LH_OK or LH_ERR_DOMAIN
// ---- launchers (k_solver.hip)
int run_solve(Ctx& c, float rho, float dt, lh_solve_stats* st);  // CGSolver<3>::Solve(liquidDensity, dt)
void build_system(Ctx& c, float rho, float dt);       // Init + BuildRHS + BuildLHS(rho, dt)
// solve A x = b on the built operator (configured precision), x -> pressure
int solve_built(Ctx& c, const float* b, const float* x0, lh_solve_stats* st, bool record_start = true);
void launch_build_rhs(Ctx& c);
void launch_build_lhs(Ctx& c, float rho, float dt);
// the same over the level-0 liquid rows of the operator built from the current liquidPhi, no
// wait: out[0] = FP64 sum, out[1] = L-inf, out[2] = count (out may be pinned host memory)
void launch_divergence_rows(Ctx& c, double* out);
// final reduction of per-block divergence partials (sum, max, count at [b], [nb + b], [2 nb + b])
void launch_div_final_on(hipStream_t s, const double* partial, int nb, double* out, int as_mean);

// message of the last failed context-free call on this thread (lh_noalloc.hip)
const char* noalloc_last_error();

// device error check helper
bool hip_ok(Ctx& c, hipError_t e, const char* what);

inline size_t div_up(size_t a, size_t b) { return (a + b - 1) / b; }

}  // namespace lh
