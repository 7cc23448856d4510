/*
 * lh_api.h -- C ABI of the MI355X liquid-step library (liblh.so).
 *
 * The reference exposes no FFI: its boundary is the C++ class surface of
 * Scene<3> (Scene.h:14-48), LiquidModel<3> (LiquidModel.h:11-61) and
 * CGSolver<3> (PressureSolver.h:81-145).  Every entry point below replaces one
 * method of that surface (cited per function).  The C++ facade in
 * include/lh/liquid.hpp and the Python binding in liquidhairinteration_amd/
 * are thin wrappers over these functions.
 *
 * Conventions
 *  - every call returns an int status (LH_OK == 0); no exception crosses the ABI;
 *    lh_last_error(ctx) returns the message of the last failure on ctx;
 *  - the context owns all device memory; host pointers are borrowed for the
 *    duration of the call (copy in / copy out);
 *  - calls on one context are serialised on the context's HIP stream and return
 *    after completion; one context per host thread;
 *  - grid arrays use the reference Array3 layout (Array.h:78-83): x-major,
 *    z fastest, index = (x*Ny + y)*Nz + z, with the shape of each field given by
 *    lh_field_shape();
 *  - particle arrays: pos[3*i+a], vel[3*i+a], C[9*i + 3*col + row] (the Eigen
 *    column-major affine matrix of Particle<3>, ParticleSystem.h:5-22).
 */
#ifndef LH_API_H
#define LH_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define LH_OK 0
#define LH_ERR_CONFIG 1        /* bad / missing JSON key (SceneParameter.h:22-43 throws) */
#define LH_ERR_OOM 2           /* device allocation failed */
#define LH_ERR_NOT_CONVERGED 3 /* PCG hit maxIterations (PressureSolver.cpp:929-931 asserts) */
#define LH_ERR_NAN 4           /* NaN/Inf alpha or beta (PressureSolver.cpp:884-887, 908-911) */
#define LH_ERR_DEVICE 5        /* HIP runtime error */
#define LH_ERR_ARG 6           /* invalid argument (sizes, ids, null pointers) */
#define LH_ERR_DOMAIN 7        /* a particle left the grid box (binning needs in-box cells) */

/* ---- field ids for lh_get_field / lh_set_field ------------------------ */
enum lh_field {
  LH_FIELD_VX = 0,          /* (N+1,N,N)   Grid<3>::vx       Grid.h:152 */
  LH_FIELD_VY = 1,          /* (N,N+1,N)   Grid<3>::vy */
  LH_FIELD_VZ = 2,          /* (N,N,N+1)   Grid<3>::vz */
  LH_FIELD_TEMPVX = 3,      /* extrapolation ping-pong buffers, Grid.h:157-159 */
  LH_FIELD_TEMPVY = 4,
  LH_FIELD_TEMPVZ = 5,
  LH_FIELD_LIQUID_WX = 6,   /* solid face openness, Grid.h:165-167 */
  LH_FIELD_LIQUID_WY = 7,
  LH_FIELD_LIQUID_WZ = 8,
  LH_FIELD_PRESSURE = 9,    /* (N,N,N) */
  LH_FIELD_LIQUID_PHI = 10, /* (N,N,N) */
  LH_FIELD_SOLID_PHI = 11,  /* (N+1,N+1,N+1) nodal */
  LH_FIELD_IS_VALID = 12,   /* (N+1,N+1,N+1) uint8, Grid.h:160 */
  LH_FIELD_RHS = 13,        /* (N,N,N) CGSolver<3>::m_rhs of the last solve */
  LH_FIELD_COUNT = 14
};

/* ---- solver / model options (mirror the JSON keys) -------------------- */
enum lh_precond { LH_PRECOND_NONE = 0, LH_PRECOND_JACOBI = 1, LH_PRECOND_MULTIGRID = 2 };
enum lh_tol_mode { LH_TOL_ABSOLUTE = 0, LH_TOL_RELATIVE = 1 };
/* reference: Q4/Q5 smoother semantics of PressureSolver.cpp:821-837 with z[0] kept
 * between applications; symmetric: z <- z + w(r-Az)/D, z[0] zeroed per application,
 * post-smoothing in reverse colour order (a fixed SPD preconditioner). */
enum lh_smoother { LH_SMOOTHER_REFERENCE = 0, LH_SMOOTHER_SYMMETRIC = 1 };

typedef struct lh_solve_stats {
  int iterations;          /* PCG iterations run */
  int converged;           /* 1 if the stop rule was met */
  int outer_iterations;    /* FP64 refinement sweeps (0 in plain FP32 mode) */
  int warm_start_used;     /* 1 if a learned initial guess was kept */
  double initial_residual; /* sqrt(sum r0^2) as the solver measures it */
  double final_residual;   /* last residual the stop rule saw */
  double rhs_norm;         /* ||b||_2 */
  double true_residual;    /* FP64 ||b - A x||_2 / ||b||_2 over liquid cells (refined mode) */
  double solve_ms;         /* device time of the solve (hip events) */
} lh_solve_stats;

typedef struct lh_step_stats {
  lh_solve_stats solve;
  double divergence_mean;  /* LiquidModel<3>::ComputeDivergence, LiquidModel.cpp:1065-1083 */
  double divergence_linf;  /* max |div| over cells with liquidPhi <= 0 */
  int fluid_cells;         /* count of cells with liquidPhi <= 0 */
  float max_velocity;      /* max_p max_a |v_pa| (LiquidModel.h:49-55) */
  float cfl_dt;            /* clamp(dx / max_velocity, 0.001, 0.1) */
  double sim_time;         /* simulated time after the step */
  double step_ms;          /* device time of the whole step */
} lh_step_stats;

typedef struct lh_ctx lh_ctx;

/* ---- lifecycle: Scene<3>(configFile) + LiquidModel<3> ctor, Scene.cpp:13-20 */
int lh_create(const char* json_path, lh_ctx** out);
int lh_create_from_string(const char* json_text, lh_ctx** out);
void lh_destroy(lh_ctx* ctx);
/* message of the last failure on ctx; with ctx == NULL, of the last failed context-free call
 * (lh_divergence_noalloc) on the calling thread */
const char* lh_last_error(const lh_ctx* ctx);
/* grid geometry: Grid<3> ctor, Grid.cpp:86-98 */
int lh_grid_info(const lh_ctx* ctx, int dims[3], float min_coord[3], float dx[3]);
int lh_field_shape(const lh_ctx* ctx, int field, int shape[3]);

/* ---- LiquidModel<3> driver surface -------------------------------------- */
int lh_init(lh_ctx* ctx);                            /* LiquidModel.cpp:506-512 */
int lh_pre_step(lh_ctx* ctx, float dt);              /* LiquidModel.cpp:514-517 */
int lh_step(lh_ctx* ctx, float dt);                  /* LiquidModel.cpp:519-528 */
int lh_post_step(lh_ctx* ctx, float dt, lh_step_stats* stats); /* LiquidModel.cpp:530-537 */
int lh_get_cfl(lh_ctx* ctx, float* dt);              /* LiquidModel.h:49-55 */
/* main.cpp:47-53 loop without rendering: nsteps x {PreStep, Step, PostStep, dt=GetCFL}.
 * *dt is the first step's dt on entry and the next dt on return. */
int lh_run_steps(lh_ctx* ctx, int nsteps, float* dt, lh_step_stats* last);

/* ---- stage entry points (public stage methods of LiquidModel.h:25-47) ---- */
int lh_compute_solid_phi(lh_ctx* ctx);               /* LiquidModel.cpp:562-611 */
/* LiquidModel<3>::SetBoundary (LiquidModel.h:57): replace the solid by a boundary
 * tree given as the JSON text of a config "Boundary" node (Rect / Sphere leaves
 * under Union / Intersection / Subtract / None, Boundary.h:27-94; at most 32
 * nodes).  Like the reference it does not recompute solidPhi: call
 * lh_compute_solid_phi, before each step for a moving solid. */
int lh_set_boundary(lh_ctx* ctx, const char* json);
int lh_clear_grid(lh_ctx* ctx);                      /* Grid.cpp:157-169 */
int lh_compute_liquid_phi(lh_ctx* ctx);              /* LiquidModel.cpp:539-560 */
int lh_p2g(lh_ctx* ctx);                             /* LiquidModel.cpp:613-721 */
int lh_apply_gravity(lh_ctx* ctx, float dt);         /* LiquidModel.cpp:898-910 */
/* CGSolver<3>::Solve(liquidDensity, dt), PressureSolver.h:104 / PressureSolver.cpp:839-934:
 * BuildRHS, BuildLHS(rho, dt) on every level, MG-PCG; writes the pressure field.  The
 * step (lh_step) passes the scene's density, as LiquidModel.cpp:830-831 does.  rho must be
 * positive and finite (LH_ERR_ARG otherwise). */
int lh_solve_pressure(lh_ctx* ctx, float rho, float dt, lh_solve_stats* stats);
int lh_velocity_update(lh_ctx* ctx, float dt);       /* LiquidModel.cpp:834-889 */
int lh_extrapolate(lh_ctx* ctx);                     /* LiquidModel.cpp:937-1063 */
int lh_g2p(lh_ctx* ctx);                             /* LiquidModel.cpp:723-823 */
int lh_advect(lh_ctx* ctx, float dt);                /* LiquidModel.cpp:912-935 */
int lh_divergence(lh_ctx* ctx, double* mean, double* linf, int* count); /* LiquidModel.cpp:1065-1083 */
/* LiquidModel.cpp:1085-1132: particles with id % correctCycle == step pushed apart by their
 * neighbours closer than 2 dx, each counted as often as the reference's hash search
 * (NeighborSearch.h:22-57, 131-136; 2 NumParticles slots) visits it: cells the two 3x3x3 regions
 * share plus slot collisions; coincident particles get a seeded counter-RNG kick; positions are
 * written after all are computed; step advances.  Runs inside lh_step after lh_advect when
 * "correctVolume": true. */
int lh_correct_volume(lh_ctx* ctx, float dt);

/* ---- solver-level access (CGSolver<3>::BuildRHS / BuildLHS / Solve split,
 * PressureSolver.cpp:682-698, 544-664, 839-934) ---------------------------- */
/* BuildRHS + BuildLHS(liquidDensity, dt) (PressureSolver.h:97-99) on all levels from the
 * current liquidPhi / liquidW / v */
int lh_build_system(lh_ctx* ctx, float rho, float dt);
/* solve A x = b on the built operator in the configured precision (FP32 MGPCG,
 * or FP32 MGPCG under the FP64 refinement loop with "precision": "refined");
 * b == NULL uses the built right-hand side; x0 (optional) is an initial guess,
 * dropped when ||b - A x0|| > ||b||.  x is written to the pressure field. */
int lh_solve_rhs(lh_ctx* ctx, const float* b, const float* x0, lh_solve_stats* stats);
/* the last solution in FP64 (refined mode; the FP32 pressure widened otherwise) */
int lh_get_solution64(lh_ctx* ctx, double* host, size_t count);
/* multigrid level l of the built operator: its dims and the number of z-rows (x, y)
 * holding a liquid cell -- the rows the smoother and the PCG vector kernels visit */
int lh_solver_level_info(lh_ctx* ctx, int level, int dims[3], int64_t* active_rows);
/* multigrid level l of the built operator: the z-segments its colour passes visit when it keeps a
 * segment list (the finest level, when its z extent is a multiple of 64 cells): the cells per
 * segment (0 without a list) and the number of segments holding a liquid cell.  No counterpart in
 * the reference: introspection of this library's work lists, for profiling (as
 * lh_solver_level_info). */
int lh_solver_level_segments(lh_ctx* ctx, int level, int* segment_cells, int64_t* segments);
/* initial guess used by the following pressure solves (NULL removes it) */
int lh_set_warm_start(lh_ctx* ctx, const float* x0, size_t count);

/* ---- learned initial guess --------------------------------------------------
 * No reference counterpart: the reference's CGSolver<3>::Init starts every solve
 * from x = 0 (PressureSolver.cpp:666-680).  A 3D convolutional encoder-decoder
 * (channels 16/32/64/96, 3x3x3 kernels, BF16 weights and activations, FP32
 * accumulation) maps (divergence / rms, cell type +1 liquid / 0 air / -1 solid)
 * on the level-0 grid to a pressure guess; the guess is restricted to the cells
 * the PCG updates and scaled by alpha = (x0.b)/(x0.A x0), and the PCG still runs
 * to the configured tolerance from it, falling back to x = 0 whenever
 * ||b - A x0|| > ||b||.  Grid dimensions must be multiples of 8.
 * Tensors handed across this ABI are channels-last: in[v * 2 + ch] with v the
 * Array3 index (z fastest) of dims[0] x dims[1] x dims[2]. */
int lh_warmstart_init(lh_ctx* ctx, uint64_t seed);      /* fresh He-initialised network */
int lh_warmstart_load(lh_ctx* ctx, const char* path);
int lh_warmstart_save(lh_ctx* ctx, const char* path);
/* on != 0: every following pressure solve starts from the network's guess */
int lh_warmstart_enable(lh_ctx* ctx, int on);
/* number of parameters; get/set them (per layer W[cout][27][cin] with tap index
 * (dx+1)*9 + (dy+1)*3 + (dz+1), then bias[cout]; layers E0 E1 E2 E3 D2 D1 D0 OUT) */
int lh_warmstart_num_params(lh_ctx* ctx, int64_t* n);
int lh_warmstart_get_params(lh_ctx* ctx, float* w, int64_t n);
int lh_warmstart_set_params(lh_ctx* ctx, const float* w, int64_t n);
/* target scale: the training target is p / rms(p over interior liquid cells) / scale */
int lh_warmstart_set_target_scale(lh_ctx* ctx, float scale);
/* network output y[v] for the input in[v * 2 + ch] */
int lh_warmstart_forward(lh_ctx* ctx, const int dims[3], const float* in, float* y);
/* loss = sum mask (y - target)^2 / sum mask and its gradient (parameter layout
 * above; grad may be NULL); with lr > 0 also one Adam step (beta 0.9 / 0.999) */
int lh_warmstart_step(lh_ctx* ctx, const int dims[3], const float* in, const float* target, const float* mask,
                      float lr, float* grad, float* loss);
/* training loss on the simulator's systems: 0 (default) masked MSE against
 * p / rms(p); 1 energy norm: 1 - (y.b)^2 / ((y.Ay)(x*.b)), the squared A-norm error of
 * the energy-scaled guess relative to that of the zero guess */
int lh_warmstart_set_loss(lh_ctx* ctx, int mode);
/* loss and gradient (grad may be NULL) on the last built system and its solved
 * pressure; with lr > 0 also one Adam step */
int lh_warmstart_train(lh_ctx* ctx, float lr, float* grad, float* loss);
/* the last built system's training pair: in (ncell x 2, BF16-rounded), target, mask */
int lh_warmstart_sample(lh_ctx* ctx, float* in, float* target, float* mask);

/* ---- data ---------------------------------------------------------------- */
int lh_num_particles(const lh_ctx* ctx, int64_t* n);
/* ParticleSystem<3>::Resize + SetPosition/SetVelocity/SetAffineMatrix
 * (ParticleSystem.cpp:10-21, ParticleSystem.h:44-49); vel / C may be NULL (zero). */
int lh_set_particles(lh_ctx* ctx, const float* pos, const float* vel, const float* C, int64_t n);
int lh_get_particles(lh_ctx* ctx, float* pos, float* vel, float* C, int64_t n);
/* ParticleSystem<3>::Resize (ParticleSystem.cpp:10-21), as Scene<3>::PlaySimulation calls it with
 * the count line of a trajectory file (Scene.cpp:34-35): the first min(count, n) particles keep
 * their state, new particles start with zero position, velocity and affine matrix (the
 * reference leaves them uninitialised) and the default radius 2 dx. */
int lh_resize_particles(lh_ctx* ctx, int64_t n);
/* Scene<3>::GetLiquidDensity (Scene.h:23): the config's LiquidParameter.density */
int lh_liquid_density(const lh_ctx* ctx, float* rho);
/* count = number of elements of the field (uint8 for LH_FIELD_IS_VALID, float otherwise) */
int lh_get_field(lh_ctx* ctx, int field, void* host, size_t count);
int lh_set_field(lh_ctx* ctx, int field, const void* host, size_t count);
/* ---- live kernel timing (HIP events on the context's stream) --------------
 * No reference counterpart (the reference has no timers; it only
 * prints residuals under _DEBUG, PressureSolver.cpp:925).  While enabled, the listed kernels are bracketed
 * by hipEventRecord pairs; lh_profile_read returns launch count and summed ms. */
enum lh_prof_kernel {
  LH_PROF_SMOOTH_L0 = 0,   /* red-black smoother on the finest level: one colour pass (one fused sweep when fused) */
  LH_PROF_P2G = 1,         /* particle-to-grid gather (the three components in one launch) */
  LH_PROF_G2P = 2,         /* grid-to-particle + advection */
  LH_PROF_LIQUID_PHI = 3,  /* liquid SDF gather */
  LH_PROF_SPMV = 4,        /* PCG A*p + dot */
  LH_PROF_NN_FWD = 5,      /* warm-start network forward (all layers) */
  LH_PROF_NN_BWD = 6,      /* warm-start network backward (all layers, gradients only) */
  LH_PROF_CONV_D0 = 7,     /* the network's largest convolution: 48 -> 16 channels at full resolution */
  LH_PROF_SMOOTH_L1 = 8,   /* red-black smoother on multigrid level 1: one colour pass (or row sweep) */
  LH_PROF_SMOOTH_L2 = 9,   /* levels 2 and coarser: a colour pass, a row sweep or two Chebyshev steps */
  LH_PROF_COUNT = 10
};
int lh_profile_enable(lh_ctx* ctx, int on);
int lh_profile_read(lh_ctx* ctx, int kernel, int* launches, double* total_ms);
/* ---- trajectory files ------------------------------------------------------
 * ParticleSystem<3>::WriteToFile (ParticleSystem.cpp:66-84): a line with the
 * particle count, then one line "x,y,z,x,y,z,...," with the default ostream
 * float format (6 significant digits); append != 0 opens in append mode.
 * LoadFromLine (ParticleSystem.cpp:86-101): reads positions from such a line
 * into the first min(n, count) particles (missing tokens are skipped). */
int lh_write_positions(lh_ctx* ctx, const char* path, int append);
int lh_load_positions_line(lh_ctx* ctx, const char* line);
/* simulated time and the tank-acceleration clock */
int lh_set_time(lh_ctx* ctx, double t);
int lh_get_time(const lh_ctx* ctx, double* t);
/* ---- snapshots -------------------------------------------------------------
 * No reference counterpart (the reference can only replay positions from its
 * trajectory file, Scene.cpp:26-45).  lh_save writes the state the next step
 * depends on: the particles (positions, velocities, affine matrices, in id
 * order), the persistent isValid field (LiquidModel.cpp:948-951 keeps it across
 * steps), the time, the CorrectVolume cycle counters and the last max |v| (so
 * lh_get_cfl gives the same dt).  lh_load restores it into a context created from
 * the same configuration (same grid; LH_ERR_ARG otherwise): the following steps
 * are bit-identical to the saved run's.  Little-endian binary, magic "LHSNAP01". */
int lh_save(lh_ctx* ctx, const char* path);
int lh_load(lh_ctx* ctx, const char* path);

/* ---- allocation-free variant ----------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  lh_divergence's
 * ComputeDivergence (LiquidModel.cpp:1065-1083) on caller-owned device arrays of a grid of
 * dims[0] x dims[1] x dims[2] cells in the Array3 layout (faces (N0+1,N1,N2), (N0,N1+1,N2),
 * (N0,N1,N2+1); phi per cell), dx the cell size.  It is enqueued on stream (a hipStream_t; NULL is
 * the null stream) and neither allocates nor waits: when the stream reaches it, out (device memory,
 * 3 doubles) holds the mean over the cells with liquidPhi <= 0, the L-inf norm and that count,
 * bit-identical to lh_divergence on the same fields.  The per-block partial sums live in one
 * 24 KiB device array per GPU (looked up on the first call on each device, kept until the process
 * ends) that every call on that GPU shares: calls that overlap on one GPU (other streams, other
 * threads) race.  Limits: 1 <= dims[a] <= 4096 and at most 2^31 - 1 cells; LH_ERR_ARG otherwise,
 * with the message in lh_last_error(NULL). */
int lh_divergence_noalloc(const float* vx, const float* vy, const float* vz, const float* lwx, const float* lwy,
                          const float* lwz, const float* phi, const int dims[3], float dx, double* out,
                          void* stream);

/* ---- the 2D model: LiquidModel<2> / CGSolver<2> --------------------------
 * The reference's other instantiation of the same classes (LiquidModel.cpp:10-474,
 * PressureSolver.cpp:9-418), selected there by DIMENSION (Definition.h:3).  A 2D
 * context is created from a config whose gridDimension / minCoord / maxCoord have
 * two entries; Boundary trees use 2D centers and extents (Sphere = circle); without
 * FluidBlocks, Init spawns the reference's lattice (spacing dx / 2 in the box shrunk
 * by 30 % per side, at most numParticles).  Particles: pos and vel 2 floats, C 4
 * floats (2x2 column-major) per particle, in spawn order.  Fields use the lh_field
 * ids that exist in 2D (v, tempV, liquidW x/y, pressure, liquidPhi, solidPhi,
 * isValid, rhs) with Array2 layout (x-major, y fastest). */
typedef struct lh2_ctx lh2_ctx;
int lh2_create(const char* json_path, lh2_ctx** out);
int lh2_create_from_string(const char* json_text, lh2_ctx** out);
void lh2_destroy(lh2_ctx* ctx);
const char* lh2_last_error(const lh2_ctx* ctx);
int lh2_init(lh2_ctx* ctx);                              /* LiquidModel.cpp:44-49 */
int lh2_pre_step(lh2_ctx* ctx, float dt);                /* :51-55 */
int lh2_step(lh2_ctx* ctx, float dt, lh_solve_stats* stats); /* :57-66 */
int lh2_get_cfl(lh2_ctx* ctx, float* dt);                /* LiquidModel.h:49-55 */
int lh2_run_steps(lh2_ctx* ctx, int n, float* dt, lh_solve_stats* last);
int lh2_clear_grid(lh2_ctx* ctx);                        /* Grid.cpp:52-61 */
int lh2_compute_liquid_phi(lh2_ctx* ctx);                /* LiquidModel.cpp:78-98 */
int lh2_compute_solid_phi(lh2_ctx* ctx);                 /* :100-125 */
int lh2_set_boundary(lh2_ctx* ctx, const char* json);    /* LiquidModel.h:57 */
int lh2_p2g(lh2_ctx* ctx);                               /* :127-186 */
int lh2_apply_gravity(lh2_ctx* ctx, float dt);           /* :306-313 */
int lh2_solve_pressure(lh2_ctx* ctx, float dt, lh_solve_stats* stats); /* PressureSolver.cpp:329-416 */
int lh2_velocity_update(lh2_ctx* ctx, float dt);         /* LiquidModel.cpp:238-291 */
int lh2_extrapolate(lh2_ctx* ctx);                       /* :341-409 */
int lh2_g2p(lh2_ctx* ctx);                               /* :188-236 */
int lh2_advect(lh2_ctx* ctx, float dt);                  /* :315-339 */
int lh2_correct_volume(lh2_ctx* ctx, float dt);          /* :428-468 */
int lh2_divergence(lh2_ctx* ctx, double* mean, double* linf, int* count); /* :411-426 */
int lh2_num_particles(const lh2_ctx* ctx, int64_t* n);
int lh2_set_particles(lh2_ctx* ctx, const float* pos, const float* vel, const float* C, int64_t n);
int lh2_get_particles(lh2_ctx* ctx, float* pos, float* vel, float* C, int64_t n);
int lh2_get_field(lh2_ctx* ctx, int field, void* host, size_t count);
int lh2_set_field(lh2_ctx* ctx, int field, const void* host, size_t count);
int lh2_set_time(lh2_ctx* ctx, double t);
int lh2_get_time(const lh2_ctx* ctx, double* t);

#ifdef __cplusplus
}
#endif
#endif /* LH_API_H */
