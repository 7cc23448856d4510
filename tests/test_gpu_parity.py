"""HIP library vs the CPU oracle, stage by stage, on identical inputs (teacher forcing).

Every stage of the liquid step has no order-dependent float reduction on the
GPU except the PCG dot products, which accumulate in FP64 and round to float
exactly like the oracle's dot_mode=1.  The expected result is therefore
bit-exact for every field (np.array_equal treats -0 == +0).
"""
import json

import numpy as np
import pytest

from conftest import config_file
from oracle import Oracle

lh = pytest.importorskip("liquidhairinteration_amd")

pytestmark = pytest.mark.gpu

F = {"vx": 0, "vy": 1, "vz": 2, "tempVx": 3, "tempVy": 4, "tempVz": 5, "liquidWx": 6, "liquidWy": 7, "liquidWz": 8,
     "pressure": 9, "liquidPhi": 10, "solidPhi": 11, "isValid": 12, "rhs": 13}


def small_dam(N=24, smoother="reference", tank=False, fused=False):
    cfg = json.load(open(config_file("dam_break_48.json")))
    cfg["gridDimension"] = [N, N, N]
    wall = 0.5 - 3.0 / N
    cfg["Boundary"]["child"]["halfExtent"] = [wall] * 3
    cfg["SolverParameter"]["smoother"] = smoother
    # 0: every level uses the one-kernel red-black sweep (default: two colour passes)
    cfg["SolverParameter"]["fusedSweepMinCells"] = 0 if fused else 1 << 40
    if smoother == "symmetric":
        cfg["SolverParameter"].update({"smoothingFactor": 1.0, "numLevel": 3, "numPreSmooth": 2, "numPostSmooth": 2})
    if tank:
        cfg["TankAcceleration"] = {"amplitude": 3.0, "frequency": 1.0, "axis": 0}
        cfg["Boundary"] = {"type": "Union", "left": cfg["Boundary"]["child"],
                           "right": {"type": "Sphere", "center": [0.5, 0.5, 0.5], "radius": 0.15}}
    return cfg


def tall_dam():
    """16 x 24 x 160 cells of 1/160 over 0.1 x 0.15 x 1: level 1 has 80-cell z-rows, swept in two
    64-cell segments (the row sweep's segment-end recomputation)"""
    cfg = small_dam(24, "symmetric")
    h = 1.0 / 160
    cfg["gridDimension"] = [16, 24, 160]
    cfg["maxCoord"] = [0.1, 0.15, 1.0]
    cfg["Boundary"]["child"]["center"] = [0.05, 0.075, 0.5]
    cfg["Boundary"]["child"]["halfExtent"] = [0.05 - 3 * h, 0.075 - 3 * h, 0.5 - 3 * h]
    cfg["FluidBlocks"] = [{"min": [0.0, 0.0, 0.0], "max": [0.1, 0.08, 0.55], "velocity": [0, 0, 1.0],
                           "velocityNoise": 0.0}]
    return cfg


def coarse_passes(cfg):
    cfg["SolverParameter"]["coarseSweep"] = "passes"
    return cfg


def coarse_tiles(cfg):
    """every coarse level swept by the tiled one-launch sweep (an option: by default coarse levels
    use the row sweep or two colour passes)"""
    cfg["SolverParameter"]["coarseSweep"] = "tiles"
    return cfg


def coarsest64(cfg):
    """the multigrid coarsest level's sweeps in FP64 (the refined mode's default), here under the
    FP32 PCG so that the oracle's restatement is compared bit for bit"""
    cfg["SolverParameter"]["coarsestPrecision"] = "fp64"
    return cfg


def chebyshev(cfg, steps, lo, levels=None):
    """the coarsest level solved by Chebyshev steps on its Jacobi-scaled system (coarsestSolver
    chebyshev, FP64; an option of this build, the reference's coarsest level is numFinalSmooth
    sweeps, PressureSolver.cpp:754-757), two steps per launch"""
    cfg["SolverParameter"].update({"coarsestPrecision": "fp64", "coarsestSolver": "chebyshev",
                                   "coarsestSteps": steps, "chebyshevLower": lo})
    if levels:
        cfg["SolverParameter"]["numLevel"] = levels
    return cfg


def precond(cfg, kind):
    """CGSolver<3>'s other preconditioners (PressureSolver.cpp:721-743): None (z = r) and Jacobi
    (z = r / Adiag on cells with a nonzero diagonal)"""
    cfg["SolverParameter"]["preconditionerType"] = kind
    cfg["SolverParameter"]["maxIterations"] = 400
    return cfg


def p2g_gather(cfg):
    """P2G as the per-lane gather (p2gForm "gather"; default "shared": a wave's records evaluated
    once and handed to the face lanes through LDS)"""
    cfg["p2gForm"] = "gather"
    return cfg


def small_collision(N=40):
    """two 0.3^3 blocks at opposite corners moving at +-2 m/s (0.4 particles per cell: P2G's tile
    length for liquid that does not fill its tiles)"""
    cfg = small_dam(N, "symmetric")
    cfg["FluidBlocks"] = [{"min": [0.0, 0.0, 0.0], "max": [0.3, 0.3, 0.3], "velocity": [2.0, 0.0, 0.0],
                           "velocityNoise": 0.05},
                          {"min": [0.7, 0.7, 0.7], "max": [1.0, 1.0, 1.0], "velocity": [-2.0, 0.0, 0.0],
                           "velocityNoise": 0.05}]
    return cfg


def odd_grid(dx):
    """17 x 23 x 31 cells of dx (odd on every axis: partial P2G and solver tiles, multigrid levels
    rounded up; dx = 1/32 takes the power-of-two paths, 1/24 the divisions)"""
    cfg = small_dam(24, "symmetric")
    N = [17, 23, 31]
    mx = [n * dx for n in N]
    cfg["gridDimension"] = N
    cfg["maxCoord"] = mx
    cfg["Boundary"]["child"]["center"] = [m / 2 for m in mx]
    cfg["Boundary"]["child"]["halfExtent"] = [m / 2 - 3 * dx for m in mx]
    cfg["FluidBlocks"] = [{"min": [0.0, 0.0, 0.0], "max": [0.5 * mx[0], 0.6 * mx[1], mx[2]],
                           "velocity": [0.3, 0.0, -0.2], "velocityNoise": 0.05}]
    return cfg


def empty_scene():
    """no fluid: zero particles (gravity still reaches the interior y-faces, Q8; nothing to solve)"""
    cfg = small_dam(24, "symmetric")
    cfg["FluidBlocks"] = []
    return cfg


def single_particle():
    """one particle in one cell"""
    cfg = small_dam(24, "symmetric")
    h = 1.0 / 24
    cfg["FluidBlocks"] = [{"min": [0.5, 0.5, 0.5], "max": [0.5 + h, 0.5 + h, 0.5 + h], "velocity": [0.5, 0.0, 0.0],
                           "velocityNoise": 0.0}]
    cfg["particlesPerCellAxis"] = 1
    return cfg


def cv_cycle(cfg, k):
    cfg["LiquidParameter"]["correctCycle"] = k
    return cfg


def same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    bad = int((a != b).sum())
    assert bad == 0, "%s: %d of %d elements differ (max |d| %.3e)" % (
        what, bad, a.size, float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max()))


CASES = {
    "reference_config_64": lambda: json.load(open(config_file("pure_liquid_3d.json"))),
    "dam24_reference_smoother": lambda: small_dam(24, "reference"),
    "dam24_symmetric_smoother": lambda: small_dam(24, "symmetric"),
    # coarse levels as two colour passes per sweep (default: one row-sweep launch)
    "dam24_symmetric_colour_passes": lambda: coarse_passes(small_dam(24, "symmetric")),
    "tall_16x24x160_row_segments": tall_dam,
    "dam24_reference_colour_passes": lambda: coarse_passes(small_dam(24, "reference")),
    "dam24_tank_sphere": lambda: small_dam(24, "symmetric", tank=True),
    "dam24_reference_fused_sweep": lambda: small_dam(24, "reference", fused=True),
    "dam40_symmetric_fused_sweep": lambda: small_dam(40, "symmetric", fused=True),
    # coarse levels in 4 x 4 x 64 tiles (partial tiles in x, y and z: 20^3 and 10^3 levels; the
    # tall grid's 80- and 40-cell rows)
    "dam40_symmetric_tile_sweep": lambda: coarse_tiles(small_dam(40, "symmetric")),
    "dam24_reference_tile_sweep": lambda: coarse_tiles(small_dam(24, "reference")),
    "tall_16x24x160_tile_sweep": lambda: coarse_tiles(tall_dam()),
    # CorrectVolume: at 24^3 (9072 particles, 18144 hash slots) most regions hold a cell whose
    # slot collides with a nearby cell (exact visit count), at 40^3 (60112 particles) few do
    # (overlap product); every second particle id moves per call at 40^3
    "dam24_correct_volume": lambda: dict(small_dam(24, "symmetric"), correctVolume=True),
    "dam40_correct_volume": lambda: cv_cycle(dict(small_dam(40, "symmetric", fused=True), correctVolume=True), 2),
    "dam24_symmetric_coarsest_fp64": lambda: coarsest64(small_dam(24, "symmetric")),
    "dam24_reference_coarsest_fp64": lambda: coarsest64(small_dam(24, "reference")),
    "dam24_precond_none": lambda: precond(small_dam(24, "reference"), "None"),
    "dam24_precond_jacobi": lambda: precond(small_dam(24, "reference"), "Jacobi"),
    "dam32_tank_sphere_jacobi": lambda: precond(small_dam(32, "symmetric", tank=True), "Jacobi"),
    # even and odd step counts (an odd count ends with a one-step launch); two levels of the tall
    # grid put the Chebyshev steps on 80-cell rows (two segments: the segment-end recomputation)
    "dam24_chebyshev_coarsest": lambda: chebyshev(small_dam(24, "symmetric"), 12, 0.02),
    "dam32_tank_sphere_chebyshev_odd": lambda: chebyshev(small_dam(32, "symmetric", tank=True), 7, 0.01),
    "tall_16x24x160_chebyshev_segments": lambda: chebyshev(tall_dam(), 9, 0.005, levels=2),
    # four two-step launches on a 10^3 coarsest level
    "dam40_chebyshev_8_steps": lambda: chebyshev(small_dam(40, "symmetric"), 8, 0.01),
    # both P2G forms at both tile lengths (liquid filling its tiles, and not)
    "dam24_tank_sphere_p2g_gather": lambda: p2g_gather(small_dam(24, "symmetric", tank=True)),
    "collision40": lambda: small_collision(40),
    "collision40_p2g_gather": lambda: p2g_gather(small_collision(40)),
    # ragged and degenerate inputs
    "odd_17x23x31_pow2_dx": lambda: odd_grid(1.0 / 32),
    "odd_17x23x31_dx_1_24": lambda: odd_grid(1.0 / 24),
    "odd_17x23x31_p2g_gather": lambda: p2g_gather(odd_grid(1.0 / 32)),
    "empty_scene": empty_scene,
    "single_particle": single_particle,
}


def same_divergence(g, o, what):
    """ComputeDivergence (LiquidModel.cpp:1065-1083): count and L-inf bit-exact; the mean's FP64
    sums differ only in order, |d mean| <= 2 eps64 L-inf"""
    mg, lg, cg = g
    mo, lo, co = o
    assert cg == co, "%s: count %d vs %d" % (what, cg, co)
    assert lg == lo, "%s: L-inf %r vs %r" % (what, lg, lo)
    assert abs(mg - mo) <= 1e-12 * max(lo, 1e-30), "%s: mean %r vs %r" % (what, mg, mo)


@pytest.mark.parametrize("case", list(CASES))
def test_stages_bit_exact(case):
    print("case", case, flush=True)
    cfg = CASES[case]()
    orc = Oracle(cfg, dot_mode=1, p2g_order=1)
    scene = lh.Scene(cfg)
    m = scene.GetLiquidModel()
    ps = m.GetParticleSystem()
    orc.init()
    m.Init()
    po, vo, co = orc.particles()
    pg, vg, cg = ps.get()
    same(pg, po, "initial positions")
    same(vg, vo, "initial velocities")
    for f in ("solidPhi", "liquidWx", "liquidWy", "liquidWz"):
        same(m.get_field(f), orc.get(F[f]), f)
    dt = np.float32(cfg.get("firstDt", 0.01))
    for step in range(3):
        po, vo, co = orc.particles()
        ps.set(po, vo, co)
        m.set_field("isValid", orc.get(F["isValid"]))
        m.time = orc.time
        orc.clear_grid()
        orc.compute_liquid_phi()
        m.PreStep(dt)
        same(m.get_field("liquidPhi"), orc.get(F["liquidPhi"]), "liquidPhi step %d" % step)
        orc.p2g()
        orc.apply_gravity(dt)
        m.Particle2Grid()
        m.ApplyGravity(dt)
        for f in ("vx", "vy", "vz"):
            same(m.get_field(f), orc.get(F[f]), "P2G+gravity %s step %d" % (f, step))
        st_o = orc.solve_pressure(dt)
        st_g = m.GetSolver().Solve(None, dt)
        assert st_g["iterations"] == st_o.iterations
        same(m.get_field("rhs"), orc.get(F["rhs"]), "rhs step %d" % step)
        same(m.get_field("pressure"), orc.get(F["pressure"]), "pressure step %d" % step)
        orc.velocity_update(dt)
        m.VelocityUpdate(dt)
        for f in ("vx", "vy", "vz"):
            same(m.get_field(f), orc.get(F[f]), "velocity update %s step %d" % (f, step))
        orc.extrapolate()
        m.Extrapolate()
        for f in ("vx", "vy", "vz", "tempVx", "tempVy", "tempVz", "isValid"):
            same(m.get_field(f), orc.get(F[f]), "extrapolate %s step %d" % (f, step))
        orc.g2p()
        m.Grid2Particle()
        pg, vg, cg = ps.get()
        po, vo, co = orc.particles()
        same(vg, vo, "G2P velocity step %d" % step)
        same(cg, co, "G2P affine step %d" % step)
        orc.advect(dt)
        m.AdvectParticle(dt)
        same(ps.get()[0], orc.particles()[0], "advected positions step %d" % step)
        same_divergence(m.ComputeDivergence(), orc.divergence(), "divergence step %d" % step)
        if cfg.get("correctVolume"):
            orc.correct_volume(dt)
            m.CorrectVolume(dt)
            same(ps.get()[0], orc.particles()[0], "CorrectVolume positions step %d" % step)
        orc.set_time(orc.time + float(dt))
        dto = np.float32(orc.get_cfl()[0])
        assert np.float32(m.GetCFL()) == dto
        dt = dto
    scene.close()


@pytest.mark.parametrize("case", ["dam24_reference_smoother", "dam24_tank_sphere", "dam40_symmetric_fused_sweep",
                                  "dam24_correct_volume", "tall_16x24x160_row_segments",
                                  "dam40_symmetric_tile_sweep", "tall_16x24x160_tile_sweep",
                                  "dam24_precond_none", "dam32_tank_sphere_jacobi",
                                  "dam24_symmetric_coarsest_fp64", "dam24_reference_coarsest_fp64",
                                  "dam24_chebyshev_coarsest", "dam32_tank_sphere_chebyshev_odd",
                                  "tall_16x24x160_chebyshev_segments", "dam40_chebyshev_8_steps",
                                  "collision40", "collision40_p2g_gather", "odd_17x23x31_pow2_dx",
                                  "odd_17x23x31_dx_1_24", "empty_scene", "single_particle"])
def test_free_running_steps_bit_exact(case):
    print("case", case, flush=True)
    cfg = CASES[case]()
    orc = Oracle(cfg, dot_mode=1, p2g_order=1)
    orc.init()
    scene = lh.Scene(cfg)
    m = scene.GetLiquidModel()
    m.Init()
    dt = np.float32(cfg.get("firstDt", 0.01))
    dtg = float(dt)
    for _ in range(6):
        st, _, _ = orc.step(dt)
        dt = np.float32(orc.get_cfl()[0])
        dtg, sst = m.RunSteps(1, dtg)
        assert sst["solve"]["iterations"] == st.iterations
        assert np.float32(dtg) == dt
        # PostStep's divergence, computed inside every step of the native loop over the liquid rows
        same_divergence((sst["divergence_mean"], sst["divergence_linf"], sst["fluid_cells"]), orc.divergence(),
                        "per-step divergence")
    pg, vg, cg = m.GetParticleSystem().get()
    po, vo, co = orc.particles()
    same(pg, po, "positions after 6 steps")
    same(vg, vo, "velocities after 6 steps")
    scene.close()


def test_dam_break_128_free_running_bit_exact():
    """the 128^3 dam break (block x in [0, 0.4], y in [0, 0.5], z in [0, 1], 8 particles per
    cell, ~3.3 M particles), 3-level symmetric MGPCG to 1e-6 relative: three free-running
    steps bit-identical to the oracle (particles, iteration counts, CFL dt, liquid mask)"""
    print("case dam128 symmetric", flush=True)
    cfg = small_dam(128, "symmetric")
    orc = Oracle(cfg, dot_mode=1, p2g_order=1)
    orc.init()
    scene = lh.Scene(cfg)
    m = scene.GetLiquidModel()
    m.Init()
    dt = np.float32(cfg.get("firstDt", 0.01))
    dtg = float(dt)
    for k in range(3):
        st, _, _ = orc.step(dt)
        dt = np.float32(orc.get_cfl()[0])
        dtg, sst = m.RunSteps(1, dtg)
        print("step", k, "iterations", st.iterations, flush=True)
        assert sst["solve"]["iterations"] == st.iterations
        assert np.float32(dtg) == dt
    same(m.get_field("liquidPhi") <= 0, orc.get(F["liquidPhi"]) <= 0, "liquid mask after 3 steps")
    pg, vg, _ = m.GetParticleSystem().get()
    po, vo, _ = orc.particles()
    same(pg, po, "positions after 3 steps")
    same(vg, vo, "velocities after 3 steps")
    # the finest colour passes ran over the z-segment list (128 = two 64-cell segments per row):
    # every row holding liquid has one or two segments holding liquid
    (_, _, d2), rows = m.GetSolver().LevelInfo(0)
    cells, segs = m.GetSolver().LevelSegments(0)
    assert d2 == 128 and cells == 64 and rows <= segs <= 2 * rows, (rows, cells, segs)
    scene.close()


def test_reference_run_known_answer():
    """LiquidData.csv frame 1 on the GPU: every particle falls by (0, -g dt^2, 0)."""
    import os
    kat = np.load(os.path.join(os.path.dirname(__file__), "golden", "liquiddata_kat.npz"))
    scene = lh.Scene(config_file("pure_liquid_3d.json"))
    m = scene.GetLiquidModel()
    m.Init()
    p0 = m.GetParticleSystem().get()[0]
    np.testing.assert_allclose(p0, kat["frame0"], rtol=0, atol=1e-6)
    m.PreStep(0.01)
    m.Step(0.01)
    m.PostStep(0.01)
    p1 = m.GetParticleSystem().get()[0]
    np.testing.assert_allclose(p1 - p0, np.tile([0.0, -9.81e-4, 0.0], (len(p0), 1)), rtol=0, atol=1e-8)
    keep = np.arange(len(p0)) % 8 != 0
    np.testing.assert_allclose(p1[keep], kat["frame1"][keep], rtol=0, atol=1e-6)
    scene.close()


def test_extrapolation_in_place_validity_chain():
    """Q6: validity is updated in place in (x,y,z) order, values come from the old field.

    Only the x-face (1,1,1) is open and carries v=1; every other face of the x
    range is closed.  In one sweep the whole range becomes valid through the
    chain of predecessors, but only faces that see (1,1,1) get a nonzero value:
    (1,1,2) sees only (1,1,1) -> 1; (1,2,1) sees (1,1,1) and its predecessor
    (1,1,2), validated earlier in the same sweep with old value 0 -> 1/2.
    A Jacobi-style sweep would give 1 for both.
    """
    cfg = small_dam(8, "reference")
    scene = lh.Scene(cfg)
    m = scene.GetLiquidModel()
    m.Init()
    orc = Oracle(cfg)
    orc.init()
    # Extrapolate returns at once without particles (LiquidModel.cpp:939): give it one
    one = np.array([[0.5, 0.5, 0.5]], np.float32)
    m.GetParticleSystem().set(one)
    orc.set_particles(one)
    N = 8
    wx = np.zeros((N + 1, N, N), np.float32)
    wx[1, 1, 1] = 1.0
    vx = np.zeros_like(wx)
    vx[1, 1, 1] = 1.0
    for tgt in (m,):
        tgt.set_field("liquidWx", wx)
        tgt.set_field("vx", vx)
        tgt.set_field("isValid", np.zeros((N + 1,) * 3, np.uint8))
    orc.set(F["liquidWx"], wx)
    orc.set(F["vx"], vx)
    orc.set(F["isValid"], np.zeros((N + 1,) * 3, np.uint8))
    m.Extrapolate()
    orc.extrapolate()
    g = m.get_field("vx")
    assert g[1, 1, 2] == 1.0 and g[1, 2, 1] == 0.5
    same(g, orc.get(F["vx"]), "extrapolated vx")
    same(m.get_field("isValid"), orc.get(F["isValid"]), "validity")
    scene.close()


def test_snapshot_resume_bit_exact(tmp_path):
    """lh_save after 3 steps, 3 more steps; a fresh context that loads the snapshot and runs
    3 steps ends bit-identical (particles, isValid, time, next dt)"""
    print("case snapshot resume, dam24 CorrectVolume", flush=True)
    cfg = CASES["dam24_correct_volume"]()
    snap = str(tmp_path / "state.lhsnap")
    sc = lh.Scene(cfg)
    m = sc.GetLiquidModel()
    m.Init()
    dt, _ = m.RunSteps(3, float(cfg.get("firstDt", 0.01)))
    m.SaveState(snap)
    dt_a, _ = m.RunSteps(3, dt)
    pa, va, ca = m.GetParticleSystem().get()
    valid_a, time_a = m.get_field("isValid"), m.time
    sc.close()
    sc = lh.Scene(cfg)
    m = sc.GetLiquidModel()
    m.Init()
    m.LoadState(snap)
    assert m.GetCFL() == dt
    dt_b, _ = m.RunSteps(3, dt)
    pb, vb, cb = m.GetParticleSystem().get()
    same(pb, pa, "positions after resume")
    same(vb, va, "velocities after resume")
    same(cb, ca, "affine matrices after resume")
    same(m.get_field("isValid"), valid_a, "isValid after resume")
    assert m.time == time_a and dt_b == dt_a
    sc.close()


def test_moving_sphere_bit_exact():
    """SetBoundary + ComputeSolidPhi before every step (a moving obstacle): the sphere of the
    tank scene moves 0.02 along x per step; solidPhi, liquidW and the step stay bit-identical"""
    import copy

    print("case moving sphere, dam24 tank", flush=True)
    cfg = small_dam(24, "symmetric", tank=True)
    orc = Oracle(cfg, dot_mode=1, p2g_order=1)
    orc.init()
    scene = lh.Scene(cfg)
    m = scene.GetLiquidModel()
    m.Init()
    dt = np.float32(cfg.get("firstDt", 0.01))
    dtg = float(dt)
    for k in range(4):
        b = copy.deepcopy(cfg["Boundary"])
        b["right"]["center"] = [0.5 + 0.02 * (k + 1), 0.5, 0.5]
        m.SetBoundary(b)
        m.ComputeSolidPhi()
        orc.set_boundary(b)
        orc.compute_solid_phi()
        for f in ("solidPhi", "liquidWx", "liquidWy", "liquidWz"):
            same(m.get_field(f), orc.get(F[f]), "%s, step %d" % (f, k))
        st, _, _ = orc.step(dt)
        dt = np.float32(orc.get_cfl()[0])
        dtg, sst = m.RunSteps(1, dtg)
        assert sst["solve"]["iterations"] == st.iterations
        assert np.float32(dtg) == dt
        pg, vg, _ = m.GetParticleSystem().get()
        po, vo, _ = orc.particles()
        same(pg, po, "positions, step %d" % k)
        same(vg, vo, "velocities, step %d" % k)
    scene.close()


def test_solve_density_argument():
    """CGSolver<3>::Solve(liquidDensity, dt) (PressureSolver.h:104) with a density other than the
    scene's: bit-exact against the oracle's Solve(500, dt).  Every operator term W dt / (rho dx^2)
    doubles exactly when rho halves, so the PCG iterates scale by exact powers of two: the
    iteration count is the same and the pressure is exactly half of the rho = 1000 solve's."""
    print("case dam24 solve density", flush=True)
    cfg = small_dam(24, "symmetric")
    orc = Oracle(cfg, dot_mode=1, p2g_order=1)
    scene = lh.Scene(cfg)
    m = scene.GetLiquidModel()
    orc.init()
    m.Init()
    assert scene.GetLiquidDensity() == 1000.0
    dt = np.float32(0.01)
    for _ in range(3):   # a few steps so the pressure has more than a hydrostatic profile
        orc.step(dt)
        dt = np.float32(orc.get_cfl()[0])
    po, vo, co = orc.particles()
    m.GetParticleSystem().set(po, vo, co)
    m.set_field("isValid", orc.get(F["isValid"]))
    orc.clear_grid()
    orc.compute_liquid_phi()
    orc.p2g()
    orc.apply_gravity(dt)
    m.PreStep(dt)
    m.Particle2Grid()
    m.ApplyGravity(dt)
    st_o = orc.solve_pressure(dt, rho=500.0)
    st_g = m.GetSolver().Solve(500.0, dt)
    assert st_g["iterations"] == st_o.iterations and st_g["converged"]
    p500 = m.get_field("pressure")
    same(p500, orc.get(F["pressure"]), "pressure with rho = 500")
    st_1000 = m.GetSolver().Solve(None, dt)
    p1000 = m.get_field("pressure")
    assert st_1000["iterations"] == st_g["iterations"]
    assert np.abs(p1000).max() > 0
    same(p500, p1000 * np.float32(0.5), "rho = 500 pressure vs half the rho = 1000 pressure")
    with pytest.raises(lh.LHError):
        m.GetSolver().Solve(-1.0, dt)
    scene.close()
