"""The 2D model (LiquidModel<2> / CGSolver<2>) on the GPU against the oracle's restatement:
every stage bit for bit with teacher forcing, and free-running steps (CorrectVolume included,
as in the reference's 2D step)."""
import copy
import json

import numpy as np
import pytest

from conftest import config_file
from oracle import Oracle2D

lh = pytest.importorskip("liquidhairinteration_amd")
pytestmark = pytest.mark.gpu


def same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    bad = int((a != b).sum())
    assert bad == 0, "%s: %d of %d elements differ (max |d| %.3e)" % (
        what, bad, a.size, float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max()))


def dam2d(N=64, smoother="reference", obstacle=False, precond="Multigrid"):
    cfg = json.load(open(config_file("dam_break_2d_64.json")))
    cfg["gridDimension"] = [N, N]
    cfg["SolverParameter"]["smoother"] = smoother
    cfg["SolverParameter"]["preconditionerType"] = precond
    if smoother == "symmetric":
        cfg["SolverParameter"].update({"smoothingFactor": 1.0, "numPreSmooth": 2, "numPostSmooth": 2,
                                       "toleranceMode": "relative"})
    if obstacle:
        cfg["Boundary"] = {"type": "Union",
                           "left": {"type": "None", "child": {"type": "Rect", "center": [0.0, 0.0],
                                                              "halfExtent": [0.9, 0.9], "useInside": True}},
                           "right": {"type": "Sphere", "center": [0.1, -0.6], "radius": 0.15}}
    return cfg


CASES = {
    "dam64_reference": lambda: dam2d(64),
    "dam64_symmetric_circle": lambda: dam2d(64, "symmetric", obstacle=True),
    "dam48_jacobi": lambda: dam2d(48, precond="Jacobi"),
}
F = {"vx": 0, "vy": 1, "tempVx": 3, "tempVy": 4, "liquidWx": 6, "liquidWy": 7, "pressure": 9, "liquidPhi": 10,
     "solidPhi": 11, "isValid": 12, "rhs": 13}


@pytest.mark.parametrize("case", list(CASES))
def test_stages_bit_exact_2d(case):
    print("case", case, flush=True)
    cfg = CASES[case]()
    orc = Oracle2D(cfg, dot_mode=1, p2g_order=1)
    orc.init()
    sc = lh.Scene2D(cfg)
    m = sc.GetLiquidModel()
    m.Init()
    ps = m.GetParticleSystem()
    pg, vg, cg = ps.get()
    po, vo, co = orc.particles()
    same(pg, po, "initial positions")
    for f in ("solidPhi", "liquidWx", "liquidWy"):
        same(m.get_field(f), orc.get(F[f]), f)
    dt = np.float32(0.01)
    for step in range(4):
        po, vo, co = orc.particles()
        ps.set(po, vo, co.reshape(-1, 4))
        m.set_field("isValid", orc.get(F["isValid"]))
        orc.clear_grid()
        orc.compute_liquid_phi()
        m.PreStep(dt)
        same(m.get_field("liquidPhi"), orc.get(F["liquidPhi"]), "liquidPhi step %d" % step)
        orc.p2g(1)
        orc.apply_gravity(dt)
        m.Particle2Grid()
        m.ApplyGravity(dt)
        for f in ("vx", "vy"):
            same(m.get_field(f), orc.get(F[f]), "P2G+gravity %s step %d" % (f, step))
        st_o = orc.solve_pressure(dt)
        st_g = m.Solve(dt)
        assert st_g["iterations"] == st_o.iterations, (st_g["iterations"], st_o.iterations)
        same(m.get_field("rhs"), orc.get(F["rhs"]), "rhs step %d" % step)
        same(m.get_field("pressure"), orc.get(F["pressure"]), "pressure step %d" % step)
        orc.velocity_update(dt)
        m.VelocityUpdate(dt)
        for f in ("vx", "vy"):
            same(m.get_field(f), orc.get(F[f]), "velocity update %s step %d" % (f, step))
        orc.extrapolate()
        m.Extrapolate()
        for f in ("vx", "vy", "tempVx", "tempVy", "isValid"):
            same(m.get_field(f), orc.get(F[f]), "extrapolate %s step %d" % (f, step))
        orc.g2p()
        m.Grid2Particle()
        pg, vg, cg = ps.get()
        po, vo, co = orc.particles()
        same(vg, vo, "G2P velocity step %d" % step)
        same(cg, co.reshape(-1, 4), "G2P affine step %d" % step)
        orc.advect(dt)
        m.AdvectParticle(dt)
        same(ps.get()[0], orc.particles()[0], "advected positions step %d" % step)
        orc.correct_volume(dt)
        m.CorrectVolume(dt)
        same(ps.get()[0], orc.particles()[0], "CorrectVolume positions step %d" % step)
        dto = np.float32(orc.get_cfl()[0])
        assert np.float32(m.GetCFL()) == dto
        dt = dto
    sc.close()


@pytest.mark.parametrize("case", ["dam64_reference", "dam64_symmetric_circle"])
def test_free_running_2d(case):
    print("case", case, flush=True)
    cfg = CASES[case]()
    orc = Oracle2D(cfg, dot_mode=1, p2g_order=1)
    orc.init()
    sc = lh.Scene2D(cfg)
    m = sc.GetLiquidModel()
    m.Init()
    dt = np.float32(0.01)
    dtg = float(dt)
    for k in range(12):
        st = orc.step(dt)
        dt = np.float32(orc.get_cfl()[0])
        dtg, sg = m.RunSteps(1, dtg)
        assert sg["iterations"] == st.iterations, k
        assert np.float32(dtg) == dt, k
    same(m.GetParticleSystem().get()[0], orc.particles()[0], "positions after 12 steps")
    same(m.get_field("isValid"), orc.get(F["isValid"]), "isValid after 12 steps")
    sc.close()
