"""The 128^3 dam break over 200 steps (block x in [0, 0.4], y in [0, 0.5], z in [0, 1], 8 particles
per cell, ~3.3 M particles; 3-level symmetric MGPCG to 1e-6 relative).

Every step's solve must converge.  At steps 0, 100 and 199 the oracle takes over the GPU's state
(particles, the persistent isValid field, time) and both run that step stage by stage: the liquid
mask and liquid phi, the transferred velocities, the right-hand side and the pressure (with the
PCG iteration count), the projected and extrapolated velocities and the particles after G2P and
advection must be bit-identical.  At those steps the projection must lower the divergence's
L-infinity norm over the liquid cells by at least 1e3 (FP32 solve to 1e-6 relative residual).
"""
import numpy as np
import pytest

from oracle import Oracle
from conftest import projection_floor
from test_gpu_parity import F, same, small_dam

lh = pytest.importorskip("liquidhairinteration_amd")
pytestmark = pytest.mark.gpu

STEPS = 200
CHECK = (0, 100, STEPS - 1)


def staged_step(orc, m, dt, k):
    orc.clear_grid()
    orc.compute_liquid_phi()
    m.PreStep(dt)
    phi = m.get_field("liquidPhi")
    same(phi <= 0, orc.get(F["liquidPhi"]) <= 0, "liquid mask, step %d" % k)
    same(phi, orc.get(F["liquidPhi"]), "liquidPhi, step %d" % k)
    orc.p2g()
    orc.apply_gravity(dt)
    m.Particle2Grid()
    m.ApplyGravity(dt)
    for f in ("vx", "vy", "vz"):
        same(m.get_field(f), orc.get(F[f]), "P2G + gravity %s, step %d" % (f, k))
    before = m.ComputeDivergence()[1]
    st_o = orc.solve_pressure(dt)
    st_g = m.GetSolver().Solve(None, dt)
    assert st_g["converged"] == 1 and st_g["iterations"] == st_o.iterations, k
    same(m.get_field("rhs"), orc.get(F["rhs"]), "rhs, step %d" % k)
    same(m.get_field("pressure"), orc.get(F["pressure"]), "pressure, step %d" % k)
    orc.velocity_update(dt)
    m.VelocityUpdate(dt)
    after = m.ComputeDivergence()[1]
    for f in ("vx", "vy", "vz"):
        same(m.get_field(f), orc.get(F[f]), "velocity update %s, step %d" % (f, k))
    orc.extrapolate()
    m.Extrapolate()
    for f in ("vx", "vy", "vz", "isValid"):
        same(m.get_field(f), orc.get(F[f]), "extrapolate %s, step %d" % (f, k))
    orc.g2p()
    m.Grid2Particle()
    orc.advect(dt)
    m.AdvectParticle(dt)
    pg, vg, cg = m.GetParticleSystem().get()
    po, vo, co = orc.particles()
    same(vg, vo, "G2P velocity, step %d" % k)
    same(cg, co, "G2P affine, step %d" % k)
    same(pg, po, "advected positions, step %d" % k)
    floor = projection_floor(m, st_g, float(dt))
    print("step %d: %d PCG iterations, divergence L-inf %.3e -> %.3e (residual + FP32 floor %.3e)"
          % (k, st_o.iterations, before, after, floor), flush=True)
    assert after <= 1e-3 * before and after <= floor, k


def test_dam_break_128_200_steps():
    print("case dam128 symmetric, %d steps" % STEPS, flush=True)
    cfg = small_dam(128, "symmetric")
    orc = Oracle(cfg, dot_mode=1, p2g_order=1)
    orc.init()
    scene = lh.Scene(cfg)
    m = scene.GetLiquidModel()
    m.Init()
    dt = float(cfg.get("firstDt", 0.01))
    iters = []
    for k in range(STEPS):
        if k in CHECK:
            pos, vel, Cm = m.GetParticleSystem().get()
            orc.set_particles(pos, vel, Cm)
            orc.set(F["isValid"], m.get_field("isValid"))
            t = m.time
            orc.set_time(t)
            staged_step(orc, m, np.float32(dt), k)
            m.time = t + dt
            dt = m.GetCFL()
        else:
            dt, st = m.RunSteps(1, dt)
            assert st["solve"]["converged"] == 1, k
            iters.append(st["solve"]["iterations"])
    print("free-running steps: PCG iterations min %d, mean %.1f, max %d" % (min(iters), np.mean(iters), max(iters)),
          flush=True)
    pos = m.GetParticleSystem().get()[0]
    assert np.isfinite(pos).all()
    scene.close()
