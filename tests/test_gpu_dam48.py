"""The CPU-runnable dam break (48^3, block x in [0, 0.4], y in [0, 0.5], z in [0, 1], 8 particles per
cell, the reference's V-cycle: 4 levels, 5/5 sweeps, 20 on the coarsest level, omega 0.5, 1e-6
relative) for its full 100 steps, free running on the GPU and in the oracle side by side
(main.cpp:47-53: PreStep, Step, PostStep, GetCFL): every step's PCG iteration count, next dt and
divergence (count and L-inf bit-exact, mean to 1e-12) agree, every solve converges, and after 100
steps the particles and the liquid mask are bit-identical."""
import json

import numpy as np
import pytest

from conftest import config_file
from oracle import Oracle

lh = pytest.importorskip("liquidhairinteration_amd")
pytestmark = pytest.mark.gpu

STEPS = 100


def test_dam_break_48_hundred_steps_bit_exact():
    print("case dam48 100 steps", flush=True)
    cfg = json.load(open(config_file("dam_break_48.json")))
    orc = Oracle(cfg, dot_mode=1, p2g_order=1)
    orc.init()
    scene = lh.Scene(cfg)
    m = scene.GetLiquidModel()
    m.Init()
    po, vo, co = orc.particles()
    pg, vg, cg = m.GetParticleSystem().get()
    assert np.array_equal(pg, po) and np.array_equal(vg, vo)
    dt = np.float32(cfg.get("firstDt", 0.01))
    dtg = float(dt)
    iters = []
    for step in range(STEPS):
        st, _, _ = orc.step(dt)
        dt = np.float32(orc.get_cfl()[0])
        dtg, sst = m.RunSteps(1, dtg)
        assert sst["solve"]["converged"] == 1 and st.converged == 1, "step %d" % step
        assert sst["solve"]["iterations"] == st.iterations, "step %d: %d vs %d PCG iterations" % (
            step, sst["solve"]["iterations"], st.iterations)
        assert np.float32(dtg) == dt, "step %d: dt %r vs %r" % (step, dtg, dt)
        mo, lo, no = orc.divergence()
        assert sst["fluid_cells"] == no and sst["divergence_linf"] == lo, "step %d divergence" % step
        assert abs(sst["divergence_mean"] - mo) <= 1e-12 * max(lo, 1e-30), "step %d divergence mean" % step
        iters.append(st.iterations)
        if step % 10 == 9:
            print("step %d: t = %.4f, %d PCG iterations" % (step + 1, orc.time, st.iterations), flush=True)
    pg, vg, cg = m.GetParticleSystem().get()
    po, vo, co = orc.particles()
    assert np.array_equal(pg, po), "positions after %d steps" % STEPS
    assert np.array_equal(vg, vo), "velocities after %d steps" % STEPS
    assert np.array_equal(cg, co), "affine matrices after %d steps" % STEPS
    assert np.array_equal(m.get_field("liquidPhi") <= 0, orc.get(10) <= 0), "liquid mask"
    print("100 steps bit-identical; PCG iterations per step mean %.1f, max %d" % (np.mean(iters), max(iters)))
    scene.close()
