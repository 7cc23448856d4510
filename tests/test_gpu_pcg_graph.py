"""SolverParameter.pcgGraph: each PCG iteration's x/r update, preconditioner and r.z replayed from
an instantiated hipGraph.  Every replay is checked against the same scene run with one launch at a
time (itself bit-exact against the oracle in test_gpu_parity.py): per step the PCG iteration count
and the next dt, and after the run the pressure, the velocities and the particles, bit for bit.
The graphs are captured once per (p buffer, rho slot, level ping-pong state) and replayed across
iterations, solves and steps, with the library's own launches and host waits between replays."""
import json

import numpy as np
import pytest

from conftest import config_file

lh = pytest.importorskip("liquidhairinteration_amd")
pytestmark = pytest.mark.gpu


def dam(N, smoother="symmetric", **solver):
    cfg = json.load(open(config_file("dam_break_48.json")))
    cfg["gridDimension"] = [N, N, N]
    cfg["Boundary"]["child"]["halfExtent"] = [0.5 - 3.0 / N] * 3
    cfg["SolverParameter"]["smoother"] = smoother
    if smoother == "symmetric":
        cfg["SolverParameter"].update({"smoothingFactor": 1.0, "numLevel": 3, "numPreSmooth": 2,
                                       "numPostSmooth": 2})
    cfg["SolverParameter"].update(solver)
    return cfg


CASES = {
    # the reference's V-cycle: colour passes on level 0, row sweeps (z / z2 ping-pong) below
    "reference_smoother": lambda: dam(24, "reference"),
    # symmetric V-cycle, FP64 Chebyshev coarsest level with 5 steps: 3 launches per application,
    # so the z64 / d64 ping-pong state alternates between applications (two graphs per p buffer)
    "chebyshev_odd_launches": lambda: dam(32, coarsestPrecision="fp64", coarsestSolver="chebyshev",
                                          coarsestSteps=5, chebyshevLower=0.01),
    # the refined (FP64 outer loop) solve of the 256^3 workloads, at 32^3
    "refined_chebyshev": lambda: dict(dam(32, coarsestSolver="chebyshev", coarsestSteps=24,
                                          chebyshevLower=0.005), precision="refined"),
    "jacobi": lambda: dam(24, "reference", preconditionerType="Jacobi", maxIterations=400),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_pcg_graph_matches_direct_launches(case):
    print("case", case, flush=True)
    cfg = CASES[case]()
    gcfg = json.loads(json.dumps(cfg))
    gcfg["SolverParameter"]["pcgGraph"] = True
    scenes, models = [], []
    for c in (cfg, gcfg):
        s = lh.Scene(c)
        m = s.GetLiquidModel()
        m.Init()
        scenes.append(s)
        models.append(m)
    dt = [float(cfg.get("firstDt", 0.01))] * 2
    for step in range(5):
        out = [m.RunSteps(1, d) for m, d in zip(models, dt)]
        (d0, s0), (d1, s1) = out
        assert s1["solve"]["iterations"] == s0["solve"]["iterations"], (step, s0["solve"], s1["solve"])
        assert s0["solve"]["converged"] == 1 and s1["solve"]["converged"] == 1
        assert np.float32(d1) == np.float32(d0), step
        dt = [d0, d1]
    for f in ("pressure", "vx", "vy", "vz"):
        a, b = models[0].get_field(f), models[1].get_field(f)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f
    for a, b in zip(models[0].GetParticleSystem().get(), models[1].GetParticleSystem().get()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    for s in scenes:
        s.close()
