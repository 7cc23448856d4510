"""Training data, training loop and evaluation of the learned initial guess.

The network (k_nn.hip) is trained on the simulator's own systems: every liquid step
of a randomised scene builds the pressure system A p = b and solves it with plain
MGPCG; the pair (b, cell types) -> converged p is one training sample, and one Adam
step is taken on it on the GPU (lh_warmstart_train) before the scene moves on.

Scenes: the unit cube with solid walls 3 cells in, as the dam-break configs; either a
"dam" (one block standing on the floor) or a "collision" (two blocks), block edges
U(0.2, 0.6), positions uniform inside the walls, block velocities U(-2, 2) m/s per
axis plus N(0, 0.05) per particle, 8 jittered particles per cell.  Particles are set
through lh_set_particles on one scene context, so the network and its Adam state stay
resident on the GPU across scenes.

    python -m liquidhairinteration_amd.warmstart train --grid 64 --pairs 20000 --out w.lhws
    python -m liquidhairinteration_amd.warmstart eval --grid 128 --weights w.lhws
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import LHError, Scene, config_path

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warmstart.lhws")


def scene_config(n, solver=None):
    """unit cube at n^3 with walls 3 cells in; 3-level symmetric MGPCG to 1e-6 relative"""
    cfg = json.load(open(config_path("dam_break_48")))
    cfg["gridDimension"] = [n, n, n]
    cfg["Boundary"]["child"]["halfExtent"] = [0.5 - 3.0 / n] * 3
    cfg["SolverParameter"].update({"smoother": "symmetric", "smoothingFactor": 1.0, "numLevel": 3,
                                   "numPreSmooth": 2, "numPostSmooth": 2, "numFinalSmooth": 20,
                                   "maxIterations": 200, "tolerance": 1e-6, "toleranceMode": "relative"})
    if solver:
        cfg["SolverParameter"].update(solver)
    cfg["FluidBlocks"] = [{"min": [0.3, 0.3, 0.3], "max": [0.5, 0.5, 0.5], "velocity": [0, 0, 0]}]
    return cfg


def random_blocks(rng, kind):
    """[(min, max, velocity)] of a random dam-break or collision scene"""
    out = []
    for b in range(1 if kind == "dam" else 2):
        size = rng.uniform(0.2, 0.6, 3)
        lo = rng.uniform(0.0, 1.0 - size)
        if kind == "dam":
            lo[1] = 0.0  # standing on the floor
        vel = rng.uniform(-2.0, 2.0, 3)
        out.append((lo, lo + size, vel))
    return out


def spawn(n, blocks, rng, ppc_axis=2, noise=0.05):
    """jittered particles (ppc_axis^3 per cell) in the blocks, clipped to the open region
    [3h, 1 - 3h]^3 (one cell of margin from the walls); overlapping blocks keep the first"""
    h = 1.0 / n
    lo_open, hi_open = 3.0 * h, 1.0 - 3.0 * h
    taken = np.zeros((n, n, n), bool)
    P, V = [], []
    k = ppc_axis
    sub = (np.stack(np.meshgrid(*[np.arange(k)] * 3, indexing="ij"), -1).reshape(-1, 3) + 0.0)
    for lo, hi, vel in blocks:
        a = np.maximum(np.ceil((np.maximum(lo, lo_open)) / h).astype(int), 0)
        b = np.minimum(np.floor((np.minimum(hi, hi_open)) / h).astype(int), n)
        if (b <= a).any():
            continue
        cells = np.stack(np.meshgrid(*[np.arange(a[i], b[i]) for i in range(3)], indexing="ij"), -1).reshape(-1, 3)
        cells = cells[~taken[cells[:, 0], cells[:, 1], cells[:, 2]]]
        taken[cells[:, 0], cells[:, 1], cells[:, 2]] = True
        pos = (cells[:, None, :] + (sub[None] + rng.random((len(cells), len(sub), 3))) / k) * h
        pos = pos.reshape(-1, 3)
        P.append(pos)
        V.append(vel[None, :] + rng.normal(0.0, noise, pos.shape))
    if not P:
        return np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32)
    return np.concatenate(P).astype(np.float32), np.concatenate(V).astype(np.float32)


class Simulator:
    """one scene context reused for many random scenes"""

    def __init__(self, n, solver=None):
        self.n = n
        self.scene = Scene(scene_config(n, solver))
        self.model = self.scene.GetLiquidModel()
        self.model.Init()
        self.ws = self.model.GetWarmStart()

    def reset(self, rng, kind):
        pos, vel = spawn(self.n, random_blocks(rng, kind), rng)
        if len(pos) == 0:
            return False
        self.model.GetParticleSystem().set(pos, vel, None)
        self.model.time = 0.0
        self.dt = 0.01
        return True

    def step(self):
        m = self.model
        m.PreStep(self.dt)
        m.Step(self.dt)
        st = m.PostStep(self.dt)
        self.dt = m.GetCFL()
        return st


def lr_at(i, total, lr0, lr1):
    """cosine decay from lr0 to lr1"""
    return lr1 + 0.5 * (lr0 - lr1) * (1.0 + math.cos(math.pi * min(1.0, i / max(1, total))))


def train(n=64, pairs=20000, steps_per_scene=40, lr0=1e-3, lr1=2e-5, seed=0, out=WEIGHTS, init=None,
          log=print, time_limit=None, loss="energy"):
    """train on `pairs` simulator systems at n^3; returns the loss history"""
    rng = np.random.default_rng(seed)
    sim = Simulator(n)
    ws = sim.ws
    if init:
        ws.load(init)
    else:
        ws.init(seed)
    ws.set_loss(loss)
    hist = []
    i = 0
    t0 = time.time()
    while i < pairs:
        kind = "dam" if rng.random() < 0.5 else "collision"
        if not sim.reset(rng, kind):
            continue
        for _ in range(steps_per_scene):
            if i >= pairs:
                break
            try:
                sim.step()
            except LHError as e:  # a scene the solver cannot finish is skipped
                log("scene skipped: %s" % e)
                break
            loss = ws.train(lr_at(i, pairs, lr0, lr1))
            hist.append(loss)
            i += 1
            if i % 500 == 0:
                log("pair %d  loss %.4g (mean of last 500 %.4g)  %.1f s" % (i, loss, np.mean(hist[-500:]),
                                                                          time.time() - t0))
                sys.stdout.flush()
            if time_limit and time.time() - t0 > time_limit:
                log("time limit reached at pair %d" % i)
                pairs = i
                break
    if out:
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        ws.save(out)
    return hist


def forward_flops(dims):
    """multiply-add FLOPs of one network forward pass on a grid of dims (2 per MAC)"""
    V0 = float(np.prod(dims))
    layers = [(2, 16, 0), (16, 32, 1), (32, 64, 2), (64, 96, 3), (160, 64, 2), (96, 32, 1), (48, 16, 0), (16, 1, 0)]
    return sum(2.0 * ci * co * 27 * V0 / 8 ** lv for ci, co, lv in layers)


def evaluate(weights=WEIGHTS, n=128, scenes=4, steps=30, seed=12345, log=print, solver=None):
    """plain vs learned MGPCG on identical built systems of held-out scenes.

    Per step: BuildRHS/BuildLHS once, then the plain solve (x0 = 0) and the learned solve
    (network inference + energy scaling + r0 inside the timed region), both to the same
    relative tolerance; the step then continues from the learned solution.  Also the
    energy error of the learned guess relative to the zero guess's (1 - (y.b)^2/((y.Ay)(x*.b)))
    and the network's forward time (HIP events) with its MFMA throughput.
    Returns (summary dict, per-solve arrays)."""
    rng = np.random.default_rng(seed)
    sim = Simulator(n, solver)
    ws = sim.ws
    ws.load(weights)
    ws.set_loss("energy")
    m = sim.model
    sv = m.GetSolver()
    rec = {"plain_iters": [], "learned_iters": [], "plain_ms": [], "learned_ms": [], "used": [], "energy": []}
    fwd_ms, d0_ms, fwd_n = 0.0, 0.0, 0
    done = 0
    while done < scenes:
        kind = "dam" if done % 2 == 0 else "collision"
        if not sim.reset(rng, kind):
            continue
        dt = 0.01
        for s in range(steps):
            m.PreStep(dt)
            m.Particle2Grid()
            m.ApplyGravity(dt)
            sv.BuildSystem(dt)
            ws.enable(False)
            sp = sv.SolveRHS()
            ws.enable(True)
            m.profile_enable(True)
            sl = sv.SolveRHS()
            k, t = m.profile_read("nn_fwd")
            _, t0 = m.profile_read("conv_d0")
            m.profile_enable(False)
            ws.enable(False)
            if s >= 1:  # the first step of a scene starts from rest
                rec["plain_iters"].append(sp["iterations"])
                rec["learned_iters"].append(sl["iterations"])
                rec["plain_ms"].append(sp["solve_ms"])
                rec["learned_ms"].append(sl["solve_ms"])
                rec["used"].append(sl["warm_start_used"])
                rec["energy"].append(ws.train(0.0))
                fwd_ms += t
                d0_ms += t0
                fwd_n += k
            m.VelocityUpdate(dt)
            m.Extrapolate()
            m.Grid2Particle()
            m.AdvectParticle(dt)
            m.PostStep(dt)
            dt = m.GetCFL()
        done += 1
    r = {k: np.asarray(v, np.float64) for k, v in rec.items()}
    fwd = fwd_ms / max(fwd_n, 1)
    d0 = d0_ms / max(fwd_n, 1)
    summary = {
        "grid": n, "solves": int(r["plain_iters"].size),
        "plain_iterations": float(r["plain_iters"].mean()), "learned_iterations": float(r["learned_iters"].mean()),
        "plain_ms": float(np.median(r["plain_ms"])), "learned_ms": float(np.median(r["learned_ms"])),
        "warm_start_used_fraction": float(r["used"].mean()),
        "energy_error_ratio_median": float(np.median(r["energy"])),
        "inference_ms": fwd,
        "inference_tflops": forward_flops((n, n, n)) / (fwd * 1e-3) / 1e12 if fwd > 0 else None,
        "conv_d0_ms": d0,
        "conv_d0_tflops": 2.0 * 48 * 16 * 27 * n ** 3 / (d0 * 1e-3) / 1e12 if d0 > 0 else None,
    }
    summary["iteration_ratio"] = summary["plain_iterations"] / max(summary["learned_iterations"], 1e-9)
    summary["time_ratio"] = summary["plain_ms"] / max(summary["learned_ms"], 1e-9)
    log(json.dumps(summary))
    return summary, r


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--grid", type=int, default=64)
    t.add_argument("--pairs", type=int, default=20000)
    t.add_argument("--steps-per-scene", type=int, default=40)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--lr-final", type=float, default=2e-5)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--init", default=None, help="continue from these weights")
    t.add_argument("--out", default=WEIGHTS)
    t.add_argument("--time-limit", type=float, default=None, help="seconds")
    t.add_argument("--loss", default="energy", choices=["energy", "mse"])
    e = sub.add_parser("eval")
    e.add_argument("--grid", type=int, default=128)
    e.add_argument("--weights", default=WEIGHTS)
    e.add_argument("--scenes", type=int, default=4)
    e.add_argument("--steps", type=int, default=30)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        train(a.grid, a.pairs, a.steps_per_scene, a.lr, a.lr_final, a.seed, a.out, a.init,
              time_limit=a.time_limit, loss=a.loss)
    else:
        evaluate(a.weights, a.grid, a.scenes, a.steps)


if __name__ == "__main__":
    main()
