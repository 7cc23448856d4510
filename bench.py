"""Liquid-step benchmark on one MI355X.

Workload (default): two-block collision on a 256^3 MAC grid (configs/collision_256.json):
two 0.3^3 blocks at opposite corners of the unit cube, 8 jittered particles per
cell, v = +-(2,0,0) + N(0, 0.05^2), solid walls 3 cells thick, MGPCG (3 levels,
2/2 red-black Gauss-Seidel sweeps, symmetric V-cycle, the coarsest level solved by 24 FP64
Chebyshev steps) to 1e-6 relative residual inside an FP64 residual loop to 1e-9.
One step = PreStep + Step + PostStep (its ComputeDivergence, mean and L-inf) + GetCFL
(main.cpp:50-53) run by the library's native loop; particles and grid stay resident in HBM for
the whole timed region.

Prints one JSON line: steps/s over K timed steps with nothing else in the timed region; per
kernel, the average launch time from HIP events around every launch over the next K steps (the
events stay out of the timed region: each pair costs microseconds of queue time) and, for the
dominant one by total time, its roofline; the CPU baseline (the oracle, timed on this host for two
steps from each of three states of the GPU run: after the warm-up, after the timed steps and after
the profiled steps) and, under "warm_start", ms and iterations per solve to 1e-6
relative residual of the learned initial guess versus plain MGPCG on identical systems of
held-out 128^3 scenes.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak, GB/s


def level_dims(cfg):
    dims = [list(cfg["gridDimension"])]
    for _ in range(1, cfg["SolverParameter"]["numLevel"]):
        dims.append([(d + 1) // 2 for d in dims[-1]])
    return dims


def coarse_form(cfg, l):
    """how coarse level l >= 1 is swept: "rows" (one k_sweep_rows launch per sweep, on levels of
    <= rowSweepMaxCells with coarseSweep "rows"), "tiles" (one k_tile_sweep launch per sweep,
    coarseSweep "tiles") or "passes" (two k_smooth colour passes)"""
    sp = cfg["SolverParameter"]
    mode = sp.get("coarseSweep", "rows")
    if mode in ("passes", "tiles"):
        return mode
    d = level_dims(cfg)[l]
    return "rows" if d[0] * d[1] * d[2] <= sp.get("rowSweepMaxCells", 1 << 18) else "passes"


def coarsest_fp64(cfg):
    """the coarsest level's sweeps in FP64 (SolverParameter.coarsestPrecision; default with "precision":
    "refined")"""
    s = cfg["SolverParameter"]
    return s.get("coarsestPrecision", "fp64" if cfg.get("precision") == "refined" else "fp32") == "fp64"


def coarsest_chebyshev(cfg):
    """Chebyshev steps of the coarsest level's solve (SolverParameter.coarsestSolver "chebyshev",
    coarsestSteps; 0: the reference's numFinalSmooth sweeps)"""
    s = cfg["SolverParameter"]
    return int(s.get("coarsestSteps", 12)) if s.get("coarsestSolver", "sweeps") == "chebyshev" else 0


def is_coarsest(cfg, l):
    return l == cfg["SolverParameter"]["numLevel"] - 1 and l > 0


def sweep_bytes(cfg, l):
    """algorithmic bytes per cell of one red-black sweep on level l: adiag + 3 face terms 16 B, r 4 B,
    z read + written 8 B (16 B when the coarsest level's z is FP64); for the coarsest level's
    Chebyshev launch (two steps): z and d read and written in FP64 32 B, r 4 B, adiag + 3 face
    terms 16 B = 52 B"""
    if is_coarsest(cfg, l) and coarsest_chebyshev(cfg):
        return 52.0
    return 36.0 if is_coarsest(cfg, l) and coarsest_fp64(cfg) else 28.0


def sweeps_per_launch(cfg, l):
    """red-black sweeps per smoother launch on coarse level l: one for k_sweep_rows and k_tile_sweep,
    half of one for a k_smooth colour pass (1 for the coarsest level's Chebyshev launches, whose
    bytes sweep_bytes counts per launch)"""
    if is_coarsest(cfg, l) and coarsest_chebyshev(cfg):
        return 1.0
    return 0.5 if coarse_form(cfg, l) == "passes" else 1.0


def algorithmic_bytes(kernel, n0, P, active, fused, per_launch=None, sweep_b=None):
    """Bytes per launch counted once per array element the kernel must touch.

    smooth_l0: one colour pass = half of a red-black sweep, whose minimum is 28 B/cell
               (Adiag + 3 face terms = 16 B, r 4 B, z read 4 B + write 4 B) -> 14 B per
               cell of the level-0 z-segments holding liquid (active["seg0"]: the pass
               visits only those), or of the level-0 rows holding liquid when the level keeps
               no segment list; the fused one-launch sweep, when selected, is dense: 28 B/cell;
    smooth_l1, smooth_l2: the same per cell of the rows holding liquid on level 1, and on
               level 2 (the coarsest of the three-level hierarchy): 28 B per sweep, times the
               sweeps of one launch (per_launch[l]: 1/2 for a colour pass, 1 for k_sweep_rows or
               k_tile_sweep), 36 B per sweep when the coarsest level's z is FP64, 52 B per
               launch of the coarsest level's two Chebyshev steps (sweep_b[l]); with more than
               three levels smooth_l2 times every level from 2 down, its bytes count level 2's;
    spmv:      the p update fused with q = A p (k_cg_p_spmv, all but the first SpMV of a solve):
               z and p read, new p and q written, 4 coefficient floats -> 32 B per cell it
               visits: of the level-0 z-segments holding liquid, or of the rows holding liquid
               without a segment list (the plain first SpMV moves 24 B);
    p2g:       the three components in one launch: 60 B/particle (position, velocity,
               affine matrix) + 4 B per face written (~3 n0 faces);
    g2p:       72 B/particle (read position, write position, velocity, C) + 3
               velocity components read around the particles (~12 B/cell);
    liquid_phi: 12 B/particle + 4 B/cell.
    active[l] = cells of the level-l rows that hold liquid; active["seg0"] = cells of the
    level-0 z-segments that hold liquid (0 without a segment list).
    """
    per_launch = per_launch or {}
    sweep_b = sweep_b or {}
    return {
        "smooth_l0": 28.0 * n0 if fused else 14.0 * (active.get("seg0") or active[0]),
        "smooth_l1": sweep_b.get(1, 28.0) * per_launch.get(1, 0.5) * active.get(1, 0),
        "smooth_l2": sweep_b.get(2, 28.0) * per_launch.get(2, 0.5) * active.get(2, 0),
        "spmv": 32.0 * (active.get("seg0") or active[0]),
        "p2g": 60.0 * P + 12.0 * n0,
        "g2p": 72.0 * P + 12.0 * n0,
        "liquid_phi": 12.0 * P + 4.0 * n0,
    }[kernel]


# instrumented kernel -> its symbol in the library (rocprofv3 kernel names); the coarse levels'
# symbols depend on the configuration (level_symbol)
KERNEL_SYMBOL = {"smooth_l0": "k_smooth<true>", "spmv": "k_cg_p_spmv", "p2g": "k_p2g", "g2p": "k_g2p_advect",
                 "liquid_phi": "k_liquid_phi"}


def kernel_symbols(cfg):
    """KERNEL_SYMBOL for this configuration: the finest colour passes run over the level's z-segment
    list (k_smooth_seg) when its z extent is a multiple of 64 cells (at most 31 segments); P2G is
    k_p2g_shared unless the config selects the plain gather ("p2gForm": "gather")"""
    sym = dict(KERNEL_SYMBOL)
    if cfg.get("p2gForm", "shared") == "shared":
        sym["p2g"] = "k_p2g_shared"
    d2 = cfg["gridDimension"][2]
    if d2 % 64 == 0 and d2 // 64 <= 31:
        sym["smooth_l0"] = "k_smooth_seg"
    return sym


def smooth_grid(d):
    """work-items of one k_smooth launch on a level of dims d (the library's smooth_blocks)"""
    H = (d[2] + 1) // 2
    R = 1 if H >= 256 else 256 // H
    return -(-max(1, min(-(-d[0] * d[1] // R), 8192)) // 8) * 8 * 256


def row_sweep_grid(d):
    """work-items of one k_sweep_rows / k_cheb_rows launch on a level of dims d (a block of five waves
    per 64-cell row segment, capped at 65536 blocks)"""
    nseg = -(-d[2] // 64)
    return max(1, min(d[0] * d[1] * nseg, 65536)) * 320


def tile_sweep_grid(d):
    """work-items of one k_tile_sweep launch on a level of dims d (4 x 4 x 64 tiles, capped)"""
    return max(1, min(-(-d[0] // 4) * -(-d[1] // 4) * -(-d[2] // 64), 8192)) * 256


def level_symbol(cfg, l):
    """(kernel symbol, grid size) of the smoother launches on coarse level l"""
    d = level_dims(cfg)[l]
    mode = 1 if cfg["SolverParameter"].get("smoother", "reference") == "symmetric" else 0
    form = coarse_form(cfg, l)
    if is_coarsest(cfg, l) and coarsest_chebyshev(cfg):
        return "k_cheb_rows<double>", row_sweep_grid(d)
    if form == "rows":
        if is_coarsest(cfg, l) and coarsest_fp64(cfg):
            return "k_sweep_rows<%d, double>" % mode, row_sweep_grid(d)
        return "k_sweep_rows<%d>" % mode, row_sweep_grid(d)
    if form == "tiles":
        return "k_tile_sweep<%d>" % mode, tile_sweep_grid(d)
    return "k_smooth<false>", smooth_grid(d)


def coarsest_label(cfg):
    """the coarsest level's solve: its sweep count, or its Chebyshev steps and interval"""
    s = cfg["SolverParameter"]
    k = coarsest_chebyshev(cfg)
    prec = "FP64" if coarsest_fp64(cfg) else "FP32"
    if k:
        return "chebyshev %d on [%g, 2] %s" % (k, float(s.get("chebyshevLower", 0.01)), prec)
    return "%d sweeps %s" % (int(s["numFinalSmooth"]), prec)


COPY_BYTES = 256 * 1024 * 1024  # the calibration copy: bytes read and bytes written per launch


def pmc_calibration_copy():
    """three launches of a copy of COPY_BYTES (read once, written once) for scaling the
    FETCH_SIZE / WRITE_SIZE counters; torch's elementwise multiply by 1 is the copy kernel"""
    import torch

    a = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    del a, b


def pmc_reduce(fetch_csv, write_csv, workload, cfg):
    """rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE counter_collection.csv files of one
    `bench.py --pmc-calibrate` run each -> profiles/pmc_traffic.json[workload]: per kernel,
    KiB per dispatch of each counter, the scale that maps the copy kernel's reading to its
    known bytes, and the scaled HBM bytes per dispatch"""
    import csv

    symbols = kernel_symbols(cfg)
    level_grid = {}
    for l in (1, 2):
        if l < cfg["SolverParameter"]["numLevel"]:
            symbols["smooth_l%d" % l], level_grid["smooth_l%d" % l] = level_symbol(cfg, l)

    def per_kernel(path, counter):
        sums, counts, copy = {}, {}, []
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            name, v = r["Kernel_Name"], float(r["Counter_Value"])
            if "MulFunctor" in name and int(r["Grid_Size"]) * 16 == COPY_BYTES:  # 4 floats per work-item
                copy.append(v)
            for k, sym in symbols.items():
                if k in level_grid and int(r["Grid_Size"]) != level_grid[k]:
                    continue
                # the symbol, alone or with template arguments (k_p2g<true>)
                if "::" + sym + "(" in name or ("<" not in sym and "::" + sym + "<" in name):
                    sums[k] = sums.get(k, 0.0) + v
                    counts[k] = counts.get(k, 0) + 1
        if not copy:
            raise SystemExit("%s: no calibration copy dispatch (run bench.py with --pmc-calibrate)" % path)
        scale = COPY_BYTES / 1024.0 / (sum(copy) / len(copy))
        return {k: (sums[k] / counts[k], counts[k]) for k in sums}, scale

    fetch, fs = per_kernel(fetch_csv, "FETCH_SIZE")
    write, ws = per_kernel(write_csv, "WRITE_SIZE")
    out = {}
    for k in sorted(set(fetch) & set(write)):
        f, n = fetch[k]
        w, _ = write[k]
        out[k] = {"kernel": symbols[k], "dispatches": n, "fetch_kib": round(f, 1), "write_kib": round(w, 1),
                  "fetch_scale": round(fs, 4), "write_scale": round(ws, 4),
                  "bytes_per_dispatch": int(round((f * fs + w * ws) * 1024))}
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    allw = json.load(open(path)) if os.path.exists(path) else {}
    allw[workload] = out
    with open(path, "w") as fh:
        json.dump(allw, fh, indent=1)
    print(json.dumps({workload: out}, indent=1))


def measured_traffic(kernel, workload, symbol):
    """HBM bytes per dispatch of `kernel` (library symbol `symbol`) on `workload` from the committed
    rocprofv3 PMC runs (FETCH_SIZE and WRITE_SIZE, each scaled by what the counter reads for a copy
    kernel of known size), or None when that pair was not measured for this symbol"""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path)).get(workload, {}).get(kernel)
    return None if rec is None or rec.get("kernel") != symbol else rec["bytes_per_dispatch"]


def host_cpu():
    """the host CPU's model name (lscpu) and logical CPU count"""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count()}


def snapshot(model):
    """the state the next step depends on: particles, the persistent isValid field, time"""
    pos, vel, Cm = model.GetParticleSystem().get()
    return {"pos": pos, "vel": vel, "C": Cm, "isValid": model.get_field("isValid"), "time": model.time}


def cpu_baseline(cfg, snapshots, steps=2, threads=0):
    """The oracle (FP32 restatement of the reference's step, OpenMP) on this host: `steps` steps
    from each snapshot of the GPU run (early, middle and late states of the benchmark window, each
    with the dt the GPU run had there); value = steps / total time.  Without snapshots (--cpu-only)
    the steps start from the workload's initial state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    from oracle import Oracle

    threads = threads or os.cpu_count() or 1
    orc = Oracle(cfg, dot_mode=1, p2g_order=1, threads=threads)
    per, total, n = [], 0.0, 0
    src = "states of the GPU run"
    if not snapshots:
        orc.init()
        snapshots = [("initial", None, cfg.get("firstDt", 0.01))]
        src = "the workload's"
    else:
        orc.compute_solid_phi()
    for name, snap, dt in snapshots:
        if snap is not None:
            orc.set_particles(snap["pos"], snap["vel"], snap["C"])
            orc.set(12, snap["isValid"])
            orc.set_time(snap["time"])
        dt = np.float32(dt)
        t0 = time.perf_counter()
        for _ in range(steps):
            orc.step(dt)
            dt = np.float32(orc.get_cfl()[0])
        el = time.perf_counter() - t0
        per.append({"state": name, "sim_time": round(snap["time"], 5) if snap is not None else 0.0,
                    "steps": steps, "s_per_step": round(el / steps, 3)})
        total += el
        n += steps
    cores = int(orc.L.orc_max_threads())
    return {"value": n / total, "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": "%d steps of the same workload, %d from each of %d %s (%s), FP32 oracle "
                      "restatement, OpenMP over %d host thread(s) (%.1f s)"
                      % (n, steps, len(snapshots), src, ", ".join(p["state"] for p in per), cores, total),
            "per_state": per, "host": host_cpu()}


MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA peak


def warm_start_eval(grid, scenes, steps):
    """ms per solve to 1e-6 relative residual, learned vs plain MGPCG, on held-out scenes"""
    from liquidhairinteration_amd import warmstart as W

    if not os.path.exists(W.WEIGHTS):
        return None
    s, _ = W.evaluate(W.WEIGHTS, n=grid, scenes=scenes, steps=steps, seed=12345, log=lambda *a: None)
    out = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in s.items()}
    out["solver"] = "MGPCG L=3 pre/post=2/2 symmetric, tol 1e-06 relative, unit-cube walls"
    out["scenes"] = "%d held-out dam/collision scenes x %d steps (first step of each excluded)" % (scenes, steps)
    if s.get("conv_d0_tflops"):
        out["mfma"] = {"kernel": "conv 48->16 at full resolution (the largest layer)", "bound": "mfma",
                       "achieved": round(s["conv_d0_tflops"], 1), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": round(s["conv_d0_tflops"] / MFMA_BF16_PEAK_TFLOPS, 4)}
    return out


def bench_2d(args, cfg):
    """the 2D model (LiquidModel<2>): steps/s of the library's native loop, and the oracle's 2D
    restatement (single-threaded) on a bounded sample of the same workload"""
    import numpy as np

    import liquidhairinteration_amd as lh

    sc = lh.Scene2D(cfg)
    m = sc.GetLiquidModel()
    m.Init()
    P = m.GetParticleSystem().NumParticles()
    dt = float(cfg.get("firstDt", 0.01))
    if args.warmup > 0:
        dt, _ = m.RunSteps(args.warmup, dt)
    t0 = time.perf_counter()
    dt, last = m.RunSteps(args.steps, dt)
    el = time.perf_counter() - t0
    sc.close()
    sp = cfg["SolverParameter"]
    out = {
        "metric": "liquid steps/s, 2D model (LiquidModel<2>)", "value": round(args.steps / el, 3),
        "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(el * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (the reference's lattice spawn)",
        "config": {"workload": os.path.basename(args.config).replace(".json", ""), "grid": cfg["gridDimension"],
                   "particles": P, "solver": "MGPCG L=%d pre/post=%d/%d coarsest=%d %s, tol %g %s" % (
                       sp["numLevel"], sp["numPreSmooth"], sp["numPostSmooth"], sp["numFinalSmooth"],
                       sp.get("smoother", "reference"), sp["tolerance"], sp.get("toleranceMode", "absolute"))},
        "last_step": {"pcg_iterations": last["iterations"], "converged": last["converged"]},
        "roofline": None, "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle2D

        orc = Oracle2D(cfg, dot_mode=1, p2g_order=1)
        orc.init()
        d = np.float32(cfg.get("firstDt", 0.01))
        for _ in range(args.warmup):  # the same warm-up steps as the GPU run, untimed
            orc.step(d)
            d = np.float32(orc.get_cfl()[0])
        n, t0 = 0, time.perf_counter()
        while n < args.cpu_steps or (time.perf_counter() - t0 < 10.0 and n < 10 * args.cpu_steps):
            orc.step(d)
            d = np.float32(orc.get_cfl()[0])
            n += 1
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(n / el, 4), "unit": "steps/s", "cores": 1, "kind": "port",
                               "sample": "%d step(s) after the warm-up steps, FP32 oracle restatement of the 2D "
                                         "model, one thread (%.1f s)" % (n, el)}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default=os.path.join(ROOT, "liquidhairinteration_amd", "configs", "collision_256.json"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2, help="oracle steps from each sampled state")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads for cpu_baseline (0: every core)")
    ap.add_argument("--cpu-only", action="store_true",
                    help="time only the CPU baseline of the workload (no GPU) and print it as one JSON line")
    ap.add_argument("--no-warm-start", action="store_true", help="skip the learned-vs-plain solve comparison")
    ap.add_argument("--learned", action="store_true",
                    help="start every pressure solve of the timed workload from the learned guess")
    ap.add_argument("--coarse-sweep", choices=("config", "rows", "tiles", "passes"), default="config",
                    help="override SolverParameter.coarseSweep (how the coarse levels are swept)")
    ap.add_argument("--pcg-graph", choices=("config", "on", "off"), default="config",
                    help="SolverParameter.pcgGraph: each PCG iteration launched as one hipGraph")
    ap.add_argument("--smoother", choices=("config", "reference"), default="config",
                    help="reference: the reference's Q4/Q5 smoother with its own solver parameters "
                         "(Config/PureLiquid3D.json: 4 levels, 5/5/20 sweeps, omega 0.5) in place of the config's")
    ap.add_argument("--ws-grid", type=int, default=128)
    ap.add_argument("--ws-scenes", type=int, default=2)
    ap.add_argument("--ws-steps", type=int, default=8)
    ap.add_argument("--pmc-calibrate", action="store_true",
                    help="launch a copy kernel of known size first (scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--pmc-reduce", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="reduce two rocprofv3 --pmc runs of this workload into profiles/pmc_traffic.json and exit")
    args = ap.parse_args()
    workload = os.path.basename(args.config).replace(".json", "")
    if args.pmc_reduce:
        pmc_reduce(args.pmc_reduce[0], args.pmc_reduce[1], workload, json.load(open(args.config)))
        return
    if args.cpu_only:
        cfg = json.load(open(args.config))
        print(json.dumps({"workload": workload, "cpu_baseline": cpu_baseline(cfg, None, args.cpu_steps,
                                                                             args.cpu_threads)}))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (see DESIGN.md, sharding).",
              file=sys.stderr)
        sys.exit(2)

    import liquidhairinteration_amd as lh

    if args.pmc_calibrate:
        pmc_calibration_copy()
    cfg = json.load(open(args.config))
    if len(cfg["gridDimension"]) == 2:
        bench_2d(args, cfg)
        return
    if args.smoother == "reference":
        # the reference's V-cycle throughout: its smoother, and numFinalSmooth sweeps on the coarsest level
        cfg["SolverParameter"].update({"smoother": "reference", "smoothingFactor": 0.5, "numLevel": 4,
                                       "numPreSmooth": 5, "numPostSmooth": 5, "numFinalSmooth": 20,
                                       "maxIterations": 400, "coarsestSolver": "sweeps"})
        workload += "_reference_smoother"
    if args.coarse_sweep != "config":
        cfg["SolverParameter"]["coarseSweep"] = args.coarse_sweep
    if args.pcg_graph != "config":
        cfg["SolverParameter"]["pcgGraph"] = args.pcg_graph == "on"
    scene = lh.Scene(cfg)
    model = scene.GetLiquidModel()
    model.Init()
    if args.learned:
        from liquidhairinteration_amd import warmstart as W

        model.GetWarmStart().load(W.WEIGHTS)
        model.GetWarmStart().enable(True)
    P = model.GetParticleSystem().NumParticles()
    n0 = 1
    for d in scene.GetGridDimension():
        n0 *= d
    dt = float(cfg.get("firstDt", 0.01))
    if args.warmup > 0:
        dt, _ = model.RunSteps(args.warmup, dt)
    want_cpu = not args.no_cpu_baseline
    snaps = [("early", snapshot(model), dt)] if want_cpu else []
    # the timed region: K steps with nothing but the step in it
    t0 = time.perf_counter()
    dt, last = model.RunSteps(args.steps, dt)
    el = time.perf_counter() - t0
    if want_cpu:
        snaps.append(("middle", snapshot(model), dt))
    # kernel timings: the next K steps with HIP events around every launch of the listed kernels
    # (each event pair costs several microseconds of queue time, so they stay out of the timed region)
    kern = ("smooth_l0", "smooth_l1", "smooth_l2", "p2g", "g2p", "liquid_phi", "spmv")
    model.profile_enable(True)
    t1 = time.perf_counter()
    dt, _ = model.RunSteps(args.steps, dt)
    el_prof = time.perf_counter() - t1
    prof = {k: model.profile_read(k) for k in kern}
    model.profile_enable(False)
    if want_cpu:
        snaps.append(("late", snapshot(model), dt))
    prof = {k: v for k, v in prof.items() if v[0]}
    # cells of the rows holding liquid per level in the last step's operator (the smoother and
    # SpMV visit those)
    active, rows = {}, {}
    for lvl in range(cfg["SolverParameter"]["numLevel"]):
        (_, _, d2), nr = model.GetSolver().LevelInfo(lvl)
        active[lvl], rows[lvl] = nr * d2, nr
    seg_cells, nseg0 = model.GetSolver().LevelSegments(0)
    active["seg0"] = seg_cells * nseg0
    fused = cfg["SolverParameter"].get("fusedSweepMinCells", 1 << 62) <= n0
    per_launch = {l: sweeps_per_launch(cfg, l) for l in range(1, cfg["SolverParameter"]["numLevel"])}
    sweep_b = {l: sweep_bytes(cfg, l) for l in range(1, cfg["SolverParameter"]["numLevel"])}

    def abytes(k):
        return algorithmic_bytes(k, n0, P, active, fused, per_launch, sweep_b)

    def rate(k):
        n, ms = prof[k]
        return abytes(k) / (ms / 1e3 / n) / 1e9 if n else None

    # dominant instrumented kernel by total device time
    dom = max(prof, key=lambda k: prof[k][1])
    launches, tot_ms = prof[dom]
    avg_s = tot_ms / 1e3 / max(launches, 1)
    ach = rate(dom)
    symbol = level_symbol(cfg, int(dom[-1]))[0] if dom in ("smooth_l1", "smooth_l2") else kernel_symbols(cfg)[dom]
    roof = {"kernel": dom, "symbol": symbol, "bound": "hbm", "achieved": round(ach, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(dom, workload, symbol), "launches": launches,
            "avg_launch_us": round(avg_s * 1e6, 2),
            "algorithmic_bytes_per_launch": abytes(dom)}
    kernels = {k: {"launches": v[0], "ms": round(v[1], 3), "pct_of_pass": round(100.0 * v[1] / (el_prof * 1e3), 1),
                   "avg_us": round(1e3 * v[1] / v[0], 2), "achieved_gbs": round(rate(k), 1)}
               for k, v in prof.items()}
    kernels["_timing"] = ("HIP events on the library's stream around every launch of these kernels, over the %d "
                          "steps that follow the timed region (%.2f ms/step with the events)"
                          % (args.steps, el_prof * 1e3 / args.steps))

    out = {
        "metric": "liquid steps/s at 256^3; ms per solve to 1e-6 residual, learned vs plain MGPCG",
        "value": round(args.steps / el, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered particle blocks)",
        "config": {"workload": workload,
                   "grid": list(scene.GetGridDimension()), "particles": P, "active_rows_per_level": rows,
                   "active_segments_level0": nseg0,
                   "solver": "MGPCG L=%d pre/post=%d/%d coarsest=%s %s, tol %g %s" % (
                       cfg["SolverParameter"]["numLevel"], cfg["SolverParameter"]["numPreSmooth"],
                       cfg["SolverParameter"]["numPostSmooth"], coarsest_label(cfg),
                       cfg["SolverParameter"].get("smoother", "reference"),
                       cfg["SolverParameter"]["tolerance"], cfg["SolverParameter"].get("toleranceMode", "absolute")),
                   "precision": cfg.get("precision", "fp32") + (
                       " (FP32 MGPCG inside an FP64 residual loop to %g)" % cfg.get("refinedTolerance", 1e-9)
                       if cfg.get("precision") == "refined" else ""),
                   "initial_guess": "learned (network, energy-scaled, zero-guess fallback)" if args.learned else "zero"},
        "roofline": roof,
        "kernels": kernels,
        "last_step": {"pcg_iterations": last["solve"]["iterations"], "converged": last["solve"]["converged"],
                      "warm_start_used": last["solve"]["warm_start_used"],
                      "solve_ms": round(last["solve"]["solve_ms"], 3), "step_ms": round(last["step_ms"], 3),
                      "divergence_linf": last["divergence_linf"], "sim_time": last["sim_time"]},
        "cpu_baseline": None,
    }
    scene.close()
    if not args.no_warm_start:
        out["warm_start"] = warm_start_eval(args.ws_grid, args.ws_scenes, args.ws_steps)
    if want_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, snaps, args.cpu_steps, args.cpu_threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
