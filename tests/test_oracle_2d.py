"""Pins the oracle's 2D model (LiquidModel<2>, CGSolver<2>) without the reference.

* spawn: ParticleSystem<2>::SpawnParticles' lattice (spacing dx / 2 from min + 0.3 (max - min)),
  truncated at numParticles with the break that leaves only the y loop;
* the box and circle SDFs at points with known distances;
* P2G: an FP64 brute force over the reference's taps (a particle reaches the faces c + d,
  d in [-1, 1]^2, around its cell c), the bin order against the particle order, and a uniform
  field reproduced exactly;
* the pressure solve: the level-0 operator restated here in numpy from liquidPhi and liquidW
  (BuildLHS, PressureSolver.cpp:96-175) and solved densely by numpy, against the oracle's PCG
  run to a tiny tolerance;
* CorrectVolume: an FP64 brute force that counts the hash search's visits (slots of the two
  3 x 3 regions, 2 P slots, two primes) independently.
"""
import json

import numpy as np

from conftest import config_file
from oracle import Oracle2D

BASE = config_file("dam_break_2d_64.json")


def cfg2(N=32, nparticles=20000, **solver):
    cfg = json.load(open(BASE))
    cfg["gridDimension"] = [N, N]
    cfg["LiquidParameter"]["numParticles"] = nparticles
    cfg["SolverParameter"].update(solver)
    return cfg


def test_spawn_lattice():
    for n in (10 ** 6, 1000):
        cfg = cfg2(32, n)
        o = Oracle2D(cfg)
        o.init()
        p, v, c = o.particles()
        dx = 2.0 / 32
        h = np.float32(0.5 * np.float32(dx))
        off = np.float32(0.3) * np.float32(2.0)  # Scalar(0.3) * (max - min), in float
        lo = np.float32(-1.0) + off
        hi = np.float32(1.0) - off
        nx = int((hi - lo) / h)
        ny = nx
        want = []
        for xi in range(nx):
            for yi in range(ny):
                if xi * ny + yi >= min(nx * ny, n):
                    break
                want.append((np.float32(xi) * h + lo, np.float32(yi) * h + lo))
        assert len(p) == min(nx * ny, n)
        np.testing.assert_array_equal(p, np.array(want, np.float32))
        assert not v.any() and not c.any()


def test_sdf_closed_forms():
    cfg = cfg2(16)
    cfg["Boundary"] = {"type": "Union",
                       "left": {"type": "Rect", "center": [0.0, 0.0], "halfExtent": [0.5, 0.25], "useInside": False},
                       "right": {"type": "Sphere", "center": [0.5, 0.5], "radius": 0.25, "useInside": False}}
    o = Oracle2D(cfg)
    o.compute_solid_phi()
    s = o.get(11)
    dx = 2.0 / 16
    x = -1.0 + dx * np.arange(17)
    X, Y = np.meshgrid(x, x, indexing="ij")
    q = np.stack([np.abs(X) - 0.5, np.abs(Y) - 0.25])
    box = np.sqrt((np.maximum(q, 0) ** 2).sum(0)) + np.minimum(q.max(0), 0)
    circ = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2) - 0.25
    np.testing.assert_allclose(s, np.minimum(box, circ), rtol=0, atol=2e-7)


def brute_p2g(pos, vel, Cm, N, mn, dx):
    """FP64 P2G over the reference's taps; returns (vx, vy) normalised"""
    def w(d):
        d = abs(d)
        return 0.75 - d * d if d < 0.5 else (0.5 * (1.5 - d) ** 2 if d < 1.5 else 0.0)

    out = []
    for a in range(2):
        shape = (N + 1, N) if a == 0 else (N, N + 1)
        ws, vs = np.zeros(shape), np.zeros(shape)
        for p, v, c in zip(pos.astype(np.float64), vel.astype(np.float64), Cm.astype(np.float64)):
            g = (p - mn) / dx
            ci = g.astype(int)
            for ddx in (-1, 0, 1):
                for ddy in (-1, 0, 1):
                    xi, yi = ci[0] + ddx, ci[1] + ddy
                    if xi < 0 or yi < 0 or xi >= shape[0] or yi >= shape[1]:
                        continue
                    cc = np.array([xi, yi], float) + (0.5 if a == 1 else 0.0) * np.array([1, 0]) \
                        + (0.5 if a == 0 else 0.0) * np.array([0, 1])
                    wt = w(g[0] - cc[0]) * w(g[1] - cc[1])
                    dpos = p - (cc * dx + mn)
                    ws[xi, yi] += wt
                    vs[xi, yi] += wt * (v[a] + c[2 * a] * dpos[0] + c[2 * a + 1] * dpos[1])
        out.append(np.where(np.abs(ws) < 1e-8, 0.0, vs / np.where(ws == 0, 1, ws)))
    return out


def test_p2g_brute_force_and_orders():
    N = 16
    cfg = cfg2(N)
    rng = np.random.default_rng(7)
    pos = (rng.random((40, 2)) * 1.2 - 0.6).astype(np.float32)
    vel = rng.normal(0, 1, (40, 2)).astype(np.float32)
    Cm = rng.normal(0, 1, (40, 4)).astype(np.float32)
    o = Oracle2D(cfg)
    o.compute_solid_phi()
    o.set_particles(pos, vel, Cm)
    res = {}
    for order in (0, 1):
        o.clear_grid()
        o.p2g(order)
        res[order] = (o.get(0), o.get(1))
    ref = brute_p2g(pos, vel, Cm, N, -1.0, 2.0 / N)
    for a in range(2):
        np.testing.assert_allclose(res[1][a], ref[a], rtol=0, atol=2e-5 * np.abs(ref[a]).max())
        np.testing.assert_allclose(res[0][a], res[1][a], rtol=2e-6, atol=1e-6)
    # a uniform field (a power of two, C = 0) is reproduced exactly where weights exist
    o.set_particles(pos, np.full_like(vel, 0.5), np.zeros_like(Cm))
    o.clear_grid()
    o.p2g(1)
    vx = o.get(0)
    assert set(np.unique(vx)) <= {0.0, 0.5} and (vx == 0.5).sum() > 50


def numpy_operator(phi, lwx, lwy, dx, dt, rho):
    """level-0 A (BuildLHS, PressureSolver.cpp:96-175) as a dense matrix over liquid cells"""
    N0, N1 = phi.shape
    cells = [(x, y) for x in range(N0) for y in range(N1) if not phi[x, y] > 0]
    idx = {c: k for k, c in enumerate(cells)}
    A = np.zeros((len(cells), len(cells)))

    def frac(l, r):
        if l < 0 and r < 0:
            return 1.0
        if l < 0 < r:
            return l / (l - r)
        if l > 0 > r:
            return r / (r - l)
        return 0.0

    for (x, y), k in idx.items():
        for nx, ny, W in ((x + 1, y, lwx[x + 1, y] if x + 1 < N0 else 0), (x - 1, y, lwx[x, y] if x >= 1 else 0),
                          (x, y + 1, lwy[x, y + 1] if y + 1 < N1 else 0), (x, y - 1, lwy[x, y] if y >= 1 else 0)):
            if not W > 0:
                continue
            t = W * dt / (rho * dx * dx)
            if phi[nx, ny] < 0:
                A[k, idx[(nx, ny)]] -= t
                A[k, k] += t
            else:
                A[k, k] += t / max(0.1, frac(phi[x, y], phi[nx, ny]))
    return A, cells


def test_solver_against_dense_solve():
    N = 24
    cfg = cfg2(N, tolerance=1e-13, maxIterations=400)
    o = Oracle2D(cfg)
    o.init()
    dt = np.float32(0.01)
    o.step(dt)  # a first step so the fluid moves
    o.clear_grid()
    o.compute_liquid_phi()
    o.p2g(1)
    o.apply_gravity(dt)
    st = o.solve_pressure(dt)
    phi, lwx, lwy, b = o.get(10), o.get(6), o.get(7), o.get(13)
    A, cells = numpy_operator(phi.astype(np.float64), lwx.astype(np.float64), lwy.astype(np.float64), 2.0 / N,
                              float(dt), 1000.0)
    rhs = np.array([b[c] for c in cells], np.float64)
    # unknowns with a zero row are not determined; solve on the rest
    keep = np.abs(np.diag(A)) > 0
    x = np.linalg.solve(A[np.ix_(keep, keep)], rhs[keep])
    p = o.get(9)
    got = np.array([p[c] for c in cells], np.float64)[keep]
    assert st.iterations > 0
    # interior cells: the reference's loop updates x on cells in [1, N-2]^2 only
    inner = np.array([1 <= c[0] <= N - 2 and 1 <= c[1] <= N - 2 for c in cells])[keep]
    err = np.linalg.norm(got[inner] - x[inner]) / np.linalg.norm(x[inner])
    assert err < 1e-4, err


def slots(cells, m):
    c = cells.astype(np.int64).astype(np.uint32)
    h = (np.uint32(73856093) * c[:, 0]) ^ (np.uint32(19349663) * c[:, 1])
    return h.view(np.int32).astype(np.int64) % m


def test_correct_volume_brute_force():
    N = 16
    cfg = cfg2(N)
    cfg["LiquidParameter"]["correctCycle"] = 1
    o = Oracle2D(cfg)
    o.compute_solid_phi()
    rng = np.random.default_rng(3)
    dx = 2.0 / N
    pos = (rng.random((200, 2)) - 0.5).astype(np.float32) * np.float32(6 * dx)
    o.set_particles(pos)
    dt = 0.004
    o.correct_volume(dt)
    q, _, _ = o.particles()
    P = len(pos)
    m = 2 * P
    p64 = pos.astype(np.float64)
    cell = np.floor((p64 + 1.0) / dx).astype(int)

    def region(c):
        lo, hi = np.clip(c - 1, 0, N - 1), np.clip(c + 1, 0, N - 1)
        return np.stack(np.meshgrid(np.arange(lo[0], hi[0] + 1), np.arange(lo[1], hi[1] + 1), indexing="ij"),
                        -1).reshape(-1, 2)

    def qw(d):
        return 0.75 - d * d if d < 0.5 else (0.5 * (1.5 - d) ** 2 if d < 1.5 else 0.0)

    ref = np.zeros_like(p64)
    for i in range(P):
        si = np.bincount(slots(region(cell[i]), m), minlength=m)
        rep = np.zeros(2)
        for j in range(P):
            if j == i:
                continue
            d = p64[i] - p64[j]
            dist = np.sqrt((d * d).sum())
            if dist >= dx:
                continue
            sj = np.bincount(slots(region(cell[j]), m), minlength=m)
            mult = int((si * sj).sum())
            rep += mult * np.sqrt(dx * dx) * qw(dist / dx) * dx * d / dist
        ref[i] = rep * dt
    got = q.astype(np.float64) - p64
    assert np.abs(ref).max() > 0
    np.testing.assert_allclose(got, ref, rtol=0, atol=2e-4 * np.abs(ref).max())
