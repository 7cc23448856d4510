"""The oracle's pressure operator and PCG against an independent sparse direct solve.

A is assembled from the oracle's level-0 Adiag / Acoef (BuildLHS,
PressureSolver.cpp:544-664) over the cells the PCG updates (interior liquid
cells); scipy's spsolve in FP64 gives the exact solution of that system.
"""
import json

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import config_file
from oracle import Oracle


def advanced(N=16, smoother="symmetric", steps=3, coarsest="fp32", **solver):
    """the system of the 4th step of a 16^3 dam break (the first 3 stepped with the FP32 coarsest
    level, so every coarsest precision and coarsest solver builds the same system)"""
    cfg = json.load(open(config_file("dam_break_48.json")))
    cfg["gridDimension"] = [N, N, N]
    cfg["Boundary"]["child"]["halfExtent"] = [0.5 - 3.0 / N] * 3
    cfg["SolverParameter"].update({"smoother": smoother, "numLevel": 2, "maxIterations": 400})
    run = Oracle(cfg, dot_mode=1, p2g_order=1)
    run.init()
    dt = np.float32(0.01)
    for _ in range(steps):
        run.step(dt)
        dt = np.float32(run.get_cfl()[0])
    cfg["SolverParameter"]["coarsestPrecision"] = coarsest
    cfg["SolverParameter"].update(solver)
    orc = Oracle(cfg, dot_mode=1, p2g_order=1)
    orc.compute_solid_phi()
    orc.set_particles(*run.particles())
    orc.set(12, run.get(12))
    orc.set_time(run.time)
    orc.clear_grid()
    orc.compute_liquid_phi()
    orc.p2g()
    orc.apply_gravity(dt)
    orc.build_system(dt)
    return orc


def assemble(orc):
    ad, co = orc.level(0)
    N = ad.shape
    phi = orc.get(10)
    idx = np.arange(ad.size).reshape(N)
    inter = np.zeros(N, bool)
    inter[1:-1, 1:-1, 1:-1] = True
    # liquid cells the PCG updates; rows with a zero diagonal (cells whose faces are all
    # closed, inside the walls) have b = 0 and stay at x = 0, so they are left out
    rows = np.flatnonzero((phi.reshape(-1) <= 0) & inter.reshape(-1) & (ad.reshape(-1) != 0))
    pos = -np.ones(ad.size, np.int64)
    pos[rows] = np.arange(len(rows))
    I, J, V = [], [], []
    offs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for r in rows:
        x, y, z = np.unravel_index(r, N)
        I.append(pos[r]); J.append(pos[r]); V.append(float(ad[x, y, z]))
        for k, (dx, dy, dz) in enumerate(offs):
            c = float(co[x, y, z, k])
            if c == 0.0:
                continue
            nb = idx[x + dx, y + dy, z + dz]
            assert pos[nb] >= 0, "coupling to a cell outside the solved set"
            I.append(pos[r]); J.append(pos[nb]); V.append(c)
    A = sp.csr_matrix((V, (I, J)), shape=(len(rows), len(rows)))
    return A, rows


def test_operator_is_symmetric_and_diagonally_dominant():
    orc = advanced()
    A, rows = assemble(orc)
    assert len(rows) > 100
    assert abs(A - A.T).max() == 0.0              # symmetric (no cell has phi == 0 here)
    d = A.diagonal()
    off = np.asarray(abs(A).sum(axis=1)).ravel() - np.abs(d)
    assert np.all(d >= off * (1 - 1e-6))          # weakly diagonally dominant -> SPD with a Dirichlet row
    assert np.any(d > off * (1 + 1e-3))           # free-surface (ghost fluid) rows are strictly dominant


CHEB = {"coarsestSolver": "chebyshev", "coarsestSteps": 7, "chebyshevLower": 0.02}


@pytest.mark.parametrize("smoother,coarsest,solver", [("symmetric", "fp32", {}), ("reference", "fp32", {}),
                                                      ("symmetric", "fp64", {}), ("reference", "fp64", {}),
                                                      ("symmetric", "fp64", CHEB)])
def test_pcg_matches_sparse_direct_solve(smoother, coarsest, solver):
    """with the coarsest level's sweeps in FP32 or FP64 (coarsestPrecision), or its Chebyshev
    solve, the FP32 MGPCG and the FP64 PCG both converge to the direct solution"""
    orc = advanced(smoother=smoother, coarsest=coarsest, **solver)
    A, rows = assemble(orc)
    b = orc.get(13).reshape(-1)
    x_direct = spla.spsolve(A.tocsc(), b[rows].astype(np.float64))
    x64, st64 = orc.solve_fp64(b.astype(np.float64), rel_tol=1e-13, max_iter=1000)
    assert st64.converged
    e64 = np.linalg.norm(x64.reshape(-1)[rows] - x_direct) / np.linalg.norm(x_direct)
    assert e64 < 1e-9
    x32, st32 = orc.solve_fp32(b)
    assert st32.converged
    e32 = np.linalg.norm(x32.reshape(-1)[rows].astype(np.float64) - x_direct) / np.linalg.norm(x_direct)
    assert e32 < 1e-4
    # x stays zero outside the interior liquid cells (PressureSolver.cpp:875-877, 891-893)
    phi = orc.get(10)
    upd = np.zeros(phi.shape, bool)
    upd[1:-1, 1:-1, 1:-1] = True
    upd &= phi <= 0
    assert np.all(x32.reshape(phi.shape)[~upd] == 0)


@pytest.mark.parametrize("smoother", ["symmetric", "reference"])
def test_fp64_coarsest_changes_only_rounding(smoother):
    """the coarsest level in FP64 is the same preconditioner up to FP32 rounding: the FP32 MGPCG's
    iteration count moves by at most one and the solutions agree to the solve's tolerance"""
    o32 = advanced(smoother=smoother, coarsest="fp32")
    o64 = advanced(smoother=smoother, coarsest="fp64")
    b = o32.get(13)
    assert np.array_equal(b, o64.get(13))
    x32, s32 = o32.solve_fp32(b)
    x64, s64 = o64.solve_fp32(b)
    assert s32.converged and s64.converged
    assert abs(s32.iterations - s64.iterations) <= 1
    assert not np.array_equal(x32, x64)           # the FP64 sweeps did run
    np.testing.assert_allclose(x64, x32, rtol=0, atol=1e-4 * np.abs(x32).max())


def coarsest_system(orc):
    """the coarsest level's operator over its cells with a non-negligible diagonal (the cells its
    solve updates), from that level's Adiag / Acoef"""
    L = orc.num_levels
    ad, co = orc.level(L - 1)
    N = ad.shape
    flat = ad.reshape(-1).astype(np.float64)
    rows = np.flatnonzero(np.abs(flat) >= 1e-8)
    pos = -np.ones(flat.size, np.int64)
    pos[rows] = np.arange(len(rows))
    I, J, V = list(range(len(rows))), list(range(len(rows))), list(flat[rows])
    offs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for r in rows:
        x, y, z = np.unravel_index(r, N)
        for k, (dx, dy, dz) in enumerate(offs):
            c = float(co[x, y, z, k])
            if c == 0.0:
                continue
            nb = np.ravel_multi_index((x + dx, y + dy, z + dz), N)
            assert pos[nb] >= 0, "coupling to a cell the coarsest solve leaves alone"
            I.append(pos[r]); J.append(pos[nb]); V.append(c)
    A = sp.csr_matrix((V, (I, J)), shape=(len(rows), len(rows))).toarray()
    return A, rows, flat.size


@pytest.mark.parametrize("K,lo", [(1, 0.02), (2, 0.02), (7, 0.02), (12, 0.01), (16, 0.005)])
def test_chebyshev_coarsest_closed_form(K, lo):
    """K Chebyshev steps from z = 0 on the Jacobi-scaled coarsest system (an option of this build;
    the reference's coarsest level is numFinalSmooth red-black sweeps, PressureSolver.cpp:754-757)
    equal the closed form z_K = D^-1/2 q_K(S) D^-1/2 r, S = D^-1/2 A D^-1/2, with
    q_K(l) = (1 - T_K((theta - l) / delta) / T_K(theta / delta)) / l on the interval [lo, 2]
    (Saad, Iterative Methods for Sparse Linear Systems, 12.3): the three-term recurrence's
    coefficients, signs and indices, and the cells it leaves at zero"""
    orc = advanced(N=24, coarsest="fp64", coarsestSolver="chebyshev", coarsestSteps=K, chebyshevLower=lo)
    A, rows, n = coarsest_system(orc)
    assert len(rows) > 50
    d = np.diag(A)
    S = A / np.sqrt(np.outer(d, d))
    lam, Q = np.linalg.eigh(S)
    assert lam.min() > 0 and lam.max() <= 2.0 + 1e-12          # the interval's upper end bounds the spectrum
    lo = float(np.float32(lo))                                   # the parameter is a float
    theta, delta = 0.5 * (2.0 + lo), 0.5 * (2.0 - lo)
    TK = np.polynomial.chebyshev.Chebyshev([0] * K + [1])
    q = (1.0 - TK((theta - lam) / delta) / TK(theta / delta)) / lam
    assert np.all(q > 0)                                         # positive definite preconditioner
    rng = np.random.default_rng(K)
    r = np.zeros(n)
    r[rows] = rng.standard_normal(len(rows))
    z = orc.coarsest_solve64(r)
    dm = 1.0 / np.sqrt(d)
    want = dm * (Q @ (q * (Q.T @ (dm * r[rows]))))
    np.testing.assert_allclose(z[rows], want, rtol=0, atol=1e-11 * np.abs(want).max())
    mask = np.ones(n, bool)
    mask[rows] = False
    assert np.all(z[mask] == 0)
    # linear in r, symmetric as an operator: u . M r == r . M u
    u = np.zeros(n)
    u[rows] = rng.standard_normal(len(rows))
    zu = orc.coarsest_solve64(u)
    assert abs(u @ z - r @ zu) <= 1e-11 * (abs(u @ z) + abs(r @ zu))
