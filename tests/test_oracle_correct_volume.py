"""Pins the oracle's CorrectVolume (LiquidModel.cpp:1085-1132) without the reference.

The reference's hash search (NeighborSearch.h:22-57, 131-136) stores every particle in the
slot of its cell, slot = (73856093 x ^ 19349663 y ^ 83492791 z) mod 2 P (32-bit wrapping
products, non-negative remainder), and queries every cell of the clamped 3x3x3 region around
the query particle.  A neighbour is therefore visited once per pair (query cell, neighbour's
region cell) that land in one slot: shared cells plus hash collisions.  These tests state
that rule and the repulsion law independently (numpy, FP64, all pairs) and compare:

* two particles in one cell: visits = sum over slots of (cells of the region in it)^2,
  which is 27 when the table is large enough that the region's cells do not collide;
* a random cloud: every selected particle's displacement against an all-pairs brute force,
  which pins the neighbour set (distance < radius, cells within 2) and the multiplicity;
* the id cycle: only particles with id % correctCycle == step move, and step advances.
"""
import json
import os

import numpy as np

from conftest import config_file
from oracle import Oracle


def cfg_open(N=16, cycle=1):
    """unit box at N^3, solid walls well away from the particles used here"""
    cfg = json.load(open(config_file("dam_break_48.json")))
    cfg["gridDimension"] = [N, N, N]
    cfg["Boundary"]["child"]["halfExtent"] = [0.5 - 1.0 / N] * 3
    cfg["LiquidParameter"]["correctCycle"] = cycle
    cfg["FluidBlocks"] = [{"min": [0.45, 0.45, 0.45], "max": [0.55, 0.55, 0.55], "velocity": [0, 0, 0]}]
    return cfg


def qw(d):
    d = abs(d)
    return np.where(d < 0.5, 0.75 - d * d, np.where(d < 1.5, 0.5 * (1.5 - d) ** 2, 0.0))


def slots(cells, m):
    """hash slot of integer cells (k, 3): wrapping 32-bit products, XOR, non-negative mod m"""
    c = cells.astype(np.int64).astype(np.uint32)
    h = (np.uint32(73856093) * c[:, 0]) ^ (np.uint32(19349663) * c[:, 1]) ^ (np.uint32(83492791) * c[:, 2])
    return h.view(np.int32).astype(np.int64) % m  # numpy's % of a negative is already >= 0


def region(cell, N):
    lo = np.clip(cell - 1, 0, N - 1)
    hi = np.clip(cell + 1, 0, N - 1)
    ax = [np.arange(lo[a], hi[a] + 1) for a in range(3)]
    return np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)


def visits(ci, cj, N, m):
    """times the search from a particle in cell ci visits one in cell cj"""
    si = np.bincount(slots(region(ci, N), m), minlength=m)
    sj = np.bincount(slots(region(cj, N), m), minlength=m)
    return int((si * sj).sum())


def brute(pos, N, dx, dt, select):
    """FP64 displacement of the selected particles: sum over neighbours closer than the
    radius r = 2 dx of visits * sqrt(dx^3) * w(dist / r) * r * dir / dist, times dt"""
    r = 2.0 * dx
    m = 2 * len(pos)
    cell = np.floor(pos / dx).astype(int)
    out = np.zeros_like(pos)
    for i in np.nonzero(select)[0]:
        d = pos[i] - pos
        dist = np.sqrt((d * d).sum(1))
        ok = (dist < r) & (np.arange(len(pos)) != i)
        mult = np.zeros(len(pos))
        for j in np.nonzero(ok)[0]:
            mult[j] = visits(cell[i], cell[j], N, m)
        w = np.sqrt(dx ** 3) * qw(dist / r)
        rep = ((mult * w * r / np.where(ok, dist, 1.0))[:, None] * d)[ok].sum(0)
        out[i] = rep * dt
    return out


def two_in_one_cell(extra):
    """two particles 0.4 dx apart along x in one cell, plus `extra` particles far outside the
    radius (they only enlarge the hash table to 2 (2 + extra) slots)"""
    N = 16
    dx = 1.0 / N
    orc = Oracle(cfg_open(N))
    orc.init()
    p = [[0.5 + 0.2 * dx, 0.5 + 0.3 * dx, 0.5 + 0.4 * dx],
         [0.5 + 0.6 * dx, 0.5 + 0.3 * dx, 0.5 + 0.4 * dx]]
    p = np.array(p + [[0.1 + 0.01 * (k % 10), 0.1 + 0.01 * (k // 10 % 10), 0.1 + 0.01 * (k // 100)]
                      for k in range(extra)], np.float32)
    orc.set_particles(p)
    dt = 0.01
    orc.correct_volume(dt)
    q, _, _ = orc.particles()
    cell = np.floor(p[0].astype(np.float64) / dx).astype(int)
    hist = np.bincount(slots(region(cell, N), 2 * len(p)))
    dist = 0.4 * dx
    r = 2 * dx
    w = np.sqrt(dx ** 3) * (0.75 - (dist / r) ** 2)
    return q[:2] - p[:2], w * r * dt, int((hist * hist).sum())


def test_two_particles_closed_form():
    # 4 slots: the 27 region cells share them, each particle is visited sum n_s^2 times
    disp, unit, mult = two_in_one_cell(0)
    assert mult > 27
    np.testing.assert_allclose(disp[0], [-mult * unit, 0, 0], rtol=1e-4, atol=1e-9)
    np.testing.assert_allclose(disp[1], [mult * unit, 0, 0], rtol=1e-4, atol=1e-9)


def test_two_particles_large_table():
    # 2 (2 + 998) = 2000 slots: cells (7,9,8) and (9,7,8) collide (29 visits); 2 (2 + 1000) slots: 27
    _, _, mult = two_in_one_cell(998)
    assert mult == 29
    disp, unit, mult = two_in_one_cell(1000)
    assert mult == 27
    np.testing.assert_allclose(disp[0], [-27 * unit, 0, 0], rtol=1e-4, atol=1e-9)
    np.testing.assert_allclose(disp[1], [27 * unit, 0, 0], rtol=1e-4, atol=1e-9)


def test_random_cloud_against_brute_force():
    N = 16
    dx = 1.0 / N
    orc = Oracle(cfg_open(N))
    orc.init()
    rng = np.random.default_rng(3)
    p = (0.5 + (rng.random((300, 3)) - 0.5) * 8 * dx).astype(np.float32)
    orc.set_particles(p)
    dt = 0.005
    orc.correct_volume(dt)
    q, _, _ = orc.particles()
    ref = brute(p.astype(np.float64), N, dx, dt, np.ones(len(p), bool))
    got = q.astype(np.float64) - p
    scale = np.abs(ref).max()
    assert scale > 0
    np.testing.assert_allclose(got, ref, rtol=0, atol=2e-4 * scale)


def test_id_cycle():
    N = 16
    dx = 1.0 / N
    orc = Oracle(cfg_open(N, cycle=4))
    orc.init()
    rng = np.random.default_rng(5)
    p = (0.5 + (rng.random((200, 3)) - 0.5) * 6 * dx).astype(np.float32)
    orc.set_particles(p)
    cur = p.copy()
    for step in range(5):
        orc.correct_volume(0.01)
        q, _, _ = orc.particles()
        moved = np.any(q != cur, axis=1)
        ids = np.arange(len(p))
        assert not moved[ids % 4 != step % 4].any()
        assert moved[ids % 4 == step % 4].mean() > 0.5
        ref = brute(cur.astype(np.float64), N, dx, 0.01, ids % 4 == step % 4)
        np.testing.assert_allclose(q - cur, ref, rtol=0, atol=2e-4 * np.abs(ref).max())
        cur = q
