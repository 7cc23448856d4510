"""Pins of the oracle's 3D stages that the HIP library is compared against bit for bit.

The GPU matches Oracle(p2g_order=1): P2G as a per-face gather over half-cell bins
(oracle/lh_oracle.cpp p2g_bin_order), with its own window of bins per face.  These tests
check that mode, and the other stages no recorded frame reaches, against what the
mathematics fixes:

* P2G (LiquidModel.cpp:613-721): an FP64 brute force over every (particle, face) pair of
  the grid -- not over the reference's tap window, so a window that drops a bin fails --
  on random particles, particles 0.05 cells inside every half-cell boundary (taps at
  |d| = 1.45 carry weight) and particles in the boundary half-cells of the box; both
  summation orders against it and against each other;
* ComputeLiquidPhi (LiquidModel.cpp:539-560): an FP64 brute-force min over the particles
  whose cell window (Point2Index +- ceil(r/dx), clamped) holds the cell;
* the tank acceleration (a body force a(t) = A sin(2 pi f t) on one axis, the sloshing-tank
  workload's extension of ApplyGravity, LiquidModel.cpp:898-910): a floating block from rest
  falls and slides by exactly (A sin(2 pi f t) dt^2, g dt^2, 0), and the faces it touches are
  gravity's faces rotated to its axis;
* ComputeDivergence (LiquidModel.cpp:1065-1083): linear velocity fields with a known
  divergence, a single face spike for the L-inf norm, the liquid test phi <= 0 (a cell with
  phi == 0 counts) and the openness weights.
"""
import json

import numpy as np
import pytest

from conftest import config_file
from oracle import Oracle

VX, VY, VZ, LWX, LWY, LWZ, PHI, WX, WY, WZ = 0, 1, 2, 6, 7, 8, 10, 100, 101, 102


def box_cfg(N=12, lo=-1.0, hi=1.0):
    cfg = json.load(open(config_file("dam_break_48.json")))
    cfg["gridDimension"] = [N, N, N]
    cfg["minCoord"] = [lo] * 3
    cfg["maxCoord"] = [hi] * 3
    c = 0.5 * (lo + hi)
    cfg["Boundary"]["child"]["center"] = [c] * 3
    cfg["Boundary"]["child"]["halfExtent"] = [0.5 * (hi - lo) * 0.8] * 3
    return cfg


def quad_w64(d):
    """quadratic B-spline (Kernel.h:12-17) in FP64"""
    a = np.abs(d)
    return np.where(a < 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))


def brute_p2g(pos, vel, Cm, N, mn, dx):
    """w_f = sum_p W(x_p - x_f), v_f = sum_p W (v_a + C.col(a) . (x_p - x_f)) / w_f over EVERY
    particle and face, in FP64 (no tap window at all)"""
    pos = pos.astype(np.float64)
    g = (pos - mn) / dx
    out = []
    for a in range(3):
        shp = [N, N, N]
        shp[a] += 1
        idx = [np.arange(s, dtype=np.float64) + (0.0 if k == a else 0.5) for k, s in enumerate(shp)]
        # per-axis weights and offsets, (P, shp[k])
        wk = [quad_w64(g[:, k:k + 1] - idx[k][None, :]) for k in range(3)]
        dk = [pos[:, k:k + 1] - (idx[k][None, :] * dx + mn) for k in range(3)]
        col = Cm[:, 3 * a:3 * a + 3].astype(np.float64)
        W = np.einsum("pi,pj,pk->ijk", wk[0], wk[1], wk[2])
        base = np.einsum("pi,pj,pk->ijk", wk[0], wk[1], wk[2] * vel[:, a:a + 1].astype(np.float64))
        aff = (np.einsum("pi,pj,pk->ijk", wk[0] * col[:, 0:1] * dk[0], wk[1], wk[2])
               + np.einsum("pi,pj,pk->ijk", wk[0], wk[1] * col[:, 1:2] * dk[1], wk[2])
               + np.einsum("pi,pj,pk->ijk", wk[0], wk[1], wk[2] * col[:, 2:3] * dk[2]))
        V = base + aff
        v = np.where(np.abs(W) < 1e-8, 0.0, V / np.where(W == 0, 1.0, W))
        out.append((W, v))
    return out


def adversarial_particles(N, mn, dx, rng):
    """particles 0.05 cells inside half-cell boundaries on every axis, and in the first / last
    half-cell of the box"""
    pts = []
    for _ in range(40):
        k = rng.integers(1, 2 * N - 1, 3)
        s = rng.choice([-0.05, 0.05], 3)
        pts.append(k / 2.0 + s)
    for _ in range(16):
        g = rng.uniform(0.0, 0.5, 3)
        hi = rng.random(3) < 0.5
        g = np.where(hi, N - g, g)
        pts.append(np.clip(g, 1e-3, N - 1e-3))
    g = np.asarray(pts)
    return (g * dx + mn).astype(np.float32)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_p2g_bin_order_against_brute_force(seed):
    N = 12
    cfg = box_cfg(N)
    o = Oracle(cfg, dot_mode=1, p2g_order=1)
    o.compute_solid_phi()
    mn, dx = float(o.mn[0]), float(o.dx[0])
    rng = np.random.default_rng(seed)
    pos = np.concatenate([(rng.random((24, 3)) * 2.0 - 1.0).astype(np.float32) * np.float32(0.999),
                          adversarial_particles(N, mn, dx, rng)])
    P = len(pos)
    vel = rng.normal(0, 1, (P, 3)).astype(np.float32)
    Cm = rng.normal(0, 2, (P, 9)).astype(np.float32)
    o.set_particles(pos, vel, Cm)
    res = {}
    for order in (0, 1):
        o.clear_grid()
        o.p2g(order)
        res[order] = [(o.get(WX + a), o.get(VX + a)) for a in range(3)]
    ref = brute_p2g(pos, vel, Cm, N, mn, dx)
    for a in range(3):
        Wr, vr = ref[a]
        for order in (0, 1):
            w, v = res[order][a]
            # weights: FP32 sums of at most ~100 terms <= 0.42; a dropped tap at |d| = 1.45 moves
            # a face's weight by >= 1.25e-3 * 0.125^2 ~ 2e-5
            np.testing.assert_allclose(w, Wr, rtol=0, atol=2e-6, err_msg="w a=%d order=%d" % (a, order))
            live = Wr > 1e-3
            scale = np.abs(vr).max() + 1.0
            np.testing.assert_allclose(v[live], vr[live], rtol=0, atol=2e-5 * scale,
                                       err_msg="v a=%d order=%d" % (a, order))
            assert not np.any(v[Wr == 0])
        # the two summation orders differ by FP32 rounding only
        np.testing.assert_allclose(res[1][a][0], res[0][a][0], rtol=2e-6, atol=1e-7)
        np.testing.assert_allclose(res[1][a][1], res[0][a][1], rtol=2e-5, atol=2e-6)
    assert sum(int((ref[a][0] > 0).sum()) for a in range(3)) > 3000


def test_p2g_bin_order_uniform_field_exact():
    """a uniform velocity (a power of two) with C = 0 comes back exactly on every face that has
    weight: sum W v / sum W = v whatever order the terms come in"""
    N = 12
    o = Oracle(box_cfg(N), dot_mode=1, p2g_order=1)
    o.compute_solid_phi()
    rng = np.random.default_rng(5)
    pos = (rng.random((300, 3)) * 1.9 - 0.95).astype(np.float32)
    vel = np.tile(np.float32([0.5, -0.25, 2.0]), (300, 1))
    o.set_particles(pos, vel, np.zeros((300, 9), np.float32))
    o.clear_grid()
    o.p2g(1)
    for a, val in enumerate((0.5, -0.25, 2.0)):
        v, w = o.get(VX + a), o.get(WX + a)
        live = np.abs(w) >= 1e-8
        assert live.sum() > 500
        np.testing.assert_array_equal(v[live], np.float32(val))
        assert not np.any(v[~live])


def brute_liquid_phi(pos, N, mn, dx, r):
    """min over particles whose tap window holds the cell of |centre - x_p| - r, FP64, capped at the
    cleared value 3 dx"""
    phi = np.full((N, N, N), 3.0 * dx, np.float64)
    c = ((pos - np.float32(mn)) / np.float32(dx)).astype(np.float32).astype(np.int64)  # Point2Index
    d = int(np.ceil(np.float32(r) / np.float32(dx)))
    ctr = (np.arange(N) + 0.5) * dx + mn
    for p in range(len(pos)):
        lo = np.maximum(c[p] - d, 0)
        hi = np.minimum(c[p] + d, N - 1)
        X, Y, Z = np.meshgrid(ctr[lo[0]:hi[0] + 1], ctr[lo[1]:hi[1] + 1], ctr[lo[2]:hi[2] + 1], indexing="ij")
        dist = np.sqrt((X - pos[p, 0]) ** 2 + (Y - pos[p, 1]) ** 2 + (Z - pos[p, 2]) ** 2) - r
        sl = phi[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1]
        np.minimum(sl, dist, out=sl)
    return phi


@pytest.mark.parametrize("seed", [0, 1])
def test_liquid_phi_against_brute_force_min(seed):
    N = 14
    o = Oracle(box_cfg(N), dot_mode=1, p2g_order=1)
    o.compute_solid_phi()
    mn, dx = float(o.mn[0]), float(o.dx[0])
    rng = np.random.default_rng(seed)
    pos = np.concatenate([(rng.random((200, 3)) * 1.98 - 0.99).astype(np.float32),
                          adversarial_particles(N, mn, dx, rng)])
    o.set_particles(pos)
    o.clear_grid()
    o.compute_liquid_phi()
    phi = o.get(PHI)
    ref = brute_liquid_phi(pos, N, mn, dx, 2.0 * dx)
    np.testing.assert_allclose(phi, ref, rtol=0, atol=1e-6)
    # the classification the solver uses (phi <= 0) agrees away from rounding distance of 0
    clear = np.abs(ref) > 1e-5
    np.testing.assert_array_equal((phi <= 0)[clear], (ref <= 0)[clear])
    assert (ref < 0).sum() > 200 and (ref == 3.0 * dx).sum() > 0


@pytest.mark.parametrize("t0,axis", [(0.25, 0), (0.75, 0), (0.25, 2), (0.125, 1)])
def test_tank_acceleration_known_answer(t0, axis):
    """the reference config's floating block from rest: gravity and the tank acceleration are
    uniform over all interior faces (including air faces, Q8), so the right-hand side is zero,
    the pressure is zero, G2P returns the uniform field (partition of unity) and every particle
    moves by exactly v dt"""
    cfg = json.load(open(config_file("pure_liquid_3d.json")))
    cfg["TankAcceleration"] = {"amplitude": 3.0, "frequency": 1.0, "axis": axis}
    o = Oracle(cfg, dot_mode=1, p2g_order=1)
    o.init()
    o.set_time(t0)
    x0 = o.particles()[0].copy()
    dt = np.float32(0.01)
    st, _, _ = o.step(dt)
    assert st.iterations == 0
    acc = 3.0 * np.sin(2.0 * np.pi * t0)            # 3, -3, 3 and 3 / sqrt(2)
    v = np.zeros(3, np.float32)
    v[1] = np.float32(-9.81) * dt
    v[axis] += np.float32(acc) * dt
    d = o.particles()[0] - x0
    np.testing.assert_allclose(d, np.tile(v * dt, (len(x0), 1)), rtol=0, atol=1e-8)
    assert abs(float(v[axis] * dt)) > 2e-4     # the tank term is resolved by the tolerance


def test_tank_acceleration_faces_are_gravity_faces_rotated():
    N = 10
    for axis in range(3):
        cfg = box_cfg(N)
        cfg["gravity"] = -1.0
        cfg["TankAcceleration"] = {"amplitude": 1.0, "frequency": 1.0, "axis": axis}
        o = Oracle(cfg, dot_mode=1, p2g_order=1)
        o.set_time(0.25)
        o.clear_grid()
        o.apply_gravity(np.float32(1.0))
        grav = o.get(VY) == -1.0
        if axis == 1:   # both on vy: -1 + 1
            assert np.all(o.get(VY)[grav] == 0.0)
            continue
        tank = o.get(VX + axis) == 1.0
        assert tank.sum() == grav.sum() == (N - 2) * (N - 1) * (N - 2)
        # gravity's mask (faces x in [1, N-2], y in [1, N-1], z in [1, N-2]) with the roles of
        # y and the tank axis exchanged
        perm = [0, 1, 2]
        perm[1], perm[axis] = perm[axis], perm[1]
        np.testing.assert_array_equal(tank, np.transpose(grav, perm))
        assert np.count_nonzero(o.get(VX + axis)) == tank.sum()


def test_divergence_closed_forms():
    N = 8
    for dot_mode in (0, 1):
        o = Oracle(box_cfg(N), dot_mode=dot_mode, p2g_order=1)
        dx = float(o.dx[0])
        sx, sy, sz = o.field_shape(0), o.field_shape(1), o.field_shape(2)
        # v = (a i, b j, c k) on the faces (index units), all faces open: every cell's divergence
        # is (a + b + c) / dx
        a_, b_, c_ = 0.5, -0.25, 1.0
        o.set(VX, a_ * np.indices(sx)[0])
        o.set(VY, b_ * np.indices(sy)[1])
        o.set(VZ, c_ * np.indices(sz)[2])
        for f, s in ((LWX, sx), (LWY, sy), (LWZ, sz)):
            o.set(f, np.ones(s))
        phi = np.full((N, N, N), 0.1, np.float32)
        phi[2:6, 1:7, 3:8] = -0.5
        phi[0, 0, 0] = 0.0                                    # phi == 0 counts (LiquidModel.cpp:1071)
        o.set(PHI, phi)
        mean, linf, cnt = o.divergence()
        assert cnt == 4 * 6 * 5 + 1
        expect = (a_ + b_ + c_) / dx
        assert mean == pytest.approx(expect, rel=1e-6)
        assert linf == pytest.approx(abs(expect), rel=1e-6)
        # one spike on an x face between two liquid cells: +s/dx in the upwind cell, -s/dx downwind,
        # zero elsewhere; a closed face (W = 0) hides it
        o.set(VX, np.zeros(sx))
        o.set(VY, np.zeros(sy))
        o.set(VZ, np.zeros(sz))
        vx = np.zeros(sx, np.float32)
        vx[3, 2, 4] = 0.75
        o.set(VX, vx)
        mean, linf, cnt = o.divergence()
        assert linf == pytest.approx(0.75 / dx, rel=1e-6)
        assert mean == pytest.approx(0.0, abs=1e-6)
        lw = np.ones(sx, np.float32)
        lw[3, 2, 4] = 0.25
        o.set(LWX, lw)
        mean, linf, cnt = o.divergence()
        assert linf == pytest.approx(0.25 * 0.75 / dx, rel=1e-6)
        # the spike on the liquid block's +x boundary face: only the liquid cell counts
        vx[3, 2, 4] = 0.0
        vx[6, 2, 4] = 0.75
        o.set(VX, vx)
        o.set(LWX, np.ones(sx))
        mean, linf, cnt = o.divergence()
        assert linf == pytest.approx(0.75 / dx, rel=1e-6)
        assert mean == pytest.approx(0.75 / dx / cnt, rel=1e-5)


def grad_w64(d):
    """unsigned derivative of the B-spline in |d| (Kernel.h:26-31, quirk Q1), FP64"""
    a = np.abs(d)
    return np.where(a < 0.5, -2.0 * a, np.where(a < 1.5, a - 1.5, 0.0))


@pytest.mark.parametrize("seed", [0, 1])
def test_g2p_against_brute_force(seed):
    """v_a = sum_f W v_f (not normalised) and C.col(a) = sum_f v_f (g_x w_y w_z, w_x g_y w_z,
    w_x w_y g_z) over every face, FP64 (LiquidModel.cpp:723-823)"""
    N = 12
    o = Oracle(box_cfg(N), dot_mode=1, p2g_order=1)
    mn, dx = float(o.mn[0]), float(o.dx[0])
    rng = np.random.default_rng(10 + seed)
    pos = np.concatenate([(rng.random((30, 3)) * 1.998 - 0.999).astype(np.float32),
                          adversarial_particles(N, mn, dx, rng)])
    o.set_particles(pos)
    fields = []
    for a in range(3):
        f = rng.normal(0, 1, o.field_shape(a)).astype(np.float32)
        o.set(VX + a, f)
        fields.append(f.astype(np.float64))
    o.g2p()
    _, vel, Cm = o.particles()
    g = (pos.astype(np.float64) - mn) / dx
    for a in range(3):
        shp = o.field_shape(a)
        idx = [np.arange(s, dtype=np.float64) + (0.0 if k == a else 0.5) for k, s in enumerate(shp)]
        wk = [quad_w64(g[:, k:k + 1] - idx[k][None, :]) for k in range(3)]
        gk = [grad_w64(g[:, k:k + 1] - idx[k][None, :]) for k in range(3)]
        f = fields[a]
        v = np.einsum("pi,pj,pk,ijk->p", wk[0], wk[1], wk[2], f)
        c = [np.einsum("pi,pj,pk,ijk->p", gk[0], wk[1], wk[2], f),
             np.einsum("pi,pj,pk,ijk->p", wk[0], gk[1], wk[2], f),
             np.einsum("pi,pj,pk,ijk->p", wk[0], wk[1], gk[2], f)]
        np.testing.assert_allclose(vel[:, a], v, rtol=0, atol=2e-6, err_msg="v a=%d" % a)
        for k in range(3):
            np.testing.assert_allclose(Cm[:, 3 * a + k], c[k], rtol=0, atol=5e-6, err_msg="C col %d row %d" % (a, k))


def test_extrapolation_in_place_validity_chain():
    """Q6 (LiquidModel.cpp:953-978): validity updated in place in (x, y, z) order, values from the
    old field.  Only the x-face (1,1,1) is open, with v = 1.  (1,1,2) sees only (1,1,1): 1.
    (1,2,1) sees (1,1,1) and (1,1,2), validated earlier in the same sweep with old value 0: 1/2.
    A Jacobi sweep would give 1 to both.  (2,1,1): the validity range starts at y, z = 1, so its
    valid neighbours are the x = 1 faces with y, z in {1, 2}, one of them carrying 1: 1/4."""
    N = 8
    o = Oracle(box_cfg(N), dot_mode=1, p2g_order=1)
    o.set_particles(np.zeros((1, 3), np.float32))   # Extrapolate returns at once without particles
    wx = np.zeros((N + 1, N, N), np.float32)
    wx[1, 1, 1] = 1.0
    o.set(LWX, wx)
    o.set(VX, wx)
    o.set(12, np.zeros((N + 1,) * 3, np.uint8))
    o.extrapolate()
    vx = o.get(VX)
    assert vx[1, 1, 2] == 1.0 and vx[1, 2, 1] == 0.5 and vx[2, 1, 1] == 0.25
    assert vx[1, 1, 1] == 1.0
