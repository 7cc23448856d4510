"""Closed-form checks of the oracle's scalar routines (Kernel.h, MathUtils, Boundary.h).

Each expected value below comes from the mathematics of the routine, not from
retyping its formula: B-spline knots, partition of unity, first moment,
finite-difference derivatives, exact linear interpolation, box-SDF geometry.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import scalar_lib


@pytest.fixture(scope="module")
def L():
    return scalar_lib()


def test_quadratic_weight_knots(L):
    # quadratic B-spline: 3/4 at the centre, 1/2 at the half-knot, 1/8 at 1, 0 at 3/2 (Kernel.h:12-17)
    for d, w in [(0.0, 0.75), (0.5, 0.5), (1.0, 0.125), (1.5, 0.0), (2.0, 0.0), (-0.5, 0.5), (-1.0, 0.125)]:
        assert L.orc_quad_w(d) == pytest.approx(w, abs=0)


def test_quadratic_weight_partition_and_moment(L):
    rng = np.random.default_rng(0)
    for x in rng.uniform(0, 1, 200):
        taps = [x - i for i in (-2, -1, 0, 1, 2, 3)]
        w = np.array([L.orc_quad_w(float(np.float32(t))) for t in taps], np.float64)
        assert w.sum() == pytest.approx(1.0, abs=2e-7)          # partition of unity
        assert (w * np.array(taps)).sum() == pytest.approx(0.0, abs=2e-7)  # first moment vanishes


def test_gradient_is_unsigned_derivative(L):
    # Kernel.h:26-31 returns dW/d|d| (sign of d dropped: quirk Q1); check against a
    # central difference of the weight in |d|, away from the knots at 0.5 and 1.5
    for a in [0.1, 0.3, 0.45, 0.6, 0.9, 1.2, 1.4]:
        h = 1e-3
        fd = (L.orc_quad_w(a + h) - L.orc_quad_w(a - h)) / (2 * h)
        assert L.orc_quad_gw(a) == pytest.approx(fd, abs=2e-3)
        assert L.orc_quad_gw(-a) == L.orc_quad_gw(a)   # unsigned
    assert L.orc_quad_gw(1.5) == 0.0 and L.orc_quad_gw(3.0) == 0.0


def test_fraction_in_region(L):
    f = L.orc_fraction_in_region
    assert f(-1.0, -2.0) == 1.0           # both inside
    assert f(1.0, 2.0) == 0.0             # both outside
    assert f(-1.0, 1.0) == 0.5            # zero crossing at the middle
    assert f(-1.0, 3.0) == 0.25           # crossing at 1/4 from the negative end
    assert f(3.0, -1.0) == 0.25
    assert f(0.0, -1.0) == 0.0 and f(-1.0, 0.0) == 0.0   # exact zeros do not count


def test_trilerp_reproduces_linear_fields(L):
    rng = np.random.default_rng(1)
    for _ in range(50):
        a, b, c, d = rng.uniform(-2, 2, 4)
        corners = [a + b * i + c * j + d * k for k in (0, 1) for j in (0, 1) for i in (0, 1)]
        v = (C.c_float * 8)(*corners)
        f1, f2, f3 = rng.uniform(0, 1, 3)
        got = L.orc_trilerp(v, f1, f2, f3)
        assert got == pytest.approx(a + b * f1 + c * f2 + d * f3, abs=5e-6)
    v = (C.c_float * 8)(*range(8))
    assert L.orc_trilerp(v, 2.0, -1.0, 0.0) == L.orc_trilerp(v, 1.0, 0.0, 0.0)   # factors clamp to [0,1]


def test_barycentric_clamping(L):
    i, f = C.c_int(), C.c_float()
    L.orc_barycentric(2.25, 0, 10, C.byref(i), C.byref(f)); assert (i.value, f.value) == (2, 0.25)
    L.orc_barycentric(-0.3, 0, 10, C.byref(i), C.byref(f)); assert (i.value, f.value) == (0, 0.0)
    L.orc_barycentric(9.5, 0, 10, C.byref(i), C.byref(f)); assert (i.value, f.value) == (8, 1.0)


def _prog(*nodes):
    return (C.c_float * (12 * len(nodes)))(*[x for n in nodes for x in n])


def rect(c, h, inside):
    return [0, c[0], c[1], c[2], h[0], h[1], h[2], 0, 1.0 if inside else 0.0, 0, 0, 0]


def sphere(c, r, inside=False):
    return [1, c[0], c[1], c[2], 0, 0, 0, r, 1.0 if inside else 0.0, 0, 0, 0]


def test_box_sdf_geometry(L):
    box = _prog(rect((0, 0, 0), (0.9, 0.9, 0.9), False))
    e = lambda p: L.orc_sdf_eval(box, 1, *p)
    assert e((0, 0, 0)) == pytest.approx(-0.9)              # deepest inside point
    assert e((0.5, 0, 0)) == pytest.approx(-0.4)
    assert e((2.0, 0, 0)) == pytest.approx(1.1)             # face region
    assert e((2.0, 2.0, 2.0)) == pytest.approx(1.1 * np.sqrt(3), rel=1e-6)  # corner region
    tank = _prog(rect((0, 0, 0), (0.9, 0.9, 0.9), True))   # useInside flips the sign (solid outside)
    assert L.orc_sdf_eval(tank, 1, 0.5, 0, 0) == pytest.approx(0.4)


def test_csg_operators(L):
    a = rect((0, 0, 0), (1, 1, 1), False)
    b = sphere((1, 0, 0), 0.5)
    p = (0.5, 0.0, 0.0)
    va = L.orc_sdf_eval(_prog(a), 1, *p)
    vb = L.orc_sdf_eval(_prog(b), 1, *p)
    assert vb == pytest.approx(0.0, abs=1e-7)
    assert L.orc_sdf_eval(_prog(a, b, [10] + [0] * 11), 3, *p) == min(va, vb)
    assert L.orc_sdf_eval(_prog(a, b, [11] + [0] * 11), 3, *p) == max(va, vb)
    assert L.orc_sdf_eval(_prog(a, b, [12] + [0] * 11), 3, *p) == max(va, -vb)
    assert L.orc_sdf_eval(_prog(a, [13] + [0] * 11), 2, *p) == va
