"""ctypes binding of the CPU oracle (oracle/liblh_oracle.so).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module; the HIP library never does.  The oracle reads the same JSON configs as
the library but through its own parser (this file), so a config bug in one is
caught by the other.

Config keys follow the reference (SceneParameter.h:9-152); the optional keys
(Boundary, FluidBlocks, particlesPerCell, seed, TankAcceleration, toleranceMode,
smoother) are the build's extensions, with defaults that reproduce the
reference (Scene.cpp:18 boundary, LiquidModel.cpp:506-512 spawn).
"""
import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblh_oracle.so")


class Params(C.Structure):
    _fields_ = [
        ("N", C.c_int * 3), ("mn", C.c_float * 3), ("mx", C.c_float * 3),
        ("gravity", C.c_float), ("density", C.c_float),
        ("max_iter", C.c_int), ("tol", C.c_float), ("tol_mode", C.c_int), ("precond", C.c_int),
        ("num_level", C.c_int), ("num_pre", C.c_int), ("num_post", C.c_int), ("num_final", C.c_int),
        ("omega", C.c_float), ("dot_mode", C.c_int), ("smoother", C.c_int), ("p2g_order", C.c_int),
        ("tank_amp", C.c_float), ("tank_freq", C.c_float), ("tank_axis", C.c_int),
        ("nsdf", C.c_int), ("sdf", C.POINTER(C.c_float)),
        ("correct_volume", C.c_int), ("correct_cycle", C.c_int), ("seed", C.c_ulonglong),
        ("num_particles", C.c_longlong), ("coarsest_fp64", C.c_int),
        ("coarsest_cheb", C.c_int), ("cheb_lo", C.c_float),
    ]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("initial_residual", C.c_double),
                ("final_residual", C.c_double), ("rhs_norm", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(Params)]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_sdf.argtypes = [vp, fp, C.c_int]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_max_threads.restype = C.c_int
        L.orc_grid_info.argtypes = [vp, ip, fp, fp]
        L.orc_field.argtypes = [vp, C.c_int, vp, C.c_longlong, C.c_int]
        L.orc_set_particles.argtypes = [vp, fp, fp, fp, C.c_longlong]
        L.orc_get_particles.argtypes = [vp, fp, fp, fp, C.c_longlong]
        L.orc_num_particles.restype = C.c_longlong
        L.orc_num_particles.argtypes = [vp]
        L.orc_set_time.argtypes = [vp, C.c_double]
        L.orc_get_time.restype = C.c_double
        L.orc_get_time.argtypes = [vp]
        L.orc_spawn_lattice.argtypes = [vp, fp, fp, C.c_longlong]
        L.orc_seed_blocks.restype = C.c_longlong
        L.orc_seed_blocks.argtypes = [vp, fp, C.c_int, C.c_int, C.c_ulonglong]
        for name in ("orc_compute_solid_phi", "orc_clear_grid", "orc_compute_liquid_phi", "orc_extrapolate",
                     "orc_g2p"):
            getattr(L, name).argtypes = [vp]
        L.orc_p2g.argtypes = [vp, C.c_int]
        L.orc_apply_gravity.argtypes = [vp, C.c_float]
        L.orc_solve_pressure.argtypes = [vp, C.c_float, C.c_float, C.POINTER(SolveStats)]
        L.orc_velocity_update.argtypes = [vp, C.c_float]
        L.orc_advect.argtypes = [vp, C.c_float]
        L.orc_correct_volume.argtypes = [vp, C.c_float]
        L.orc_get_cfl.restype = C.c_float
        L.orc_get_cfl.argtypes = [vp, fp]
        L.orc_divergence.argtypes = [vp, dp, dp, ip]
        L.orc_solid_phi_at.restype = C.c_float
        L.orc_solid_phi_at.argtypes = [vp, C.c_float, C.c_float, C.c_float]
        L.orc_solid_gradient.argtypes = [vp, C.c_float, C.c_float, C.c_float, fp]
        L.orc_step.argtypes = [vp, C.c_float, C.POINTER(SolveStats), dp, dp]
        L.orc_build_system.argtypes = [vp, C.c_float, C.c_float]
        L.orc_get_level.argtypes = [vp, C.c_int, fp, fp, ip]
        L.orc_num_levels.restype = C.c_int
        L.orc_num_levels.argtypes = [vp]
        L.orc_solve_fp32.argtypes = [vp, fp, fp, C.POINTER(SolveStats)]
        L.orc_solve_fp64.argtypes = [vp, dp, dp, C.c_double, C.c_int, C.POINTER(SolveStats)]
        L.orc_apply_a64.argtypes = [vp, dp, dp]
        L.orc_coarsest_solve64.argtypes = [vp, dp, dp]
        L.orc_coarsest_solve64.restype = C.c_int
        _lib = L
    return _lib


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ----------------------------------------------------------------- config
def load_config(path_or_dict):
    if isinstance(path_or_dict, dict):
        return path_or_dict
    with open(path_or_dict) as f:
        return json.load(f)


def sdf_program(node):
    """Postfix program (12 floats/node) of a Boundary JSON tree (Boundary.h:27-94)."""
    out = []

    def emit(n):
        t = n["type"]
        if t in ("Rect", "Sphere"):
            c = list(n.get("center", [0, 0, 0])) + [0.0] * 3  # 2D trees: z ignored
            h = list(n.get("halfExtent", [0, 0, 0])) + [0.0] * 3
            out.append([0 if t == "Rect" else 1, c[0], c[1], c[2], h[0], h[1], h[2],
                        n.get("radius", 0.0), 1.0 if n.get("useInside", False) else 0.0, 0, 0, 0])
            return
        kinds = {"Union": 10, "Intersection": 11, "Subtract": 12, "Substract": 12, "None": 13}
        if t not in kinds:
            raise ValueError("unknown boundary type %r" % t)
        if t == "None":
            emit(n["child"])
        else:
            emit(n["left"])
            emit(n["right"])
        out.append([kinds[t]] + [0.0] * 11)

    emit(node)
    return np.asarray(out, dtype=np.float32).reshape(-1)


DEFAULT_BOUNDARY = {"type": "None", "child": {"type": "Rect", "center": [0, 0, 0],
                                              "halfExtent": [0.9, 0.9, 0.9], "useInside": True}}


def params_from_config(cfg, dot_mode=1, p2g_order=1, smoother=None):
    s = cfg["SolverParameter"]
    lq = cfg["LiquidParameter"]
    p = Params()
    for a in range(3):
        p.N[a], p.mn[a], p.mx[a] = 1, 0.0, 1.0
    for a in range(len(cfg["gridDimension"])):
        p.N[a] = int(cfg["gridDimension"][a])
        p.mn[a] = float(cfg["minCoord"][a])
        p.mx[a] = float(cfg["maxCoord"][a])
    p.gravity = float(cfg["gravity"])
    p.density = float(lq["density"])
    p.max_iter = int(s["maxIterations"])
    p.tol = float(s["tolerance"])
    p.tol_mode = 1 if s.get("toleranceMode", "absolute") == "relative" else 0
    pt = s["preconditionerType"]
    p.precond = {"None": 0, "Jacobi": 1, "Multigrid": 2}[pt]
    p.num_level = int(s["numLevel"])
    p.num_pre = int(s["numPreSmooth"])
    p.num_post = int(s["numPostSmooth"])
    p.num_final = int(s["numFinalSmooth"])
    p.omega = float(s["smoothingFactor"])
    p.dot_mode = dot_mode
    sm = smoother if smoother is not None else s.get("smoother", "reference")
    p.smoother = 1 if sm == "symmetric" else 0
    p2g = p2g_order
    p.p2g_order = p2g
    tank = cfg.get("TankAcceleration")
    if tank:
        p.tank_amp = float(tank.get("amplitude", 0.0))
        p.tank_freq = float(tank.get("frequency", 1.0))
        p.tank_axis = int(tank.get("axis", 0))
    else:
        p.tank_amp, p.tank_freq, p.tank_axis = 0.0, 0.0, -1
    p.correct_volume = 1 if cfg.get("correctVolume", False) else 0
    p.correct_cycle = int(lq["correctCycle"])
    p.seed = int(cfg.get("seed", 0))
    p.num_particles = int(lq["numParticles"])
    p.coarsest_fp64 = 1 if coarsest_precision(cfg) == "fp64" else 0
    if s.get("coarsestSolver", "sweeps") == "chebyshev":
        p.coarsest_cheb = int(s.get("coarsestSteps", 12))
        p.cheb_lo = float(s.get("chebyshevLower", 0.01))
    return p


def coarsest_precision(cfg):
    """SolverParameter.coarsestPrecision; default fp64 with "precision": "refined", else fp32"""
    s = cfg["SolverParameter"]
    return s.get("coarsestPrecision", "fp64" if cfg.get("precision") == "refined" else "fp32")


class Oracle:
    """Stage-level handle on the CPU restatement of LiquidModel<3> / CGSolver<3>."""

    FIELD_SHAPES = None

    def __init__(self, cfg, dot_mode=1, p2g_order=1, smoother=None, threads=None):
        self.cfg = load_config(cfg)
        self.L = lib()
        if threads is not None:
            self.L.orc_set_threads(int(threads))
        prm = params_from_config(self.cfg, dot_mode, p2g_order, smoother)
        self._sdf = sdf_program(self.cfg.get("Boundary", DEFAULT_BOUNDARY))
        prm.nsdf = len(self._sdf) // 12
        prm.sdf = _fptr(self._sdf)
        self.prm = prm
        self.h = self.L.orc_create(C.byref(prm))
        dims = (C.c_int * 3)()
        mn = (C.c_float * 3)()
        dx = (C.c_float * 3)()
        self.L.orc_grid_info(self.h, dims, mn, dx)
        self.N = tuple(dims)
        self.mn = np.array(list(mn), np.float32)
        self.dx = np.array(list(dx), np.float32)

    def __del__(self):
        try:
            self.L.orc_destroy(self.h)
        except Exception:
            pass

    # --- shapes follow Grid<3> (Grid.cpp:86-98)
    def field_shape(self, f):
        N = self.N
        return {0: (N[0] + 1, N[1], N[2]), 1: (N[0], N[1] + 1, N[2]), 2: (N[0], N[1], N[2] + 1),
                3: (N[0] + 1, N[1], N[2]), 4: (N[0], N[1] + 1, N[2]), 5: (N[0], N[1], N[2] + 1),
                6: (N[0] + 1, N[1], N[2]), 7: (N[0], N[1] + 1, N[2]), 8: (N[0], N[1], N[2] + 1),
                9: N, 10: N, 11: (N[0] + 1, N[1] + 1, N[2] + 1), 12: (N[0] + 1, N[1] + 1, N[2] + 1), 13: N,
                100: (N[0] + 1, N[1], N[2]), 101: (N[0], N[1] + 1, N[2]), 102: (N[0], N[1], N[2] + 1)}[f]

    def get(self, f):
        shp = self.field_shape(f)
        a = np.empty(shp, np.uint8 if f == 12 else np.float32)
        rc = self.L.orc_field(self.h, f, a.ctypes.data, a.size, 0)
        assert rc == 0
        return a

    def set(self, f, a):
        shp = self.field_shape(f)
        a = np.ascontiguousarray(a, np.uint8 if f == 12 else np.float32).reshape(shp)
        rc = self.L.orc_field(self.h, f, a.ctypes.data, a.size, 1)
        assert rc == 0

    def set_particles(self, pos, vel=None, Cm=None):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        n = pos.shape[0]
        vel = np.zeros((n, 3), np.float32) if vel is None else np.ascontiguousarray(vel, np.float32).reshape(n, 3)
        Cm = np.zeros((n, 9), np.float32) if Cm is None else np.ascontiguousarray(Cm, np.float32).reshape(n, 9)
        self.L.orc_set_particles(self.h, _fptr(pos), _fptr(vel), _fptr(Cm), n)

    def particles(self):
        n = self.L.orc_num_particles(self.h)
        pos = np.empty((n, 3), np.float32)
        vel = np.empty((n, 3), np.float32)
        Cm = np.empty((n, 9), np.float32)
        self.L.orc_get_particles(self.h, _fptr(pos), _fptr(vel), _fptr(Cm), n)
        return pos, vel, Cm

    @property
    def num_particles(self):
        return int(self.L.orc_num_particles(self.h))

    # --- LiquidModel<3>::Init (LiquidModel.cpp:506-512) with the build's optional seeding
    def init(self):
        self.L.orc_compute_solid_phi(self.h)
        blocks = self.cfg.get("FluidBlocks")
        if blocks:
            arr = []
            for b in blocks:
                v = b.get("velocity", [0, 0, 0])
                arr += list(b["min"]) + list(b["max"]) + list(v) + [float(b.get("velocityNoise", 0.0))]
            arr = np.asarray(arr, np.float32)
            k = int(self.cfg.get("particlesPerCellAxis", 2))
            self.L.orc_seed_blocks(self.h, _fptr(arr), len(blocks), k, int(self.cfg.get("seed", 0)))
        else:
            mn = np.array(self.cfg["minCoord"], np.float32)
            mx = np.array(self.cfg["maxCoord"], np.float32)
            off = np.float32(0.3) * (mx - mn)
            lo = (mn + off).astype(np.float32)
            hi = (mx - off).astype(np.float32)
            cap = int(self.cfg["LiquidParameter"]["numParticles"])
            self.L.orc_spawn_lattice(self.h, _fptr(lo), _fptr(hi), cap)

    def set_time(self, t):
        self.L.orc_set_time(self.h, float(t))

    @property
    def time(self):
        return float(self.L.orc_get_time(self.h))

    # --- stages
    def set_boundary(self, node):
        """LiquidModel<3>::SetBoundary with a config "Boundary" node"""
        prog = sdf_program(node)
        self.L.orc_set_sdf(self.h, _fptr(prog), len(prog) // 12)

    def compute_solid_phi(self):
        self.L.orc_compute_solid_phi(self.h)

    def clear_grid(self):
        self.L.orc_clear_grid(self.h)

    def compute_liquid_phi(self):
        self.L.orc_compute_liquid_phi(self.h)

    def p2g(self, order=None):
        rc = self.L.orc_p2g(self.h, self.prm.p2g_order if order is None else order)
        assert rc == 0, rc

    def apply_gravity(self, dt):
        self.L.orc_apply_gravity(self.h, dt)

    def solve_pressure(self, dt, rho=None):
        """CGSolver<3>::Solve(rho, dt); rho=None: the scene's density"""
        st = SolveStats()
        self.L.orc_solve_pressure(self.h, self.prm.density if rho is None else rho, dt, C.byref(st))
        return st

    def velocity_update(self, dt):
        self.L.orc_velocity_update(self.h, dt)

    def extrapolate(self):
        self.L.orc_extrapolate(self.h)

    def g2p(self):
        self.L.orc_g2p(self.h)

    def advect(self, dt):
        self.L.orc_advect(self.h, dt)

    def correct_volume(self, dt):
        """LiquidModel<3>::CorrectVolume (LiquidModel.cpp:1085-1132) on the current particles"""
        rc = self.L.orc_correct_volume(self.h, dt)
        assert rc == 0, rc

    def get_cfl(self):
        mv = C.c_float()
        dt = self.L.orc_get_cfl(self.h, C.byref(mv))
        return float(dt), float(mv.value)

    def divergence(self):
        m, l, c = C.c_double(), C.c_double(), C.c_int()
        self.L.orc_divergence(self.h, C.byref(m), C.byref(l), C.byref(c))
        return m.value, l.value, c.value

    def pre_step(self, dt):
        self.clear_grid()
        self.compute_liquid_phi()

    def step(self, dt):
        st = SolveStats()
        m, l = C.c_double(), C.c_double()
        rc = self.L.orc_step(self.h, dt, C.byref(st), C.byref(m), C.byref(l))
        assert rc == 0, rc
        return st, m.value, l.value

    # --- solver access
    def build_system(self, dt, rho=None):
        self.L.orc_build_system(self.h, self.prm.density if rho is None else rho, dt)

    def level(self, l):
        dims = (C.c_int * 3)()
        self.L.orc_get_level(self.h, l, None, None, dims)
        d = tuple(dims)
        ad = np.empty(d, np.float32)
        co = np.empty(d + (6,), np.float32)
        self.L.orc_get_level(self.h, l, _fptr(ad), _fptr(co), dims)
        return ad, co

    @property
    def num_levels(self):
        return int(self.L.orc_num_levels(self.h))

    def solve_fp32(self, b):
        b = np.ascontiguousarray(b, np.float32)
        x = np.zeros_like(b)
        st = SolveStats()
        self.L.orc_solve_fp32(self.h, _fptr(b), _fptr(x), C.byref(st))
        return x, st

    def solve_fp64(self, b, rel_tol=1e-12, max_iter=2000):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b)
        st = SolveStats()
        self.L.orc_solve_fp64(self.h, _dptr(b), _dptr(x), rel_tol, max_iter, C.byref(st))
        return x, st

    def apply_a64(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        self.L.orc_apply_a64(self.h, _dptr(x), _dptr(y))
        return y

    def coarsest_solve64(self, r):
        """the multigrid coarsest level's solve of r (its configured sweeps or Chebyshev steps) in
        FP64 from z = 0, on the system of the last build"""
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros_like(r)
        assert self.L.orc_coarsest_solve64(self.h, _dptr(r), _dptr(z)) == 0
        return z


def scalar_lib():
    """The oracle's scalar helpers (Kernel.h, MathUtils, Boundary.h) for closed-form tests."""
    L = lib()
    fp = C.POINTER(C.c_float)
    for name in ("orc_quad_w", "orc_quad_gw"):
        getattr(L, name).restype = C.c_float
        getattr(L, name).argtypes = [C.c_float]
    L.orc_fraction_in_region.restype = C.c_float
    L.orc_fraction_in_region.argtypes = [C.c_float, C.c_float]
    L.orc_trilerp.restype = C.c_float
    L.orc_trilerp.argtypes = [fp, C.c_float, C.c_float, C.c_float]
    L.orc_sdf_eval.restype = C.c_float
    L.orc_sdf_eval.argtypes = [fp, C.c_int, C.c_float, C.c_float, C.c_float]
    L.orc_barycentric.argtypes = [C.c_float, C.c_int, C.c_int, C.POINTER(C.c_int), fp]
    return L


class Oracle2D:
    """Stage-level handle on the CPU restatement of LiquidModel<2> / CGSolver<2>
    (LiquidModel.cpp:10-474, PressureSolver.cpp:9-418)."""

    def __init__(self, cfg, dot_mode=1, p2g_order=1):
        self.cfg = load_config(cfg)
        assert len(self.cfg["gridDimension"]) == 2
        self.L = lib()
        L = self.L
        vp, fp, dp, ip = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.orc2_create.restype = vp
        L.orc2_create.argtypes = [C.POINTER(Params)]
        for name, args in (("orc2_destroy", [vp]), ("orc2_init", [vp]), ("orc2_clear_grid", [vp]),
                           ("orc2_compute_liquid_phi", [vp]), ("orc2_compute_solid_phi", [vp]),
                           ("orc2_p2g", [vp, C.c_int]), ("orc2_apply_gravity", [vp, C.c_float]),
                           ("orc2_velocity_update", [vp, C.c_float]), ("orc2_extrapolate", [vp]),
                           ("orc2_g2p", [vp]), ("orc2_advect", [vp, C.c_float]),
                           ("orc2_correct_volume", [vp, C.c_float]), ("orc2_set_time", [vp, C.c_double]),
                           ("orc2_set_sdf", [vp, fp, C.c_int]), ("orc2_divergence", [vp, dp, dp, ip]),
                           ("orc2_particles", [vp, fp, fp, fp, C.c_longlong, C.c_int])):
            getattr(L, name).argtypes = args
            getattr(L, name).restype = None
        L.orc2_solve_pressure.argtypes = [vp, C.c_float, C.POINTER(SolveStats)]
        L.orc2_step.argtypes = [vp, C.c_float, C.POINTER(SolveStats)]
        L.orc2_field.argtypes = [vp, C.c_int, vp, C.c_size_t, C.c_int]
        L.orc2_get_cfl.restype = C.c_float
        L.orc2_get_cfl.argtypes = [vp, fp]
        L.orc2_num_particles.restype = C.c_longlong
        L.orc2_num_particles.argtypes = [vp]
        L.orc2_time.restype = C.c_double
        L.orc2_time.argtypes = [vp]
        prm = params_from_config(self.cfg, dot_mode=dot_mode, p2g_order=p2g_order)
        self._sdf = sdf_program(self.cfg.get("Boundary", DEFAULT_BOUNDARY))
        prm.nsdf = len(self._sdf) // 12
        prm.sdf = _fptr(self._sdf)
        self.N = [int(v) for v in self.cfg["gridDimension"]]
        self.h = L.orc2_create(C.byref(prm))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc2_destroy(self.h)
            self.h = None

    def field_shape(self, f):
        N = self.N
        return {0: (N[0] + 1, N[1]), 1: (N[0], N[1] + 1), 3: (N[0] + 1, N[1]), 4: (N[0], N[1] + 1),
                6: (N[0] + 1, N[1]), 7: (N[0], N[1] + 1), 9: tuple(N), 10: tuple(N), 11: (N[0] + 1, N[1] + 1),
                12: (N[0] + 1, N[1] + 1), 13: tuple(N)}[f]

    def get(self, f):
        a = np.empty(self.field_shape(f), np.uint8 if f == 12 else np.float32)
        assert self.L.orc2_field(self.h, f, a.ctypes.data, a.size, 0) == 0
        return a

    def set(self, f, a):
        a = np.ascontiguousarray(a, np.uint8 if f == 12 else np.float32).reshape(self.field_shape(f))
        assert self.L.orc2_field(self.h, f, a.ctypes.data, a.size, 1) == 0

    def init(self):
        self.L.orc2_init(self.h)

    def particles(self):
        n = int(self.L.orc2_num_particles(self.h))
        pos = np.empty((n, 2), np.float32)
        vel = np.empty((n, 2), np.float32)
        Cm = np.empty((n, 2, 2), np.float32)  # [particle, column, row]
        self.L.orc2_particles(self.h, _fptr(pos), _fptr(vel), _fptr(Cm), n, 0)
        return pos, vel, Cm

    def set_particles(self, pos, vel=None, Cm=None):
        pos = np.ascontiguousarray(pos, np.float32)
        v = None if vel is None else np.ascontiguousarray(vel, np.float32)
        c = None if Cm is None else np.ascontiguousarray(Cm, np.float32)
        self.L.orc2_particles(self.h, _fptr(pos), None if v is None else _fptr(v), None if c is None else _fptr(c),
                              len(pos), 1)
        self._keep = (pos, v, c)

    def set_boundary(self, node):
        prog = sdf_program(node)
        self.L.orc2_set_sdf(self.h, _fptr(prog), len(prog) // 12)

    def compute_solid_phi(self):
        self.L.orc2_compute_solid_phi(self.h)

    def clear_grid(self):
        self.L.orc2_clear_grid(self.h)

    def compute_liquid_phi(self):
        self.L.orc2_compute_liquid_phi(self.h)

    def p2g(self, order=1):
        self.L.orc2_p2g(self.h, order)

    def apply_gravity(self, dt):
        self.L.orc2_apply_gravity(self.h, float(dt))

    def solve_pressure(self, dt):
        st = SolveStats()
        self.L.orc2_solve_pressure(self.h, float(dt), C.byref(st))
        return st

    def velocity_update(self, dt):
        self.L.orc2_velocity_update(self.h, float(dt))

    def extrapolate(self):
        self.L.orc2_extrapolate(self.h)

    def g2p(self):
        self.L.orc2_g2p(self.h)

    def advect(self, dt):
        self.L.orc2_advect(self.h, float(dt))

    def correct_volume(self, dt):
        self.L.orc2_correct_volume(self.h, float(dt))

    def get_cfl(self):
        mv = C.c_float()
        return float(self.L.orc2_get_cfl(self.h, C.byref(mv))), float(mv.value)

    def divergence(self):
        mean, linf, cnt = C.c_double(), C.c_double(), C.c_int()
        self.L.orc2_divergence(self.h, C.byref(mean), C.byref(linf), C.byref(cnt))
        return mean.value, linf.value, cnt.value

    @property
    def time(self):
        return float(self.L.orc2_time(self.h))

    def set_time(self, t):
        self.L.orc2_set_time(self.h, float(t))

    def step(self, dt):
        st = SolveStats()
        self.L.orc2_step(self.h, float(dt), C.byref(st))
        return st
