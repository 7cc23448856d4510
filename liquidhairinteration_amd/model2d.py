"""Scene<2> / LiquidModel<2> / CGSolver<2> over the lh2_* C ABI (include/lh_api.h).

The reference's 2D instantiation of the liquid model (LiquidModel.cpp:10-474,
PressureSolver.cpp:9-418), run on the GPU.  Same method names as the 3D facade; particle
arrays are (n, 2) positions and velocities and (n, 4) affine matrices (2x2, column-major).
"""
import ctypes as C
import json
import os

import numpy as np

LH_OK = 0
FIELDS = {"vx": 0, "vy": 1, "tempVx": 3, "tempVy": 4, "liquidWx": 6, "liquidWy": 7, "pressure": 9,
          "liquidPhi": 10, "solidPhi": 11, "isValid": 12, "rhs": 13}

_bound = False


def _lib():
    from . import LHError, SolveStats, load_library  # noqa: F401

    global _bound
    L = load_library()
    if _bound:
        return L
    vp, fp, dp, ip = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int)
    st = C.POINTER(SolveStats)
    sigs = {
        "lh2_create": [C.c_char_p, C.POINTER(vp)], "lh2_create_from_string": [C.c_char_p, C.POINTER(vp)],
        "lh2_init": [vp], "lh2_pre_step": [vp, C.c_float], "lh2_step": [vp, C.c_float, st],
        "lh2_get_cfl": [vp, fp], "lh2_run_steps": [vp, C.c_int, fp, st], "lh2_clear_grid": [vp],
        "lh2_compute_liquid_phi": [vp], "lh2_compute_solid_phi": [vp], "lh2_set_boundary": [vp, C.c_char_p],
        "lh2_p2g": [vp], "lh2_apply_gravity": [vp, C.c_float], "lh2_solve_pressure": [vp, C.c_float, st],
        "lh2_velocity_update": [vp, C.c_float], "lh2_extrapolate": [vp], "lh2_g2p": [vp],
        "lh2_advect": [vp, C.c_float], "lh2_correct_volume": [vp, C.c_float], "lh2_divergence": [vp, dp, dp, ip],
        "lh2_num_particles": [vp, C.POINTER(C.c_int64)], "lh2_set_particles": [vp, fp, fp, fp, C.c_int64],
        "lh2_get_particles": [vp, fp, fp, fp, C.c_int64], "lh2_get_field": [vp, C.c_int, vp, C.c_size_t],
        "lh2_set_field": [vp, C.c_int, vp, C.c_size_t], "lh2_set_time": [vp, C.c_double],
        "lh2_get_time": [vp, dp],
    }
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.lh2_destroy.argtypes = [vp]
    L.lh2_destroy.restype = None
    L.lh2_last_error.argtypes = [vp]
    L.lh2_last_error.restype = C.c_char_p
    _bound = True
    return L


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


class Scene2D:
    """Scene<2> (Scene.h:13-47): owns the 2D liquid model built from a config path or dict"""

    def __init__(self, config):
        from . import LHError

        self.L = _lib()
        h = C.c_void_p()
        if isinstance(config, dict):
            rc = self.L.lh2_create_from_string(json.dumps(config).encode(), C.byref(h))
        else:
            rc = self.L.lh2_create(os.fspath(config).encode(), C.byref(h))
        if rc != LH_OK:
            msg = self.L.lh2_last_error(h).decode() if h.value else "create failed"
            if h.value:
                self.L.lh2_destroy(h)
            raise LHError(rc, msg)
        self.h = h
        cfg = config if isinstance(config, dict) else json.load(open(config))
        self.N = [int(v) for v in cfg["gridDimension"]]
        self._model = LiquidModel2D(self)

    def check(self, rc):
        from . import LHError

        if rc != LH_OK:
            raise LHError(rc, self.L.lh2_last_error(self.h).decode())

    def GetLiquidModel(self):
        return self._model

    def GetGridDimension(self):
        return tuple(self.N)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.lh2_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ParticleSystem2D:
    def __init__(self, scene):
        self._s = scene

    def NumParticles(self):
        n = C.c_int64()
        self._s.check(self._s.L.lh2_num_particles(self._s.h, C.byref(n)))
        return int(n.value)

    def get(self):
        n = self.NumParticles()
        pos, vel, Cm = np.empty((n, 2), np.float32), np.empty((n, 2), np.float32), np.empty((n, 4), np.float32)
        self._s.check(self._s.L.lh2_get_particles(self._s.h, _fp(pos), _fp(vel), _fp(Cm), n))
        return pos, vel, Cm

    def set(self, pos, vel=None, Cm=None):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 2)
        n = pos.shape[0]
        vel = None if vel is None else np.ascontiguousarray(vel, np.float32).reshape(n, 2)
        Cm = None if Cm is None else np.ascontiguousarray(Cm, np.float32).reshape(n, 4)
        self._s.check(self._s.L.lh2_set_particles(self._s.h, _fp(pos), _fp(vel), _fp(Cm), n))


class LiquidModel2D:
    """LiquidModel<2> (LiquidModel.h:11-61)"""

    def __init__(self, scene):
        self._s = scene
        self._ps = ParticleSystem2D(scene)

    def _c(self, rc):
        self._s.check(rc)

    def Init(self):
        self._c(self._s.L.lh2_init(self._s.h))

    def PreStep(self, dt):
        self._c(self._s.L.lh2_pre_step(self._s.h, float(dt)))

    def Step(self, dt):
        from . import SolveStats

        st = SolveStats()
        self._c(self._s.L.lh2_step(self._s.h, float(dt), C.byref(st)))
        return st.as_dict()

    def PostStep(self, dt):
        return None

    def GetCFL(self):
        v = C.c_float()
        self._c(self._s.L.lh2_get_cfl(self._s.h, C.byref(v)))
        return float(v.value)

    def RunSteps(self, n, dt):
        from . import SolveStats

        d = C.c_float(float(dt))
        st = SolveStats()
        self._c(self._s.L.lh2_run_steps(self._s.h, int(n), C.byref(d), C.byref(st)))
        return float(d.value), st.as_dict()

    def GetParticleSystem(self):
        return self._ps

    def ComputeLiquidPhi(self):
        self._c(self._s.L.lh2_compute_liquid_phi(self._s.h))

    def ComputeSolidPhi(self):
        self._c(self._s.L.lh2_compute_solid_phi(self._s.h))

    def SetBoundary(self, node):
        text = node if isinstance(node, str) else json.dumps(node)
        self._c(self._s.L.lh2_set_boundary(self._s.h, text.encode()))

    def ClearGrid(self):
        self._c(self._s.L.lh2_clear_grid(self._s.h))

    def Particle2Grid(self):
        self._c(self._s.L.lh2_p2g(self._s.h))

    def ApplyGravity(self, dt):
        self._c(self._s.L.lh2_apply_gravity(self._s.h, float(dt)))

    def SolvePressure(self, dt):
        """CGSolver<2>::Solve, then the velocity update"""
        st = self.Solve(dt)
        self.VelocityUpdate(dt)
        return st

    def Solve(self, dt):
        from . import SolveStats

        st = SolveStats()
        self._c(self._s.L.lh2_solve_pressure(self._s.h, float(dt), C.byref(st)))
        return st.as_dict()

    def VelocityUpdate(self, dt):
        self._c(self._s.L.lh2_velocity_update(self._s.h, float(dt)))

    def Extrapolate(self):
        self._c(self._s.L.lh2_extrapolate(self._s.h))

    def Grid2Particle(self):
        self._c(self._s.L.lh2_g2p(self._s.h))

    def AdvectParticle(self, dt):
        self._c(self._s.L.lh2_advect(self._s.h, float(dt)))

    def CorrectVolume(self, dt):
        self._c(self._s.L.lh2_correct_volume(self._s.h, float(dt)))

    def ComputeDivergence(self):
        mean, linf, cnt = C.c_double(), C.c_double(), C.c_int()
        self._c(self._s.L.lh2_divergence(self._s.h, C.byref(mean), C.byref(linf), C.byref(cnt)))
        return mean.value, linf.value, cnt.value

    def field_shape(self, name):
        N = self._s.N
        f = FIELDS[name]
        return {0: (N[0] + 1, N[1]), 1: (N[0], N[1] + 1), 3: (N[0] + 1, N[1]), 4: (N[0], N[1] + 1),
                6: (N[0] + 1, N[1]), 7: (N[0], N[1] + 1), 9: tuple(N), 10: tuple(N), 11: (N[0] + 1, N[1] + 1),
                12: (N[0] + 1, N[1] + 1), 13: tuple(N)}[f]

    def get_field(self, name):
        a = np.empty(self.field_shape(name), np.uint8 if name == "isValid" else np.float32)
        self._c(self._s.L.lh2_get_field(self._s.h, FIELDS[name], a.ctypes.data, a.size))
        return a

    def set_field(self, name, arr):
        a = np.ascontiguousarray(arr, np.uint8 if name == "isValid" else np.float32).reshape(self.field_shape(name))
        self._c(self._s.L.lh2_set_field(self._s.h, FIELDS[name], a.ctypes.data, a.size))

    @property
    def time(self):
        t = C.c_double()
        self._c(self._s.L.lh2_get_time(self._s.h, C.byref(t)))
        return t.value

    @time.setter
    def time(self, t):
        self._c(self._s.L.lh2_set_time(self._s.h, float(t)))
