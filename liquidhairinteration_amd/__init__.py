"""MI355X liquid step: Python binding of liblh.so (the C ABI in include/lh_api.h).

The classes mirror the reference's C++ surface so that code written against it
reads the same:

    scene = Scene("liquidhairinteration_amd/configs/pure_liquid_3d.json")   # Scene<3>(configFile)
    model = scene.GetLiquidModel()                                          # LiquidModel<3>&
    model.Init()
    dt = 0.01
    while running:                                                          # main.cpp:47-53
        model.PreStep(dt); model.Step(dt); model.PostStep(dt)
        dt = model.GetCFL()

Errors raise LHError (a RuntimeError, as the reference's std::runtime_error /
json::out_of_range).  All compute runs in the HIP library; there is no CPU
fallback: importing works without a GPU, but creating a Scene needs liblh.so
and a visible MI355X.
"""
import ctypes as C
import json
import os

import numpy as np

__all__ = ["Scene", "LiquidModel", "CGSolver", "Scene2D", "LiquidModel2D", "LHError", "library_path", "load_library", "FIELDS",
           "SolveStats", "StepStats", "config_path"]

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "liblh.so"

LH_OK = 0
LH_ERR_ARG = 6
ERROR_NAMES = {1: "config", 2: "out of memory", 3: "not converged", 4: "NaN detected", 5: "device", 6: "argument",
               7: "domain"}

FIELDS = {"vx": 0, "vy": 1, "vz": 2, "tempVx": 3, "tempVy": 4, "tempVz": 5, "liquidWx": 6, "liquidWy": 7,
          "liquidWz": 8, "pressure": 9, "liquidPhi": 10, "solidPhi": 11, "isValid": 12, "rhs": 13}


class LHError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s error (%d): %s" % (ERROR_NAMES.get(code, "unknown"), code, msg))
        self.code = code


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("outer_iterations", C.c_int),
                ("warm_start_used", C.c_int), ("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("rhs_norm", C.c_double), ("true_residual", C.c_double), ("solve_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StepStats(C.Structure):
    _fields_ = [("solve", SolveStats), ("divergence_mean", C.c_double), ("divergence_linf", C.c_double),
                ("fluid_cells", C.c_int), ("max_velocity", C.c_float), ("cfl_dt", C.c_float),
                ("sim_time", C.c_double), ("step_ms", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "solve"}
        d["solve"] = self.solve.as_dict()
        return d


def library_path():
    return os.path.join(HERE, _LIB_NAME)


def config_path(name):
    """Path of a config shipped in liquidhairinteration_amd/configs."""
    return os.path.join(HERE, "configs", name if name.endswith(".json") else name + ".json")


_lib = None


def load_library():
    """Load liblh.so (built by __graft_entry__.build() / `make -C liquidhairinteration_amd/csrc`)."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise LHError(5, "HIP library %s is not built; run __graft_entry__.build()" % path)
    L = C.CDLL(path)
    vp, fp, dp, ip = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int)
    sigs = {
        "lh_create": ([C.c_char_p, C.POINTER(vp)], C.c_int),
        "lh_create_from_string": ([C.c_char_p, C.POINTER(vp)], C.c_int),
        "lh_destroy": ([vp], None),
        "lh_last_error": ([vp], C.c_char_p),
        "lh_grid_info": ([vp, ip, fp, fp], C.c_int),
        "lh_field_shape": ([vp, C.c_int, ip], C.c_int),
        "lh_init": ([vp], C.c_int),
        "lh_pre_step": ([vp, C.c_float], C.c_int),
        "lh_step": ([vp, C.c_float], C.c_int),
        "lh_post_step": ([vp, C.c_float, C.POINTER(StepStats)], C.c_int),
        "lh_get_cfl": ([vp, fp], C.c_int),
        "lh_run_steps": ([vp, C.c_int, fp, C.POINTER(StepStats)], C.c_int),
        "lh_compute_solid_phi": ([vp], C.c_int),
        "lh_clear_grid": ([vp], C.c_int),
        "lh_compute_liquid_phi": ([vp], C.c_int),
        "lh_p2g": ([vp], C.c_int),
        "lh_apply_gravity": ([vp, C.c_float], C.c_int),
        "lh_solve_pressure": ([vp, C.c_float, C.c_float, C.POINTER(SolveStats)], C.c_int),
        "lh_velocity_update": ([vp, C.c_float], C.c_int),
        "lh_extrapolate": ([vp], C.c_int),
        "lh_g2p": ([vp], C.c_int),
        "lh_advect": ([vp, C.c_float], C.c_int),
        "lh_divergence": ([vp, dp, dp, ip], C.c_int),
        "lh_correct_volume": ([vp, C.c_float], C.c_int),
        "lh_num_particles": ([vp, C.POINTER(C.c_int64)], C.c_int),
        "lh_set_particles": ([vp, fp, fp, fp, C.c_int64], C.c_int),
        "lh_get_particles": ([vp, fp, fp, fp, C.c_int64], C.c_int),
        "lh_get_field": ([vp, C.c_int, vp, C.c_size_t], C.c_int),
        "lh_set_field": ([vp, C.c_int, vp, C.c_size_t], C.c_int),
        "lh_set_time": ([vp, C.c_double], C.c_int),
        "lh_profile_enable": ([vp, C.c_int], C.c_int),
        "lh_profile_read": ([vp, C.c_int, ip, dp], C.c_int),
        "lh_get_time": ([vp, dp], C.c_int),
        "lh_build_system": ([vp, C.c_float, C.c_float], C.c_int),
        "lh_resize_particles": ([vp, C.c_int64], C.c_int),
        "lh_divergence_noalloc": ([vp, vp, vp, vp, vp, vp, vp, ip, C.c_float, vp, vp], C.c_int),
        "lh_liquid_density": ([vp, fp], C.c_int),
        "lh_solve_rhs": ([vp, fp, fp, C.POINTER(SolveStats)], C.c_int),
        "lh_get_solution64": ([vp, dp, C.c_size_t], C.c_int),
        "lh_set_warm_start": ([vp, fp, C.c_size_t], C.c_int),
        "lh_solver_level_info": ([vp, C.c_int, ip, C.POINTER(C.c_int64)], C.c_int),
        "lh_solver_level_segments": ([vp, C.c_int, ip, C.POINTER(C.c_int64)], C.c_int),
        "lh_write_positions": ([vp, C.c_char_p, C.c_int], C.c_int),
        "lh_save": ([vp, C.c_char_p], C.c_int),
        "lh_set_boundary": ([vp, C.c_char_p], C.c_int),
        "lh_load": ([vp, C.c_char_p], C.c_int),
        "lh_load_positions_line": ([vp, C.c_char_p], C.c_int),
        "lh_warmstart_init": ([vp, C.c_uint64], C.c_int),
        "lh_warmstart_load": ([vp, C.c_char_p], C.c_int),
        "lh_warmstart_save": ([vp, C.c_char_p], C.c_int),
        "lh_warmstart_enable": ([vp, C.c_int], C.c_int),
        "lh_warmstart_num_params": ([vp, C.POINTER(C.c_int64)], C.c_int),
        "lh_warmstart_get_params": ([vp, fp, C.c_int64], C.c_int),
        "lh_warmstart_set_params": ([vp, fp, C.c_int64], C.c_int),
        "lh_warmstart_set_target_scale": ([vp, C.c_float], C.c_int),
        "lh_warmstart_forward": ([vp, ip, fp, fp], C.c_int),
        "lh_warmstart_step": ([vp, ip, fp, fp, fp, C.c_float, fp, fp], C.c_int),
        "lh_warmstart_train": ([vp, C.c_float, fp, fp], C.c_int),
        "lh_warmstart_set_loss": ([vp, C.c_int], C.c_int),
        "lh_warmstart_sample": ([vp, fp, fp, fp], C.c_int),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _fptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


class _Handle:
    def __init__(self, config):
        self.L = load_library()
        h = C.c_void_p()
        if isinstance(config, dict):
            import json
            rc = self.L.lh_create_from_string(json.dumps(config).encode(), C.byref(h))
        else:
            rc = self.L.lh_create(os.fspath(config).encode(), C.byref(h))
        self.h = h
        if rc != LH_OK:
            msg = self.L.lh_last_error(h).decode() if h.value else "create failed"
            if h.value:
                self.L.lh_destroy(h)
            self.h = C.c_void_p()
            raise LHError(rc, msg)

    def check(self, rc):
        if rc != LH_OK:
            raise LHError(rc, self.L.lh_last_error(self.h).decode())
        return rc

    def close(self):
        if self.h and self.h.value:
            self.L.lh_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ParticleSystem:
    """Host view of ParticleSystem<3> (ParticleSystem.h:24-61): copies in and out of HBM."""

    def __init__(self, handle):
        self._h = handle

    def NumParticles(self):
        n = C.c_int64()
        self._h.check(self._h.L.lh_num_particles(self._h.h, C.byref(n)))
        return int(n.value)

    def get(self):
        n = self.NumParticles()
        pos = np.empty((n, 3), np.float32)
        vel = np.empty((n, 3), np.float32)
        Cm = np.empty((n, 9), np.float32)
        self._h.check(self._h.L.lh_get_particles(self._h.h, _fptr(pos), _fptr(vel), _fptr(Cm), n))
        return pos, vel, Cm

    def set(self, pos, vel=None, Cm=None):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        n = pos.shape[0]
        vel = None if vel is None else np.ascontiguousarray(vel, np.float32).reshape(n, 3)
        Cm = None if Cm is None else np.ascontiguousarray(Cm, np.float32).reshape(n, 9)
        self._h.check(self._h.L.lh_set_particles(self._h.h, _fptr(pos), _fptr(vel), _fptr(Cm), n))

    def GetPositions(self):
        return self.get()[0]

    def GetVelocities(self):
        return self.get()[1]

    def GetAffineMatrices(self):
        """(n, 3, 3) with [i, row, col] (the Eigen column-major matrix of each particle)."""
        Cm = self.get()[2]
        return Cm.reshape(-1, 3, 3).transpose(0, 2, 1)

    def WriteToFile(self, filename, isAppend=False):
        """ParticleSystem::WriteToFile (ParticleSystem.cpp:66-84): count line, then 'x,y,z,...,'."""
        self._h.check(self._h.L.lh_write_positions(self._h.h, os.fspath(filename).encode(), 1 if isAppend else 0))

    def LoadFromLine(self, line):
        """ParticleSystem::LoadFromLine (ParticleSystem.cpp:86-101), first token skipped as in the reference."""
        self._h.check(self._h.L.lh_load_positions_line(self._h.h, line.encode()))

    def Resize(self, size):
        """ParticleSystem::Resize (ParticleSystem.cpp:10-21): the first min(size, n) particles keep their
        state, new ones start at zero with the default radius."""
        self._h.check(self._h.L.lh_resize_particles(self._h.h, int(size)))


class CGSolver:
    """CGSolver<3> (PressureSolver.h:81-145) bound to a scene's grid."""

    def __init__(self, handle):
        self._h = handle
        self.last_stats = None

    def _density(self, liquidDensity):
        if liquidDensity is not None:
            return float(liquidDensity)
        rho = C.c_float()
        self._h.check(self._h.L.lh_liquid_density(self._h.h, C.byref(rho)))
        return rho.value

    def Solve(self, liquidDensity=None, dt=0.01):
        """Solve(liquidDensity, dt) (PressureSolver.h:104): BuildRHS, BuildLHS(rho, dt), MG-PCG;
        writes grid.pressure.  liquidDensity=None takes the scene's density, which is what
        LiquidModel<3>::SolvePressure passes (LiquidModel.cpp:830-831)."""
        st = SolveStats()
        rc = self._h.L.lh_solve_pressure(self._h.h, self._density(liquidDensity), float(dt), C.byref(st))
        self.last_stats = st.as_dict()
        if rc not in (LH_OK, 3):
            self._h.check(rc)
        return self.last_stats

    # split access: BuildRHS + BuildLHS, then solves on that operator
    def BuildSystem(self, dt, liquidDensity=None):
        """BuildRHS + BuildLHS(liquidDensity, dt) on every level (PressureSolver.h:97-99)"""
        self._h.check(self._h.L.lh_build_system(self._h.h, self._density(liquidDensity), float(dt)))

    def SolveRHS(self, b=None, x0=None):
        """Solve A x = b (b=None: the built RHS) from x0 (optional); x -> pressure. Returns stats."""
        b = None if b is None else np.ascontiguousarray(b, np.float32)
        x0 = None if x0 is None else np.ascontiguousarray(x0, np.float32)
        st = SolveStats()
        rc = self._h.L.lh_solve_rhs(self._h.h, _fptr(b), _fptr(x0), C.byref(st))
        self.last_stats = st.as_dict()
        if rc not in (LH_OK, 3):
            self._h.check(rc)
        return self.last_stats

    def GetSolution64(self, shape):
        x = np.empty(shape, np.float64)
        self._h.check(self._h.L.lh_get_solution64(self._h.h, x.ctypes.data_as(C.POINTER(C.c_double)), x.size))
        return x

    def LevelInfo(self, level):
        """(dims, active z-rows) of multigrid level `level` of the built operator"""
        dims = (C.c_int * 3)()
        rows = C.c_int64()
        self._h.check(self._h.L.lh_solver_level_info(self._h.h, int(level), dims, C.byref(rows)))
        return tuple(dims), int(rows.value)

    def LevelSegments(self, level):
        """(cells per z-segment, segments holding liquid) of level `level`'s colour-pass work list;
        (0, 0) when the level keeps no segment list"""
        cells = C.c_int()
        n = C.c_int64()
        self._h.check(self._h.L.lh_solver_level_segments(self._h.h, int(level), C.byref(cells), C.byref(n)))
        return int(cells.value), int(n.value)

    def SetWarmStart(self, x0):
        if x0 is None:
            self._h.check(self._h.L.lh_set_warm_start(self._h.h, None, 0))
        else:
            x0 = np.ascontiguousarray(x0, np.float32)
            self._h.check(self._h.L.lh_set_warm_start(self._h.h, _fptr(x0), x0.size))


class WarmStart:
    """Learned initial guess of the pressure solve (no reference counterpart: the
    reference's CGSolver<3>::Init starts from x = 0, PressureSolver.cpp:666-680).

    A 3D convolutional encoder-decoder (16/32/64/96 channels, 3x3x3 kernels, BF16,
    FP32 accumulation, trained and run by the library's own HIP kernels).  Arrays are
    channels-last numpy arrays of the grid shape: inputs [d0, d1, d2, 2], outputs
    [d0, d1, d2].
    """

    def __init__(self, handle):
        self._h = handle

    def init(self, seed=0):
        self._h.check(self._h.L.lh_warmstart_init(self._h.h, int(seed)))

    def load(self, path):
        self._h.check(self._h.L.lh_warmstart_load(self._h.h, os.fspath(path).encode()))

    def save(self, path):
        self._h.check(self._h.L.lh_warmstart_save(self._h.h, os.fspath(path).encode()))

    def enable(self, on=True):
        """use the network's guess in every following pressure solve"""
        self._h.check(self._h.L.lh_warmstart_enable(self._h.h, 1 if on else 0))

    def num_params(self):
        n = C.c_int64()
        self._h.check(self._h.L.lh_warmstart_num_params(self._h.h, C.byref(n)))
        return int(n.value)

    def get_params(self):
        w = np.empty(self.num_params(), np.float32)
        self._h.check(self._h.L.lh_warmstart_get_params(self._h.h, _fptr(w), w.size))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._h.check(self._h.L.lh_warmstart_set_params(self._h.h, _fptr(w), w.size))

    def set_target_scale(self, s):
        self._h.check(self._h.L.lh_warmstart_set_target_scale(self._h.h, float(s)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        dims = (C.c_int * 3)(*x.shape[:3])
        y = np.empty(x.shape[:3], np.float32)
        self._h.check(self._h.L.lh_warmstart_forward(self._h.h, dims, _fptr(x), _fptr(y)))
        return y

    def step(self, x, target, mask, lr=0.0, want_grad=True):
        """loss and gradient for (x, target, mask); with lr > 0 also one Adam step"""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        m = np.ascontiguousarray(mask, np.float32)
        dims = (C.c_int * 3)(*x.shape[:3])
        g = np.empty(self.num_params(), np.float32) if want_grad else None
        loss = C.c_float()
        self._h.check(self._h.L.lh_warmstart_step(self._h.h, dims, _fptr(x), _fptr(t), _fptr(m), float(lr), _fptr(g),
                                                  C.byref(loss)))
        return float(loss.value), g

    LOSSES = {"mse": 0, "energy": 1}

    def set_loss(self, mode):
        """'mse' (p / rms(p) target) or 'energy' (A-norm error of the energy-scaled guess)"""
        self._h.check(self._h.L.lh_warmstart_set_loss(self._h.h, self.LOSSES.get(mode, mode)))

    def train(self, lr, want_grad=False):
        """loss (and gradient) on the last built system and its solved pressure; with lr > 0
        also one Adam step.  Returns the loss, or (loss, grad) with want_grad."""
        loss = C.c_float()
        g = np.empty(self.num_params(), np.float32) if want_grad else None
        self._h.check(self._h.L.lh_warmstart_train(self._h.h, float(lr), _fptr(g), C.byref(loss)))
        return (float(loss.value), g) if want_grad else float(loss.value)

    def sample(self, shape):
        """(input [.., 2], target, mask) of the last built system"""
        x = np.empty(tuple(shape) + (2,), np.float32)
        t = np.empty(shape, np.float32)
        m = np.empty(shape, np.float32)
        self._h.check(self._h.L.lh_warmstart_sample(self._h.h, _fptr(x), _fptr(t), _fptr(m)))
        return x, t, m


class LiquidModel:
    """LiquidModel<3> (LiquidModel.h:11-82) on the GPU."""

    def __init__(self, handle):
        self._h = handle
        self._ps = ParticleSystem(handle)
        self._solver = CGSolver(handle)
        self._warm = WarmStart(handle)
        self.last_step = None

    def GetWarmStart(self):
        return self._warm

    # LiquidModel.cpp:506-537
    def Init(self):
        self._h.check(self._h.L.lh_init(self._h.h))

    def PreStep(self, dt):
        self._h.check(self._h.L.lh_pre_step(self._h.h, float(dt)))

    def Step(self, dt):
        rc = self._h.L.lh_step(self._h.h, float(dt))
        if rc not in (LH_OK, 3):
            self._h.check(rc)

    def PostStep(self, dt):
        st = StepStats()
        self._h.check(self._h.L.lh_post_step(self._h.h, float(dt), C.byref(st)))
        self.last_step = st.as_dict()
        return self.last_step

    def GetCFL(self):
        v = C.c_float()
        self._h.check(self._h.L.lh_get_cfl(self._h.h, C.byref(v)))
        return float(v.value)

    def SetBoundary(self, node):
        """LiquidModel<3>::SetBoundary (LiquidModel.h:57) with a config "Boundary" node (dict or
        JSON text); call ComputeSolidPhi afterwards (each step for a moving solid)"""
        text = node if isinstance(node, str) else json.dumps(node)
        self._h.check(self._h.L.lh_set_boundary(self._h.h, text.encode()))

    def SaveState(self, path):
        """snapshot of what the next step depends on (particles, isValid, time, CorrectVolume
        counters, last max |v|); no reference counterpart"""
        self._h.check(self._h.L.lh_save(self._h.h, os.fspath(path).encode()))

    def LoadState(self, path):
        """restore a SaveState snapshot into a model built from the same configuration"""
        self._h.check(self._h.L.lh_load(self._h.h, os.fspath(path).encode()))

    def RunSteps(self, nsteps, dt):
        """nsteps x {PreStep, Step, PostStep, dt = GetCFL()} in one native call; returns (next dt, stats)."""
        d = C.c_float(float(dt))
        st = StepStats()
        self._h.check(self._h.L.lh_run_steps(self._h.h, int(nsteps), C.byref(d), C.byref(st)))
        return float(d.value), st.as_dict()

    # public stage methods, LiquidModel.h:25-47
    def ComputeLiquidPhi(self):
        self._h.check(self._h.L.lh_compute_liquid_phi(self._h.h))

    def ComputeSolidPhi(self):
        self._h.check(self._h.L.lh_compute_solid_phi(self._h.h))

    def ClearGrid(self):
        self._h.check(self._h.L.lh_clear_grid(self._h.h))

    def Particle2Grid(self):
        self._h.check(self._h.L.lh_p2g(self._h.h))

    def Grid2Particle(self):
        self._h.check(self._h.L.lh_g2p(self._h.h))

    def ApplyGravity(self, dt):
        self._h.check(self._h.L.lh_apply_gravity(self._h.h, float(dt)))

    def SolvePressure(self, dt):
        # LiquidModel<3>::SolvePressure: Solve(scene density, dt), then the velocity update
        st = self._solver.Solve(None, dt)
        self._h.check(self._h.L.lh_velocity_update(self._h.h, float(dt)))
        return st

    def Projection(self, dt):
        return self.SolvePressure(dt)

    def VelocityUpdate(self, dt):
        self._h.check(self._h.L.lh_velocity_update(self._h.h, float(dt)))

    def Extrapolate(self):
        self._h.check(self._h.L.lh_extrapolate(self._h.h))

    def AdvectParticle(self, dt):
        self._h.check(self._h.L.lh_advect(self._h.h, float(dt)))

    def CorrectVolume(self, dt):
        """LiquidModel<3>::CorrectVolume (LiquidModel.cpp:1085-1132)"""
        self._h.check(self._h.L.lh_correct_volume(self._h.h, float(dt)))

    def ComputeDivergence(self):
        """Mean weighted divergence over liquid cells (LiquidModel.cpp:1065-1083); also returns L-inf and count."""
        m, l, c = C.c_double(), C.c_double(), C.c_int()
        self._h.check(self._h.L.lh_divergence(self._h.h, C.byref(m), C.byref(l), C.byref(c)))
        return m.value, l.value, c.value

    def GetParticleSystem(self):
        return self._ps

    def GetSolver(self):
        return self._solver

    # grid access (Grid<3> public members)
    def field_shape(self, name):
        f = FIELDS[name] if isinstance(name, str) else int(name)
        s = (C.c_int * 3)()
        self._h.check(self._h.L.lh_field_shape(self._h.h, f, s))
        return tuple(s)

    def get_field(self, name):
        f = FIELDS[name] if isinstance(name, str) else int(name)
        shp = self.field_shape(f)
        a = np.empty(shp, np.uint8 if f == FIELDS["isValid"] else np.float32)
        self._h.check(self._h.L.lh_get_field(self._h.h, f, a.ctypes.data, a.size))
        return a

    def set_field(self, name, arr):
        f = FIELDS[name] if isinstance(name, str) else int(name)
        shp = self.field_shape(f)
        a = np.ascontiguousarray(arr, np.uint8 if f == FIELDS["isValid"] else np.float32).reshape(shp)
        self._h.check(self._h.L.lh_set_field(self._h.h, f, a.ctypes.data, a.size))

    # live kernel timing (HIP events on the library's stream)
    PROF_KERNELS = {"smooth_l0": 0, "p2g": 1, "g2p": 2, "liquid_phi": 3, "spmv": 4, "nn_fwd": 5, "nn_bwd": 6,
                    "conv_d0": 7, "smooth_l1": 8, "smooth_l2": 9}

    def profile_enable(self, on=True):
        self._h.check(self._h.L.lh_profile_enable(self._h.h, 1 if on else 0))

    def profile_read(self, kernel):
        k = self.PROF_KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
        n, ms = C.c_int(), C.c_double()
        self._h.check(self._h.L.lh_profile_read(self._h.h, k, C.byref(n), C.byref(ms)))
        return int(n.value), float(ms.value)

    @property
    def time(self):
        t = C.c_double()
        self._h.check(self._h.L.lh_get_time(self._h.h, C.byref(t)))
        return t.value

    @time.setter
    def time(self, t):
        self._h.check(self._h.L.lh_set_time(self._h.h, float(t)))


class Scene:
    """Scene<3>(configFile) (Scene.h:14-48): parses the JSON config and owns the liquid model."""

    def __init__(self, config):
        self._h = _Handle(config)
        dims = (C.c_int * 3)()
        mn = (C.c_float * 3)()
        dx = (C.c_float * 3)()
        self._h.check(self._h.L.lh_grid_info(self._h.h, dims, mn, dx))
        self.grid_dimension = tuple(dims)
        self.min_coord = np.array(list(mn), np.float32)
        self.dx = np.array(list(dx), np.float32)
        self._model = LiquidModel(self._h)

    def GetLiquidModel(self):
        return self._model

    def GetGridDimension(self):
        return self.grid_dimension

    def GetLiquidDensity(self):
        """Scene<3>::GetLiquidDensity (Scene.h:23)"""
        return self._model.GetSolver()._density(None)

    def PlaySimulation(self, filename, on_frame=None):
        """Scene<3>::PlaySimulation (Scene.cpp:26-45) without the renderer: the first line's count
        resizes the particle system, then every following line goes through LoadFromLine (count
        lines too, as in the reference: they change nothing past the skipped first token) and
        on_frame(particle_system) is called where the reference renders.  Returns the number of
        lines replayed."""
        ps = self._model.GetParticleSystem()
        with open(filename) as f:
            first = f.readline()
            if not first:
                raise LHError(LH_ERR_ARG, "empty trajectory file %s" % filename)
            ps.Resize(int(first.strip()))      # std::stoul(line)
            n = 0
            for line in f:
                ps.LoadFromLine(line.rstrip("\n"))
                if on_frame is not None:
                    on_frame(ps)
                n += 1
        return n

    def GetMinCoord(self):
        return self.min_coord

    def close(self):
        self._h.close()


from .model2d import LiquidModel2D, Scene2D  # noqa: E402  (the 2D instantiation, LiquidModel<2>)
