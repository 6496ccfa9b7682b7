"""ctypes mirror of include/ewg_capi.h and a thin context wrapper.

The same wrapper drives any library that exports the ewg_* entry points under a
given symbol prefix; the package itself only ever loads libewg.so (the HIP build).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

EWG_OK = 0
EWG_NUMERICAL_ISSUE = 1
EWG_NOT_CONVERGED = 2
EWG_INVALID_ARGUMENT = -1
EWG_DEVICE_ERROR = -2
EWG_INVALID_STATE = -3

EWG_LAW_VAN_GENUCHTEN = 0
EWG_LAW_BROOKS_COREY = 1

EWG_BC_NO_FLOW = 0
EWG_BC_FREE_FLOW = 1
EWG_BC_MASS_RATE = 2

EWG_KERNEL_LINEARIZE = 0
EWG_KERNEL_SPMV = 1
EWG_KERNEL_ILU_APPLY = 2
EWG_KERNEL_ILU_FACTOR = 3
EWG_KERNEL_COPY = 5
EWG_KERNEL_COLD = 0x100  # or-ed into a kernel id: the caches hold other data before each launch


class GridDesc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double)]


class FluidDesc(C.Structure):
    _fields_ = [("rho", C.c_double * 2), ("mu", C.c_double * 2), ("gravity", C.c_double * 3)]


class MaterialDesc(C.Structure):
    _fields_ = [("law", C.c_int32), ("nregions", C.c_int32),
                ("region_params", C.POINTER(C.c_double))]


class Params(C.Structure):
    _fields_ = [("newton_tolerance", C.c_double),
                ("newton_max_error", C.c_double),
                ("newton_target_iterations", C.c_int32),
                ("newton_max_iterations", C.c_int32),
                ("linear_tolerance", C.c_double),
                ("linear_abs_tolerance", C.c_double),
                ("linear_max_error", C.c_double),
                ("linear_max_iterations", C.c_int32),
                ("enable_storage_cache", C.c_int32),
                ("recycle_first_iteration_storage", C.c_int32),
                ("max_time_step_divisions", C.c_int32),
                ("min_time_step_size", C.c_double),
                ("max_time_step_size", C.c_double),
                ("precond_fp32", C.c_int32),
                ("ilu_schedule", C.c_int32),
                ("kernel_events", C.c_int32),
                ("linearize_fp32", C.c_int32),
                ("warm_start", C.c_int32),
                ("local_linearizer", C.c_int32),
                ("numeric_difference_method", C.c_int32),
                ("operator_fp32", C.c_int32)]


class RestartClock(C.Structure):
    _fields_ = [("episode_idx", C.c_int32), ("episode_start_time", C.c_double),
                ("episode_length", C.c_double), ("start_time", C.c_double),
                ("time", C.c_double), ("time_step_idx", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("newton_iterations", C.c_int64),
                ("linear_iterations", C.c_int64),
                ("linear_solves", C.c_int64),
                ("time_steps", C.c_int64),
                ("failed_steps", C.c_int64),
                ("last_newton_iterations", C.c_int32),
                ("last_linear_iterations", C.c_int32),
                ("last_newton_error", C.c_double),
                ("time_linearize_s", C.c_double),
                ("time_solve_s", C.c_double),
                ("time_update_s", C.c_double),
                ("precond_retries", C.c_int64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def default_params(**overrides) -> Params:
    """Reference defaults (newtonmethod.hh:150-155, fvbasediscretization.hh:128-262,
    parallelbasebackend.hh:449-514, parallelbicgstabbackend.hh:49)."""
    p = Params()
    p.newton_tolerance = 1e-8
    p.newton_max_error = 1e100
    p.newton_target_iterations = 10
    p.newton_max_iterations = 18
    p.linear_tolerance = 1e-3
    p.linear_abs_tolerance = -1.0
    p.linear_max_error = 1e7
    p.linear_max_iterations = 1000
    p.enable_storage_cache = 1
    p.recycle_first_iteration_storage = 1
    p.max_time_step_divisions = 10
    p.min_time_step_size = 0.0
    p.max_time_step_size = float("inf")
    p.precond_fp32 = 0
    p.ilu_schedule = 0
    p.kernel_events = 0
    p.linearize_fp32 = 0
    p.warm_start = 0
    p.local_linearizer = 0
    p.numeric_difference_method = 1
    p.operator_fp32 = 0
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)


def _dptr(a):
    return None if a is None else a.ctypes.data_as(_D)


def _iptr(a):
    return None if a is None else a.ctypes.data_as(_I)


def bind(lib: C.CDLL, prefix: str) -> C.CDLL:
    """Declare the argument/return types of the <prefix>_* entry points."""
    vp = C.c_void_p

    def f(name, res, *args):
        fn = getattr(lib, f"{prefix}_{name}")
        fn.restype = res
        fn.argtypes = list(args)

    f("create", C.c_int, C.POINTER(GridDesc), C.POINTER(FluidDesc), C.POINTER(MaterialDesc),
      C.POINTER(Params), C.POINTER(vp))
    f("destroy", None, vp)
    f("last_error", C.c_char_p, vp)
    f("set_params", C.c_int, vp, C.POINTER(Params))
    f("num_cells", C.c_int64, vp)
    f("num_blocks", C.c_int64, vp)
    f("num_boundary_faces", C.c_int64, vp)
    f("set_static", C.c_int, vp, _D, _D, _I, _D)
    f("set_boundary", C.c_int, vp, _I, _D)
    f("set_solution", C.c_int, vp, C.c_int, _D)
    f("get_solution", C.c_int, vp, C.c_int, _D)
    f("advance_time_level", C.c_int, vp)
    f("get_pattern", C.c_int, vp, _I, _I)
    f("linearize", C.c_int, vp, C.c_double, C.c_int, _D, _D)
    f("set_system", C.c_int, vp, _D, _D)
    f("solve", C.c_int, vp, _D, _I)
    f("newton", C.c_int, vp, C.c_double, _I, _D)
    f("run", C.c_int, vp, C.c_double, C.c_double, C.c_double, _I)
    f("get_stats", C.c_int, vp, C.POINTER(Stats))
    return lib


class SolverError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"status {code}: {msg}")
        self.code = code


class Context:
    """Owns one <prefix>_ctx built from a problem description (ewg.problems.Problem)."""

    def __init__(self, lib: C.CDLL, prefix: str, problem, params: Params | None = None):
        self.lib = lib
        self.prefix = prefix
        self.problem = problem
        self._fn = lambda name: getattr(lib, f"{prefix}_{name}")
        self.params = params if params is not None else problem.params
        grid = GridDesc(problem.dim, problem.nx, problem.ny, problem.nz,
                        problem.lx, problem.ly, problem.lz)
        fluid = FluidDesc((C.c_double * 2)(*problem.rho), (C.c_double * 2)(*problem.mu),
                          (C.c_double * 3)(*problem.gravity))
        self._regions = np.ascontiguousarray(problem.region_params, dtype=np.float64).ravel()
        mat = MaterialDesc(problem.law, len(self._regions) // 4, _dptr(self._regions))
        h = C.c_void_p()
        rc = self._fn("create")(C.byref(grid), C.byref(fluid), C.byref(mat),
                               C.byref(self.params), C.byref(h))
        if rc != 0 or not h.value:
            raise SolverError(rc, f"{prefix}_create failed")
        self.h = h
        self.N = int(self._fn("num_cells")(h))
        self.nnzb = int(self._fn("num_blocks")(h))
        self.B = int(self._fn("num_boundary_faces")(h))
        self._check(self._fn("set_static")(
            h, _dptr(problem.K), _dptr(problem.phi), _iptr(problem.region),
            _dptr(problem.pe)))
        self._check(self._fn("set_boundary")(h, _iptr(problem.bc_type), _dptr(problem.bc_val)))
        self.set_solution(problem.x0, 0)
        self.set_solution(problem.x0, 1)

    # ------------------------------------------------------------------ helpers
    def _check(self, rc, allow=()):
        if rc != 0 and rc not in allow:
            raise SolverError(rc, self.last_error())
        return rc

    def last_error(self) -> str:
        s = self._fn("last_error")(self.h)
        return s.decode() if s else ""

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self._fn("destroy")(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ API
    def set_params(self, params: Params):
        self.params = params
        self._check(self._fn("set_params")(self.h, C.byref(params)))

    def set_solution(self, x, time_idx=0):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        assert x.size == 2 * self.N
        self._check(self._fn("set_solution")(self.h, time_idx, _dptr(x)))

    def solution(self, time_idx=0):
        x = np.empty(2 * self.N)
        self._check(self._fn("get_solution")(self.h, time_idx, _dptr(x)))
        return x

    def advance_time_level(self):
        self._check(self._fn("advance_time_level")(self.h))

    def pattern(self):
        rp = np.empty(self.N + 1, dtype=np.int32)
        ci = np.empty(self.nnzb, dtype=np.int32)
        self._check(self._fn("get_pattern")(self.h, _iptr(rp), _iptr(ci)))
        return rp, ci

    def linearize(self, dt, newton_iter=0, want=True):
        r = np.empty(2 * self.N) if want else None
        j = np.empty(4 * self.nnzb) if want else None
        self._check(self._fn("linearize")(self.h, dt, newton_iter, _dptr(r), _dptr(j)))
        return r, j

    def set_system(self, resid=None, jac=None):
        r = None if resid is None else np.ascontiguousarray(resid, dtype=np.float64)
        j = None if jac is None else np.ascontiguousarray(jac, dtype=np.float64)
        self._check(self._fn("set_system")(self.h, _dptr(r), _dptr(j)))

    def solve(self):
        dx = np.empty(2 * self.N)
        it = C.c_int32(0)
        rc = self._fn("solve")(self.h, _dptr(dx), C.byref(it))
        return rc, dx, it.value

    def newton(self, dt):
        it = C.c_int32(0)
        err = C.c_double(0)
        rc = self._fn("newton")(self.h, dt, C.byref(it), C.byref(err))
        return rc, it.value, err.value

    def run(self, t0, t_end, dt0):
        steps = C.c_int32(0)
        rc = self._fn("run")(self.h, t0, t_end, dt0, C.byref(steps))
        self._check(rc)
        return steps.value

    def stats(self):
        s = Stats()
        self._check(self._fn("get_stats")(self.h, C.byref(s)))
        return s.as_dict()


def package_dir() -> str:
    return os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
