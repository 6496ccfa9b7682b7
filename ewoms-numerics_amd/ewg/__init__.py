"""ewg -- Python view of the MI355X-native ECFV two-phase Newton step (libewg.so).

The compute path is the HIP library built from csrc/; this module only loads it and
passes host arrays through its C ABI. Importing load() on a machine without the built
library raises: there is no CPU fallback in the product.
"""
from __future__ import annotations

import ctypes as C
import os

from .capi import Context, Params, SolverError, Stats, bind, default_params  # noqa: F401
from . import problems  # noqa: F401

_LIB = None


def library_path() -> str:
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    return os.path.join(here, "lib", "libewg.so")


def load():
    """Load libewg.so (the HIP build); raises if it has not been built."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(f"libewg.so not built at {path}; run __graft_entry__.build()")
        lib = C.CDLL(path)
        bind(lib, "ewg")
        vp = C.c_void_p
        lib.ewg_time_kernel.restype = C.c_int
        lib.ewg_time_kernel.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double)]
        lib.ewg_reset_stats.restype = C.c_int
        lib.ewg_reset_stats.argtypes = [vp]
        lib.ewg_kernel_samples.restype = C.c_int
        lib.ewg_kernel_samples.argtypes = [vp, C.c_int, C.POINTER(C.c_int64),
                                           C.POINTER(C.c_double)]
        lib.ewg_device_info.restype = C.c_int
        _D = C.POINTER(C.c_double)
        from .capi import RestartClock
        lib.ewg_restart_file_name.restype = C.c_int
        lib.ewg_restart_file_name.argtypes = [C.c_char_p, C.c_char_p, C.c_double, C.c_int32,
                                              C.c_char_p, C.c_int32]
        lib.ewg_write_vtk.restype = C.c_int
        lib.ewg_write_vtk.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_double]
        for name in ("write_restart", "read_restart"):
            fn = getattr(lib, "ewg_" + name)
            fn.restype = C.c_int
            fn.argtypes = [vp, C.c_char_p, C.POINTER(RestartClock)]
        for name, args in (("get_preconditioner", [vp, _D]),
                           ("apply_preconditioner", [vp, _D, _D]),
                           ("apply_operator", [vp, _D, _D])):
            fn = getattr(lib, "ewg_" + name)
            fn.restype = C.c_int
            fn.argtypes = args
        lib.ewg_device_info.argtypes = [C.c_char_p, C.c_int]
        _LIB = lib
    return _LIB


class GpuContext(Context):
    """libewg context for a problem (see ewg.problems)."""

    def __init__(self, problem, params: Params | None = None):
        super().__init__(load(), "ewg", problem, params)

    def time_kernel(self, which: int, reps: int) -> float:
        ms = C.c_double(0.0)
        self._check(self.lib.ewg_time_kernel(self.h, which, reps, C.byref(ms)))
        return ms.value

    def reset_stats(self):
        self._check(self.lib.ewg_reset_stats(self.h))

    def preconditioner(self):
        """ILU(0) factors of the current Jacobian in BCRS order (ewg_get_preconditioner)."""
        import numpy as np
        lu = np.empty(4 * self.nnzb)
        self._check(self.lib.ewg_get_preconditioner(self.h, lu.ctypes.data_as(C.POINTER(C.c_double))))
        return lu

    def apply_preconditioner(self, d):
        import numpy as np
        d = np.ascontiguousarray(d, dtype=np.float64).ravel()
        v = np.empty(2 * self.N)
        P = C.POINTER(C.c_double)
        self._check(self.lib.ewg_apply_preconditioner(self.h, d.ctypes.data_as(P), v.ctypes.data_as(P)))
        return v

    def apply_operator(self, x):
        import numpy as np
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        y = np.empty(2 * self.N)
        P = C.POINTER(C.c_double)
        self._check(self.lib.ewg_apply_operator(self.h, x.ctypes.data_as(P), y.ctypes.data_as(P)))
        return y

    def write_restart(self, path, **clock):
        """Restart file of solution(0) with the given clock fields (ewg_write_restart)."""
        from .capi import RestartClock
        k = RestartClock(**clock)
        self._check(self.lib.ewg_write_restart(self.h, str(path).encode(), C.byref(k)))

    def read_restart(self, path):
        """Load a restart file (solution(0) = solution(1) = its state); returns its clock."""
        from .capi import RestartClock
        k = RestartClock()
        self._check(self.lib.ewg_read_restart(self.h, str(path).encode(), C.byref(k)))
        return {n: getattr(k, n) for n, _ in k._fields_}

    def write_vtk(self, output_dir, sim_name, t):
        """VTK output of solution(0) (ewg_write_vtk)."""
        self._check(self.lib.ewg_write_vtk(self.h, str(output_dir).encode(), sim_name.encode(), t))

    def kernel_samples(self, which: int):
        """(count, total_ms) of the event-timed launches of kernel `which` (needs
        params.kernel_events = 1)."""
        n = C.c_int64(0)
        ms = C.c_double(0.0)
        self._check(self.lib.ewg_kernel_samples(self.h, which, C.byref(n), C.byref(ms)))
        return n.value, ms.value


def restart_file_name(output_dir, sim_name, t, rank=0) -> str:
    """The reference's restart file name (Restart::restartFileName_)."""
    buf = C.create_string_buffer(4096)
    rc = load().ewg_restart_file_name(str(output_dir).encode(), sim_name.encode(), t, rank, buf,
                                      4096)
    if rc != 0:
        raise ValueError("restart file name too long")
    return buf.value.decode()


def device_info() -> str:
    buf = C.create_string_buffer(512)
    load().ewg_device_info(buf, 512)
    return buf.value.decode()
