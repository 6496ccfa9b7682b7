"""GPU: ewg_apply_preconditioner_noalloc (the preconditioner application that allocates
nothing; its sweeps' scratch is a fixed array per device) against ewg_apply_preconditioner
on the same factors and right-hand sides: bit for bit, on two calls in a row, the second
on a smaller grid, for the FP64 and FP32 tile-wavefront sweeps and the FP32 hyperplane
sweeps of a 2D grid. A grid beyond the scratch is rejected with EWG_INVALID_ARGUMENT.
Device buffers come from the HIP runtime libewg itself links (plumbing only)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bind():
    import ewg
    lib = ewg.load()
    fn = lib.ewg_apply_preconditioner_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    return fn


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime libewg.so is linked against."""

    def __init__(self):
        import os
        lib = None
        for name in ("libamdhip64.so.7", "libamdhip64.so", os.path.join("/opt/rocm/lib", "libamdhip64.so")):
            try:
                lib = C.CDLL(name)
                break
            except OSError:
                continue
        assert lib is not None, "libamdhip64.so not found"
        lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        lib.hipFree.argtypes = [C.c_void_p]
        lib.hipDeviceSynchronize.argtypes = []
        self.lib = lib

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def upload(self, dst, a):
        assert self.lib.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0

    def download(self, a, src):
        assert self.lib.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, a.nbytes, 2) == 0

    def free(self, p):
        self.lib.hipFree(p)


def _problem(shape):
    from ewg import problems
    if shape == "lens":
        return problems.lens_problem()
    nx, ny, nz = shape
    return problems.heterogeneous_3d(nx, ny, nz, seed=sum(shape), patch=4)


def _apply_noalloc(fn, g, d):
    """v = the no-allocation apply of d, through device buffers of the test's own."""
    hip = _Hip()
    n = d.nbytes
    d_dev, v_dev = hip.alloc(n), hip.alloc(n)
    try:
        hip.upload(d_dev, d)
        hip.upload(v_dev, np.full_like(d, np.nan))
        rc = fn(g.h, d_dev, v_dev)
        assert rc == 0, (rc, g.last_error())
        v = np.empty_like(d)
        hip.download(v, v_dev)
    finally:
        hip.free(d_dev)
        hip.free(v_dev)
    return v


def _check_pair(fn, shape, fp32, seed):
    import ewg
    from ewg import problems
    P = _problem(shape)
    print(f"no-alloc apply: shape {shape} fp32 {fp32}", flush=True)
    g = ewg.GpuContext(P, ewg.default_params(precond_fp32=fp32))
    g.set_solution(problems.perturbed_state(P, seed=seed, dS=0.2), 0)
    g.linearize(P.dt, 0, want=False)
    g.preconditioner()                       # factors of the current Jacobian
    d = np.random.default_rng(seed).standard_normal(2 * P.N)
    v_ref = g.apply_preconditioner(d)
    v = _apply_noalloc(fn, g, d)
    assert np.array_equal(v, v_ref), np.abs(v - v_ref).max()
    g.close()


@pytest.mark.parametrize("fp32", [0, 1])
def test_noalloc_apply_matches_apply_larger_then_smaller(fp32):
    fn = _bind()
    _check_pair(fn, (32, 24, 16), fp32, seed=3)
    _check_pair(fn, (20, 13, 7), fp32, seed=4)


def test_noalloc_apply_hyperplane_2d_larger_then_smaller():
    fn = _bind()
    _check_pair(fn, "lens", 1, seed=5)
    from ewg import problems
    import ewg
    P = problems.lens_problem(nx=12, ny=8)
    g = ewg.GpuContext(P, ewg.default_params(precond_fp32=1))
    g.linearize(P.dt, 0, want=False)
    g.preconditioner()
    d = np.random.default_rng(6).standard_normal(2 * P.N)
    v_ref = g.apply_preconditioner(d)
    assert np.array_equal(_apply_noalloc(fn, g, d), v_ref)


def test_noalloc_apply_rejects_a_grid_beyond_the_scratch():
    import ewg
    from ewg.capi import EWG_INVALID_ARGUMENT
    fn = _bind()
    P = _problem((72, 72, 512))      # tile edges 21.2 MB, intermediate 43.8 MB
    g = ewg.GpuContext(P, ewg.default_params())
    rc = fn(g.h, C.c_void_p(1), C.c_void_p(1))   # rejected before any device access
    assert rc == EWG_INVALID_ARGUMENT
    assert "too large" in g.last_error()
