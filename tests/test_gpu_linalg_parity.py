"""Direct GPU parity of the linear-algebra kernels against the oracle, one operation at a
time, through the C ABI: the operator apply (one SpMV), the ILU(0) factors and one
preconditioner apply, on the GPU's own Jacobian (so each check isolates one kernel).

Bars: FP64 results within 1e-12 relative (entry-wise with a 1e-13 * max floor); the FP32
factors are the FP64 factors rounded to float (within one float rounding), and the FP32
apply equals the oracle's float apply of those factors within 1e-6 relative. The FP32
operator (params.operator_fp32) equals the oracle's FP64 product with the Jacobian's
blocks rounded to float within 1e-12.
Shapes include partial 8x8 tiles (x, y not multiples of 8), thin grids (one z-layer,
three cells across) and the 2D lens grid.
"""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

SHAPES = [(20, 13, 7), (32, 24, 16), (37, 21, 5), (17, 4, 1), (3, 11, 2)]


def _system(P, params, seed):
    """GPU and oracle contexts of P, the GPU linearized at a perturbed state; returns the
    GPU's Jacobian in BCRS order."""
    import ewg
    from ewg import problems
    from oracle.pyoracle import OracleContext
    g = ewg.GpuContext(P, params)
    o = OracleContext(P, params)
    x = problems.perturbed_state(P, seed=seed, dS=0.2)
    g.set_solution(x, 0)
    _, J = g.linearize(P.dt, 0)
    rp_g, ci_g = g.pattern()
    rp_o, ci_o = o.pattern()
    assert np.array_equal(rp_g, rp_o) and np.array_equal(ci_g, ci_o)
    return g, o, J


def _problem(shape):
    from ewg import problems
    if shape == "lens":
        return problems.lens_problem()
    nx, ny, nz = shape
    return problems.heterogeneous_3d(nx, ny, nz, seed=sum(shape), patch=4)


@pytest.mark.parametrize("shape", SHAPES + ["lens"])
def test_operator_apply_matches_oracle(shape):
    import ewg
    P = _problem(shape)
    print(f"spmv shape {shape}", flush=True)
    g, o, J = _system(P, ewg.default_params(), seed=3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(2 * P.N) * np.tile([1e4, 0.1], P.N)
    yg = g.apply_operator(x)
    yo = o.spmv(J, x)
    print("bitwise:", np.array_equal(yg, yo), flush=True)
    assert rel_err(yg, yo) < 1e-12


@pytest.mark.parametrize("shape", SHAPES + ["lens"])
@pytest.mark.parametrize("schedule", [0, 1])
def test_ilu0_factors_and_apply_match_oracle(shape, schedule):
    import ewg
    if shape == "lens" and schedule == 0:
        pytest.skip("2D grids run the hyperplane schedule only")
    P = _problem(shape)
    print(f"ilu0 fp64 shape {shape} schedule {schedule}", flush=True)
    g, o, J = _system(P, ewg.default_params(ilu_schedule=schedule), seed=5)
    LUg = g.preconditioner()
    LUo = o.ilu0_factor(J)
    print("factors bitwise:", np.array_equal(LUg, LUo), flush=True)
    assert rel_err(LUg, LUo) < 1e-12
    rng = np.random.default_rng(6)
    d = rng.standard_normal(2 * P.N)
    vg = g.apply_preconditioner(d)
    vo = o.ilu0_apply(LUo, d)
    print("apply bitwise:", np.array_equal(vg, vo), flush=True)
    assert rel_err(vg, vo) < 1e-12
    # a second apply with the same factors and another right-hand side (the sweeps'
    # tickets and edge buffers are reused)
    d2 = rng.standard_normal(2 * P.N)
    assert rel_err(g.apply_preconditioner(d2), o.ilu0_apply(LUo, d2)) < 1e-12


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("schedule", [0, 1])
def test_fp32_preconditioner_matches_float_oracle(shape, schedule):
    import ewg
    P = _problem(shape)
    print(f"ilu0 fp32 shape {shape} schedule {schedule}", flush=True)
    g, o, J = _system(P, ewg.default_params(ilu_schedule=schedule, precond_fp32=1), seed=7)
    LUg = g.preconditioner()
    LUo = o.ilu0_factor(J)
    LU32 = LUo.astype(np.float32)
    # the stored factors are the FP64 factors rounded once to float
    assert np.all(LUg == LUg.astype(np.float32).astype(np.float64))
    assert rel_err(LUg, LU32.astype(np.float64), floor_frac=1e-7) <= 2.0 ** -23
    rng = np.random.default_rng(8)
    d = rng.standard_normal(2 * P.N)
    vg = g.apply_preconditioner(d)
    vo32 = o.ilu0_apply32(LUg.astype(np.float32), d)
    print("fp32 apply bitwise:", np.array_equal(vg, vo32), flush=True)
    assert rel_err(vg, vo32, floor_frac=1e-6) < 1e-6
    # and its distance from the FP64 apply, reported against the FP64 oracle
    vo64 = o.ilu0_apply(LUo, d)
    dev = np.abs(vg - vo64).max() / np.abs(vo64).max()
    print(f"fp32 apply vs fp64 oracle: {dev:.3e} of max |v|", flush=True)
    assert dev < 1e-4


@pytest.mark.parametrize("shape", SHAPES + ["lens"])
def test_fp32_operator_apply_matches_rounded_jacobian(shape):
    """BiCGStab's operator with params.operator_fp32: the product with the Jacobian's
    blocks rounded once to float, in FP64 arithmetic and the FP64 product's order, i.e.
    the oracle's SpMV of the rounded blocks."""
    import ewg
    P = _problem(shape)
    print(f"fp32 operator shape {shape}", flush=True)
    g, o, J = _system(P, ewg.default_params(operator_fp32=1), seed=9)
    rng = np.random.default_rng(10)
    x = rng.standard_normal(2 * P.N) * np.tile([1e4, 0.1], P.N)
    yg = g.apply_operator(x)
    yo = o.spmv(J.astype(np.float32).astype(np.float64), x)
    print("bitwise:", np.array_equal(yg, yo), flush=True)
    assert rel_err(yg, yo) < 1e-12
    # it is not the FP64 product
    assert not np.array_equal(yg, o.spmv(J, x))


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("fp32_precond", [0, 1])
def test_fp32_operator_solves_agree_between_schedules(shape, fp32_precond):
    """The solve's FP32 copy of the Jacobian comes from the wavefront layout's operand
    packing (schedule 0) or from a rounding pass (schedule 1): the two BiCGStab solves of
    the same system are bit-identical (same copy, same preconditioner, same kernels)."""
    import ewg
    from ewg import problems
    P = _problem(shape)
    x = problems.perturbed_state(P, seed=11, dS=0.2)
    out = []
    for schedule in (0, 1):
        g = ewg.GpuContext(P, ewg.default_params(operator_fp32=1, precond_fp32=fp32_precond,
                                                 ilu_schedule=schedule))
        g.set_solution(x, 0)
        g.linearize(P.dt, 0, want=False)
        rc, dx, its = g.solve()
        print(f"shape {shape} fp32_precond {fp32_precond} schedule {schedule}: rc {rc} "
              f"iterations {its}", flush=True)
        assert rc == 0
        out.append((dx, its))
        g.close()
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])
