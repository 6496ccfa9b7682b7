"""GPU parity at the benchmark's own size: the 3D heterogeneous 128x128x64 case
(1,048,576 cells; 256 tiles of the ILU(0) wavefront, every tile hand-over of the timed
runs), the GPU against the oracle on the same inputs. The oracle's C kernels take about a
second per linearization and some seconds for the first time step at this size, so the
comparisons are direct, with the bars of the small cases:

* residual and Jacobian within 1e-12 relative (entry-wise, 1e-13 * max floor);
* FP64 factors, one apply and one SpMV within 1e-12; FP32 factors the FP64 ones rounded
  once, the FP32 apply within 1e-6 of the oracle's float apply of those factors;
* the first time step from the initial state (its 85 BiCGStab iterations are most of the
  benchmark's 20-step run): the same Newton iterations, saturations within 1e-8 and
  BiCGStab iterations within 2 % (the dot products sum in another order);
* the reduced-precision forms against the same FP64 oracle: FP32 face fluxes within 1e-5
  of the largest residual and Jacobian entry, and the first time step of FP32 fluxes with
  the FP32 preconditioner with saturations within 1e-6 and Newton iterations within one.
"""
import functools

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

SHAPE = (128, 128, 64)


def _problem():
    from ewg import problems
    return problems.heterogeneous_3d(*SHAPE)  # the benchmark's case (its default seed)


def _step_params(**extra):
    import ewg
    return ewg.default_params(newton_tolerance=1e-6, linear_tolerance=1e-3,
                              max_time_step_size=_problem().dt, **extra)


@functools.lru_cache(maxsize=1)
def _oracle_first_step():
    """The oracle's first time step of the case (FP64): Newton and BiCGStab iterations and
    the state after it (computed once per session)."""
    from oracle.pyoracle import OracleContext
    P = _problem()
    o = OracleContext(P, _step_params())
    rc, it, _ = o.newton(P.dt)
    assert rc == 0
    return it, o.stats()["linear_iterations"], o.solution(0)


@pytest.mark.parametrize("newton_iter", [0, 1])
def test_full_size_linearization_matches_oracle(newton_iter):
    import ewg
    from ewg import problems
    from oracle.pyoracle import OracleContext
    P = _problem()
    print(f"linearization {SHAPE} newton_iter {newton_iter}", flush=True)
    gpu, ora = ewg.GpuContext(P), OracleContext(P)
    x = problems.perturbed_state(P, seed=31)
    x1 = problems.perturbed_state(P, seed=32, dS=0.1)
    for c in (gpu, ora):
        c.set_solution(x1, 1)
        c.set_solution(x1, 0)
        c.linearize(P.dt, 0, want=False)  # fills the storage cache from x1
        c.set_solution(x, 0)
    rg, jg = gpu.linearize(P.dt, newton_iter)
    ro, jo = ora.linearize(P.dt, newton_iter)
    print("residual bitwise:", np.array_equal(rg, ro), "jacobian bitwise:",
          np.array_equal(jg, jo), flush=True)
    assert rel_err(rg, ro) < 1e-12
    assert rel_err(jg, jo) < 1e-12


@pytest.mark.parametrize("fp32", [0, 1])
def test_full_size_ilu0_and_spmv_match_oracle(fp32):
    import ewg
    from ewg import problems
    from oracle.pyoracle import OracleContext
    P = _problem()
    print(f"ilu0 and spmv {SHAPE} precond_fp32 {fp32}", flush=True)
    prm = ewg.default_params(precond_fp32=fp32)
    g, o = ewg.GpuContext(P, prm), OracleContext(P, prm)
    g.set_solution(problems.perturbed_state(P, seed=33, dS=0.2), 0)
    _, J = g.linearize(P.dt, 0)
    LUg = g.preconditioner()
    LUo = o.ilu0_factor(J)
    rng = np.random.default_rng(34)
    d = rng.standard_normal(2 * P.N)
    vg = g.apply_preconditioner(d)
    if fp32:
        assert np.all(LUg == LUg.astype(np.float32).astype(np.float64))
        assert rel_err(LUg, LUo.astype(np.float32).astype(np.float64), floor_frac=1e-7) <= 2.0 ** -23
        vo = o.ilu0_apply32(LUg.astype(np.float32), d)
        print("fp32 apply bitwise:", np.array_equal(vg, vo), flush=True)
        assert rel_err(vg, vo, floor_frac=1e-6) < 1e-6
    else:
        print("factors bitwise:", np.array_equal(LUg, LUo), flush=True)
        assert rel_err(LUg, LUo) < 1e-12
        vo = o.ilu0_apply(LUo, d)
        print("apply bitwise:", np.array_equal(vg, vo), flush=True)
        assert rel_err(vg, vo) < 1e-12
    x = rng.standard_normal(2 * P.N) * np.tile([1e4, 0.1], P.N)
    assert rel_err(g.apply_operator(x), o.spmv(J, x)) < 1e-12


def test_full_size_first_time_step_matches_oracle():
    import ewg
    P = _problem()
    print(f"first time step {SHAPE}", flush=True)
    g = ewg.GpuContext(P, _step_params())
    rg, itg, _ = g.newton(P.dt)
    ito, lito, xo = _oracle_first_step()
    litg = g.stats()["linear_iterations"]
    print("newton", itg, ito, "bicgstab", litg, lito, flush=True)
    assert rg == 0
    assert itg == ito
    assert abs(litg - lito) <= 0.02 * lito
    dS = np.abs(g.solution(0)[1::2] - xo[1::2]).max()
    print(f"max |dS| {dS:.3e}", flush=True)
    assert dS < 1e-8


def test_full_size_fp32_flux_linearization_close_to_fp64():
    import ewg
    from ewg import problems
    from oracle.pyoracle import OracleContext
    P = _problem()
    print(f"fp32 flux linearization {SHAPE}", flush=True)
    g32, ora = ewg.GpuContext(P, ewg.default_params(linearize_fp32=1)), OracleContext(P)
    x = problems.perturbed_state(P, seed=35)
    x1 = problems.perturbed_state(P, seed=36, dS=0.1)
    for c in (g32, ora):
        c.set_solution(x1, 1)
        c.set_solution(x1, 0)
        c.linearize(P.dt, 0, want=False)
        c.set_solution(x, 0)
    r32, j32 = g32.linearize(P.dt, 1)
    ro, jo = ora.linearize(P.dt, 1)
    er = np.abs(r32 - ro).max() / np.abs(ro).max()
    ej = np.abs(j32 - jo).max() / np.abs(jo).max()
    print(f"fp32 flux: residual {er:.3e}, jacobian {ej:.3e} (of the largest entry)", flush=True)
    assert er < 1e-5 and ej < 1e-5
    assert er > 0.0 or ej > 0.0  # the FP32 form is really taken


def test_full_size_mixed_first_time_step_close_to_fp64_oracle():
    import ewg
    P = _problem()
    print(f"first time step, FP32 fluxes and FP32 preconditioner {SHAPE}", flush=True)
    g = ewg.GpuContext(P, _step_params(linearize_fp32=1, precond_fp32=1))
    rg, itg, _ = g.newton(P.dt)
    ito, lito, xo = _oracle_first_step()
    xg = g.solution(0)
    dS = np.abs(xg[1::2] - xo[1::2]).max()
    dp = np.abs(xg[0::2] - xo[0::2]).max() / np.abs(xo[0::2]).max()
    print(f"newton {itg} vs {ito}, bicgstab {g.stats()['linear_iterations']} vs {lito}, "
          f"max |dS| {dS:.3e}, max |dp| / max |p| {dp:.3e}", flush=True)
    assert rg == 0
    assert abs(itg - ito) <= 1
    assert dS < 1e-6
