"""Newton / time-loop semantics of the oracle (NewtonMethod::apply, FvBaseProblem::
timeIntegration, Simulator::run) and their failure paths, forced from the host."""
import numpy as np
import pytest


def _lens(**kw):
    from ewg import problems
    return problems.lens_problem(**kw)


def test_newton_converges_quadratically_on_lens():
    from oracle.pyoracle import OracleContext
    P = _lens(nx=24, ny=16)
    o = OracleContext(P, _with(P.params, newton_tolerance=1e-11, linear_tolerance=1e-12,
                               linear_abs_tolerance=0.0))
    rc, it, err = o.newton(P.dt)
    e = o.newton_history()
    assert rc == 0 and len(e) == it and e[-1] <= 1e-11
    # exact Jacobian: in the asymptotic regime e_{k+1} <= C e_k^2, with C estimated from
    # the preceding pair, until the rounding floor is reached
    checked = 0
    for k in range(1, len(e) - 1):
        if e[k] < 1e-4 * e[0] and e[k + 1] > 1e-11:
            C_ = e[k] / e[k - 1] ** 2
            assert e[k + 1] <= 10 * C_ * e[k] ** 2
            checked += 1
    assert checked >= 2


def _with(params, **kw):
    import copy
    p = copy.copy(params)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_fixed_dt_run_step_count_and_iterations():
    from oracle.pyoracle import OracleContext
    P = _lens()
    o = OracleContext(P)
    n = o.run(0.0, 3000.0, P.dt)
    its, dts = o.step_log()
    assert n == 12 and np.all(dts == 250.0)
    assert its[0] >= its[-1] >= 2
    s = o.stats()
    assert s["newton_iterations"] == its.sum()
    assert s["failed_steps"] == 0


def test_adaptive_dt_follows_suggestion():
    """With no MaxTimeStepSize the next dt is suggestTimeStepSize(dt) (newtonmethod.hh:527-544)."""
    from oracle.pyoracle import OracleContext
    P = _lens()
    params = _with(P.params, max_time_step_size=float("inf"))
    o = OracleContext(P, params)
    o.run(0.0, 6000.0, P.dt)
    its, dts = o.step_log()
    for k in range(len(dts) - 2):
        n = its[k]
        tgt = 10
        sug = dts[k] / (1 + (n - tgt) / tgt) if n > tgt else dts[k] * (1 + (tgt - n) / tgt / 1.2)
        assert dts[k + 1] == pytest.approx(sug, rel=1e-14) or dts[k + 1] < sug  # last step clipped


def test_linear_solver_failure_halves_dt_until_give_up():
    from oracle.pyoracle import OracleContext
    from ewg.capi import SolverError
    P = _lens(nx=12, ny=8)
    params = _with(P.params, linear_max_iterations=1, linear_tolerance=1e-14)
    o = OracleContext(P, params)
    with pytest.raises(SolverError) as e:
        o.run(0.0, 1000.0, P.dt)
    assert e.value.code == 2
    assert o.stats()["failed_steps"] == 10
    # the solution was restored to the previous time level after every failure
    assert np.array_equal(o.solution(0), P.x0.ravel())


def test_newton_max_error_is_a_numerical_issue():
    from oracle.pyoracle import OracleContext
    P = _lens(nx=12, ny=8)
    o = OracleContext(P, _with(P.params, newton_max_error=1e-30))
    rc, it, err = o.newton(P.dt)
    assert rc == 1 and it == 20  # failed_(): numIterations = 2 * target


def test_non_finite_state_fails_the_step():
    from oracle.pyoracle import OracleContext
    P = _lens(nx=12, ny=8)
    o = OracleContext(P)
    x = P.x0.copy()
    x[5, 0] = np.nan
    o.set_solution(x, 0)
    rc, it, err = o.newton(P.dt)
    assert rc != 0
