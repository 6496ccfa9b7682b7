"""Committed golden fixtures (tests/golden/, written by tests/golden/make_golden.py from the
oracle): the oracle must keep reproducing them exactly, and the HIP path must match them
to the parity bars (identical Newton iteration counts per step, saturations within 1e-8,
residual and Jacobian within 1e-12 relative)."""
import json
import os

import numpy as np
import pytest

from helpers import rel_err

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def _cases():
    from ewg import problems
    return {
        "lens48x32_run.json": problems.lens_problem(),
        "het3d_12x10x8_run.json": problems.heterogeneous_3d(12, 10, 8, seed=7, patch=4, steps=5),
    }


@pytest.mark.parametrize("name", ["lens48x32_run.json", "het3d_12x10x8_run.json"])
def test_oracle_reproduces_golden_run(name):
    from oracle.pyoracle import OracleContext
    g = _load(name)
    P = _cases()[name]
    ora = OracleContext(P)
    steps = ora.run(0.0, g["t_end"], P.dt)
    its, dts = ora.step_log()
    st = ora.stats()
    assert steps == g["steps"]
    assert its.tolist() == g["newton_iterations_per_step"]
    assert np.allclose(dts, g["dt_per_step"], rtol=0, atol=0)
    for k, v in g["totals"].items():
        assert st[k] == v, k
    x = ora.solution(0)
    S, p = x[1::2], x[0::2]
    f = g["final"]
    assert S.sum() == pytest.approx(f["sum_S"], rel=1e-13)
    assert p.sum() == pytest.approx(f["sum_p"], rel=1e-13)
    assert [float(S[i]) for i in f["probe_cells"]] == pytest.approx(f["probe_S"], rel=1e-13)


def test_oracle_reproduces_golden_linearization():
    from ewg import problems
    from oracle.pyoracle import OracleContext
    d = np.load(os.path.join(GOLDEN, "lens6x5_linearize.npz"))
    P = problems.lens_problem(nx=6, ny=5)
    ora = OracleContext(P)
    ora.set_solution(d["x_prev"], 1)
    ora.set_solution(d["x_prev"], 0)
    ora.linearize(float(d["dt"]), 0, want=False)
    ora.set_solution(d["x"], 0)
    r, j = ora.linearize(float(d["dt"]), 1)
    assert np.array_equal(r, d["residual"]) and np.array_equal(j, d["jacobian"])


@pytest.mark.gpu
def test_gpu_matches_golden_linearization():
    import ewg
    from ewg import problems
    d = np.load(os.path.join(GOLDEN, "lens6x5_linearize.npz"))
    P = problems.lens_problem(nx=6, ny=5)
    gpu = ewg.GpuContext(P)
    gpu.set_solution(d["x_prev"], 1)
    gpu.set_solution(d["x_prev"], 0)
    gpu.linearize(float(d["dt"]), 0, want=False)
    gpu.set_solution(d["x"], 0)
    r, j = gpu.linearize(float(d["dt"]), 1)
    rp, ci = gpu.pattern()
    assert np.array_equal(rp, d["rowptr"]) and np.array_equal(ci, d["colidx"])
    assert rel_err(r, d["residual"]) < 1e-12
    assert rel_err(j, d["jacobian"]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["lens48x32_run.json", "het3d_12x10x8_run.json"])
def test_gpu_matches_golden_run(name):
    """Step by step with the fixture's time step sizes: the same Newton iteration count
    in every step, the same totals (BiCGStab counts within 2 %: the reductions sum in a
    different order), and the same final fields."""
    import ewg
    g = _load(name)
    P = _cases()[name]
    gpu = ewg.GpuContext(P)
    its = []
    for dt in g["dt_per_step"]:
        rc, it, err = gpu.newton(dt)
        assert rc == 0, (len(its), rc, err)
        its.append(it)
        gpu.advance_time_level()
    print(name, "newton iterations per step", its, flush=True)
    assert its == g["newton_iterations_per_step"]
    st = gpu.stats()
    tot = g["totals"]
    assert st["newton_iterations"] == tot["newton_iterations"]
    assert st["linear_solves"] == tot["linear_solves"]
    assert abs(st["linear_iterations"] - tot["linear_iterations"]) <= 0.02 * tot["linear_iterations"]
    x = gpu.solution(0)
    S, p = x[1::2], x[0::2]
    f = g["final"]
    assert np.max(np.abs(np.array([S[i] for i in f["probe_cells"]]) - f["probe_S"])) < 1e-8
    assert abs(S.sum() - f["sum_S"]) < 1e-8 * S.size
    assert abs(p.sum() - f["sum_p"]) <= 1e-9 * abs(f["sum_p"])
