"""GPU: the finite-difference local linearizer (FvBaseFdLocalLinearizer,
fvbasefdlocallinearizer.hh:131-515; params.local_linearizer = 1) against the oracle's, and
against the AD Jacobian at the full 128x128x64 size, where the CPU oracle is too slow to
give the Jacobian: an independent check of the AD derivatives at scale.

Bars: residual within 1e-12 relative of the oracle (it is the AD path's value); FD
Jacobian within 1e-6 relative of the oracle's FD Jacobian (entry-wise, 1e-9 * max floor:
differences of residuals that agree to the last bits, divided by eps ~ 1e-10); GPU FD vs
GPU AD below 2e-3 of each block column (the differences' truncation and round-off, the
bar the oracle's own FD-vs-AD test holds)."""
import numpy as np
import pytest

from helpers import rel_err
from test_oracle_linearization import block_column_error

pytestmark = pytest.mark.gpu


def _cases():
    from ewg import problems
    return [("lens48x32", problems.lens_problem()),
            ("het3d_20x13x7", problems.heterogeneous_3d(20, 13, 7, seed=17, patch=4)),
            ("het3d_32x24x16", problems.heterogeneous_3d(32, 24, 16, seed=12, patch=8))]


@pytest.mark.parametrize("case", range(3))
@pytest.mark.parametrize("method", [1, 0, -1])
def test_fd_linearize_parity(case, method):
    import ewg
    from ewg import problems
    from oracle.pyoracle import OracleContext
    name, P = _cases()[case]
    print(f"fd parity {name} method {method}", flush=True)
    prm = ewg.default_params(local_linearizer=1, numeric_difference_method=method)
    gpu, ora = ewg.GpuContext(P, prm), OracleContext(P, prm)
    x = problems.perturbed_state(P, seed=5 + case)
    x1 = problems.perturbed_state(P, seed=50 + case, dS=0.1)
    for c in (gpu, ora):
        c.set_solution(x1, 1)
        c.set_solution(x1, 0)
        c.linearize(P.dt, 0, want=False)
        c.set_solution(x, 0)
    for it in (1, 0):
        rg, jg = gpu.linearize(P.dt, it)
        ro, jo = ora.linearize(P.dt, it)
        print("residual bitwise:", np.array_equal(rg, ro), "jacobian bitwise:",
              np.array_equal(jg, jo), flush=True)
        assert rel_err(rg, ro) < 1e-12
        assert rel_err(jg, jo, floor_frac=1e-9) < 1e-6


def test_fd_newton_lens_matches_oracle():
    """Newton steps with the FD Jacobian: the same iteration counts as the oracle's FD
    Newton, saturations within 1e-8."""
    import ewg
    from ewg import problems
    from oracle.pyoracle import OracleContext
    P = problems.lens_problem()
    prm = ewg.default_params(newton_tolerance=1e-8, linear_tolerance=1e-10, local_linearizer=1)
    gpu, ora = ewg.GpuContext(P, prm), OracleContext(P, prm)
    for step in range(3):
        rcg, ig, _ = gpu.newton(P.dt)
        rco, io, _ = ora.newton(P.dt)
        print(f"step {step}: gpu {rcg} {ig}, oracle {rco} {io}", flush=True)
        assert rcg == rco == 0 and ig == io
        gpu.advance_time_level()
        ora.advance_time_level()
    assert np.abs(gpu.solution(0)[1::2] - ora.solution(0)[1::2]).max() < 1e-8


def _ad_fd_pair(P, x, x1):
    import ewg
    out = []
    for ll in (0, 1):
        c = ewg.GpuContext(P, ewg.default_params(newton_tolerance=1e-6, local_linearizer=ll))
        c.set_solution(x1, 1)
        c.set_solution(x1, 0)
        c.linearize(P.dt, 0, want=False)
        c.set_solution(x, 0)
        out.append(c.linearize(P.dt, 1))
        c.close()
    return out


def test_fd_vs_ad_jacobian_full_size():
    """128x128x64 at a perturbed state (random pressures and saturations: no face sits
    at an upstream tie and no cell at a material-law kink within the deflections): the
    FD Jacobian is within 2e-3 of the AD Jacobian in every block column."""
    from ewg import problems
    P = problems.heterogeneous_3d()
    x1 = problems.perturbed_state(P, seed=61, dS=0.1)
    x = problems.perturbed_state(P, seed=62, dS=0.2)
    (r_ad, J_ad), (r_fd, J_fd) = _ad_fd_pair(P, x, x1)
    err = block_column_error(J_fd, J_ad)
    print(f"128x128x64 perturbed: residual bitwise {np.array_equal(r_fd, r_ad)}, FD vs AD "
          f"{err:.3e} of the block columns", flush=True)
    assert np.array_equal(r_fd, r_ad)
    assert err < 2e-3


def test_fd_vs_ad_jacobian_full_size_after_two_steps():
    """128x128x64 after two time steps of the run (DNAPL below the injection patch, the
    rest at S_w = 1 near hydrostatic balance): faces whose potential difference is within
    a deflection of zero switch their upstream cell under the FD deflection, so there the
    difference quotient mixes the two cells' mobilities (the reference's FD linearizer
    does the same); everywhere else FD and AD agree to 2e-3. The switching block columns
    are counted and must stay a small minority."""
    import ewg
    from ewg import problems
    P = problems.heterogeneous_3d()
    run = ewg.GpuContext(P, ewg.default_params(newton_tolerance=1e-6, max_time_step_size=P.dt))
    run.run(0.0, 2 * P.dt, P.dt)
    x, x1 = run.solution(0), run.solution(1)
    run.close()
    (r_ad, J_ad), (r_fd, J_fd) = _ad_fd_pair(P, x, x1)
    assert np.array_equal(r_fd, r_ad)
    Jb, Rb = J_fd.reshape(-1, 2, 2), J_ad.reshape(-1, 2, 2)
    scale = np.maximum(np.abs(Rb).max(axis=1, keepdims=True), 1e-300)
    col = (np.abs(Jb - Rb) / scale).max(axis=1)          # [block][pv]
    bad = col > 2e-3
    frac = bad.mean()
    print(f"128x128x64 after two steps: {bad.sum()} of {bad.size} block columns beyond 2e-3 "
          f"({frac:.2e}); the rest within {col[~bad].max():.3e}", flush=True)
    assert frac < 1e-3
