"""GPU parity: libewg (HIP, through its C ABI) against the CPU oracle on the same inputs.

Bars: residual and Jacobian within 1e-12 relative (entry-wise, with a 1e-13 * max floor
for cancellations), ILU(0) factors and sweeps within 1e-12, identical Newton iteration
counts, saturations within 1e-8.
"""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu


def _pair(P, params=None):
    import ewg
    from oracle.pyoracle import OracleContext
    return ewg.GpuContext(P, params), OracleContext(P, params)


def _lin_cases():
    from ewg import problems
    return [
        ("lens48x32", problems.lens_problem()),
        ("lens_ref_alpha", problems.lens_problem(nx=24, ny=16, outer_alpha=3.7e-3, mu_n=5.7e-4)),
        ("het3d_12x10x8", problems.heterogeneous_3d(12, 10, 8, seed=11, patch=4)),
        ("het3d_32x24x16", problems.heterogeneous_3d(32, 24, 16, seed=12, patch=8)),
    ]


@pytest.mark.parametrize("case", range(4))
@pytest.mark.parametrize("newton_iter", [0, 1])
def test_linearize_parity(case, newton_iter):
    from ewg import problems
    name, P = _lin_cases()[case]
    print(f"case {name} newton_iter {newton_iter}", flush=True)
    gpu, ora = _pair(P)
    x = problems.perturbed_state(P, seed=5 + case)
    x1 = problems.perturbed_state(P, seed=50 + case, dS=0.1)
    for c in (gpu, ora):
        c.set_solution(x1, 1)
        c.set_solution(x1, 0)
        c.linearize(P.dt, 0, want=False)   # fills the storage cache from x1
        c.set_solution(x, 0)
    rg, jg = gpu.linearize(P.dt, newton_iter)
    ro, jo = ora.linearize(P.dt, newton_iter)
    assert rel_err(rg, ro) < 1e-12
    assert rel_err(jg, jo) < 1e-12
    rp_g, ci_g = gpu.pattern()
    rp_o, ci_o = ora.pattern()
    assert np.array_equal(rp_g, rp_o) and np.array_equal(ci_g, ci_o)


@pytest.mark.parametrize("shape", [(20, 13, 7), (37, 21, 5), (33, 17, 9), (48, 40, 19)])
@pytest.mark.parametrize("h", [1.0, 0.75])
def test_fused_3d_linearization_is_bitwise(shape, h):
    """The fused 3D pass (16x16 column tiles marching in z in 6-plane chunks, intensive
    quantities in LDS) gives the oracle's residual and Jacobian bit for bit on shapes with
    partial tiles and partial chunks, at Newton iterations 0 and 1: with a power-of-two
    spacing (the faces' geometry evaluated once per direction) and with another one (the
    geometry evaluated per face)."""
    import ewg
    from ewg import problems
    P = problems.heterogeneous_3d(*shape, seed=sum(shape), patch=4, h=h)
    gpu, ora = _pair(P)
    x = problems.perturbed_state(P, seed=3, dS=0.2)
    x1 = problems.perturbed_state(P, seed=4, dS=0.1)
    for c in (gpu, ora):
        c.set_solution(x1, 1)
        c.set_solution(x, 0)
    for it in (0, 1):
        rg, jg = gpu.linearize(P.dt, it)
        ro, jo = ora.linearize(P.dt, it)
        print(f"{shape} h {h} newton_iter {it}: residual bitwise {np.array_equal(rg, ro)}, "
              f"Jacobian bitwise {np.array_equal(jg, jo)}", flush=True)
        assert np.array_equal(rg, ro) and np.array_equal(jg, jo)
    gpu.close()
    ora.close()


def test_linear_solve_parity():
    from ewg import problems
    P = problems.lens_problem()
    gpu, ora = _pair(P)
    x = problems.perturbed_state(P, seed=9, dS=0.05, dp=50.0)
    for c in (gpu, ora):
        c.set_solution(P.x0, 1)
        c.set_solution(P.x0, 0)
        c.linearize(P.dt, 0, want=False)
        c.set_solution(x, 0)
        c.linearize(P.dt, 1, want=False)
    rcg, dxg, itg = gpu.solve()
    rco, dxo, ito = ora.solve()
    print("iterations gpu/oracle", itg, ito, flush=True)
    assert rcg == 0 and rco == 0
    assert abs(itg - ito) <= 1
    # both solves meet relTol 1e-10: the updates agree far below the Newton tolerance
    assert np.max(np.abs(dxg - dxo)) <= 1e-6 * np.max(np.abs(dxo))


def test_newton_lens_first_steps():
    from ewg import problems
    P = problems.lens_problem()
    gpu, ora = _pair(P)
    its_g, its_o = [], []
    for step in range(4):
        rcg, ig, eg = gpu.newton(P.dt)
        rco, io, eo = ora.newton(P.dt)
        print(f"step {step}: gpu rc {rcg} it {ig} err {eg:.3e} | oracle rc {rco} it {io} err {eo:.3e}",
              flush=True)
        assert rcg == 0 and rco == 0
        its_g.append(ig)
        its_o.append(io)
        gpu.advance_time_level()
        ora.advance_time_level()
    assert its_g == its_o
    dS = np.abs(gpu.solution(0)[1::2] - ora.solution(0)[1::2]).max()
    assert dS < 1e-8


def test_run_lens_3000s():
    """The reference's own test length (CMakeLists.txt:73-74: --end-time=3000)."""
    from ewg import problems
    P = problems.lens_problem()
    gpu, ora = _pair(P)
    ng = gpu.run(0.0, 3000.0, P.dt)
    no = ora.run(0.0, 3000.0, P.dt)
    sg, so = gpu.stats(), ora.stats()
    print("gpu", sg, "\noracle", so, flush=True)
    assert ng == no
    assert sg["newton_iterations"] == so["newton_iterations"]
    xg, xo = gpu.solution(0), ora.solution(0)
    assert np.abs(xg[1::2] - xo[1::2]).max() < 1e-8
    assert rel_err(xg[0::2], xo[0::2]) < 1e-9


def test_run_het3d_small():
    from ewg import problems
    P = problems.heterogeneous_3d(24, 20, 12, seed=3, patch=6, steps=5)
    gpu, ora = _pair(P)
    ng = gpu.run(0.0, P.t_end, P.dt)
    no = ora.run(0.0, P.t_end, P.dt)
    sg, so = gpu.stats(), ora.stats()
    print("gpu", sg, "\noracle", so, flush=True)
    assert ng == no
    assert sg["newton_iterations"] == so["newton_iterations"]
    xg, xo = gpu.solution(0), ora.solution(0)
    assert np.abs(xg[1::2] - xo[1::2]).max() < 1e-8


def test_fp32_preconditioner_close_to_fp64():
    import ewg
    from ewg import problems
    from oracle.pyoracle import OracleContext
    P = problems.heterogeneous_3d(24, 20, 12, seed=4, patch=6, steps=3)
    p32 = ewg.default_params(newton_tolerance=1e-6, linear_tolerance=1e-3,
                             max_time_step_size=P.dt, precond_fp32=1)
    g32 = ewg.GpuContext(P, p32)
    o64 = OracleContext(P)
    g32.run(0.0, P.t_end, P.dt)
    o64.run(0.0, P.t_end, P.dt)
    dS = np.abs(g32.solution(0)[1::2] - o64.solution(0)[1::2]).max()
    print("max |dS| fp32-precond vs fp64 oracle", dS, g32.stats(), flush=True)
    assert dS < 1e-6


@pytest.mark.parametrize("shape", [(20, 13, 7), (32, 24, 16), (8, 8, 3), (37, 21, 5), (9, 3, 1),
                                   (3, 11, 2), (40, 3, 9)])
@pytest.mark.parametrize("fp32", [0, 1])
def test_wavefront_ilu_matches_hyperplane_schedule(shape, fp32):
    """The tile-wavefront ILU(0) schedule (one persistent launch per factorization and
    sweep, factors in the tiles' wavefront layout) and the one-launch-per-hyperplane
    schedule run the same arithmetic in the same order: the whole BiCGStab solve is
    bit-identical (partial tiles included, and thin grids: one z-layer, a single row of
    tiles in x or y, three cells across)."""
    import ewg
    from ewg import problems
    nx, ny, nz = shape
    P = problems.heterogeneous_3d(nx, ny, nz, seed=21, patch=4)
    x = problems.perturbed_state(P, seed=22, dS=0.05, dp=50.0)
    out = {}
    for sched in (1, 0):
        print(f"shape {shape} fp32 {fp32} schedule {sched}", flush=True)
        prm = ewg.default_params(newton_tolerance=1e-6, linear_tolerance=1e-8,
                                 precond_fp32=fp32, ilu_schedule=sched)
        g = ewg.GpuContext(P, prm)
        g.set_solution(x, 0)
        g.linearize(P.dt, 0, want=False)
        out[sched] = g.solve()
    rc1, dx1, it1 = out[1]
    assert rc1 == 0
    for sched in (0,):
        rc, dx, it = out[sched]
        assert rc == 0 and it == it1, (sched, rc, it, it1)
        assert np.array_equal(dx, dx1), sched


@pytest.mark.parametrize("case", [0, 3])
def test_fp32_flux_linearization_close_to_fp64(case):
    """FP32 face fluxes (pressure differences in FP64, accumulation into the FP64 residual
    and Jacobian) against the FP64 oracle: both agree to a few FP32 roundings of the flux
    terms (bound: 1e-5 of the largest entry), and the FP64 path is untouched."""
    import ewg
    from ewg import problems
    from oracle.pyoracle import OracleContext
    name, P = _lin_cases()[case]
    print(f"case {name}", flush=True)
    g32 = ewg.GpuContext(P, ewg.default_params(linearize_fp32=1))
    g64 = ewg.GpuContext(P)
    ora = OracleContext(P)
    x = problems.perturbed_state(P, seed=5 + case)
    x1 = problems.perturbed_state(P, seed=50 + case, dS=0.1)
    for c in (g32, g64, ora):
        c.set_solution(x1, 1)
        c.set_solution(x1, 0)
        c.linearize(P.dt, 0, want=False)
        c.set_solution(x, 0)
    r32, j32 = g32.linearize(P.dt, 1)
    r64, j64 = g64.linearize(P.dt, 1)
    ro, jo = ora.linearize(P.dt, 1)
    er = np.abs(r32 - ro).max() / np.abs(ro).max()
    ej = np.abs(j32 - jo).max() / np.abs(jo).max()
    print(f"fp32 flux: residual {er:.3e}, jacobian {ej:.3e} (of the largest entry)", flush=True)
    assert er < 1e-5 and ej < 1e-5
    assert er > 0.0 or ej > 0.0   # the FP32 form is really taken
    assert rel_err(r64, ro) < 1e-12 and rel_err(j64, jo) < 1e-12


@pytest.mark.parametrize("form", ["fp32_flux", "mixed", "mixed_operator"])
def test_mixed_precision_steps_close_to_fp64_oracle(form):
    """FP32 face fluxes (alone, with the FP32 preconditioner, and with the FP32
    preconditioner and the FP32 copy of the Jacobian in BiCGStab's operator products)
    against the FP64 oracle, both at Newton tolerance 1e-6, every time step: saturations
    within 1e-6, pressures within 1e-6 relative, Newton iterations within one of the
    oracle's."""
    import ewg
    from ewg import problems
    from oracle.pyoracle import OracleContext
    P = problems.heterogeneous_3d(24, 20, 12, seed=4, patch=6, steps=4)
    extra = {"linearize_fp32": 1}
    if form in ("mixed", "mixed_operator"):
        extra["precond_fp32"] = 1
    if form == "mixed_operator":
        extra["operator_fp32"] = 1
    g = ewg.GpuContext(P, ewg.default_params(newton_tolerance=1e-6, **extra))
    o = OracleContext(P, ewg.default_params(newton_tolerance=1e-6))
    for step in range(4):
        kg0, ko0 = g.stats()["linear_iterations"], o.stats()["linear_iterations"]
        rcg, ig, _ = g.newton(P.dt)
        rco, io, _ = o.newton(P.dt)
        g.advance_time_level()
        o.advance_time_level()
        xg, xo = g.solution(0), o.solution(0)
        dS = np.abs(xg[1::2] - xo[1::2]).max()
        dp = np.abs(xg[0::2] - xo[0::2]).max() / np.abs(xo[0::2]).max()
        kg = g.stats()["linear_iterations"] - kg0
        ko = o.stats()["linear_iterations"] - ko0
        print(f"{form} step {step}: newton {ig} vs {io}, bicgstab {kg} vs {ko}, "
              f"max|dS| {dS:.3e}, max|dp|/max|p| {dp:.3e}", flush=True)
        assert rcg == 0 and rco == 0
        assert abs(ig - io) <= 1
        assert dS < 1e-6 and dp < 1e-6


def test_run_lens_with_step_halving():
    """The time loop with failing Newton solves (NewtonMaxIterations = 4 with the
    reference's continuation rule, dt0 = 2000 s): dt is halved and the step redone from
    solution(1) (fvbaseproblem.hh:507-553, fvbasediscretization.hh:1370-1387) exactly as
    in the oracle: the same accepted steps, halvings and linearizations."""
    import ewg
    from ewg import problems
    P = problems.lens_problem()
    prm = ewg.default_params(newton_tolerance=1e-8, linear_tolerance=1e-10,
                             newton_max_iterations=4)
    gpu, ora = _pair(P, prm)
    ng = gpu.run(0.0, 4000.0, 2000.0)
    no = ora.run(0.0, 4000.0, 2000.0)
    sg, so = gpu.stats(), ora.stats()
    print("gpu", sg, "\noracle", so, flush=True)
    assert so["failed_steps"] >= 5
    assert ng == no
    assert sg["failed_steps"] == so["failed_steps"]
    assert sg["time_steps"] == so["time_steps"]
    assert sg["newton_iterations"] == so["newton_iterations"]
    xg, xo = gpu.solution(0), ora.solution(0)
    assert np.abs(xg[1::2] - xo[1::2]).max() < 1e-8


@pytest.mark.parametrize("case", ["max_error", "linear_solver"])
def test_newton_failures_match_oracle(case):
    """The Newton loop's failure exits: NewtonMaxError exceeded at the first
    linearization (decided on the device: numerical issue), and a linear solve that does
    not converge within LinearSolverMaxIterations = 1 (not converged); both report
    numIterations = 2 * NewtonTargetIterations and the error the oracle reports."""
    import ewg
    from ewg import problems
    P = problems.lens_problem()
    kw = ({"newton_max_error": 1e-9} if case == "max_error"
          else {"linear_max_iterations": 1})
    prm = ewg.default_params(linear_tolerance=1e-10, **kw)
    gpu, ora = _pair(P, prm)
    rcg, itg, eg = gpu.newton(4 * P.dt)
    rco, ito, eo = ora.newton(4 * P.dt)
    print(case, "gpu", rcg, itg, eg, "oracle", rco, ito, eo, flush=True)
    assert rcg == rco and itg == ito == 2 * prm.newton_target_iterations
    assert rcg == (ewg.capi.EWG_NUMERICAL_ISSUE if case == "max_error"
                   else ewg.capi.EWG_NOT_CONVERGED)
    assert abs(eg - eo) <= 1e-10 * abs(eo)
