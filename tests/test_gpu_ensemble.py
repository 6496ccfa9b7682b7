"""GPU: the ensemble solver (include/ewg_ensemble.h) against single contexts and the oracle.

Each member of an ensemble must reproduce what its own context does with the same
parameters: the same Newton iterations and the same states up to the summation order of
the BiCGStab dot products (saturations within 1e-7: BiCGStab iterates that differ in the
last bits can stop one iteration apart, within the Newton tolerance); with the FP64
preconditioner and Newton tolerance 1e-8 the members match the CPU oracle like the single
context does (identical Newton iterations, saturations within 1e-8).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(16, 12, 8, 31), (24, 16, 8, 32), (32, 8, 6, 33), (16, 16, 16, 34)]


def _problems(steps=3):
    from ewg import problems
    return [problems.heterogeneous_3d(nx, ny, nz, seed=s, patch=4, steps=steps)
            for nx, ny, nz, s in SHAPES]


def _single_runs(probs, params, steps):
    import ewg
    out = []
    for P in probs:
        c = ewg.GpuContext(P, params)
        its = 0
        for _ in range(steps):
            rc, it, _ = c.newton(P.dt)
            assert rc == 0
            its += it
            c.advance_time_level()
        out.append((its, c.stats()["linear_iterations"], c.solution(0)))
        c.close()
    return out


@pytest.mark.parametrize("mode,schedule", [("batched", 0), ("batched", 1), ("persistent", 0)])
def test_members_match_single_contexts(mode, schedule):
    """schedule 0: the batched mode's tile-wavefront sweeps; 1: its hyperplane sweeps."""
    import ewg
    from ewg import ensemble
    probs = _problems()
    prm = ensemble.member_params(ilu_schedule=schedule)
    ctxs = [ewg.GpuContext(P, prm) for P in probs]
    ens = ensemble.Ensemble(ctxs)
    ens.set_mode(ensemble.BATCHED if mode == "batched" else ensemble.PERSISTENT, 512)
    print(mode, "schedule", schedule, "ensemble of", [(P.nx, P.ny, P.nz) for P in probs],
          flush=True)
    sec = ens.run(probs[0].dt, 3)
    ref = _single_runs(probs, prm, 3)
    for i, P in enumerate(probs):
        st = ens.stats(i)
        its, lits, x = ref[i]
        xe = ctxs[i].solution(0)
        dS = np.abs(xe[1::2] - x[1::2]).max()
        print(f"member {i}: {st}; single: newton {its} bicgstab {lits}; max |dS| {dS:.2e}",
              flush=True)
        assert st["status"] == 0 and st["steps"] == 3
        assert st["newton_iterations"] == its
        assert abs(st["linear_iterations"] - lits) <= 2
        assert dS < 1e-7
    print(f"3 steps in {sec:.3f} s", flush=True)


@pytest.mark.parametrize("mode,schedule", [("batched", 0), ("batched", 1), ("persistent", 0)])
def test_fp64_members_match_the_oracle(mode, schedule):
    import ewg
    from ewg import ensemble
    from oracle.pyoracle import OracleContext
    probs = _problems()[:2]
    prm = ensemble.member_params(precond_fp32=0, newton_tolerance=1e-8, ilu_schedule=schedule)
    ctxs = [ewg.GpuContext(P, prm) for P in probs]
    ens = ensemble.Ensemble(ctxs)
    ens.set_mode(ensemble.BATCHED if mode == "batched" else ensemble.PERSISTENT, 512)
    ens.run(probs[0].dt, 3)
    for i, P in enumerate(probs):
        o = OracleContext(P, ewg.default_params(newton_tolerance=1e-8))
        its = 0
        for _ in range(3):
            rc, it, _ = o.newton(P.dt)
            assert rc == 0
            its += it
            o.advance_time_level()
        st = ens.stats(i)
        dS = np.abs(ctxs[i].solution(0)[1::2] - o.solution(0)[1::2]).max()
        print(f"member {i}: ensemble newton {st['newton_iterations']}, oracle {its}; "
              f"max |dS| {dS:.2e}", flush=True)
        assert st["newton_iterations"] == its
        assert dS < 1e-8


def test_benchmark_size_members_match_the_oracle():
    """Members of the benchmark ensemble's largest and smallest sizes (96x96x48: 144 tiles
    of the wavefront, and 32x32x32) in one batched ensemble, FP64 preconditioner, Newton
    tolerance 1e-8, two steps: identical Newton iterations and saturations within 1e-8 of
    the oracle, as the small members."""
    import ewg
    from ewg import ensemble, problems
    from oracle.pyoracle import OracleContext
    probs = [problems.ensemble_member(0, 7, sizes=((96, 96, 48),), steps=2),
             problems.ensemble_member(1, 7, sizes=((32, 32, 32),), steps=2)]
    prm = ensemble.member_params(precond_fp32=0, newton_tolerance=1e-8)
    ctxs = [ewg.GpuContext(P, prm) for P in probs]
    ens = ensemble.Ensemble(ctxs)
    ens.set_mode(ensemble.BATCHED, 0)
    print("ensemble of", [(P.nx, P.ny, P.nz) for P in probs], flush=True)
    ens.run(probs[0].dt, 2)
    for i, P in enumerate(probs):
        o = OracleContext(P, ewg.default_params(newton_tolerance=1e-8))
        its = 0
        for _ in range(2):
            rc, it, _ = o.newton(P.dt)
            assert rc == 0
            its += it
            o.advance_time_level()
        st = ens.stats(i)
        dS = np.abs(ctxs[i].solution(0)[1::2] - o.solution(0)[1::2]).max()
        print(f"member {i} {(P.nx, P.ny, P.nz)}: ensemble newton {st['newton_iterations']}, "
              f"oracle {its}; max |dS| {dS:.2e}", flush=True)
        assert st["status"] == 0 and st["steps"] == 2
        assert st["newton_iterations"] == its
        assert dS < 1e-8


@pytest.mark.parametrize("mode,schedule", [("batched", 0), ("batched", 1), ("persistent", 0)])
def test_fp32_preconditioner_members_close_to_the_fp64_oracle(mode, schedule):
    """The ensemble's own form (FP32 ILU(0) inside FP64 BiCGStab, Newton tolerance 1e-6)
    against the FP64 oracle at the same tolerance, step by step: saturations within 1e-6
    and Newton iterations within one of the oracle's in every step."""
    import ewg
    from ewg import ensemble
    from oracle.pyoracle import OracleContext
    probs = _problems()[:3]
    prm = ensemble.member_params(precond_fp32=1, newton_tolerance=1e-6, ilu_schedule=schedule)
    ctxs = [ewg.GpuContext(P, prm) for P in probs]
    ens = ensemble.Ensemble(ctxs)
    ens.set_mode(ensemble.BATCHED if mode == "batched" else ensemble.PERSISTENT, 512)
    oracles = [OracleContext(P, ewg.default_params(newton_tolerance=1e-6)) for P in probs]
    for step in range(3):
        before = [ens.stats(i)["newton_iterations"] for i in range(len(probs))]
        ens.run(probs[0].dt, 1)
        for i, (P, o) in enumerate(zip(probs, oracles)):
            rc, it, _ = o.newton(P.dt)
            assert rc == 0
            o.advance_time_level()
            st = ens.stats(i)
            ie = st["newton_iterations"] - before[i]
            dS = np.abs(ctxs[i].solution(0)[1::2] - o.solution(0)[1::2]).max()
            print(f"{mode} schedule {schedule} step {step} member {i}: newton {ie} vs oracle "
                  f"{it}; max |dS| {dS:.2e}", flush=True)
            assert st["status"] == 0
            assert abs(ie - it) <= 1
            assert dS < 1e-6


@pytest.mark.parametrize("mode", ["batched", "persistent"])
def test_failing_member_is_isolated(mode):
    """A member whose Newton solve fails (here: NewtonMaxError below its first error) is
    retried with halved steps (two halvings allowed here), then keeps its state and
    stops; the others are unaffected."""
    import ewg
    from ewg import ensemble
    probs = _problems()
    prm = ensemble.member_params()
    bad = ensemble.member_params(newton_max_error=1e-9, max_time_step_divisions=2)
    ctxs = [ewg.GpuContext(P, bad if i == 1 else prm) for i, P in enumerate(probs)]
    x_before = ctxs[1].solution(0)
    ens = ensemble.Ensemble(ctxs)
    ens.set_mode(ensemble.BATCHED if mode == "batched" else ensemble.PERSISTENT, 512)
    ens.run(probs[0].dt, 2)
    ref = _single_runs([probs[0], probs[2]], prm, 2)
    st = [ens.stats(i) for i in range(4)]
    print(st, flush=True)
    assert st[1]["status"] == ewg.capi.EWG_NUMERICAL_ISSUE and st[1]["steps"] == 0
    assert st[1]["step_divisions"] == 2 and st[1]["newton_iterations"] == 3
    assert np.array_equal(ctxs[1].solution(0), x_before)
    for i, (its, _, x) in zip((0, 2), ref):
        assert st[i]["status"] == 0 and st[i]["steps"] == 2 and st[i]["step_divisions"] == 0
        assert st[i]["newton_iterations"] == its
        assert np.abs(ctxs[i].solution(0)[1::2] - x[1::2]).max() < 1e-7


@pytest.mark.parametrize("mode", ["batched", "persistent"])
def test_tuner_and_graph_replay(mode):
    """The tuner's trial steps leave the state as it was; one run of three steps and
    three runs of one step (graph replays; the same launch shape) give identical
    states."""
    import ewg
    from ewg import ensemble
    probs = _problems()
    prm = ensemble.member_params()
    t_ctx = [ewg.GpuContext(P, prm) for P in probs]
    et = ensemble.Ensemble(t_ctx)
    x0 = [c.solution(0) for c in t_ctx]
    t = et.tune(probs[0].dt)
    print("tuning", t, flush=True)
    assert t["mode"] in ("batched", "persistent")
    assert all(c["ms"] > 0 for c in t["candidates"])
    assert sorted({c["replay"] for c in t["candidates"] if c["mode"] == "batched"}) == [4, 8, 16, 32]
    assert all(np.array_equal(c.solution(0), x) for c, x in zip(t_ctx, x0))
    assert all(et.stats(i)["steps"] == 0 for i in range(4))
    tb = et.tune(probs[0].dt, modes=ensemble.TUNE_BATCHED)
    assert tb["mode"] == "batched" and tb["replay"] in (4, 8, 16, 32)
    assert all((c["ms"] > 0) == (c["mode"] == "batched") for c in tb["candidates"])
    assert all(np.array_equal(c.solution(0), x) for c, x in zip(t_ctx, x0))
    a = [ewg.GpuContext(P, prm) for P in probs]
    b = [ewg.GpuContext(P, prm) for P in probs]
    ea, eb = ensemble.Ensemble(a), ensemble.Ensemble(b)
    m = ensemble.BATCHED if mode == "batched" else ensemble.PERSISTENT
    ea.set_mode(m, 512)
    eb.set_mode(m, 512)
    if mode == "batched":
        # the replay depth changes only what is queued past convergence
        ea.set_replay(2)
        eb.set_replay(32)
    ea.run(probs[0].dt, 3)
    for _ in range(3):
        eb.run(probs[0].dt, 1)
    for i in range(4):
        assert ea.stats(i)["steps"] == eb.stats(i)["steps"] == 3
        assert np.array_equal(a[i].solution(0), b[i].solution(0))


def test_forced_sweep_failure_is_isolated_and_recovers():
    """The host makes member 1's first batched solve start with its sweep error flag set
    (as if a wavefront wait had run out; every wait stays bounded): that solve breaks
    down, the member's step is retried as two half steps after its tile edge buffers are
    reset, and it then matches a single context stepped the same way; the other members
    match their single runs exactly as without the failure."""
    import ewg
    from ewg import ensemble
    probs = _problems()
    prm = ensemble.member_params(precond_fp32=1)
    ctxs = [ewg.GpuContext(P, prm) for P in probs]
    ens = ensemble.Ensemble(ctxs)
    ens.set_mode(ensemble.BATCHED, 512)
    ens.force_sweep_failure(1, 1)
    dt = probs[0].dt
    ens.run(dt, 2)
    st = [ens.stats(i) for i in range(4)]
    print(st, flush=True)
    ref = _single_runs([probs[0], probs[2], probs[3]], prm, 2)
    for i, (its, _, x) in zip((0, 2, 3), ref):
        assert st[i]["status"] == 0 and st[i]["steps"] == 2 and st[i]["step_divisions"] == 0
        assert st[i]["newton_iterations"] == its
        assert np.abs(ctxs[i].solution(0)[1::2] - x[1::2]).max() < 1e-7
    # member 1: the failed attempt's linearization, then dt/2 twice, then dt
    c = ewg.GpuContext(probs[1], prm)
    its = 1
    for h in (dt / 2, dt / 2, dt):
        rc, it, _ = c.newton(h)
        assert rc == 0
        its += it
        c.advance_time_level()
    assert st[1]["status"] == 0 and st[1]["steps"] == 2 and st[1]["step_divisions"] == 1
    assert st[1]["newton_iterations"] == its
    dS = np.abs(ctxs[1].solution(0)[1::2] - c.solution(0)[1::2]).max()
    print(f"member 1 vs single context with the halved first step: max |dS| {dS:.2e}",
          flush=True)
    assert dS < 1e-7


def test_members_with_the_fp32_operator_are_refused():
    """The ensemble's operator products read each member's FP64 Jacobian: a member set up
    for the FP32 operator copy (params.operator_fp32) is refused at creation."""
    import ewg
    from ewg import ensemble
    probs = _problems()[:2]
    ctxs = [ewg.GpuContext(probs[0], ensemble.member_params()),
            ewg.GpuContext(probs[1], ensemble.member_params(operator_fp32=1))]
    with pytest.raises(RuntimeError, match=f"failed: {ewg.capi.EWG_INVALID_ARGUMENT}"):
        ensemble.Ensemble(ctxs)
    for c in ctxs:
        c.close()


def _rough_member():
    """Realization 21 of the seed-3 ensemble family (96x96x48, sigma_lnK 1.87, correlation
    length 4.7 cells): with FP32 factors one of its 40 solves fails (BiCGStab stalls at
    1,000 iterations), with FP64 factors none does."""
    from ewg import problems
    return problems.ensemble_member(21, 3, steps=40)


def test_fp64_retry_of_a_failed_fp32_solve_single_context():
    """precond_fp32 = 2: the solve that fails with FP32 factors is redone with FP64 factors
    (kept for the later solves), and the 40 steps need no halving; with precond_fp32 = 1
    the same failure halves a step."""
    import ewg
    from ewg import ensemble
    P = _rough_member()
    out = {}
    for level in (1, 2):
        c = ewg.GpuContext(P, ensemble.member_params(max_time_step_size=P.dt, precond_fp32=level))
        n = c.run(0.0, 40 * P.dt, P.dt)
        out[level] = (n, c.stats())
        print(f"precond_fp32={level}: {n} steps, {out[level][1]}", flush=True)
        c.close()
    assert out[1][1]["failed_steps"] >= 1 and out[1][1]["precond_retries"] == 0
    assert out[2][0] == 40 and out[2][1]["failed_steps"] == 0
    assert out[2][1]["precond_retries"] == 1


@pytest.mark.parametrize("schedule", [0, 1])
def test_fp64_retry_in_the_batched_ensemble(schedule):
    """A member whose FP32-preconditioned solve fails redoes it, and its later solves, with
    FP64 factors (the FP64 wavefront batches, or its hyperplane factors unrounded): no
    halving, and it stays within the Newton tolerance's reach of a single context with
    precond_fp32 = 2 (their BiCGStab dot products sum in different orders, and on this
    field the FP32-preconditioned iterations amplify that); the other members match their
    single contexts as without retries."""
    import ewg
    from ewg import ensemble
    steps = 40
    probs = [_rough_member()] + _problems(steps=steps)[:2]
    prm = ensemble.member_params(precond_fp32=2, ilu_schedule=schedule)
    ctxs = [ewg.GpuContext(P, prm) for P in probs]
    ens = ensemble.Ensemble(ctxs)
    ens.set_mode(ensemble.BATCHED, 512)
    ens.run(probs[0].dt, steps)
    st = [ens.stats(i) for i in range(len(probs))]
    print(st, flush=True)
    ref = _single_runs(probs, prm, steps)
    for i, (its, kits, x) in enumerate(ref):
        dS = np.abs(ctxs[i].solution(0)[1::2] - x[1::2]).max()
        print(f"member {i}: newton {st[i]['newton_iterations']} vs {its}, bicgstab "
              f"{st[i]['linear_iterations']} vs {kits}, max |dS| {dS:.2e}", flush=True)
        assert st[i]["status"] == 0 and st[i]["steps"] == steps and st[i]["step_divisions"] == 0
        if i == 0:
            assert abs(st[i]["newton_iterations"] - its) <= 3
            assert dS < 1e-5
        else:
            assert st[i]["newton_iterations"] == its
            assert abs(st[i]["linear_iterations"] - kits) <= 2
            assert dS < 1e-7
    assert st[0]["precond_retries"] == 1
    assert st[1]["precond_retries"] == 0 and st[2]["precond_retries"] == 0


def test_forced_sweep_failure_is_redone_with_fp64_factors():
    """With precond_fp32 = 2 the forced breakdown of member 1's first solve is redone with
    FP64 factors (kept for its later solves) instead of halving the step: two whole steps,
    one retry, and the member close to a single context with FP64 factors from the start
    (its later solves are the same; its first solve stops at another iterate)."""
    import ewg
    from ewg import ensemble
    probs = _problems()
    prm = ensemble.member_params()
    ctxs = [ewg.GpuContext(P, prm) for P in probs]
    ens = ensemble.Ensemble(ctxs)
    ens.set_mode(ensemble.BATCHED, 512)
    ens.force_sweep_failure(1, 1)
    ens.run(probs[0].dt, 2)
    st = [ens.stats(i) for i in range(4)]
    print(st, flush=True)
    assert st[1]["status"] == 0 and st[1]["steps"] == 2 and st[1]["step_divisions"] == 0
    assert st[1]["precond_retries"] == 1
    assert all(st[i]["precond_retries"] == 0 for i in (0, 2, 3))
    (its, _, x), = _single_runs([probs[1]], ensemble.member_params(precond_fp32=0), 2)
    dS = np.abs(ctxs[1].solution(0)[1::2] - x[1::2]).max()
    print(f"member 1 vs the FP64-preconditioned single context: max |dS| {dS:.2e}", flush=True)
    assert dS < 1e-6


def test_graph_and_direct_launches_agree():
    """The batched run replaying its iteration graph and launching the iteration's kernels
    directly give identical states and counts."""
    import ewg
    from ewg import ensemble
    probs = _problems()
    out = []
    for graph in (True, False):
        ctxs = [ewg.GpuContext(P, ensemble.member_params()) for P in probs]
        ens = ensemble.Ensemble(ctxs)
        ens.set_mode(ensemble.BATCHED, 512)
        ens.set_graph(graph)
        ens.run(probs[0].dt, 2)
        out.append(([ens.stats(i) for i in range(4)], [c.solution(0) for c in ctxs]))
    for i in range(4):
        assert out[0][0][i] == out[1][0][i]
        assert np.array_equal(out[0][1][i], out[1][1][i])


def test_compact_rounds_agree_with_full_rounds():
    """Rounds launched over the running members only (compact maps, direct launches) and
    over all members (the iteration graph) give identical states and counts: every round
    compact, the default (compact while the running members hold at most half of the
    cells; the members' sizes differ by a factor of 40) and none."""
    import ewg
    from ewg import ensemble, problems
    probs = [problems.heterogeneous_3d(nx, ny, nz, seed=s, patch=4, steps=2)
             for nx, ny, nz, s in ((40, 32, 24, 61), (8, 8, 12, 62), (12, 8, 8, 63), (8, 16, 8, 64))]
    out = []
    for compact in ("always", 1, 0):
        print(f"compact {compact}", flush=True)
        ctxs = [ewg.GpuContext(P, ensemble.member_params()) for P in probs]
        ens = ensemble.Ensemble(ctxs)
        ens.set_mode(ensemble.BATCHED, 512)
        ens.set_compact(compact)
        ens.run(probs[0].dt, 2)
        out.append(([ens.stats(i) for i in range(4)], [c.solution(0) for c in ctxs]))
        assert all(s["status"] == 0 and s["steps"] == 2 for s in out[-1][0])
    for k in (1, 2):
        for i in range(4):
            assert out[0][0][i] == out[k][0][i], (k, i)
            assert np.array_equal(out[0][1][i], out[k][1][i]), (k, i)
