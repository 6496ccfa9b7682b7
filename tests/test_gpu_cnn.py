"""GPU: the warm-start network's kernels (BF16 operands, FP32 accumulation on the matrix
cores) against the CPU FP64 restatement oracle/cnn_ref.py, its training, and its use as
the Newton initial guess.

Bars: against the reference that rounds to BF16 where the GPU stores BF16, outputs within
2e-3 and gradients within 1e-2 (norm-wise per layer, FP32 summation order); against plain
FP64 (BF16-rounded weights and inputs only) within 3e-2 and 5e-2.
"""
import numpy as np
import pytest

from oracle import cnn_ref as R

pytestmark = pytest.mark.gpu

SHAPE = (16, 6, 5)  # nx, ny, nz of one sample


def _setup(batch, seed):
    from ewg.cnn import DeviceArray, WarmStartNet, bf16_bits
    nx, ny, nz = SHAPE
    rng = np.random.default_rng(seed)
    params = R.init_params(seed=seed, scale=2.0)
    params = (params + rng.normal(0.0, 0.05, params.size)).astype(np.float32)
    feat = np.zeros((batch, nz, ny, nx, 8))
    feat[..., :5] = rng.normal(size=(batch, nz, ny, nx, 5))
    net = WarmStartNet(nx, ny, nz, max_batch=batch, params=params)
    fdev = DeviceArray.from_numpy(bf16_bits(feat).reshape(batch, -1))
    return net, params, R.bf16_round(feat), fdev, rng


def _layer_slices():
    out, o = [], 0
    for li, (ci, co) in enumerate(R.LAYERS):
        out.append((f"W{li}", slice(o, o + ci * co * 27)))
        o += ci * co * 27
        out.append((f"b{li}", slice(o, o + co)))
        o += co
    return out


def test_forward_matches_cpu_reference():
    net, params, feat, fdev, _ = _setup(2, 11)
    print(f"forward {SHAPE} batch 2", flush=True)
    y = net.forward(fdev, 2).numpy().reshape(feat.shape[:-1] + (2,))
    y16, _ = R.forward(params, feat, bf16=True)
    y64, _ = R.forward(R.bf16_round(params), feat, bf16=False)
    scale = np.abs(y64).max()
    e16 = np.abs(y - y16).max() / scale
    e64 = np.abs(y - y64).max() / scale
    print(f"forward: vs bf16-emulating reference {e16:.2e}, vs FP64 {e64:.2e}", flush=True)
    assert e16 < 2e-3 and e64 < 3e-2


def test_gradients_match_cpu_reference():
    from ewg.cnn import DeviceArray
    net, params, feat, fdev, rng = _setup(2, 12)
    print(f"gradients {SHAPE} batch 2", flush=True)
    target = rng.normal(size=feat.shape[:-1] + (2,)).astype(np.float32)
    tdev = DeviceArray.from_numpy(target.reshape(2, -1))
    loss, g = net.gradients(fdev, tdev, 2)
    l16, g16 = R.loss_and_grad(params, feat, target, bf16=True)
    l64, g64 = R.loss_and_grad(R.bf16_round(params), feat, target, bf16=False)
    assert abs(loss - l16) < 1e-3 * l16 and abs(loss - l64) < 3e-2 * l64
    for name, sl in _layer_slices():
        n16 = np.linalg.norm(g16[sl])
        e16 = np.linalg.norm(g[sl] - g16[sl]) / n16
        e64 = np.linalg.norm(g[sl] - g64[sl]) / np.linalg.norm(g64[sl])
        print(f"{name}: |g| {n16:.3e}  vs bf16-emulating {e16:.2e}  vs FP64 {e64:.2e}",
              flush=True)
        assert e16 < 1e-2 and e64 < 5e-2, name


def test_crop_inference_matches_whole_grid_inside():
    """The warm start evaluates the network on a box of the grid (the active region grown by
    8 cells): with zero padding at the box faces, every cell at least 5 cells (the reach of
    the five 3x3x3 layers) inside a face that is not the grid's own gets the whole-grid
    output bit for bit, since each output's accumulation does not depend on the other
    cells of its tile."""
    from ewg.cnn import DeviceArray, WarmStartNet, bf16_bits
    nx, ny, nz = 64, 40, 24
    x0, x1, y0, y1, z0, z1 = 16, 48, 0, 30, 6, 24  # the box touches the grid at y- and z+
    rng = np.random.default_rng(21)
    params = (R.init_params(seed=21, scale=2.0) + rng.normal(0.0, 0.05, R.NUM_PARAMS))
    params = params.astype(np.float32)
    feat = np.zeros((nz, ny, nx, 8))
    feat[..., :5] = rng.normal(size=(nz, ny, nx, 5))
    full = WarmStartNet(nx, ny, nz, max_batch=1, params=params)
    y_full = full.forward(DeviceArray.from_numpy(bf16_bits(feat).reshape(1, -1)), 1)
    y_full = y_full.numpy().reshape(nz, ny, nx, 2)
    fc = np.ascontiguousarray(feat[z0:z1, y0:y1, x0:x1])
    crop = WarmStartNet(x1 - x0, y1 - y0, z1 - z0, max_batch=1, params=params)
    y_crop = crop.forward(DeviceArray.from_numpy(bf16_bits(fc).reshape(1, -1)), 1)
    y_crop = y_crop.numpy().reshape(z1 - z0, y1 - y0, x1 - x0, 2)
    # inside: 5 cells from the box's interior faces (none at y- and z+, the grid's faces)
    zi, yi, xi = slice(5, z1 - z0), slice(0, y1 - y0 - 5), slice(5, x1 - x0 - 5)
    box_full = y_full[z0:z1, y0:y1, x0:x1]
    inner_crop, inner_full = y_crop[zi, yi, xi], box_full[zi, yi, xi]
    print(f"crop {x1 - x0}x{y1 - y0}x{z1 - z0} of {nx}x{ny}x{nz}: inner {inner_crop.shape[:3]}, "
          f"max |diff| {np.abs(inner_crop - inner_full).max():.3e}; whole box "
          f"{np.abs(y_crop - box_full).max():.3e}", flush=True)
    assert np.array_equal(inner_crop, inner_full)
    # the zero padding does change the outputs next to the box's interior faces
    assert not np.array_equal(y_crop, box_full)


def test_training_fits_a_teacher_network():
    """Adam with the library's kernels drives the loss on a fixed batch down: targets are
    the FP64 outputs of another network of the same shape."""
    from ewg.cnn import DeviceArray
    net, _, feat, fdev, _ = _setup(4, 13)
    teacher = R.init_params(seed=99, scale=2.0)
    target = R.forward(teacher, feat)[0].astype(np.float32)
    tdev = DeviceArray.from_numpy(target.reshape(4, -1))
    net.set_params(R.init_params(seed=5))
    losses = [net.train_step(fdev, tdev, 4, 3e-3) for _ in range(150)]
    print(f"teacher fit: loss {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
    assert losses[-1] < 0.2 * losses[0]


def _member(steps):
    from ewg import problems
    return problems.heterogeneous_3d(32, 32, 16, seed=4242, sigma_lnK=1.0, corr_cells=6.0,
                                     rate=1e-3, patch=4, steps=steps)


def test_zero_network_reproduces_the_plain_newton_path():
    """A network that predicts no update gives the plain guess; the storage of the old
    level then comes from solution(1) instead of the first iteration: the same values, so
    the run is bit-identical to the plain one."""
    import ewg
    from ewg import cnn
    P = _member(4)
    plain = ewg.GpuContext(P, ewg.default_params(newton_tolerance=1e-6))
    warm = ewg.GpuContext(P, ewg.default_params(newton_tolerance=1e-6, warm_start=1))
    net = cnn.WarmStartNet(P.nx, P.ny, P.nz)
    net.set_norm({"p_ref": 1e5, "p_scale": 1e5, "lnk_ref": np.log(1e-12), "lnk_scale": 1.0,
                  "phi_ref": 0.25, "phi_scale": 0.02, "dt_scale": 500.0, "dp_scale": 1.0,
                  "ds_scale": 1.0})
    cnn.attach(warm, net)
    for step in range(4):
        a = plain.newton(P.dt)
        b = warm.newton(P.dt)
        print(f"step {step}: plain {a}, warm {b}", flush=True)
        assert a[0] == 0 and b[0] == 0 and a[1] == b[1]
        plain.advance_time_level()
        warm.advance_time_level()
    assert np.array_equal(plain.solution(0), warm.solution(0))
    st = cnn.warm_stats(warm)
    assert st["accepted"] == 4 and st["rejected"] == 0
    # the four predictions are timed by events around them on the stream; a reset clears
    # the total, including the last prediction's pending time
    assert 0.0 < st["predict_s"] < 10.0
    warm.reset_stats()
    assert cnn.warm_stats(warm)["predict_s"] == 0.0


def test_rejected_prediction_falls_back_to_the_plain_guess():
    """A prediction whose residual exceeds NewtonMaxError makes the solve fail; the step
    is redone from the plain guess and matches the plain run bit for bit. (Here: initial
    errors 0.06 and 1e-3 from the plain guess, 3.9 with 2e5 Pa added everywhere.)"""
    import ewg
    from ewg import cnn
    P = _member(3)
    prm = dict(newton_tolerance=1e-6, newton_max_error=0.5)
    plain = ewg.GpuContext(P, ewg.default_params(**prm))
    warm = ewg.GpuContext(P, ewg.default_params(warm_start=1, **prm))
    net = cnn.WarmStartNet(P.nx, P.ny, P.nz)
    params = np.zeros(cnn.NUM_PARAMS, dtype=np.float32)
    params[89120] = 1.0   # output bias: dp = 1 * dp_scale everywhere
    net.set_params(params)
    net.set_norm({"p_ref": 1e5, "p_scale": 1e5, "lnk_ref": np.log(1e-12), "lnk_scale": 1.0,
                  "phi_ref": 0.25, "phi_scale": 0.02, "dt_scale": 500.0, "dp_scale": 2e5,
                  "ds_scale": 1.0})
    cnn.attach(warm, net)
    for step in range(3):
        a = plain.newton(P.dt)
        b = warm.newton(P.dt)
        print(f"step {step}: plain {a}, warm {b}", flush=True)
        assert a[0] == 0 and b[0] == 0 and a[1] == b[1]
        plain.advance_time_level()
        warm.advance_time_level()
    assert np.array_equal(plain.solution(0), warm.solution(0))
    st = cnn.warm_stats(warm)
    assert st["accepted"] == 0 and st["rejected"] == 3


def test_trained_warm_start_on_a_held_out_member():
    """Train on pairs harvested from plain runs of ensemble members, then step a held-out
    member: each step is taken by both contexts from the same state (the plain run's), the
    one from the prediction, the other from the plain guess. Every step converges from the
    prediction (none is rejected), none takes more Newton iterations than from the plain
    guess, and the two results of a step agree within 1e-6 in saturation."""
    import ewg
    from ewg import cnn
    log = lambda m: print(m, flush=True)
    P = _member(12)
    pred, info = cnn.train_warm_start(P.nx, P.ny, P.nz, n_pairs=400, epochs=40, seed=3,
                                      steps_per_member=20)
    log(f"training: {info}")
    assert info["loss_last"] < 0.1 * info["loss_first"]
    plain = ewg.GpuContext(P, ewg.default_params(newton_tolerance=1e-6))
    warm = ewg.GpuContext(P, ewg.default_params(newton_tolerance=1e-6, warm_start=1))
    cnn.attach(warm, pred)
    for step in range(12):
        x = plain.solution(0)
        warm.set_solution(x, 0)
        warm.set_solution(x, 1)
        a = plain.newton(P.dt)
        e_plain = cnn.warm_stats(plain)["last_initial_error"]
        b = warm.newton(P.dt)
        e_warm = cnn.warm_stats(warm)["last_initial_error"]
        dS = np.abs(plain.solution(0)[1::2] - warm.solution(0)[1::2]).max()
        plain.advance_time_level()
        log(f"step {step}: newton plain {a[1]} warm {b[1]}; initial error plain {e_plain:.3e} "
            f"warm {e_warm:.3e}; max |dS| vs plain {dS:.2e}")
        assert a[0] == 0 and b[0] == 0
        assert b[1] <= a[1]
        assert dS < 1e-6
    st = cnn.warm_stats(warm)
    log(f"warm start: {st}")
    assert st["accepted"] == 12 and st["rejected"] == 0


def test_trained_warm_start_saves_work_on_the_1m_cell_case():
    """The 128x128x64 heterogeneous case, 20 steps of 500 s at Newton tolerance 1e-6 with
    the FP32 preconditioner and face fluxes, each step taken from the plain run's state
    twice: from the plain guess and from the network's (trained on 2000 pairs of other
    realizations). The prediction removes the steps whose second Newton iteration just
    misses the tolerance, and with them most of the BiCGStab work after the first step;
    every step's saturations agree with the plain step's within 1e-6 (the two solves stop
    at different iterates inside the Newton tolerance; over whole trajectories these
    differences add up, which the benchmark reports separately)."""
    import ewg
    from ewg import cnn, problems
    log = lambda m: print(m, flush=True)
    P = problems.heterogeneous_3d(128, 128, 64)
    pred, info = cnn.train_warm_start(P.nx, P.ny, P.nz, n_pairs=2000, epochs=30, seed=2)
    log(f"training: {info}")
    ctxs = {}
    for form in ("plain", "warm"):
        prm = ewg.default_params(newton_tolerance=1e-6, max_time_step_size=P.dt,
                                 precond_fp32=1, linearize_fp32=1,
                                 warm_start=1 if form == "warm" else 0)
        ctxs[form] = ewg.GpuContext(P, prm)
    plain, warm = ctxs["plain"], ctxs["warm"]
    cnn.attach(warm, pred)
    counts = {"plain": ([], []), "warm": ([], [])}
    worst = 0.0
    for step in range(20):
        x = plain.solution(0)
        warm.set_solution(x, 0)
        warm.set_solution(x, 1)
        for form, ctx in ctxs.items():
            k0 = ctx.stats()["linear_iterations"]
            rc, it, err = ctx.newton(P.dt)
            assert rc == 0
            counts[form][0].append(it)
            counts[form][1].append(ctx.stats()["linear_iterations"] - k0)
        dS = np.abs(plain.solution(0)[1::2] - warm.solution(0)[1::2]).max()
        worst = max(worst, dS)
        log(f"step {step}: newton {counts['plain'][0][-1]} / {counts['warm'][0][-1]}, "
            f"bicgstab {counts['plain'][1][-1]} / {counts['warm'][1][-1]}, max |dS| {dS:.2e}")
        assert dS < 1e-6
        plain.advance_time_level()
    st = cnn.warm_stats(warm)
    log(f"warm start: {st}")
    assert st["accepted"] == 20 and st["rejected"] == 0
    for ctx in ctxs.values():
        ctx.close()
    np_, kp = sum(counts["plain"][0]), sum(counts["plain"][1])
    nw, kw = sum(counts["warm"][0]), sum(counts["warm"][1])
    log(f"newton {np_} -> {nw}, bicgstab {kp} -> {kw}, max |dS| of a step {worst:.2e}")
    assert nw <= np_ and kw < 0.75 * kp
