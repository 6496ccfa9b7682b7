"""Pins the CPU FP64 restatement of the warm-start network (oracle/cnn_ref.py), which the
GPU gradient checks compare against: the convolution against scipy's N-d correlation, the
backward pass against central finite differences, BF16 rounding against torch's, and the
flat parameter layout."""
import numpy as np
import pytest

from oracle import cnn_ref as R


def test_parameter_count_and_layout():
    assert R.NUM_PARAMS == 89122
    p = np.arange(R.NUM_PARAMS, dtype=np.float64)
    layers = R.unpack(p)
    # layer by layer: W [out][in][kz][ky][kx] then b [out]
    assert layers[0][0].shape == (32, 5, 3, 3, 3) and layers[0][0][0, 0, 0, 0, 1] == 1
    assert layers[0][1][0] == 32 * 5 * 27
    assert layers[1][0].flat[0] == 32 * 5 * 27 + 32
    assert layers[4][0].shape == (2, 32, 3, 3, 3) and layers[4][1][-1] == R.NUM_PARAMS - 1
    assert np.array_equal(R.pack(layers), p)


def test_conv3d_matches_scipy_correlate():
    from scipy.signal import correlate
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 5, 4, 6, 3))
    W = rng.normal(size=(4, 3, 3, 3, 3))
    b = rng.normal(size=4)
    out = R.conv3d(x, W, b)
    ref = np.zeros_like(out)
    for bb in range(2):
        for o in range(4):
            ref[bb, ..., o] = b[o] + sum(correlate(x[bb, ..., i], W[o, i], mode="same")
                                         for i in range(3))
    assert np.abs(out - ref).max() < 1e-12


def test_conv3d_transpose_is_adjoint():
    """<conv(x), g> = <x, conv_T(g)> and the weight gradient is d<conv(x), g>/dW."""
    rng = np.random.default_rng(2)
    x = rng.normal(size=(2, 3, 4, 5, 3))
    g = rng.normal(size=(2, 3, 4, 5, 4))
    W = rng.normal(size=(4, 3, 3, 3, 3))
    lhs = np.sum(R.conv3d(x, W) * g)
    assert abs(lhs - np.sum(x * R.conv3d_T(g, W))) < 1e-9 * abs(lhs)
    assert abs(lhs - np.sum(W * R.conv3d_wgrad(g, x))) < 1e-9 * abs(lhs)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    params = R.init_params(seed=4, scale=3.0).astype(np.float64)
    params += rng.normal(0.0, 0.05, params.size)  # nonzero biases: mixed relu patterns
    feat = rng.normal(size=(2, 3, 3, 4, 8))
    target = rng.normal(size=(2, 3, 3, 4, 2))
    loss, g = R.loss_and_grad(params, feat, target)
    # the residual branches are active somewhere (else their gradients are trivially 0)
    _, c = R.forward(params, feat)
    assert all(np.any(c["u"][l] > 0) for l in (1, 2, 3))
    idx = np.concatenate([rng.choice(R.NUM_PARAMS, 60, replace=False),
                          [0, 4320, 4351, 32031, 87391, 89120, 89121]])
    h = 1e-6
    worst = 0.0
    for k in idx:
        pp, pm = params.copy(), params.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.loss_and_grad(pp, feat, target)[0] - R.loss_and_grad(pm, feat, target)[0]) / (2 * h)
        worst = max(worst, abs(fd - g[k]) / (abs(fd) + 1e-6 * np.abs(g).max()))
    assert worst < 1e-5, worst


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    a = np.concatenate([rng.normal(size=5000) * 10.0 ** rng.integers(-8, 8, 5000),
                        [0.0, -0.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 65504.0]])
    ours = R.bf16_round(a)
    theirs = torch.tensor(a, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(ours, theirs)


def test_features_formula():
    norm = {"p_ref": 1e5, "p_scale": 2e5, "lnk_ref": np.log(1e-12), "lnk_scale": 1.0,
            "phi_ref": 0.25, "phi_scale": 0.02, "dt_scale": 500.0}
    x = np.array([1e5, 0.8, 3e5, 0.1])
    f = R.features(x, np.array([1e-12, np.e * 1e-12]), np.array([0.25, 0.27]), 250.0, norm)
    assert np.allclose(f, [[0.0, 0.2, 0.0, 0.0, 0.5], [1.0, 0.9, 1.0, 1.0, 0.5]])
