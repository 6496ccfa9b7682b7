"""CPU FP64 restatement of the warm-start network (test infrastructure, like the rest of
oracle/: only tests/, smoke() and the benchmark's CPU leg use it).

The network has no counterpart in the reference (eWoms starts Newton from the previous
time level, newtonmethod.hh:308-517); this file fixes what the HIP kernels of
ewoms-numerics_amd/csrc/ewg_cnn.hip compute, in plain numpy float64:

    h0 = relu(conv0(f) + b0)                       5 -> 32 channels
    h_l = h_{l-1} + relu(conv_l(h_{l-1}) + b_l)    l = 1..3, 32 -> 32
    y = conv4(h3) + b4                             32 -> 2
    loss = 1/(2 n) sum_cells sum_c (y - t)^2

every conv 3x3x3 with zero padding 1, as torch.nn.Conv3d (cross-correlation), on
channels-last arrays [batch][z][y][x][channel]. With bf16=True the values the GPU keeps in
BF16 are rounded the same way (round to nearest even): the weights, the hidden states h_l,
the gradient of the loss with respect to y, and the masked gradients entering each
layer's weight gradient and backward-data product; everything else is exact here (the
GPU accumulates in FP32).

Pinned by tests/test_cnn_ref.py: the convolution against scipy.signal.correlate, the
backward pass against central finite differences of the FP64 loss, the layout of the flat
parameter array against its documented order.
"""
from __future__ import annotations

import numpy as np

LAYERS = [(5, 32), (32, 32), (32, 32), (32, 32), (32, 2)]  # (in, out) channels
NUM_PARAMS = sum(ci * co * 27 + co for ci, co in LAYERS)   # 89,122


def unpack(params):
    """Flat FP32/FP64 parameter array -> [(W [out][in][3][3][3], b [out])] per layer."""
    params = np.asarray(params, dtype=np.float64)
    assert params.size == NUM_PARAMS
    out, o = [], 0
    for ci, co in LAYERS:
        W = params[o:o + co * ci * 27].reshape(co, ci, 3, 3, 3)
        o += co * ci * 27
        b = params[o:o + co]
        o += co
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def init_params(seed=0, scale=1.0):
    """He-normal weights (std sqrt(2 / (27 in))) times `scale`, zero biases; the output
    layer starts at a tenth of that."""
    rng = np.random.default_rng(seed)
    layers = []
    for li, (ci, co) in enumerate(LAYERS):
        std = scale * np.sqrt(2.0 / (27 * ci)) * (0.1 if li == len(LAYERS) - 1 else 1.0)
        layers.append((rng.normal(0.0, std, (co, ci, 3, 3, 3)), np.zeros(co)))
    return pack(layers).astype(np.float32)


def bf16_round(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _pad(x):
    return np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))


def conv3d(x, W, b=None):
    """out[..., o] = sum_{k, i} x[z+kz-1, y+ky-1, x+kx-1, i] W[o, i, kz, ky, kx] + b[o]."""
    B, nz, ny, nx, _ = x.shape
    xp = _pad(x)
    out = np.zeros((B, nz, ny, nx, W.shape[0]))
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                out += xp[:, kz:kz + nz, ky:ky + ny, kx:kx + nx, :] @ W[:, :, kz, ky, kx].T
    if b is not None:
        out += b
    return out


def conv3d_T(g, W):
    """Gradient of conv3d with respect to its input: gx[..., i] =
    sum_{k, o} g[z-kz+1, y-ky+1, x-kx+1, o] W[o, i, kz, ky, kx]."""
    B, nz, ny, nx, _ = g.shape
    gp = _pad(g)
    out = np.zeros((B, nz, ny, nx, W.shape[1]))
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                out += gp[:, 2 - kz:2 - kz + nz, 2 - ky:2 - ky + ny, 2 - kx:2 - kx + nx, :] @ \
                    W[:, :, kz, ky, kx]
    return out


def conv3d_wgrad(g, x):
    """dW[o, i, kz, ky, kx] = sum_cells g[cell][o] x[cell + (kz-1, ky-1, kx-1)][i]."""
    B, nz, ny, nx, co = g.shape
    xp = _pad(x)
    gf = g.reshape(-1, co)
    dW = np.zeros((co, x.shape[-1], 3, 3, 3))
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                xs = xp[:, kz:kz + nz, ky:ky + ny, kx:kx + nx, :].reshape(-1, x.shape[-1])
                dW[:, :, kz, ky, kx] = gf.T @ xs
    return dW


def forward(params, feat, bf16=False):
    """feat [B][nz][ny][nx][>=5] -> (y [B][nz][ny][nx][2], cache)."""
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    layers = [(rnd(W), np.asarray(b, dtype=np.float64)) for W, b in unpack(params)]
    f = rnd(np.asarray(feat, dtype=np.float64)[..., :5])
    W0, b0 = layers[0]
    h = [rnd(np.maximum(conv3d(f, W0, b0), 0.0))]
    u = [None]
    for l in range(1, 4):
        W, b = layers[l]
        ul = np.maximum(conv3d(h[l - 1], W, b), 0.0)
        u.append(ul)
        h.append(rnd(h[l - 1] + ul))
    W4, b4 = layers[4]
    y = conv3d(h[3], W4, b4)
    return y, {"f": f, "h": h, "u": u, "layers": layers, "bf16": bf16}


def loss_and_grad(params, feat, target, bf16=False):
    """(loss, flat gradient) of 1/(2 n) sum (y - t)^2, n = cells of the batch."""
    y, c = forward(params, feat, bf16)
    rnd = bf16_round if bf16 else (lambda a: a)
    n = y.shape[0] * y.shape[1] * y.shape[2] * y.shape[3]
    d = y - np.asarray(target, dtype=np.float64)
    loss = 0.5 * float(np.sum(d * d)) / n
    layers = c["layers"]
    grads = [None] * 5
    gy = rnd(d / n)
    grads[4] = (conv3d_wgrad(gy, c["h"][3]), gy.reshape(-1, 2).sum(axis=0))
    gh = conv3d_T(gy, layers[4][0])
    for l in range(3, 0, -1):
        gz = rnd(gh * (c["u"][l] > 0))
        grads[l] = (conv3d_wgrad(gz, c["h"][l - 1]), gz.reshape(-1, 32).sum(axis=0))
        gh = gh + conv3d_T(gz, layers[l][0])
    gz = rnd(gh * (c["h"][0] > 0))
    grads[0] = (conv3d_wgrad(gz, c["f"]), gz.reshape(-1, 32).sum(axis=0))
    return loss, pack(grads)


def features(x, K, phi, dt, norm):
    """Network inputs of one state, as the library's feature kernel forms them (before
    the BF16 rounding): ((p_w - p_ref)/p_scale, S_n, (ln K - lnk_ref)/lnk_scale,
    (phi - phi_ref)/phi_scale, dt/dt_scale) per cell."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 2)
    return np.stack([(x[:, 0] - norm["p_ref"]) / norm["p_scale"], 1.0 - x[:, 1],
                     (np.log(K) - norm["lnk_ref"]) / norm["lnk_scale"],
                     (np.asarray(phi) - norm["phi_ref"]) / norm["phi_scale"],
                     np.full(x.shape[0], dt / norm["dt_scale"])], axis=1)
