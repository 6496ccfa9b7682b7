"""ewg.cnn -- the learned Newton warm start (include/ewg_cnn.h): binding, training data
harvested from the simulator's own runs, and training with the library's kernels.

Device memory for features, targets and outputs comes from the library itself
(DeviceArray over ewg_device_alloc); every computation is a call into libewg.so.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import load, problems
from .capi import default_params

NUM_PARAMS = 89122
_vp = C.c_void_p


class Norm(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("p_ref", "p_scale", "lnk_ref", "lnk_scale",
                                           "phi_ref", "phi_scale", "dt_scale", "dp_scale",
                                           "ds_scale", "dp_min", "ds_min")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_BOUND = False


def _lib():
    global _BOUND
    lib = load()
    if not _BOUND:
        F = C.POINTER(C.c_float)
        D = C.POINTER(C.c_double)
        I = C.POINTER(C.c_int32)
        sig = {
            "ewg_cnn_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                         C.POINTER(_vp)]),
            "ewg_cnn_destroy": (None, [_vp]),
            "ewg_cnn_last_error": (C.c_char_p, [_vp]),
            "ewg_cnn_set_params": (C.c_int, [_vp, F]),
            "ewg_cnn_get_params": (C.c_int, [_vp, F]),
            "ewg_cnn_set_norm": (C.c_int, [_vp, C.POINTER(Norm)]),
            "ewg_cnn_get_norm": (C.c_int, [_vp, C.POINTER(Norm)]),
            "ewg_cnn_features": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int64, C.c_double, _vp]),
            "ewg_cnn_forward": (C.c_int, [_vp, _vp, C.c_int32, _vp]),
            "ewg_cnn_gradients": (C.c_int, [_vp, _vp, _vp, C.c_int64, I, C.c_int32, D, F]),
            "ewg_cnn_train_step": (C.c_int, [_vp, _vp, _vp, C.c_int64, I, C.c_int32,
                                             C.c_float, D]),
            "ewg_cnn_last_forward_ms": (C.c_int, [_vp, D]),
            "ewg_cnn_last_inference_cells": (C.c_int64, [_vp]),
            "ewg_device_alloc": (C.c_int, [C.c_int64, C.POINTER(_vp)]),
            "ewg_device_free": (C.c_int, [_vp]),
            "ewg_device_upload": (C.c_int, [_vp, _vp, C.c_int64]),
            "ewg_device_download": (C.c_int, [_vp, _vp, C.c_int64]),
            "ewg_set_warm_start": (C.c_int, [_vp, _vp]),
            "ewg_warm_start_stats": (C.c_int, [_vp, C.POINTER(C.c_int64),
                                               C.POINTER(C.c_int64), D, D]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _BOUND = True
    return lib


class DeviceArray:
    """A device allocation of the library (ewg_device_alloc) with a numpy dtype and shape;
    views of rows share the allocation."""

    def __init__(self, shape, dtype, _base=None, _offset=0):
        self.lib = _lib()
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._base = _base
        if _base is None:
            p = _vp()
            if self.lib.ewg_device_alloc(max(self.nbytes, 1), C.byref(p)) != 0:
                raise MemoryError(f"ewg_device_alloc({self.nbytes}) failed")
            self._ptr = p.value
        else:
            self._ptr = _base._ptr + _offset

    @classmethod
    def from_numpy(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    @property
    def ptr(self):
        return C.c_void_p(self._ptr)

    def row(self, i):
        """View of row i of the first axis."""
        n = self.shape[0]
        if not 0 <= i < n:
            raise IndexError(i)
        per = self.nbytes // n
        return DeviceArray(self.shape[1:], self.dtype, _base=self._base or self, _offset=i * per)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        if self.lib.ewg_device_upload(self.ptr, a.ctypes.data_as(_vp), self.nbytes) != 0:
            raise RuntimeError("ewg_device_upload failed")

    def numpy(self):
        a = np.empty(self.shape, dtype=self.dtype)
        if self.lib.ewg_device_download(a.ctypes.data_as(_vp), self.ptr, self.nbytes) != 0:
            raise RuntimeError("ewg_device_download failed")
        return a

    def __del__(self):
        if self._base is None and getattr(self, "_ptr", None):
            try:
                self.lib.ewg_device_free(C.c_void_p(self._ptr))
            except Exception:
                pass
            self._ptr = None


LAYERS = [(5, 32), (32, 32), (32, 32), (32, 32), (32, 2)]  # (in, out) channels


def init_params(seed=0):
    """He-normal weights, std sqrt(2 / (27 in)) (a tenth of that for the output layer),
    zero biases, in the flat layout of ewg_cnn_set_params."""
    rng = np.random.default_rng(seed)
    parts = []
    for li, (ci, co) in enumerate(LAYERS):
        std = math.sqrt(2.0 / (27 * ci)) * (0.1 if li == len(LAYERS) - 1 else 1.0)
        parts += [rng.normal(0.0, std, co * ci * 27), np.zeros(co)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == NUM_PARAMS
    return p


def bf16_bits(a):
    """float array -> BF16 bit patterns (uint16), round to nearest even."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


class WarmStartNet:
    """The network for samples of nx*ny*nz cells, batches of up to max_batch samples."""

    def __init__(self, nx, ny, nz, max_batch=1, params=None, norm=None):
        self.lib = _lib()
        self.nx, self.ny, self.nz, self.max_batch = nx, ny, nz, max_batch
        self.ncell = nx * ny * nz
        h = _vp()
        rc = self.lib.ewg_cnn_create(nx, ny, nz, max_batch, C.byref(h))
        if rc != 0 or not h.value:
            raise RuntimeError(f"ewg_cnn_create({nx}, {ny}, {nz}, {max_batch}) failed: {rc}")
        self.h = h
        if params is not None:
            self.set_params(params)
        if norm is not None:
            self.set_norm(norm)

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.ewg_cnn_last_error(self.h)
            raise RuntimeError(f"ewg_cnn: status {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.ewg_cnn_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ parameters
    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == NUM_PARAMS
        self._check(self.lib.ewg_cnn_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def params(self):
        p = np.empty(NUM_PARAMS, dtype=np.float32)
        self._check(self.lib.ewg_cnn_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))
        return p

    def set_norm(self, norm: dict):
        n = Norm(**{"dp_min": 0.0, "ds_min": 0.0, **norm})
        self._check(self.lib.ewg_cnn_set_norm(self.h, C.byref(n)))

    def norm(self) -> dict:
        n = Norm()
        self._check(self.lib.ewg_cnn_get_norm(self.h, C.byref(n)))
        return n.as_dict()

    # ------------------------------------------------------------------ compute
    def features(self, x, K, phi, dt, out: DeviceArray | None = None):
        """BF16 features [cells][8] (DeviceArray of uint16) of a host state."""
        K = np.ascontiguousarray(K, dtype=np.float64).ravel()
        n = K.size
        xd = DeviceArray.from_numpy(np.asarray(x, dtype=np.float64).ravel())
        Kd = DeviceArray.from_numpy(K)
        pd = DeviceArray.from_numpy(np.asarray(phi, dtype=np.float64).ravel())
        assert xd.shape[0] == 2 * n and pd.shape[0] == n
        if out is None:
            out = DeviceArray((n, 8), np.uint16)
        assert out.nbytes == n * 16
        self._check(self.lib.ewg_cnn_features(self.h, xd.ptr, Kd.ptr, pd.ptr, n, float(dt),
                                              out.ptr))
        return out

    def forward(self, feat: DeviceArray, batch=1):
        y = DeviceArray((batch * self.ncell, 2), np.float32)
        self._check(self.lib.ewg_cnn_forward(self.h, feat.ptr, batch, y.ptr))
        return y

    def last_forward_ms(self):
        ms = C.c_double(0)
        self._check(self.lib.ewg_cnn_last_forward_ms(self.h, C.byref(ms)))
        return ms.value

    def last_inference_cells(self):
        """cells of the last warm-start prediction's active region"""
        return int(self.lib.ewg_cnn_last_inference_cells(self.h))

    @staticmethod
    def _idx(idx):
        if idx is None:
            return None, None
        a = np.ascontiguousarray(idx, dtype=np.int32)
        return a, a.ctypes.data_as(C.POINTER(C.c_int32))

    def gradients(self, feat: DeviceArray, target: DeviceArray, batch=1, idx=None):
        g = np.empty(NUM_PARAMS, dtype=np.float32)
        loss = C.c_double(0)
        keep, ip = self._idx(idx)
        self._check(self.lib.ewg_cnn_gradients(self.h, feat.ptr, target.ptr, feat.shape[0],
                                               ip, batch, C.byref(loss),
                                               g.ctypes.data_as(C.POINTER(C.c_float))))
        return loss.value, g

    def train_step(self, feat: DeviceArray, target: DeviceArray, batch, lr, idx=None):
        loss = C.c_double(0)
        keep, ip = self._idx(idx)
        self._check(self.lib.ewg_cnn_train_step(self.h, feat.ptr, target.ptr, feat.shape[0],
                                                ip, batch, float(lr), C.byref(loss)))
        return loss.value


def attach(ctx, net: WarmStartNet | None):
    """Let ctx's Newton driver start from net's prediction (params.warm_start = 1)."""
    lib = _lib()
    rc = lib.ewg_set_warm_start(ctx.h, net.h if net is not None else None)
    if rc != 0:
        raise RuntimeError(f"ewg_set_warm_start: {rc}: {ctx.last_error()}")


def warm_stats(ctx):
    lib = _lib()
    a, r = C.c_int64(0), C.c_int64(0)
    ps, e0 = C.c_double(0), C.c_double(0)
    rc = lib.ewg_warm_start_stats(ctx.h, C.byref(a), C.byref(r), C.byref(ps), C.byref(e0))
    if rc != 0:
        raise RuntimeError(f"ewg_warm_start_stats: {rc}")
    return {"accepted": a.value, "rejected": r.value, "predict_s": ps.value,
            "last_initial_error": e0.value}


# ---------------------------------------------------------------------- training data
def harvest(n_pairs, steps_per_member=20, seed=1, crop=(32, 32, 32), skip_first=True,
            top_fraction=0.75, log=None):
    """(state, converged update) pairs from plain FP64 runs of ensemble members, each cut
    to a crop of crop cells (a fraction top_fraction of the crops at the top centre,
    around the injection patch, the others anywhere). The first step of each run, where
    the pressure field settles from the initial condition, is left out by default: the
    state alone does not tell it from later steps far from the plume. Returns host arrays:
    x_old [n][cells][2], K, phi [n][cells], dt [n], dx [n][cells][2] (update of p_w, S_w)."""
    from . import GpuContext
    cx, cy, cz = crop
    rng = np.random.default_rng(seed + 7919)
    X, Kc, Pc, DT, DX = [], [], [], [], []
    k = 0
    while len(X) < n_pairs:
        P = problems.ensemble_member(k, seed, sizes=problems.ENSEMBLE_SIZES[:3],
                                     steps=steps_per_member)
        k += 1
        ctx = GpuContext(P, default_params(newton_tolerance=1e-6, max_time_step_size=P.dt))
        K3 = P.K.reshape(P.nz, P.ny, P.nx)
        F3 = P.phi.reshape(P.nz, P.ny, P.nx)
        for step in range(steps_per_member):
            x_old = ctx.solution(0)
            rc, it, err = ctx.newton(P.dt)
            if rc != 0:
                break
            ctx.advance_time_level()
            x_new = ctx.solution(0)
            if skip_first and step == 0:
                continue
            if rng.random() < top_fraction:  # the injection patch: top centre
                ox, oy, oz = (P.nx - cx) // 2, (P.ny - cy) // 2, P.nz - cz
            else:
                ox = int(rng.integers(0, P.nx - cx + 1))
                oy = int(rng.integers(0, P.ny - cy + 1))
                oz = int(rng.integers(0, P.nz - cz + 1))
            sl = (slice(oz, oz + cz), slice(oy, oy + cy), slice(ox, ox + cx))
            xo = x_old.reshape(P.nz, P.ny, P.nx, 2)[sl]
            xn = x_new.reshape(P.nz, P.ny, P.nx, 2)[sl]
            X.append(np.ascontiguousarray(xo).reshape(-1, 2))
            DX.append(np.ascontiguousarray(xn - xo).reshape(-1, 2))
            Kc.append(np.ascontiguousarray(K3[sl]).ravel())
            Pc.append(np.ascontiguousarray(F3[sl]).ravel())
            DT.append(P.dt)
            if len(X) >= n_pairs:
                break
        ctx.close()
        if log:
            log(f"harvest: member {k - 1} {P.nx}x{P.ny}x{P.nz} -> {len(X)} pairs")
    return {"x": np.stack(X), "K": np.stack(Kc), "phi": np.stack(Pc), "dt": np.array(DT),
            "dx": np.stack(DX)}


def fit_norm(data):
    """Feature and target scaling from a harvested set."""
    p = data["x"][..., 0]
    dp = data["dx"][..., 0]
    dsn = -data["dx"][..., 1]
    return {"p_ref": float(p.mean()), "p_scale": float(p.std() or 1.0),
            "lnk_ref": math.log(1e-12), "lnk_scale": 1.0,
            "phi_ref": 0.25, "phi_scale": 0.02, "dt_scale": 500.0,
            "dp_scale": float(np.sqrt(np.mean(dp * dp)) or 1.0),
            "ds_scale": float(np.sqrt(np.mean(dsn * dsn)) or 1.0)}


def device_dataset(net: WarmStartNet, data):
    """Features (BF16, with net's scaling) and scaled targets of every pair, on the device:
    DeviceArrays [pairs][cells][8] (uint16) and [pairs][cells][2] (float32)."""
    nrm = net.norm()
    n, cells = data["x"].shape[0], data["x"].shape[1]
    feats = DeviceArray((n, cells, 8), np.uint16)
    for i in range(n):
        net.features(data["x"][i], data["K"][i], data["phi"][i], data["dt"][i], out=feats.row(i))
    t = np.stack([data["dx"][..., 0] / nrm["dp_scale"], -data["dx"][..., 1] / nrm["ds_scale"]],
                 axis=-1).astype(np.float32)
    return feats, DeviceArray.from_numpy(t)


def train(net: WarmStartNet, feats: DeviceArray, targets: DeviceArray, epochs, batch,
          lr=1e-3, seed=0, log=None):
    """Adam over shuffled batches of pairs; returns the mean loss of each epoch."""
    rng = np.random.default_rng(seed)
    n = feats.shape[0]
    hist = []
    for ep in range(epochs):
        perm = rng.permutation(n).astype(np.int32)
        tot, cnt = 0.0, 0
        for b0 in range(0, n - batch + 1, batch):
            tot += net.train_step(feats, targets, batch, lr, idx=perm[b0:b0 + batch])
            cnt += 1
        hist.append(tot / max(cnt, 1))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.5f}")
    return hist


def train_warm_start(nx, ny, nz, n_pairs=2000, epochs=30, seed=2, batch=8, lr=2e-3,
                     crop=(32, 32, 16), steps_per_member=40, use_pressure=False, ds_min=1e-4,
                     log=None):
    """Harvest n_pairs training pairs from ensemble members (seed `seed`), train a network
    on crops, and return the network for nx*ny*nz grids with the trained parameters plus a
    summary (pairs, times, losses).

    The guess takes the predicted saturation update where |dS_n| >= ds_min and, unless
    use_pressure, leaves the pressure as it is: the pressure updates of a step are small
    (of order 1 Pa) but spread over the whole domain, and the smooth residual left by
    their prediction error is the kind of error ILU(0)-BiCGStab reduces slowest (measured
    on the 128x128x64 case: 169 BiCGStab iterations over 40 steps with the pressure part,
    127 without, 231 from the plain guess)."""
    import time
    t0 = time.perf_counter()
    data = harvest(n_pairs, steps_per_member=steps_per_member, seed=seed, crop=crop)
    t1 = time.perf_counter()
    net = WarmStartNet(*crop, max_batch=batch)
    net.set_params(init_params(seed))
    net.set_norm(fit_norm(data))
    feats, targets = device_dataset(net, data)
    hist = train(net, feats, targets, epochs, batch, lr=lr, seed=seed, log=log)
    t2 = time.perf_counter()
    nrm = net.norm()
    nrm["ds_min"] = ds_min
    if not use_pressure:
        nrm["dp_scale"] = 0.0
    infer = WarmStartNet(nx, ny, nz, max_batch=1, params=net.params(), norm=nrm)
    info = {"pairs": int(len(data["x"])), "crop": list(crop), "epochs": epochs, "batch": batch,
            "steps_per_member": steps_per_member, "use_pressure": use_pressure,
            "ds_min": ds_min,
            "harvest_s": t1 - t0, "train_s": t2 - t1, "loss_first": hist[0],
            "loss_last": hist[-1], "train_steps": epochs * (len(data["x"]) // batch)}
    net.close()
    return infer, info
