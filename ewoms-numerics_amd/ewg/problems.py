"""Synthetic problem descriptions (inputs only; no physics is evaluated here).

lens_problem        -- the 2D DNAPL lens of tests/problems/lensproblem.hh (48x32 cells)
heterogeneous_3d    -- 3D log-normal permeability, Brooks-Corey, top-patch DNAPL injection
                       (128x128x64 cells at full size)

Every array is float64/int32, cells in lexicographic order (x fastest) and boundary faces
grouped by side (x-, x+, y-, y+, z-, z+), each side ordered lexicographically over the
remaining two axes -- the layout of include/ewg_capi.h.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from .capi import (EWG_BC_FREE_FLOW, EWG_BC_MASS_RATE, EWG_BC_NO_FLOW,
                   EWG_LAW_BROOKS_COREY, EWG_LAW_VAN_GENUCHTEN, Params, default_params)


@dataclasses.dataclass
class Problem:
    name: str
    dim: int
    nx: int
    ny: int
    nz: int
    lx: float
    ly: float
    lz: float
    rho: tuple
    mu: tuple
    gravity: tuple
    law: int
    region_params: np.ndarray     # [nregions, 4]
    K: np.ndarray                 # [N]
    phi: np.ndarray               # [N]
    region: np.ndarray            # [N] int32
    pe: np.ndarray | None         # [N] or None
    bc_type: np.ndarray           # [B] int32
    bc_val: np.ndarray            # [B, 2]
    x0: np.ndarray                # [N, 2] (p_w, S_w)
    params: Params
    t_end: float
    dt: float
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def N(self):
        return self.nx * self.ny * self.nz

    @property
    def h(self):
        return (self.lx / self.nx, self.ly / self.ny, (self.lz / self.nz) if self.dim == 3 else 1.0)


def _side_sizes(dim, nx, ny, nz):
    s = [ny * nz, ny * nz, nx * nz, nx * nz]
    s += [nx * ny, nx * ny] if dim == 3 else [0, 0]
    return s


def boundary_faces(dim, nx, ny, nz, lx, ly, lz):
    """Face-centre coordinates of every boundary face in the C-ABI order.

    Cell centres are (i + 0.5) * h and the face of side + is at (n_a) * h, the
    expressions the oracle and the device kernels evaluate."""
    h = (lx / nx, ly / ny, (lz / nz) if dim == 3 else 1.0)
    n = (nx, ny, nz)
    out = []
    for side in range(2 * dim):
        a = side >> 1
        plus = side & 1
        other = [d for d in range(3) if d != a]
        # lexicographic over the remaining axes: the lower axis runs fastest
        i1 = np.arange(n[other[0]])
        i2 = np.arange(n[other[1]])
        g2, g1 = np.meshgrid(i2, i1, indexing="ij")
        pos = np.zeros((g1.size, 3))
        pos[:, other[0]] = (g1.ravel() + 0.5) * h[other[0]]
        pos[:, other[1]] = (g2.ravel() + 0.5) * h[other[1]]
        pos[:, a] = (n[a] if plus else 0) * h[a]
        out.append(pos)
    return out


def cell_centres(dim, nx, ny, nz, lx, ly, lz):
    h = (lx / nx, ly / ny, (lz / nz) if dim == 3 else 1.0)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.stack([(i.ravel() + 0.5) * h[0], (j.ravel() + 0.5) * h[1],
                     (k.ravel() + 0.5) * h[2]], axis=1)


def lens_problem(nx=48, ny=32, outer_alpha=3.7e-4, mu_n=9e-4, end_time=3.0e4, dt=250.0,
                 newton_tol=1e-8, linear_tol=1e-10, fixed_dt=True) -> Problem:
    """Config 0: tests/problems/lensproblem.hh:135-166, :245-288, :357-591.

    outer_alpha defaults to 3.7e-4 1/Pa; the reference file sets 0.0037
    (lensproblem.hh:275) -- pass outer_alpha=3.7e-3 for that variant."""
    lx, ly, lz = 6.0, 4.0, 1.0
    eps = 3e-6
    rho_w, rho_n = 1000.0, 1460.0
    g = (0.0, -9.81, 0.0)
    X = cell_centres(2, nx, ny, 1, lx, ly, lz)
    ll, ur = (1.0, 2.0), (4.0, 3.0)
    in_lens = np.ones(X.shape[0], bool)
    for d in range(2):  # isInLens_ (lensproblem.hh:596-604)
        in_lens &= ~((X[:, d] < ll[d] - eps) | (X[:, d] > ur[d] + eps))
    region = in_lens.astype(np.int32)           # 0 = outer, 1 = lens
    K = np.where(in_lens, 9.05e-12, 4.6e-10)
    phi = np.full(X.shape[0], 0.4)
    regions = np.array([[outer_alpha, 4.7, 0.05, 0.0],
                        [4.5e-4, 7.3, 0.18, 0.0]])
    # initial condition (lensproblem.hh:545-578): hydrostatic pw, Sw = 1
    ymax = ly
    depth = ymax - X[:, 1]
    pw = 1e5 - rho_w * g[1] * depth
    x0 = np.stack([pw, np.ones_like(pw)], axis=1)
    # boundaries (lensproblem.hh:464-533, :606-623)
    faces = boundary_faces(2, nx, ny, 1, lx, ly, lz)
    btype, bval = [], []
    for side, pos in enumerate(faces):
        t = np.full(pos.shape[0], EWG_BC_NO_FLOW, np.int32)
        v = np.zeros((pos.shape[0], 2))
        left = pos[:, 0] < 0.0 + eps
        right = pos[:, 0] > lx - eps
        height = ymax - 0.0
        fdepth = ymax - pos[:, 1]
        alpha = 1 + 1.5 / height
        pw_left = 1e5 - alpha * rho_w * g[1] * fdepth
        pw_right = 1e5 - rho_w * g[1] * fdepth
        t[left | right] = EWG_BC_FREE_FLOW
        v[left, 0] = pw_left[left]
        v[right, 0] = pw_right[right]
        v[left | right, 1] = 1.0
        lam = (lx - pos[:, 0]) / (lx - 0.0)
        upper = pos[:, 1] > ymax - eps
        inlet = upper & (0.5 < lam) & (lam < 2.0 / 3.0) & ~(left | right)
        t[inlet] = EWG_BC_MASS_RATE
        v[inlet, 0] = 0.0
        v[inlet, 1] = -0.04
        btype.append(t)
        bval.append(v)
    params = default_params(newton_tolerance=newton_tol, linear_tolerance=linear_tol)
    if fixed_dt:
        params.max_time_step_size = dt
    return Problem("lens", 2, nx, ny, 1, lx, ly, lz, (rho_w, rho_n), (1e-3, mu_n), g,
                   EWG_LAW_VAN_GENUCHTEN, regions, K, phi, region, None,
                   np.concatenate(btype), np.concatenate(bval), x0, params, end_time, dt,
                   meta={"outer_alpha": outer_alpha, "mu_n": mu_n})


def gaussian_field(shape, corr_cells, rng):
    """Unit-variance stationary Gaussian field with Gaussian covariance
    C(r) = exp(-(r/l)^2), l in cells, by spectral synthesis on the periodic grid:
    FFT of white noise times sqrt of the spectral density exp(-(k l)^2 / 4)."""
    nz, ny, nx = shape
    w = rng.standard_normal(shape)
    kz = 2 * np.pi * np.fft.fftfreq(nz)
    ky = 2 * np.pi * np.fft.fftfreq(ny)
    kx = 2 * np.pi * np.fft.rfftfreq(nx)
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    amp = np.exp(-k2 * corr_cells ** 2 / 8.0)  # sqrt(exp(-k^2 l^2 / 4))
    f = np.fft.irfftn(np.fft.rfftn(w) * amp, s=shape, axes=(0, 1, 2))
    f -= f.mean()
    f /= f.std()
    return f


def heterogeneous_3d(nx=128, ny=128, nz=64, seed=20261203, sigma_lnK=1.0, corr_cells=8.0,
                     h=1.0, rate=1e-3, patch=16, mu_n=9e-4, dt=500.0, steps=40,
                     newton_tol=1e-6, linear_tol=1e-3) -> Problem:
    """Config 1: 3D heterogeneous DNAPL injection (same two-phase physics as the lens).

    Choices the configuration leaves open (recorded in meta): cubic cells of size h, z
    vertical with gravity -9.81, free-flow Dirichlet (S_w = 1) on the x- and x+ sides as in
    the lens problem (x+: hydrostatic; x-: hydrostatic scaled by 1 + 1.5/H, so water flows
    towards x+), no-flow elsewhere except the injection patch on top, Brooks-Corey lambda = 2 with
    Swr = 0.05, Snr = 0 and entry pressure pe = 1e4 * sqrt((phi/K) / (0.25/1e-12))."""
    lx, ly, lz = nx * h, ny * h, nz * h
    rho_w, rho_n = 1000.0, 1460.0
    g = (0.0, 0.0, -9.81)
    rng = np.random.default_rng(seed)
    G = gaussian_field((nz, ny, nx), corr_cells, rng).ravel()
    K = np.exp(math.log(1e-12) + sigma_lnK * G)
    phi = np.clip(0.25 + 0.02 * rng.standard_normal(nx * ny * nz), 0.15, 0.35)
    pe = 1e4 * np.sqrt((phi / K) / (0.25 / 1e-12))
    region = np.zeros(nx * ny * nz, np.int32)
    regions = np.array([[1e4, 2.0, 0.05, 0.0]])
    X = cell_centres(3, nx, ny, nz, lx, ly, lz)
    depth = lz - X[:, 2]
    pw = 1e5 - rho_w * g[2] * depth
    x0 = np.stack([pw, np.ones_like(pw)], axis=1)
    faces = boundary_faces(3, nx, ny, nz, lx, ly, lz)
    btype, bval = [], []
    for side, pos in enumerate(faces):
        t = np.full(pos.shape[0], EWG_BC_NO_FLOW, np.int32)
        v = np.zeros((pos.shape[0], 2))
        if side in (0, 1):
            # lens-style free-flow sides (lensproblem.hh:472-499): hydrostatic on x+,
            # hydrostatic scaled by alpha = 1 + 1.5/H on x-, driving a regional water flow
            t[:] = EWG_BC_FREE_FLOW
            fdepth = lz - pos[:, 2]
            alpha = 1 + 1.5 / lz
            v[:, 0] = (1e5 - alpha * rho_w * g[2] * fdepth) if side == 0 else \
                (1e5 - rho_w * g[2] * fdepth)
            v[:, 1] = 1.0
        if side == 5:
            i = np.floor(pos[:, 0] / h).astype(int)
            j = np.floor(pos[:, 1] / h).astype(int)
            lo_x, lo_y = nx // 2 - patch // 2, ny // 2 - patch // 2
            m = (i >= lo_x) & (i < lo_x + patch) & (j >= lo_y) & (j < lo_y + patch)
            t[m] = EWG_BC_MASS_RATE
            v[m, 1] = -rate
        btype.append(t)
        bval.append(v)
    params = default_params(newton_tolerance=newton_tol, linear_tolerance=linear_tol)
    params.max_time_step_size = dt
    return Problem("heterogeneous_3d", 3, nx, ny, nz, lx, ly, lz, (rho_w, rho_n), (1e-3, mu_n),
                   g, EWG_LAW_BROOKS_COREY, regions, K, phi, region, pe,
                   np.concatenate(btype), np.concatenate(bval), x0, params, dt * steps, dt,
                   meta={"seed": seed, "sigma_lnK": sigma_lnK, "corr_cells": corr_cells,
                         "h": h, "rate": rate, "patch": patch, "steps": steps})


ENSEMBLE_SIZES = ((32, 32, 32), (48, 48, 48), (64, 64, 32), (96, 96, 48))


def ensemble_member(k, seed, sizes=ENSEMBLE_SIZES, steps=40) -> Problem:
    """Realization k of the ensemble family of config 3 (the config-1 physics): grid size
    drawn uniformly from `sizes`, sigma_lnK ~ U[0.5, 2], correlation length ~ U[4, 12]
    cells, injection rate ~ U[5e-4, 2e-3] kg/(m^2 s) over a top centre patch of nx/8 x
    nx/8 cells, its own seed for the permeability and porosity fields."""
    rng = np.random.default_rng([seed, k])
    nx, ny, nz = sizes[int(rng.integers(len(sizes)))]
    return heterogeneous_3d(nx, ny, nz, seed=int(rng.integers(1, 2**31 - 1)),
                            sigma_lnK=float(rng.uniform(0.5, 2.0)),
                            corr_cells=float(rng.uniform(4.0, 12.0)),
                            rate=float(rng.uniform(5e-4, 2e-3)), patch=max(2, nx // 8),
                            steps=steps)


def perturbed_state(problem: Problem, seed=7, dS=0.3, dp=500.0):
    """A non-trivial state for linearization parity: saturation and pressure perturbed
    away from the initial hydrostatic state (keeps S_w in (0.02, 1.05))."""
    rng = np.random.default_rng(seed)
    x = problem.x0.copy()
    x[:, 0] += dp * rng.standard_normal(problem.N)
    x[:, 1] = np.clip(1.0 - dS * rng.random(problem.N) * 2.0, 0.02, 1.05)
    return x
