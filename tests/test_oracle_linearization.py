"""Pins of the oracle's linearization against what the mathematics fixes.

- AD Jacobian vs central differences of the oracle residual (all columns on small grids);
- global mass balance: sum_i V_i R_i = storage change + boundary terms (interior face
  fluxes cancel pairwise; fvbasediscretization.hh:969-1080 checks the same invariant);
- hydrostatic equilibrium: zero residual in a closed domain at rest;
- 1D single-phase TPFA: a linear pressure profile between two Dirichlet faces is the exact
  discrete solution, with the half-cell distance to the boundary face.
A plausible slip (dropped face, wrong sign, wrong area, wrong boundary distance, wrong
upwind side) fails at least one of these.
"""
import numpy as np
import pytest

from helpers import bcrs_to_dense


def _oracle(P, params=None):
    from oracle.pyoracle import OracleContext
    return OracleContext(P, params)


def _residual_at(o, P, x, dt, x1):
    o.set_solution(x1, 1)
    o.set_solution(x1, 0)
    o.linearize(dt, 0, want=False)      # storage cache of the previous level from x1
    o.set_solution(x, 0)
    r, _ = o.linearize(dt, 1)
    return r


@pytest.mark.parametrize("which", ["lens", "het3d"])
def test_jacobian_matches_central_differences(which):
    from ewg import problems
    if which == "lens":
        P = problems.lens_problem(nx=6, ny=5)
    else:
        P = problems.heterogeneous_3d(4, 3, 3, seed=2, patch=2)
    x1 = problems.perturbed_state(P, seed=1, dS=0.1)
    x = problems.perturbed_state(P, seed=2, dS=0.2)
    o = _oracle(P)
    r0 = _residual_at(o, P, x, P.dt, x1)
    _, J = o.linearize(P.dt, 1)
    rp, ci = o.pattern()
    A = bcrs_to_dense(rp, ci, J, P.N)
    xf = x.ravel().copy()
    for col in range(2 * P.N):
        h = 1e-3 if col % 2 == 0 else 1e-7
        xp, xm = xf.copy(), xf.copy()
        xp[col] += h
        xm[col] -= h
        fd = (_residual_at(o, P, xp, P.dt, x1) - _residual_at(o, P, xm, P.dt, x1)) / (2 * h)
        scale = max(np.abs(A[:, col]).max(), 1e-300)
        # central differences: O(h^2) truncation; kinks of the regularized laws and upwind
        # switches are avoided by the perturbed state (S_w well inside (0, 1))
        assert np.abs(fd - A[:, col]).max() <= 2e-5 * scale, col
    assert np.isfinite(r0).all()


def _closed_box(P):
    P.bc_type[:] = 0
    P.bc_val[:] = 0.0
    return P


def test_hydrostatic_equilibrium_is_at_rest():
    from ewg import problems
    P = _closed_box(problems.heterogeneous_3d(6, 5, 4, seed=4))
    o = _oracle(P)
    r = _residual_at(o, P, P.x0, P.dt, P.x0)
    # scale: one face's water flux for a 1 Pa potential difference
    flux_scale = P.K.max() / P.mu[0] * P.rho[0] / (P.h[0] ** 2)
    assert np.abs(r).max() <= 1e-4 * flux_scale


def test_global_mass_balance():
    from ewg import problems
    P = problems.lens_problem(nx=10, ny=8)
    o = _oracle(P)
    x1 = problems.perturbed_state(P, seed=3, dS=0.1)
    x = problems.perturbed_state(P, seed=4, dS=0.15)
    r = _residual_at(o, P, x, P.dt, x1).reshape(-1, 2)
    V = P.h[0] * P.h[1]
    M = o.storage(x).reshape(-1, 2)
    M1 = o.storage(x1).reshape(-1, 2)
    storage = (M - M1).sum(axis=0) * V / P.dt
    # boundary: mass-rate faces plus free-flow fluxes; with the free-flow faces switched
    # off the balance closes on the inlet rate alone
    Pc = problems.lens_problem(nx=10, ny=8)
    Pc.bc_type[Pc.bc_type == 1] = 0
    oc = _oracle(Pc)
    rc = _residual_at(oc, Pc, x, Pc.dt, x1).reshape(-1, 2)
    inlet = Pc.bc_type == 2
    area = P.h[0]
    boundary = (Pc.bc_val[inlet].sum(axis=0)) * area
    total = (rc.sum(axis=0)) * V
    assert np.allclose(total, storage + boundary, rtol=1e-10, atol=1e-12 * np.abs(storage).max())
    # with the free-flow faces on, the balance differs exactly by the boundary fluxes:
    # both residuals agree on every cell that touches no free-flow face
    free = np.zeros(P.N, bool)
    free[np.arange(P.N) % P.nx == 0] = True
    free[np.arange(P.N) % P.nx == P.nx - 1] = True
    assert np.array_equal(r[~free], rc[~free])


def test_single_phase_linear_profile_is_discrete_solution():
    """1D water flow between two Dirichlet faces (x- at pL, x+ at pR), no gravity, S_w = 1:
    p(x) = pL + (pR - pL) x / L at the cell centres zeroes the residual, and the face
    flux equals K/mu rho (pL - pR)/L."""
    from ewg import problems
    P = problems.heterogeneous_3d(8, 3, 3, seed=5)
    P.gravity = (0.0, 0.0, 0.0)
    P.K[:] = 2e-12
    P.bc_type[:] = 0
    B0 = P.ny * P.nz
    pL, pR = 2.0e5, 1.0e5
    P.bc_type[:B0] = 1
    P.bc_val[:B0] = (pL, 1.0)
    P.bc_type[B0:2 * B0] = 1
    P.bc_val[B0:2 * B0] = (pR, 1.0)
    L = P.lx
    xc = (np.arange(P.N) % P.nx + 0.5) * P.h[0]
    x = np.stack([pL + (pR - pL) * xc / L, np.ones(P.N)], axis=1)
    o = _oracle(P)
    r = _residual_at(o, P, x, P.dt, x)
    q = 2e-12 / P.mu[0] * P.rho[0] * (pL - pR) / L   # mass flux [kg/(m^2 s)]
    assert np.abs(r).max() <= 1e-9 * q
    # a cell-wise constant pressure is not a solution: the first cell sees the full jump
    x2 = x.copy()
    x2[:, 0] = pR
    r2 = _residual_at(o, P, x2, P.dt, x2).reshape(-1, 2)
    assert r2[0, 0] < 0 and abs(r2[0, 0]) > 1e3 * abs(r).max() + 0.5 * q / P.h[0]


def test_neighbour_columns_are_minus_face_derivatives():
    """Column i of row j (j a neighbour of i) is -(dF_ij/dx_i)/V_j: the sum over a column
    of V-weighted blocks equals the derivative of the element's boundary + storage terms
    alone (interior face contributions cancel within the column)."""
    from ewg import problems
    P = _closed_box(problems.heterogeneous_3d(5, 4, 3, seed=6))
    o = _oracle(P)
    x = problems.perturbed_state(P, seed=8, dS=0.2)
    _residual_at(o, P, x, P.dt, x)
    _, J = o.linearize(P.dt, 1)
    rp, ci = o.pattern()
    A = bcrs_to_dense(rp, ci, J, P.N)
    colsum = A.sum(axis=0)  # V is uniform: sum_j V_j dR_j/dx_i = V * colsum
    # closed box: column sums = storage derivative only: d/dS_w of (phi S rho)/dt
    phi = P.phi
    expect = np.zeros(2 * P.N)
    expect[1::2] = phi * (P.rho[0] - P.rho[1]) / P.dt
    assert np.allclose(colsum, expect, rtol=1e-9, atol=1e-12 * np.abs(A).max())


def test_threaded_linearization_is_bitwise_identical():
    """The element loop over worker threads (ThreadsPerProcess) gives the single-thread
    residual and Jacobian bit for bit: every element writes its own row entry and column."""
    from ewg import problems
    P = problems.heterogeneous_3d(9, 7, 5, seed=9, patch=2)
    x = problems.perturbed_state(P, seed=10, dS=0.2)
    out = []
    for threads in (1, 3):
        o = _oracle(P)
        o.set_threads(threads)
        out.append(_residual_at(o, P, x, P.dt, problems.perturbed_state(P, seed=11, dS=0.1)))
        out.append(o.linearize(P.dt, 1)[1])
    assert np.array_equal(out[0], out[2]) and np.array_equal(out[1], out[3])


def _fd_vs_ad(P, method, seeds=(1, 2)):
    """Residual and Jacobian of the AD and the FD local linearizer at the same state."""
    import ewg
    from ewg import problems
    x1 = problems.perturbed_state(P, seed=seeds[0], dS=0.1)
    x = problems.perturbed_state(P, seed=seeds[1], dS=0.2)
    out = []
    for ll in (0, 1):
        o = _oracle(P, ewg.default_params(local_linearizer=ll, numeric_difference_method=method))
        out.append((_residual_at(o, P, x, P.dt, x1), o.linearize(P.dt, 1)[1]))
    return out


def block_column_error(J, J_ref):
    """max over blocks and primary variables of |J - J_ref| / max_eq |J_ref| of the
    block's column (2x2 row-major blocks)."""
    Jb, Rb = J.reshape(-1, 2, 2), J_ref.reshape(-1, 2, 2)
    scale = np.maximum(np.abs(Rb).max(axis=1, keepdims=True), 1e-300)
    return float((np.abs(Jb - Rb) / scale).max())


@pytest.mark.parametrize("method", [1, 0, -1])
@pytest.mark.parametrize("which", ["lens", "het3d"])
def test_fd_local_linearizer_matches_ad(which, method):
    """FvBaseFdLocalLinearizer semantics (fvbasefdlocallinearizer.hh:370-470): the
    residual is the undeflected one, bit for bit the AD residual; the Jacobian differs
    from the AD Jacobian by the differences' truncation and round-off only (eps = 9.1e-11
    for the saturation, 9.1e-11 p_ref / 10 for the pressure): below 2e-3 of each block
    column, with the reference's forward (+1), central (0) and backward (-1) differences."""
    from ewg import problems
    P = (problems.lens_problem(nx=12, ny=8) if which == "lens"
         else problems.heterogeneous_3d(8, 6, 5, seed=3, patch=2))
    (r_ad, J_ad), (r_fd, J_fd) = _fd_vs_ad(P, method)
    assert np.array_equal(r_fd, r_ad)
    err = block_column_error(J_fd, J_ad)
    print(which, method, err)
    assert err < 2e-3
    # a dropped derivative term is far outside that bar: half of the saturation column
    J_bad = J_ad.copy().reshape(-1, 2, 2)
    J_bad[:, :, 1] *= 0.5
    assert block_column_error(J_fd, J_bad.ravel()) > 0.5
