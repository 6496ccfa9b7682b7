"""Known-answer pins of the oracle's face permeability, upstream choice and regularized
material-law branches: the parts that closed-form residual checks (hydrostatic rest, mass
balance, FD Jacobians) leave open because a plausible mistake in them keeps every such
invariant.

- Harmonic face permeability (common/multiphasebaseproblem.hh:108-124): single-phase flow
  through a column of alternating permeability between two Dirichlet faces. The exact
  series-resistance (Darcy) solution sampled at the cell centres zeroes the discrete
  residual when the face permeability is the harmonic mean of the two cells and the
  boundary half-cell uses the cell's own K (darcyfluxmodule.hh:414-446); the discrete
  solution of an arithmetic-mean scheme leaves a large residual.
- Upstream mobility (darcyfluxmodule.hh:302-323, strict ``> 0``; boundary :428): hand-
  computed two-phase fluxes and Jacobian entries on a row of three cells, with water and
  DNAPL flowing in opposite directions, an exact potential tie (interior = focus
  upstream), and a free-flow face with inflow (boundary state upstream) and outflow.
- Regularized branches (ewoms-material Regularized{VanGenuchten,BrooksCorey} as the
  oracle states them): values and slopes at Se = 0.005, 0.995 and 1.02 from the stated
  rules, with the end cubic found by solving its four end conditions as a linear system.

Expected values come from the textbook closed forms (Darcy series resistance; Brooks-Corey
and Burdine; van Genuchten and Mualem/Parker), evaluated here with numpy, and their
derivatives by complex steps.
"""
import math

import numpy as np
import pytest

from helpers import bcrs_to_dense


def _oracle(P):
    from oracle.pyoracle import OracleContext
    return OracleContext(P)


def _residual_and_jacobian(P, x):
    """Residual and dense Jacobian at x with the previous level equal to x (zero storage
    term), Newton iteration 1."""
    o = _oracle(P)
    o.set_solution(x, 1)
    o.set_solution(x, 0)
    o.linearize(P.dt, 0, want=False)
    r, J = o.linearize(P.dt, 1)
    rp, ci = o.pattern()
    return r.reshape(-1, 2), bcrs_to_dense(rp, ci, J, P.N)


# ------------------------------------------------------------------ harmonic mean
def _series_column(K_even, K_odd, nx=8):
    from ewg import problems
    P = problems.heterogeneous_3d(nx, 3, 3, seed=5)
    P.gravity = (0.0, 0.0, 0.0)
    ix = np.arange(P.N) % P.nx
    P.K = np.where(ix % 2 == 0, K_even, K_odd).astype(np.float64)
    P.bc_type[:] = 0
    B0 = P.ny * P.nz
    pL, pR = 2.0e5, 1.0e5
    P.bc_type[:B0] = 1
    P.bc_val[:B0] = (pL, 1.0)
    P.bc_type[B0:2 * B0] = 1
    P.bc_val[B0:2 * B0] = (pR, 1.0)
    return P, pL, pR


def _centre_pressures(P, pL, pR, face_resistance):
    """Cell-centre pressures of steady 1D flow along x for the given resistance of each
    interior face (per unit area and viscosity) and half-cell boundary resistances."""
    h = P.h[0]
    Kx = P.K[:P.nx]
    half = h / (2.0 * Kx)
    faces = [face_resistance(Kx[i], Kx[i + 1]) for i in range(P.nx - 1)]
    total = half[0] + sum(faces) + half[-1]
    u_mu = (pL - pR) / total           # u * mu
    p = np.empty(P.nx)
    p[0] = pL - u_mu * half[0]
    for i in range(1, P.nx):
        p[i] = p[i - 1] - u_mu * faces[i - 1]
    col = np.stack([np.tile(p, P.ny * P.nz), np.ones(P.N)], axis=1)
    return col, u_mu


def test_harmonic_face_permeability_series_column():
    h_over = lambda K1, K2, h: h / 2.0 / K1 + h / 2.0 / K2   # noqa: E731
    P, pL, pR = _series_column(1e-12, 1e-10)
    h = P.h[0]
    # the exact continuum solution: the face between cells i and i+1 is the series of two
    # half cells, resistance h/(2 K_i) + h/(2 K_i+1)
    x, u_mu = _centre_pressures(P, pL, pR, lambda a, b: h_over(a, b, h))
    r, _ = _residual_and_jacobian(P, x)
    q = P.rho[0] * u_mu / P.mu[0]                 # mass flux [kg/(m^2 s)]
    scale = q / h                                 # one face flux per cell volume
    assert np.abs(r).max() <= 1e-9 * scale
    # an arithmetic-mean face permeability has its own discrete solution; at it the
    # oracle's residual is of the order of the flux itself
    xa, _ = _centre_pressures(P, pL, pR, lambda a, b: h / (0.5 * (a + b)))
    ra, _ = _residual_and_jacobian(P, xa)
    assert np.abs(ra).max() >= 0.1 * scale
    # and the homogeneous column's linear profile is not the heterogeneous solution
    xl = x.copy()
    xl[:, 0] = np.tile(pL + (pR - pL) * (np.arange(P.nx) + 0.5) / P.nx, P.ny * P.nz)
    rl, _ = _residual_and_jacobian(P, xl)
    assert np.abs(rl).max() >= 0.1 * scale


# ------------------------------------------------------------------ upstream choice
PE, LAM, SWR = 1.0e4, 2.0, 0.05


def _bc_laws():
    den = 1.0 - SWR

    def se(S):
        return (S - SWR) / den

    def pc(S):
        return PE * se(S) ** (-1.0 / LAM)

    def krw(S):
        return se(S) ** ((2.0 + 3.0 * LAM) / LAM)

    def krn(S):
        s = se(S)
        return (1.0 - s) ** 2 * (1.0 - s ** ((2.0 + LAM) / LAM))

    def d(f, S):      # complex-step derivative
        return (f(S + 1e-30j)).imag / 1e-30
    return pc, krw, krn, d


def _row_of_three(K0=1e-12, K1=3e-12, K2=2e-12):
    """2D row of three cells (2 m x 4 m each), no gravity, no-flow boundaries, one
    Brooks-Corey region."""
    from ewg import problems
    from ewg.capi import EWG_LAW_BROOKS_COREY
    P = problems.lens_problem(nx=3, ny=1)
    P.gravity = (0.0, 0.0, 0.0)
    P.law = EWG_LAW_BROOKS_COREY
    P.region_params = np.array([[PE, LAM, SWR, 0.0]])
    P.region = np.zeros(P.N, np.int32)
    P.pe = None
    P.K = np.array([K0, K1, K2])
    P.phi = np.full(P.N, 0.3)
    P.bc_type[:] = 0
    P.bc_val[:] = 0.0
    return P


def _face_consts(P):
    hx, hy = P.h[0], P.h[1]
    return hx, hy, hx * hy   # spacing, face area, cell volume


def test_upstream_mobility_opposed_phases():
    pc, krw, krn, d = _bc_laws()
    P = _row_of_three()
    mu_w, mu_n = P.mu
    rho_w, rho_n = P.rho
    h, A, V = _face_consts(P)
    Kf = 2 * P.K[0] * P.K[1] / (P.K[0] + P.K[1])
    S0, S1 = 0.3, 0.9
    x = np.array([[1e5, S0], [1e5 + 100.0, S1], [1e5 + 100.0, S1]])
    r, J = _residual_and_jacobian(P, x)
    # water: potential rises towards cell 1, so cell 1 is upstream (exterior of focus 0);
    # DNAPL: p_n1 - p_n0 = 100 + pc(S1) - pc(S0) < 0, so cell 0 is upstream
    gw = (x[1, 0] - x[0, 0]) / h
    gn = ((x[1, 0] + pc(S1)) - (x[0, 0] + pc(S0))) / h
    assert gw > 0 and gn < 0
    lw1, ln0 = krw(S1) / mu_w, krn(S0) / mu_n
    Fw = -lw1 * Kf * gw * rho_w
    Fn = -ln0 * Kf * gn * rho_n
    np.testing.assert_allclose(r[0], [Fw * A / V, Fn * A / V], rtol=1e-12)
    # the focus column of row 0: the water flux carries no saturation derivative (its
    # mobility is the neighbour's), the DNAPL flux carries dlambda_n/dS and dpc/dS
    dt = P.dt
    phi = P.phi[0]
    c = Kf * A / (h * V)
    exp00 = np.array([
        [lw1 * rho_w * c, phi * rho_w / dt],
        [ln0 * rho_n * c,
         -phi * rho_n / dt + (-d(krn, S0) / mu_n * Kf * gn * rho_n
                              + ln0 * Kf * d(pc, S0) * rho_n / h) * A / V],
    ])
    np.testing.assert_allclose(J[0:2, 0:2], exp00, rtol=1e-10)
    # column 0 of row 1 (cell 0's context): minus the face derivatives
    exp10 = -np.array([[lw1 * rho_w * c, 0.0], exp00[1] - [0.0, -phi * rho_n / dt]])
    np.testing.assert_allclose(J[2:4, 0:2], exp10, rtol=1e-10, atol=1e-30)
    # a downwind choice for either phase moves these values by orders of magnitude
    assert abs(krw(S0) / krw(S1) - 1) > 0.5 and abs(krn(S1) / krn(S0) - 1) > 0.5


def test_upstream_tie_takes_the_focus_cell():
    """grad Phi . n == 0 exactly (equal water pressures, no gravity): the upstream DOF is
    the interior one, i.e. the focus of each element context, so row 1 / column 0 carries
    cell 0's water mobility and row 0 / column 1 cell 1's."""
    pc, krw, krn, d = _bc_laws()
    P = _row_of_three(K0=1e-12, K1=1e-12, K2=1e-12)
    mu_w = P.mu[0]
    rho_w = P.rho[0]
    h, A, V = _face_consts(P)
    S0, S1 = 0.3, 0.9
    x = np.array([[1e5, S0], [1e5, S1], [1e5, S1]])
    r, J = _residual_and_jacobian(P, x)
    assert r[0, 0] == 0.0                      # no water flux at the tie
    c = 1e-12 * rho_w * A / (h * V)
    lw0, lw1 = krw(S0) / mu_w, krw(S1) / mu_w
    assert abs(J[2, 0] - (-lw0 * c)) <= 1e-12 * lw0 * c      # d R1_w / d p_w0
    assert abs(J[0, 2] - (-lw1 * c)) <= 1e-12 * lw1 * c      # d R0_w / d p_w1
    assert abs(lw0 / lw1 - 1) > 0.5


def test_boundary_upstream_by_potential():
    """Free-flow face on x- of cell 0 (darcyfluxmodule.hh:414-446): inflow of water takes
    the boundary state's mobility (kr at the boundary saturation with cell 0's law),
    outflow of DNAPL takes the cell's; K is the cell's own, the distance half a cell."""
    pc, krw, krn, d = _bc_laws()
    P = _row_of_three(K0=2e-12, K1=2e-12, K2=2e-12)
    mu_w, mu_n = P.mu
    rho_w, rho_n = P.rho
    h, A, V = _face_consts(P)
    S0, SB = 0.3, 0.9
    pB = 1e5 + 1000.0
    P.bc_type[0] = 1
    P.bc_val[0] = (pB, SB)
    x = np.array([[1e5, S0], [1e5, S0], [1e5, S0]])
    r, _ = _residual_and_jacobian(P, x)
    # gradient along x from the cell centre to the face centre (h/2 to the left)
    gw = (pB - x[0, 0]) * (-2.0 / h)
    gn = ((pB + pc(SB)) - (x[0, 0] + pc(S0))) * (-2.0 / h)
    # grad . n with n = -e_x: water flows in (boundary upstream), DNAPL flows out
    assert -gw > 0 and -gn < 0
    Fw = (krw(SB) / mu_w) * 2e-12 * gw * rho_w
    Fn = (krn(S0) / mu_n) * 2e-12 * gn * rho_n
    np.testing.assert_allclose(r[0], [Fw * A / V, Fn * A / V], rtol=1e-12)
    assert Fw < 0 and Fn > 0


# ------------------------------------------------------------------ regularization
def _mat_ctx(law, params):
    from ewg import problems
    from oracle.pyoracle import OracleContext
    P = problems.heterogeneous_3d(3, 3, 3, seed=1, patch=1)
    P.law = law
    P.region_params = np.asarray([params], dtype=np.float64)
    P.pe = None
    return OracleContext(P)


@pytest.mark.parametrize("params", [[3.7e-4, 4.7, 0.05, 0.0], [4.5e-4, 7.3, 0.18, 0.02]])
def test_regularized_van_genuchten_known_values(params):
    """pc linear below Se = 0.01 with the slope there; on [0.99, 1] the cubic with
    value and slope of the law at 0.99, value 0 and slope m1 = -2 pc(0.99)/0.01 at 1;
    beyond 1 the line of slope m1 through (1, 0). kr is the closed form inside (0, 1)."""
    alpha, n, swr, snr = params
    m = 1.0 - 1.0 / n
    den = 1.0 - swr - snr
    o = _mat_ctx(0, params)

    def pc(se):
        return (se ** (-1.0 / m) - 1.0) ** (1.0 / n) / alpha

    def dpc(se):
        return pc(se + 1e-30j).imag / 1e-30

    lo, hi = 0.01, 0.99
    m1 = -2.0 * pc(hi) / (1.0 - hi)
    # cubic a + b s + c s^2 + e s^3 in s = Se - 0.99 from its four end conditions
    sL = 1.0 - hi
    M = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [1, sL, sL ** 2, sL ** 3], [0, 1, 2 * sL, 3 * sL ** 2]],
                 dtype=np.float64)
    coef = np.linalg.solve(M, [pc(hi), dpc(hi), 0.0, m1])
    cubic = lambda s: coef[0] + coef[1] * s + coef[2] * s ** 2 + coef[3] * s ** 3  # noqa: E731
    dcubic = lambda s: coef[1] + 2 * coef[2] * s + 3 * coef[3] * s ** 2  # noqa: E731
    expect = {
        0.005: (pc(lo) + dpc(lo) * (0.005 - lo), dpc(lo)),
        0.995: (cubic(0.995 - hi), dcubic(0.995 - hi)),
        1.02: (m1 * 0.02, m1),
    }
    for se, (v, slope) in expect.items():
        out = o.material(0, swr + se * den)
        assert abs(out[0] - v) <= 1e-10 * abs(v), (se, out[0], v)
        assert abs(out[1] - slope / den) <= 1e-8 * abs(slope / den), (se, out[1], slope / den)
    # kr: the closed forms inside (0, 1), clamped beyond
    for se in (0.005, 0.995):
        out = o.material(0, swr + se * den)
        krw = math.sqrt(se) * (1 - (1 - se ** (1 / m)) ** m) ** 2
        krn = (1 - se) ** (1 / 3) * (1 - se ** (1 / m)) ** (2 * m)
        assert abs(out[2] - krw) <= 1e-12 * max(krw, 1e-300)
        assert abs(out[4] - krn) <= 1e-12 * max(krn, 1e-300)
    out = o.material(0, swr + 1.02 * den)
    assert (out[2], out[3], out[4], out[5]) == (1.0, 0.0, 0.0, 0.0)


@pytest.mark.parametrize("params", [[1e4, 2.0, 0.05, 0.0], [2.5e3, 3.5, 0.1, 0.05]])
def test_regularized_brooks_corey_known_values(params):
    """pc linear below Se = 0.01 with the slope there; beyond Se = 1 the line through
    (1, pe) with the slope of the law at 1; kr clamped beyond 1."""
    pe, lam, swr, snr = params
    den = 1.0 - swr - snr
    o = _mat_ctx(1, params)
    lo = 0.01
    pc_lo = pe * lo ** (-1.0 / lam)
    s_lo = pe * (-1.0 / lam) * lo ** (-1.0 / lam - 1.0)
    s_one = pe * (-1.0 / lam)
    expect = {
        0.005: (pc_lo + s_lo * (0.005 - lo), s_lo),
        0.5: (pe * 0.5 ** (-1.0 / lam), pe * (-1.0 / lam) * 0.5 ** (-1.0 / lam - 1.0)),
        1.02: (pe + s_one * 0.02, s_one),
    }
    for se, (v, slope) in expect.items():
        out = o.material(0, swr + se * den)
        assert abs(out[0] - v) <= 1e-12 * abs(v), (se, out[0], v)
        assert abs(out[1] - slope / den) <= 1e-10 * abs(slope / den), (se, out[1], slope / den)
    out = o.material(0, swr + 0.005 * den)
    assert abs(out[2] - 0.005 ** ((2 + 3 * lam) / lam)) <= 1e-12 * 0.005 ** ((2 + 3 * lam) / lam)
    krn = (1 - 0.005) ** 2 * (1 - 0.005 ** ((2 + lam) / lam))
    assert abs(out[4] - krn) <= 1e-12
    out = o.material(0, swr + 1.02 * den)
    assert (out[2], out[3], out[4], out[5]) == (1.0, 0.0, 0.0, 0.0)
