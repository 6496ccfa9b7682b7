"""Pins of the oracle's block ILU(0), SpMV and BiCGStab against brute force / numpy.

Block ILU(0) is checked against a dense, masked, textbook block ILU(0) (Saad, Iterative
Methods for Sparse Linear Systems, Alg. 10.4 in block form); on a 1D chain ILU(0) is the
exact block LU, so the preconditioned BiCGStab converges at the first half step.
"""
import numpy as np
import pytest

from helpers import bcrs_to_dense


def _ctx(nx, ny, nz, dim):
    from ewg import problems
    from oracle.pyoracle import OracleContext
    if dim == 2:
        P = problems.lens_problem(nx=nx, ny=ny)
    else:
        P = problems.heterogeneous_3d(nx, ny, nz, seed=1, patch=1)
    return P, OracleContext(P)


def _random_matrix(o, P, seed, dominance=3.0):
    rng = np.random.default_rng(seed)
    rp, ci = o.pattern()
    A = rng.standard_normal(4 * len(ci))
    for i in range(P.N):
        for p in range(rp[i], rp[i + 1]):
            if ci[p] == i:
                blk = A[4 * p:4 * p + 4].reshape(2, 2)
                blk += np.eye(2) * dominance * (rp[i + 1] - rp[i])
    return rp, ci, A


def _dense_block_ilu0(rp, ci, A, N):
    blocks = {}
    for i in range(N):
        for p in range(rp[i], rp[i + 1]):
            blocks[(i, ci[p])] = A[4 * p:4 * p + 4].reshape(2, 2).copy()
    for i in range(N):
        for k in range(i):
            if (i, k) not in blocks:
                continue
            blocks[(i, k)] = blocks[(i, k)] @ np.linalg.inv(blocks[(k, k)])
            for j in range(k + 1, N):
                if (i, j) in blocks and (k, j) in blocks:
                    blocks[(i, j)] = blocks[(i, j)] - blocks[(i, k)] @ blocks[(k, j)]
    return blocks


@pytest.mark.parametrize("shape", [(5, 4, 1, 2), (3, 3, 3, 3), (4, 3, 2, 3), (2, 3, 1, 2)])
def test_ilu0_matches_dense_block_ilu0(shape):
    nx, ny, nz, dim = shape
    P, o = _ctx(nx, ny, nz, dim)
    rp, ci, A = _random_matrix(o, P, seed=sum(shape))
    LU = o.ilu0_factor(A)
    ref = _dense_block_ilu0(rp, ci, A, P.N)
    for i in range(P.N):
        for p in range(rp[i], rp[i + 1]):
            j = ci[p]
            got = LU[4 * p:4 * p + 4].reshape(2, 2)
            want = np.linalg.inv(ref[(i, j)]) if i == j else ref[(i, j)]
            assert np.allclose(got, want, rtol=1e-11, atol=1e-12 * np.abs(want).max()), (i, j)


@pytest.mark.parametrize("shape", [(5, 4, 1, 2), (3, 3, 3, 3)])
def test_ilu0_apply_inverts_the_factors(shape):
    nx, ny, nz, dim = shape
    P, o = _ctx(nx, ny, nz, dim)
    rp, ci, A = _random_matrix(o, P, seed=7)
    LU = o.ilu0_factor(A)
    N = P.N
    Lm = np.eye(2 * N)
    Um = np.zeros((2 * N, 2 * N))
    for i in range(N):
        for p in range(rp[i], rp[i + 1]):
            j = ci[p]
            blk = LU[4 * p:4 * p + 4].reshape(2, 2)
            if j < i:
                Lm[2 * i:2 * i + 2, 2 * j:2 * j + 2] = blk
            elif j == i:
                Um[2 * i:2 * i + 2, 2 * i:2 * i + 2] = np.linalg.inv(blk)
            else:
                Um[2 * i:2 * i + 2, 2 * j:2 * j + 2] = blk
    d = np.random.default_rng(1).standard_normal(2 * N)
    v = o.ilu0_apply(LU, d)
    assert np.allclose(Lm @ Um @ v, d, rtol=1e-11, atol=1e-11)


def test_structured_ilu0_only_updates_the_diagonal():
    """For nx, ny >= 3 the factorization leaves every off-diagonal upper block equal to A
    and every lower block equal to A_ij inv(D_j): the property the device factorization
    relies on."""
    P, o = _ctx(4, 3, 3, 3)
    rp, ci, A = _random_matrix(o, P, seed=3)
    LU = o.ilu0_factor(A)
    diag = {}
    for i in range(P.N):
        for p in range(rp[i], rp[i + 1]):
            if ci[p] == i:
                diag[i] = LU[4 * p:4 * p + 4].reshape(2, 2)
    for i in range(P.N):
        for p in range(rp[i], rp[i + 1]):
            j = ci[p]
            a = A[4 * p:4 * p + 4].reshape(2, 2)
            got = LU[4 * p:4 * p + 4].reshape(2, 2)
            if j > i:
                assert np.array_equal(got, a)
            elif j < i:
                assert np.allclose(got, a @ diag[j], rtol=1e-13, atol=1e-15)


def test_spmv_matches_dense():
    P, o = _ctx(4, 3, 2, 3)
    rp, ci, A = _random_matrix(o, P, seed=5)
    x = np.random.default_rng(2).standard_normal(2 * P.N)
    Ad = bcrs_to_dense(rp, ci, A, P.N)
    assert np.allclose(o.spmv(A, x), Ad @ x, rtol=1e-13, atol=1e-13)


def test_bicgstab_matches_numpy_solve():
    from ewg import problems
    from oracle.pyoracle import OracleContext
    P = problems.lens_problem(nx=6, ny=5)
    params = problems.default_params(linear_tolerance=1e-13, linear_abs_tolerance=0.0)
    o = OracleContext(P, params)
    rp, ci, A = _random_matrix(o, P, seed=11, dominance=1.0)
    b = np.random.default_rng(4).standard_normal(2 * P.N)
    rc, x, its = o.bicgstab(A, b)
    assert rc == 0
    ref = np.linalg.solve(bcrs_to_dense(rp, ci, A, P.N), b)
    assert np.allclose(x, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


def test_bicgstab_relative_criterion_is_max_norm_of_recurrence_residual():
    from ewg import problems
    from oracle.pyoracle import OracleContext
    P = problems.lens_problem(nx=8, ny=6)
    params = problems.default_params(linear_tolerance=1e-3, linear_abs_tolerance=0.0)
    o = OracleContext(P, params)
    rp, ci, A = _random_matrix(o, P, seed=12, dominance=0.6)
    b = np.random.default_rng(5).standard_normal(2 * P.N)
    rc, x, its = o.bicgstab(A, b)
    assert rc == 0
    res = b - bcrs_to_dense(rp, ci, A, P.N) @ x
    # the recurrence residual meets 1e-3 * max|b|; the true residual agrees to rounding
    assert np.abs(res).max() <= 1.0001e-3 * np.abs(b).max()


def test_chain_ilu0_is_exact():
    from ewg import problems
    from oracle.pyoracle import OracleContext
    P = problems.lens_problem(nx=9, ny=1)
    params = problems.default_params(linear_tolerance=1e-12, linear_abs_tolerance=0.0)
    o = OracleContext(P, params)
    rp, ci, A = _random_matrix(o, P, seed=13)
    b = np.random.default_rng(6).standard_normal(2 * P.N)
    rc, x, its = o.bicgstab(A, b)
    assert rc == 0 and its == 0
    ref = np.linalg.solve(bcrs_to_dense(rp, ci, A, P.N), b)
    assert np.allclose(x, ref, rtol=1e-12, atol=1e-12)


def test_bicgstab_zero_rhs_converges_immediately():
    from ewg import problems
    from oracle.pyoracle import OracleContext
    P = problems.lens_problem(nx=5, ny=4)
    o = OracleContext(P)
    rp, ci, A = _random_matrix(o, P, seed=14)
    rc, x, its = o.bicgstab(A, np.zeros(2 * P.N))
    assert rc == 0 and its == 0 and not x.any()
