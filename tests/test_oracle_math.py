"""Pins of the oracle's elementary functions and material laws against closed forms.

The oracle (and the device code) evaluate log/exp/pow with their own fma-based
algorithms so that host and device produce identical bits; these tests tie those
functions to the C library, and the material laws to the published formulas
(van Genuchten 1980 / Mualem; Brooks-Corey 1964 / Burdine; EffToAbsLaw scaling).
"""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def ora():
    from oracle import pyoracle
    return pyoracle


def test_log_exp_against_libm(ora):
    xs = np.exp(np.linspace(-700, 700, 20001))
    worst = 0.0
    for x in xs:
        ref = math.log(x)
        if ref != 0.0:
            worst = max(worst, abs(ora.log_(x) - ref) / abs(ref))
    assert worst < 4e-16
    for x in (1.0, 2.0, 0.5, 1e-300, 5e-324, 1.7976931348623157e308):
        assert abs(ora.log_(x) - math.log(x)) <= 4e-16 * max(1.0, abs(math.log(x)))
    ys = np.linspace(-700, 700, 20001)
    worst = max(abs(ora.exp_(y) - math.exp(y)) / math.exp(y) for y in ys)
    assert worst < 4e-16
    assert ora.exp_(0.0) == 1.0 and ora.log_(1.0) == 0.0


def test_pow_against_libm(ora):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(1e-4, 1.0, 2000), rng.uniform(1.0, 50.0, 500)])
    ys = rng.uniform(-8, 8, xs.size)
    worst = max(abs(ora.pow_(x, y) - x ** y) / (x ** y) for x, y in zip(xs, ys))
    assert worst < 2e-14
    assert ora.pow_(0.0, 2.0) == 0.0 and ora.pow_(1.0, -3.7) == 1.0
    assert math.isinf(ora.pow_(0.0, -1.0))


def _ctx(law, regions, N=27):
    from ewg import problems
    from oracle.pyoracle import OracleContext
    P = problems.heterogeneous_3d(3, 3, 3, seed=1, patch=1)
    P.law = law
    P.region_params = np.asarray(regions, dtype=np.float64)
    P.pe = None
    return OracleContext(P)


def test_van_genuchten_closed_forms():
    alpha, n, swr, snr = 4.5e-4, 7.3, 0.18, 0.05
    o = _ctx(0, [[alpha, n, swr, snr]])
    m = 1.0 - 1.0 / n
    for sw in np.linspace(0.2, 0.94, 57):
        out = o.material(0, sw)
        se = (sw - swr) / (1 - swr - snr)
        assert abs(out[6] - se) < 1e-15
        if 0.01 <= se <= 0.99:
            pc = (se ** (-1 / m) - 1) ** (1 / n) / alpha
            assert abs(out[0] - pc) <= 1e-13 * pc
        krw = math.sqrt(se) * (1 - (1 - se ** (1 / m)) ** m) ** 2
        krn = (1 - se) ** (1 / 3) * (1 - se ** (1 / m)) ** (2 * m)
        assert abs(out[2] - krw) <= 1e-13 * max(krw, 1e-300)
        assert abs(out[4] - krn) <= 1e-13 * max(krn, 1e-300)


def test_brooks_corey_closed_forms():
    pe, lam, swr, snr = 2.5e3, 2.0, 0.05, 0.0
    o = _ctx(1, [[pe, lam, swr, snr]])
    for sw in np.linspace(0.07, 0.999, 61):
        out = o.material(0, sw)
        se = (sw - swr) / (1 - swr - snr)
        assert abs(out[0] - pe * se ** (-1 / lam)) <= 1e-13 * pe * se ** (-1 / lam)
        assert abs(out[2] - se ** ((2 + 3 * lam) / lam)) <= 1e-13
        krn = (1 - se) ** 2 * (1 - se ** ((2 + lam) / lam))
        assert abs(out[4] - krn) <= 1e-13
    # per-cell entry pressure overrides the region value
    out = o.material(0, 0.5, pe=7e3)
    se = (0.5 - swr) / (1 - swr)
    assert abs(out[0] - 7e3 * se ** -0.5) <= 1e-12 * out[0]


@pytest.mark.parametrize("law,params", [(0, [3.7e-4, 4.7, 0.05, 0.0]),
                                         (0, [4.5e-4, 7.3, 0.18, 0.0]),
                                         (1, [1e4, 2.0, 0.05, 0.0])])
def test_material_derivatives_and_regularization(law, params):
    o = _ctx(law, [params])
    swr, snr = params[2], params[3]
    den = 1 - swr - snr
    # analytic derivatives against central differences, away from the branch points
    for se in (0.003, 0.05, 0.3, 0.7, 0.98, 0.995, 1.02):
        sw = swr + se * den
        h = 1e-7 * den
        a = o.material(0, sw)
        p, m_ = o.material(0, sw + h), o.material(0, sw - h)
        for k in (0, 2, 4):
            fd = (p[k] - m_[k]) / (2 * h)
            assert abs(a[k + 1] - fd) <= 1e-5 * max(1.0, abs(fd)), (se, k, a[k + 1], fd)
    # continuity (C0 for kr, C1 for pc) at the regularization thresholds
    for se0 in ((0.01, 0.99, 1.0) if law == 0 else (0.01, 1.0)):
        sw0 = swr + se0 * den
        lo, hi = o.material(0, sw0 - 1e-12), o.material(0, sw0 + 1e-12)
        assert abs(lo[0] - hi[0]) <= 1e-6 * max(1.0, abs(lo[0]))
        assert abs(lo[1] - hi[1]) <= 1e-4 * max(1.0, abs(lo[1]))
    # limits: pc(Se=1) = 0 for van Genuchten, = entry pressure for Brooks-Corey
    one = o.material(0, swr + den)
    assert (abs(one[0]) < 1e-9) if law == 0 else abs(one[0] - params[0]) < 1e-9 * params[0]
    # relative permeabilities clamp outside (0, 1)
    below, above = o.material(0, swr - 0.01), o.material(0, 1.0 + 0.01)
    assert below[2] == 0.0 and below[4] == 1.0 and above[2] == 1.0 and above[4] == 0.0
    assert below[3] == 0.0 and above[5] == 0.0
