"""Pins the preprocessing oracle (oracle/preproc.py) on the CPU."""
import numpy as np
import pytest

from oracle import preproc


@pytest.mark.parametrize("n,deg", [(2, 1), (7, 3), (1000, 4), (4097, 2)])
def test_legendre_design_matches_numpy_and_closed_forms(n, deg):
    D = preproc.legendre_design(n, deg)
    x = np.arange(n) + 0.0
    x -= x.mean()
    x /= x.max()
    ref = np.polynomial.legendre.legvander(x, deg)
    assert np.max(np.abs(D - ref)) < 1e-13
    if deg >= 3:
        assert np.max(np.abs(D[:, 2] - (3 * x ** 2 - 1) / 2)) < 1e-13
        assert np.max(np.abs(D[:, 3] - (5 * x ** 3 - 3 * x) / 2)) < 1e-13


def _problem(ndet, n, seed):
    rng = np.random.default_rng(seed)
    cm = np.cumsum(rng.standard_normal(n)) * 0.1 + 3.0
    x = np.linspace(-1, 1, n)
    drift = rng.standard_normal((ndet, 3)) @ np.stack([np.ones(n), x, (3 * x ** 2 - 1) / 2])
    gains = 1 + 0.05 * rng.standard_normal(ndet)
    return gains[:, None] * cm[None, :] + drift, cm, drift


@pytest.mark.parametrize("ord,cm_ord", [(2, 1), (3, 2), (0, 1)])
def test_fit_matches_independent_least_squares(ord, cm_ord):
    dat, _, _ = _problem(9, 3001, 1)
    dat += 0.01 * np.random.default_rng(2).standard_normal(dat.shape)
    dd, pred2, cm = preproc.fit_cm_plus_poly(dat, ord, cm_ord, full_out=True)
    assert np.array_equal(cm, np.median(dat, axis=0))
    n = dat.shape[1]
    mm = np.concatenate([preproc.legendre_design(n, ord), preproc.legendre_design(n, cm_ord - 1) * cm[:, None]], 1)
    sol = np.linalg.lstsq(mm, dat.T, rcond=None)[0]
    assert np.allclose(dd, dat - (mm[:, :ord + 1] @ sol[:ord + 1]).T, rtol=0, atol=1e-9)
    assert np.allclose(pred2, (mm[:, ord + 1:] @ sol[ord + 1:]).T, rtol=0, atol=1e-9)
    # residual dat - pred1 - pred2 is orthogonal to every design column
    res = dat - (dat - dd) - pred2
    assert np.max(np.abs(mm.T @ res.T)) < 1e-7 * np.abs(dat).max() * n


def _median_known(n, seed=3):
    """Detector 0 is the common mode m(t); detectors 1-2 (3-4) add (subtract)
    quadratic drifts that stay positive on [-1, 1], so detector 0 is the
    median at every sample."""
    rng = np.random.default_rng(seed)
    m = np.cumsum(rng.standard_normal(n)) * 0.1
    x = np.linspace(-1, 1, n)
    d = [2 + 0.5 * x + 0.3 * x ** 2, 3 - x + 0.2 * x ** 2, 1.5 + 0.2 * x - 0.1 * x ** 2, 2.5 + x ** 2]
    dat = np.stack([m, m + d[0], m + d[1], m - d[2], m - d[3]])
    return dat, m


def test_fit_is_exact_when_the_median_is_known():
    dat, m = _median_known(3000)
    assert np.array_equal(np.median(dat, axis=0), m)
    dd, pred2, cm = preproc.fit_cm_plus_poly(dat, 2, 1, full_out=True)
    # every detector is m plus a drift of order <= 2: removing the drifts
    # leaves m, and the fitted common mode is m itself (coefficient 1)
    assert np.max(np.abs(dd - m[None, :])) < 1e-10
    assert np.max(np.abs(pred2 - m[None, :])) < 1e-10


@pytest.mark.parametrize("n,fac,k", [(1001, 10, 7), (4097, 4, 100), (257, 2, 60)])
def test_downsample_cosine_and_constant(n, fac, k):
    t = np.arange(n)
    nn = int(n / fac)
    x = np.cos(np.pi * k * t / (n - 1))
    y = preproc.downsample_array(x, fac)
    assert y.shape == (nn,)
    assert np.max(np.abs(y - np.cos(np.pi * k * np.arange(nn) / (nn - 1)))) < 1e-11
    c = preproc.downsample_array(np.full((2, n), 2.5), fac)
    assert np.max(np.abs(c - 2.5)) < 1e-12
