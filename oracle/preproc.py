"""Oracle: TOD preprocessing in FP64 (test infrastructure).

Functions, what they follow, and what pins them (tests/test_oracle_preproc.py):
  legendre_design   -- minkasi/tods/processing.py:554-559 (xx normalised to
                       [-1, 1], legvander); the Legendre values by the
                       three-term recurrence.  Pinned by
                       numpy.polynomial.legendre.legvander and the closed
                       forms P2 = (3x^2 - 1)/2, P3 = (5x^3 - 3x)/2.
  fit_cm_plus_poly  -- processing.py:510-579 with _linfit_2mat :28-39 (normal
                       equations through an explicit inverse).  Pinned by an
                       independent least-squares solve (numpy.linalg.lstsq on
                       the same design), by exactness on data that is a pure
                       drift plus a known common mode, and by the residual
                       being orthogonal to every design column.
  downsample_array  -- minkasi/tools/array_ops.py:17-42 (DCT-I, keep
                       int(n / fac) coefficients, DCT-I of that length,
                       / 2(n-1)).  Pinned by downsampling a pure cosine
                       cos(pi k t / (n-1)), 0 < k < nn - 1 (its DCT-I is
                       (n-1) delta_jk), which must return cos(pi k t' / (nn-1))
                       on the nn-point grid, and a constant, which must stay
                       that constant.
"""
import numpy as np

from .noise import fft_r2r


def legendre_design(n, deg):
    """n x (deg + 1): P_k(x_t), x_t = (t - (n-1)/2) / ((n-1)/2)."""
    h = 0.5 * (n - 1)
    x = (np.arange(n, dtype=np.float64) - h) / h
    out = np.empty((n, deg + 1))
    out[:, 0] = 1.0
    if deg >= 1:
        out[:, 1] = x
    for k in range(2, deg + 1):
        out[:, k] = (out[:, k - 1] * x * (2 * k - 1) - out[:, k - 2] * (k - 1)) / k
    return out


def fit_cm_plus_poly(dat, ord=2, cm_ord=1, niter=1, full_out=False):
    """Drift-subtracted data (and, with full_out, the fitted common mode and
    the median common mode), FP64."""
    dat = np.asarray(dat, dtype=np.float64)
    ndet, n = dat.shape
    pmat = legendre_design(n, ord)
    cmp = legendre_design(n, cm_ord - 1)
    dd = dat.copy()
    pred2 = np.zeros(n)
    cm = np.zeros(n)
    for _ in range(niter):
        cm = np.median(dd, axis=0)
        cm_mat = cmp * cm[:, None]
        mm = np.concatenate([pmat, cm_mat], axis=1)
        fitp = np.linalg.inv(mm.T @ mm) @ (mm.T @ dat.T)
        pred1 = (pmat @ fitp[:ord + 1]).T
        pred2 = (cm_mat @ fitp[ord + 1:]).T
        dd = dat - pred1
    if full_out:
        return dd, pred2, cm
    return dd


def downsample_array(arr, fac):
    """DCT-I downsampling along the last axis."""
    arr = np.asarray(arr, dtype=np.float64)
    n = arr.shape[-1]
    nn = int(n / fac)
    return fft_r2r(fft_r2r(arr)[..., :nn]) / (2 * (n - 1))
