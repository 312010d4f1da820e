"""Pins the oracle's DCT-I, smoothing and NoiseSmoothedSVD (CPU only)."""
import numpy as np
import pytest

from oracle import core, noise


def circular_smooth(P, g):
    """Closed form: circular convolution (period 2(n-1)) of even extensions."""
    n = P.shape[1]
    m = n - 1
    N = 2 * m
    ext = lambda a: np.concatenate([a, a[..., 1:m][..., ::-1]], axis=-1)  # length N
    Pe, ge = ext(P), ext(g)
    out = np.zeros_like(P)
    for k in range(n):
        idx = (k - np.arange(N)) % N
        out[:, k] = (Pe * ge[idx][None, :]).sum(axis=1)
    return out


def test_fft_r2r_matches_definition():
    x = np.random.default_rng(0).standard_normal((4, 33))
    assert np.allclose(noise.fft_r2r(x), core.dct1_direct(x), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n,fwhm", [(40, 10.0), (101, 50.0), (9, 50.0)])
def test_smooth_many_vecs_is_circular_convolution(n, fwhm):
    rng = np.random.default_rng(n)
    P = np.abs(rng.standard_normal((3, n))) + 0.1
    ref = noise.smooth_many_vecs(P, fwhm)
    g = noise.gaussian_taps(n, fwhm)
    assert np.allclose(ref, circular_smooth(P, g), rtol=1e-10, atol=1e-12)
    # smoothing a constant returns the constant (g is r2r-normalised)
    c = noise.smooth_many_vecs(np.full((1, n), 3.0), fwhm)
    assert np.allclose(c, 3.0, rtol=1e-10)


def _tiny_noise(seed=0, ndet=4, n=24, fwhm=5.0):
    rng = np.random.default_rng(seed)
    cm = np.cumsum(rng.standard_normal(n)) * 0.3
    d = rng.standard_normal((ndet, n)) * 0.1 + cm[None, :]
    return noise.NoiseSmoothedSVD(d, fwhm=fwhm), rng


def test_apply_noise_equals_dense_matrix_construction():
    nz, rng = _tiny_noise()
    ndet, n = nz.v.shape[0], nz.mywt.shape[1]
    A = noise.dense_operator(nz, n)
    x = rng.standard_normal((ndet, n))
    assert np.allclose(nz.apply_noise(x).ravel(), A @ x.ravel(), rtol=1e-10, atol=1e-10)
    # a dropped edge halving or a transposed rotation must be caught
    bad = nz.apply_noise(x)
    bad[:, 0] *= 2
    assert not np.allclose(bad.ravel(), A @ x.ravel(), rtol=1e-6)


def test_apply_noise_symmetric_psd_and_kills_constants():
    nz, rng = _tiny_noise(1, ndet=5, n=40)
    ndet, n = nz.v.shape[0], nz.mywt.shape[1]
    A = noise.dense_operator(nz, n)
    assert np.allclose(A, A.T, rtol=1e-9, atol=1e-9 * np.abs(A).max())
    ev = np.linalg.eigvalsh(0.5 * (A + A.T))
    assert ev.min() > -1e-9 * ev.max()
    const = np.outer(rng.standard_normal(ndet), np.ones(n))
    assert np.abs(nz.apply_noise(const)).max() < 1e-9 * np.abs(A).max()


def test_noise_invariant_to_singular_vector_signs():
    rng = np.random.default_rng(3)
    d = rng.standard_normal((4, 50))
    a = noise.NoiseSmoothedSVD(d, fwhm=5.0)
    u, s, _ = np.linalg.svd(d, False)
    flip = u * np.array([1, -1, 1, -1])[None, :]
    b = noise.NoiseSmoothedSVD(d, fwhm=5.0, u_in=flip)
    x = rng.standard_normal((4, 50))
    assert np.allclose(a.apply_noise(x), b.apply_noise(x), rtol=1e-9, atol=1e-12)


def test_get_det_weights_formula():
    nz, _ = _tiny_noise(2)
    w = nz.get_det_weights()
    ref = 2 * np.einsum("ij,j,ji->i", nz.vT, nz.mywt.sum(axis=1), nz.v)
    assert np.allclose(w, ref)


def _single_bin_closed_form(smooth, fwhm, n=2001, k0=1000):
    """smooth(delta at bin k0) against the Gaussian the FWHM defines, with no
    use of the taps' construction: sigma = fwhm / sqrt(8 ln 2), the profile
    exp(-j^2 / (2 sigma^2)) around k0 with half maximum at +-fwhm/2 bins,
    and peak 1/(sqrt(2 pi) sigma), since the r2r-normalised taps sum to the
    Gaussian's integral (Poisson summation: relative error about
    2 exp(-2 pi^2 sigma^2), below 1e-100 here).  The mirror image of the
    bin lies 2(n-1) - 2 k0 = 2000 bins away and does not contribute."""
    P = np.zeros((1, n))
    P[0, k0] = 1.0
    s = np.asarray(smooth(P, fwhm))[0]
    sig = fwhm / np.sqrt(8 * np.log(2))
    j = np.arange(-int(3 * fwhm), int(3 * fwhm) + 1)
    prof = s[k0 + j] / s[k0]
    assert np.allclose(prof, np.exp(-0.5 * (j / sig) ** 2), rtol=0, atol=1e-13)
    h = int(round(fwhm / 2))
    assert abs(s[k0 + h] / s[k0] - 0.5) < 1e-13 and abs(s[k0 - h] / s[k0] - 0.5) < 1e-13
    assert abs(s[k0] * np.sqrt(2 * np.pi) * sig - 1.0) < 1e-12
    # a width used as sigma (or sigma as the width) would fail both checks
    wrong = np.exp(-0.5 * (h / fwhm) ** 2)
    assert abs(wrong - 0.5) > 0.3


@pytest.mark.parametrize("fwhm", [10.0, 50.0])
def test_smoothing_width_closed_form(fwhm):
    """Pins the FWHM -> sigma conversion of smooth_many_vecs
    (minkasi/tools/smooth.py:66-78) by a closed form, independently of
    gaussian_taps."""
    _single_bin_closed_form(noise.smooth_many_vecs, fwhm)
