"""Oracle: DCT-I, smoothing and NoiseSmoothedSVD in FP64 (test infrastructure).

Functions, what they follow, and what pins them (tests/test_oracle_noise.py):
  fft_r2r           -- minkasi/tools/py_fft.py:206-285 (scipy.fft.dct type 1,
                       the reference's own no-FFTW path); pinned by the
                       brute-force definition (oracle/oracle.c or_dct1_direct),
                       DCT-I^2 = 2(n-1) I and the delta/ones known answers.
  smooth_many_vecs  -- minkasi/tools/smooth.py:49-78, restated line by line;
                       pinned by the closed form (circular convolution of the
                       even extensions) on small n, by smoothing a constant
                       (must return the constant), and, for the Gaussian's
                       width, by a single-bin spectrum that must come back as
                       exp(-j^2 / 2 sigma^2) with sigma = fwhm / sqrt(8 ln 2),
                       half maximum at +-fwhm/2 and peak 1/(sqrt(2 pi) sigma)
                       (test_smoothing_width_closed_form).
  NoiseSmoothedSVD  -- minkasi/mapmaking/noise.py:584-772 with
                       svd(full_matrices=False) (the full_matrices=True call in
                       the reference builds an n x n factor and cannot run at
                       these sizes; u is identical).  Pinned by: symmetry
                       <a, N^-1 b> = <N^-1 a, b>; positive semi-definiteness;
                       N^-1 . const = 0 (mywt[:,0] = 0); invariance to
                       singular-vector sign flips; a dense assembly on tiny
                       inputs equal to vT diag-free construction
                       vT . D . C W C . v written as explicit matrices.
"""
import numpy as np
import scipy.fft


def fft_r2r(dat):
    """DCT-I along the last axis, unnormalised (FFTW REDFT00)."""
    return scipy.fft.dct(np.asarray(dat, dtype=np.float64), type=1, axis=-1)


def gaussian_taps(n, fwhm):
    """The r2r-normalised Gaussian of smooth_many_vecs (smooth.py:68-73)."""
    x = np.arange(n)
    sig = fwhm / np.sqrt(8 * np.log(2))
    to_conv = np.exp(-0.5 * (x / sig) ** 2)
    tot = to_conv[0] + to_conv[-1] + 2 * to_conv[1:-1].sum()
    return to_conv / tot


def smooth_many_vecs(vecs, fwhm=20):
    """smooth.py:49-78."""
    vecs = np.asarray(vecs, dtype=np.float64)
    n = vecs.shape[1]
    to_conv = gaussian_taps(n, fwhm)
    to_conv_ft = fft_r2r(to_conv)
    xtrans = fft_r2r(vecs)
    xtrans *= to_conv_ft
    back = fft_r2r(xtrans)
    return back / (2 * (n - 1))


class NoiseSmoothedSVD:
    """noise.py:584-772 (fit_powlaw unsupported: it needs fitting.power_spectrum)."""

    def __init__(self, dat_use, fwhm=50.0, prewhiten=False, fit_powlaw=False, u_in=None):
        dat_use = np.asarray(dat_use, dtype=np.float64)
        if fit_powlaw:
            raise NotImplementedError("fit_powlaw is outside the mapmaking inner loop")
        if prewhiten:
            noisevec = np.median(np.abs(np.diff(dat_use, axis=1)), axis=1)
            dat_use = dat_use / (np.repeat([noisevec], dat_use.shape[1], axis=0).transpose())
            self.noisevec = noisevec.copy()
        else:
            self.noisevec = None
        if u_in is None:
            u, s, _ = np.linalg.svd(dat_use, False)
            ndet = u.shape[0]
        else:
            u = u_in
            assert u.shape[0] == u.shape[1]
            ndet = u.shape[0]
        self.v = np.zeros([ndet, ndet])
        self.v[:] = u.transpose()
        if u_in is None:
            self.vT = self.v.T
        else:
            self.vT = np.linalg.inv(self.v)
        dat_rot = np.dot(self.v, dat_use)
        dat_trans = fft_r2r(dat_rot)
        spec_smooth = smooth_many_vecs(dat_trans ** 2, fwhm)
        spec_smooth[:, 1:] = 1.0 / spec_smooth[:, 1:]
        spec_smooth[:, 0] = 0
        self.mywt = spec_smooth

    def apply_noise(self, dat):
        dat = np.asarray(dat, dtype=np.float64)
        noisemat = None
        if self.noisevec is not None:
            noisemat = np.repeat([self.noisevec], dat.shape[1], axis=0).transpose()
            dat = dat / noisemat
        dat_rot = np.dot(self.v, dat)
        datft = fft_r2r(dat_rot)
        nn = datft.shape[1]
        datft = datft * self.mywt[:, :nn]
        dat_rot = fft_r2r(datft)
        dd = np.dot(self.vT, dat_rot)
        dd[:, 0] = 0.5 * dd[:, 0]
        dd[:, -1] = 0.5 * dd[:, -1]
        if noisemat is not None:
            dd = dd / noisemat
        return dd

    def get_det_weights(self):
        mode_wt = np.sum(self.mywt, axis=1)
        tmp = np.dot(self.vT, np.dot(np.diag(mode_wt), self.v))
        return np.diag(tmp).copy() * 2.0


def dense_operator(noise, n):
    """N^-1 as an explicit (ndet*n) x (ndet*n) matrix for tiny inputs, built
    from matrices (no FFT): kron(vT, I) . H . kron(I, C W_r C) . kron(v, I)
    with C the DCT-I matrix and H the edge halving."""
    ndet = noise.v.shape[0]
    m = n - 1
    j = np.arange(n)
    C = np.cos(np.pi * np.outer(j, j) / m)
    C[:, 1:-1] *= 2.0
    out = np.zeros((ndet * n, ndet * n))
    Rv = np.kron(noise.v, np.eye(n))
    Rt = np.kron(noise.vT, np.eye(n))
    F = np.zeros((ndet * n, ndet * n))
    for r in range(ndet):
        F[r * n:(r + 1) * n, r * n:(r + 1) * n] = C @ np.diag(noise.mywt[r, :n]) @ C
    h = np.ones(n)
    h[0] = h[-1] = 0.5
    H = np.kron(np.eye(ndet), np.diag(h))
    out = H @ Rt @ F @ Rv
    return out
