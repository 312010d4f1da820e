"""Gaussian smoothing of spectra on the GPU (minkasi/tools/smooth.py:49-99).

smooth_many_vecs computes DCT-I(DCT-I(x) * DCT-I(g)) / (2(n-1)), which is
the circular convolution of the even extensions of x and g; the GPU kernel
evaluates that convolution directly over the Gaussian's non-negligible taps.
"""
import numpy as np
import torch

from .. import device as dev
from ..lib import mkgpu

_TAP_SIGMAS = 13.0  # exp(-0.5*13^2) ~ 2e-37: taps beyond this are below FP64 resolution of the sum


def gaussian_taps(n, fwhm):
    """r2r-normalised Gaussian exactly as smooth.py:68-73 builds it."""
    x = np.arange(n)
    sig = fwhm / np.sqrt(8 * np.log(2))
    to_conv = np.exp(-0.5 * (x / sig) ** 2)
    tot = to_conv[0] + to_conv[-1] + 2 * to_conv[1:-1].sum()
    return to_conv / tot


def tap_range(n, fwhm):
    """(t_lo, t_hi, taps) for the direct circular convolution."""
    m = n - 1
    g = gaussian_taps(n, fwhm)
    sig = fwhm / np.sqrt(8 * np.log(2))
    T = int(np.ceil(_TAP_SIGMAS * sig))
    if T >= m:
        return -(m - 1), m, g
    return -T, T, g[: T + 1]


def smooth_rows(x, fwhm, square=False, invert=False, want_f32=False):
    """Device FP64 rows -> smoothed FP64 rows (and optional FP32 copy)."""
    nrow, n = x.shape
    t_lo, t_hi, g = tap_range(n, fwhm)
    gd = torch.from_numpy(np.ascontiguousarray(g)).to(x.device)
    out = torch.empty_like(x)
    out32 = torch.empty(x.shape, dtype=torch.float32, device=x.device) if want_f32 else None
    flags = (1 if square else 0) | (2 if invert else 0)
    mkgpu.check("mk_smooth_rows_f64", dev.ptr(x, torch.float64, "x"), out.data_ptr(),
                out32.data_ptr() if out32 is not None else None, nrow, n, gd.data_ptr(), t_lo, t_hi, flags,
                dev.stream_ptr())
    return (out, out32) if want_f32 else out


def smooth_many_vecs(vecs, fwhm=20):
    is_t = isinstance(vecs, torch.Tensor)
    x = vecs if is_t else torch.from_numpy(np.ascontiguousarray(vecs, dtype=np.float64))
    x = x.to(dev.device(), torch.float64).contiguous()
    out = smooth_rows(x, fwhm)
    return out if is_t else out.cpu().numpy()


def smooth_vec(vec, fwhm=20):
    return smooth_many_vecs(np.asarray(vec)[None, :], fwhm)[0]
