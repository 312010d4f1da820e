"""Array helpers of minkasi/tools/array_ops.py on the GPU DCT-I engine."""
import numpy as np

from .fft import fft_r2r


def downsample_array_r2r(arr, fac, axis=-1):
    """Downsample by `fac` through the DCT-I (array_ops.py:17-42): the first
    int(n / fac) DCT-I coefficients transformed back with a DCT-I of that
    length and divided by 2 (n - 1).  Like the reference, the transforms run
    along the last axis and `axis` only selects where the coefficients are
    cut.  Device tensors stay on the device."""
    n = arr.shape[axis]
    nn = int(n / fac)
    ft = fft_r2r(arr)
    if hasattr(ft, "narrow"):
        ft = ft.narrow(axis if axis >= 0 else ft.dim() + axis, 0, nn).contiguous()
    else:
        ft = np.take(ft, indices=range(0, nn), axis=axis)
    return fft_r2r(ft) / (2 * (n - 1))


def downsample_vec_r2r(vec, fac):
    return downsample_array_r2r(vec, fac)
