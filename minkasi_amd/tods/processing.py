"""TOD preprocessing: truncation to an FFT-friendly length.

truncate_tod follows minkasi/tods/processing.py:648-681.  The TOD keeps
ndata_new = max(find_good_fft_lens(ndata - 1, primes)) + 1 samples (nothing
happens when that is not shorter), and every array of the info dict with an
axis of length ndata is cut to its first ndata_new entries along the first
such axis.  Device tensors are cut on the device (a contiguous copy, as
np.take makes on the host).

find_good_fft_lens restates minkasi/tools/py_fft.py:300-395: the products of
the given primes enumerated in log2 space.  A product is kept when its log2
is at most log2(n + 0.5) + 1e-3; every partial product without the last
prime's factor must stay below log2(n + 0.5).  The 1e-3 slack means a
product slightly above n can be returned, in which case truncate_tod leaves
the TOD as it is, like the reference.
"""
import math

import numpy as np


def _products(lim, logs, start, out):
    # products of the remaining primes times 2**start, log2 <= lim (+ slack on the last)
    step = logs[0]
    if len(logs) == 1:
        for v in np.arange(start, lim + 1e-3, step):
            out.append(v)
        return
    for v in np.arange(start, lim, step):
        _products(lim, logs[1:], v, out)


def find_good_fft_lens(n, primes=(2, 3, 5, 7)):
    """Sorted array of the products of `primes` up to about n."""
    lim = math.log2(n + 0.5)
    vals = []
    _products(lim, np.log2(np.asarray(primes, dtype=float)), 0.0, vals)
    return np.sort(np.asarray(np.round(2.0 ** np.asarray(vals)), dtype=np.int64))


def good_length(ndata, primes=(2, 3, 5, 7, 11)):
    """The TOD length truncate_tod keeps for ndata samples."""
    new = int(find_good_fft_lens(ndata - 1, primes).max()) + 1
    return new if new < ndata else ndata


def truncate_tod(tod_info, primes=(2, 3, 5, 7, 11)):
    """Cut the TOD in place to a good FFT length (processing.py:648-681).
    Requires 'dat_calib' (2-D); returns the new length (or None without it)."""
    if "dat_calib" not in tod_info:
        return None
    dat = tod_info["dat_calib"]
    if len(dat.shape) != 2:
        raise ValueError("dat_calib not 2d")
    ndata = int(dat.shape[1])
    new = good_length(ndata, primes)
    if new >= ndata:
        return ndata
    for key in list(tod_info.keys()):
        a = tod_info[key]
        shape = getattr(a, "shape", None)
        if shape is None or len(shape) == 0:
            continue
        axes = [i for i, s in enumerate(shape) if s == ndata]
        if not axes:
            continue
        if hasattr(a, "narrow"):  # torch tensor, any device
            tod_info[key] = a.narrow(axes[0], 0, new).contiguous()
        else:
            tod_info[key] = np.take(a, indices=range(0, new), axis=axes[0])
    return new
