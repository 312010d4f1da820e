"""DCT-I on the GPU with the reference's fft_r2r interface.

Counterpart of minkasi/tools/fft.py:290-344 (FFTW many_fft_r2r_1d) and
minkasi/tools/py_fft.py:254-285 (scipy.fft.dct per row).  Only kind 1
(REDFT00) exists, as in the reference's FFTW wrapper.  Device tensors are
transformed in place on the device; numpy arrays are uploaded, transformed
and returned as numpy arrays of the same dtype.
"""
import numpy as np
import torch

from .. import device as dev
from ..lib import mkgpu


def _dct1_tensor(x, out=None, max_work_bytes=1 << 30):
    """Row-wise DCT-I of a 2-D device tensor (FP32 or FP64)."""
    if x.dim() != 2:
        raise ValueError("expected a 2-D tensor")
    nrow, n = x.shape
    if out is None:
        out = torch.empty_like(x)
    if nrow == 0:
        return out
    eb = 8 if x.dtype == torch.float64 else 4
    if x.dtype not in (torch.float32, torch.float64):
        raise TypeError("fft_r2r supports float32 and float64")
    per_row = int(mkgpu.lib.mk_dct1_workspace_bytes(1, n, eb)) - int(mkgpu.lib.mk_dct1_workspace_bytes(0, n, eb))
    chunk = max(1, min(nrow, max_work_bytes // max(per_row, 1)))
    nb = int(mkgpu.lib.mk_dct1_workspace_bytes(chunk, n, eb))
    work = torch.empty(nb // 8 + 64, dtype=torch.float64, device=x.device)
    name = "mk_dct1_f64" if eb == 8 else "mk_dct1_f32"
    xc = x.contiguous()
    for r0 in range(0, nrow, chunk):
        r1 = min(nrow, r0 + chunk)
        mkgpu.check(name, xc[r0:r1].data_ptr(), out[r0:r1].data_ptr(), r1 - r0, n, work.data_ptr(),
                    work.numel() * 8, dev.stream_ptr())
    return out


def fft_r2r(dat, datft=None, kind=1):
    """Many 1-D DCT-I (FFTW_REDFT00) transforms along the last axis."""
    if kind != 1:
        raise ValueError("Error: only kind 1, FFTW_REDFT00 - DCT-I transforms are supported with FFTW")
    if isinstance(dat, torch.Tensor):
        x = dat if dat.dim() == 2 else dat.reshape(1, -1)
        out = _dct1_tensor(x.to(dev.device()))
        out = out.reshape(dat.shape)
        if datft is not None:
            datft.copy_(out)
            return datft
        return out
    a = np.asarray(dat)
    x = torch.from_numpy(np.ascontiguousarray(np.atleast_2d(a))).to(dev.device())
    if x.dtype not in (torch.float32, torch.float64):
        x = x.to(torch.float64)
    res = _dct1_tensor(x).cpu().numpy().reshape(a.shape).astype(a.dtype if a.dtype in (np.float32, np.float64) else np.float64)
    if datft is not None:
        datft[...] = res
        return datft
    return res


def fft_r2r_1d(dat, kind=1):
    return fft_r2r(np.asarray(dat).reshape(-1), kind=kind)


def set_threaded(n=-1):
    """FFTW thread control has no GPU counterpart; kept for API compatibility."""
    return None


def read_wisdom(double_file=".fftw_wisdom", single_file=".fftwf_wisdom"):
    """Counterpart of loading FFTW wisdom (minkasi/tools/fft.py:345-364):
    enables the autotuner (minkasi_amd.tune) from its cache, measuring the
    FFT column lengths once if the cache has none for this device.  The file
    names are FFTW's and are not used; the cache lives at
    minkasi_amd.tune.cache_path()."""
    from .. import tune
    return tune.enable()


def write_wisdom(double_file=".fftw_wisdom", single_file=".fftwf_wisdom"):
    """Saves the active autotuner's cache (minkasi/tools/fft.py:367-384)."""
    from .. import tune
    t = tune.active()
    if t is not None:
        t.save()
    return None
