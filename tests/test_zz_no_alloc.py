"""mk_dct1_f64_noalloc (the DCT-I that allocates nothing, with the library's
static 64 MiB scratch) against mk_dct1_f64 with a caller workspace, through
the C ABI: bit for bit on two calls in a row, the second smaller, and a
clear error above the scratch size.  Named to run after every other test."""
import ctypes

import numpy as np
import pytest
import scipy.fft as sf

pytestmark = pytest.mark.gpu


def _dct_pair(x):
    import torch
    from minkasi_amd import device
    from minkasi_amd.lib import mkgpu
    nrow, n = x.shape
    y_ref = torch.empty_like(x)
    y_na = torch.full_like(x, float("nan"))
    nb = int(mkgpu.lib.mk_dct1_workspace_bytes(nrow, n, 8))
    work = torch.empty(nb // 8 + 64, dtype=torch.float64, device=x.device)
    assert work.data_ptr() % 256 == 0
    sp = device.stream_ptr()
    mkgpu.check("mk_dct1_f64", x.data_ptr(), y_ref.data_ptr(), nrow, n, work.data_ptr(), nb, sp)
    mkgpu.check("mk_dct1_f64_noalloc", x.data_ptr(), y_na.data_ptr(), nrow, n, sp)
    torch.cuda.synchronize()
    return y_ref, y_na


def test_dct1_noalloc_equals_workspace_version():
    import torch
    rng = np.random.default_rng(42)
    # two calls in a row, the second smaller; the first nearly fills the scratch
    for nrow, n in [(30, 65536), (5, 40001), (3, 17)]:
        x = torch.from_numpy(rng.standard_normal((nrow, n))).cuda()
        y_ref, y_na = _dct_pair(x)
        assert torch.equal(y_na, y_ref), (nrow, n)
        ref = sf.dct(x.cpu().numpy(), type=1, axis=-1)
        e = np.abs(y_na.cpu().numpy() - ref).max() / np.abs(ref).max()
        print("noalloc DCT-I %d x %d: bitwise equal to the workspace call, %.2e from scipy" % (nrow, n, e))
        assert e < 1e-12


def test_dct1_noalloc_rejects_sizes_above_scratch():
    import torch
    from minkasi_amd import device
    from minkasi_amd.lib import mkgpu
    nrow, n = 31, 65536
    assert mkgpu.lib.mk_dct1_workspace_bytes(nrow, n, 8) > 64 << 20
    x = torch.zeros((nrow, n), dtype=torch.float64, device="cuda")
    y = torch.full_like(x, 3.0)
    st = mkgpu.fn("mk_dct1_f64_noalloc")(x.data_ptr(), y.data_ptr(), nrow, n, device.stream_ptr())
    assert st == mkgpu.MK_ERR_ARG
    msg = mkgpu.last_error()
    assert "static" in msg and "scratch" in msg, msg
    torch.cuda.synchronize()
    assert bool((y == 3.0).all())       # nothing was launched
