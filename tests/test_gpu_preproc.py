"""Device TOD preprocessing (fit_cm_plus_poly, downsample_tod) against the
FP64 oracle (oracle/preproc.py) and numpy."""
import numpy as np
import pytest

from oracle import preproc as opre
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    from minkasi_amd import device
    from minkasi_amd.lib import mkgpu
    return torch, device, mkgpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_median_cols_equals_numpy(env, dtype):
    torch, device, mkgpu = env
    rng = np.random.default_rng(0)
    sfx = "f64" if dtype == np.float64 else "f32"
    for nrow in (1, 2, 3, 63, 64, 65, 100, 128, 511, 512, 513, 1000, 1024):
        for ncol in (1, 5, 777):
            ld = ncol + 3
            a = rng.standard_normal((nrow, ld)).astype(dtype)
            a[:, ::7] = np.round(a[:, ::7])  # ties
            at = device.to_device(a)
            out = torch.empty(ncol, dtype=at.dtype, device=at.device)
            mkgpu.check("mk_median_cols_" + sfx, at.data_ptr(), nrow, ncol, ld, out.data_ptr(), device.stream_ptr())
            ref = np.median(a[:, :ncol], axis=0)
            assert np.array_equal(out.cpu().numpy(), ref.astype(dtype)), (nrow, ncol)


def _data(ndet, n, seed):
    rng = np.random.default_rng(seed)
    cm = np.cumsum(rng.standard_normal(n)) * 0.05
    x = np.linspace(-1, 1, n)
    drift = rng.standard_normal((ndet, 3)) @ np.stack([np.ones(n), x, x ** 2])
    return (1 + 0.05 * rng.standard_normal((ndet, 1))) * cm + drift + 0.1 * rng.standard_normal((ndet, n))


@pytest.mark.parametrize("ndet,n,ord,cm_ord,niter", [(64, 20001, 2, 1, 1), (215, 9000, 3, 2, 2), (7, 1001, 0, 1, 1),
                                                      (512, 40000, 2, 1, 1)])
def test_fit_cm_plus_poly_matches_oracle(env, ndet, n, ord, cm_ord, niter):
    torch, device, mkgpu = env
    from minkasi_amd.tods.processing import fit_cm_plus_poly
    dat = _data(ndet, n, ndet + n)
    dd_r, p2_r, cm_r = opre.fit_cm_plus_poly(dat, ord, cm_ord, niter, full_out=True)
    dd, p2, cm = fit_cm_plus_poly(device.to_device(dat), ord, cm_ord, niter, full_out=True)
    assert rel_err(dd.cpu().numpy(), dd_r) < 1e-10
    assert rel_err(p2.cpu().numpy(), p2_r) < 1e-10
    assert np.array_equal(cm.cpu().numpy(), cm_r) or niter > 1
    if niter == 1:
        d32 = fit_cm_plus_poly(device.to_device(dat.astype(np.float32)), ord, cm_ord, niter)
        ref32 = opre.fit_cm_plus_poly(dat.astype(np.float32).astype(np.float64), ord, cm_ord, niter)
        assert d32.dtype == torch.float32
        assert rel_err(d32.cpu().numpy(), ref32) < 1e-5
    host = fit_cm_plus_poly(dat, ord, cm_ord, niter)  # numpy in, FP64 numpy out
    assert isinstance(host, np.ndarray) and host.dtype == np.float64
    assert rel_err(host, dd_r) < 1e-10


def test_fit_cm_plus_poly_exact_when_median_is_known(env):
    torch, device, mkgpu = env
    from minkasi_amd.tods.processing import fit_cm_plus_poly
    rng = np.random.default_rng(4)
    n = 5000
    m = np.cumsum(rng.standard_normal(n)) * 0.1
    x = np.linspace(-1, 1, n)
    dat = np.stack([m, m + 2 + 0.5 * x + 0.3 * x ** 2, m + 3 - x, m - 1.5 - 0.2 * x ** 2, m - 2.5 - x ** 2])
    dd, p2, cm = fit_cm_plus_poly(device.to_device(dat), 2, 1, full_out=True)
    assert np.array_equal(cm.cpu().numpy(), m)
    assert np.max(np.abs(dd.cpu().numpy() - m)) < 1e-10
    assert np.max(np.abs(p2.cpu().numpy() - m)) < 1e-10


def test_downsample_tod_matches_oracle(env):
    torch, device, mkgpu = env
    from minkasi_amd.tods.processing import downsample_tod
    rng = np.random.default_rng(5)
    ndet, n = 6, 30001
    dat = rng.standard_normal((ndet, n))
    ra = np.linspace(0, 1, n)
    info = {"dat_calib": device.to_device(dat), "ra_b": ra.copy(), "dra": rng.standard_normal(ndet), "dt": 0.01}
    downsample_tod(info, 10)
    nn = n // 10
    assert abs(info["dt"] - 0.1) < 1e-15
    assert tuple(info["dat_calib"].shape) == (ndet, nn)
    assert rel_err(info["dat_calib"].cpu().numpy(), opre.downsample_array(dat, 10)) < 1e-12
    assert info["ra_b"].shape == (nn,) and rel_err(info["ra_b"], opre.downsample_array(ra, 10)) < 1e-12
    assert info["dra"].shape == (ndet,)
    d32 = {"dat_calib": device.to_device(dat.astype(np.float32)), "dt": 1.0}
    downsample_tod(d32, 4)
    assert rel_err(d32["dat_calib"].cpu().numpy(), opre.downsample_array(dat.astype(np.float32), 4)) < 2e-6
