"""GPU DCT-I (Bluestein) and the batched C W C noise filter against scipy's
DCT-I, the reference's own no-FFTW path (minkasi/tools/py_fft.py:239)."""
import ctypes

import numpy as np
import pytest
import scipy.fft as sf

from tests.helpers import max_rel_err, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    from minkasi_amd import device
    from minkasi_amd.lib import mkgpu
    return torch, device, mkgpu


def gpu_dct1(env, x, dtype):
    torch, device, mkgpu = env
    nrow, n = x.shape
    eb = 8 if dtype == np.float64 else 4
    nb = int(mkgpu.lib.mk_dct1_workspace_bytes(nrow, n, eb))
    work = torch.empty(nb // 8 + 64, dtype=torch.float64, device=device.device())
    xt = device.to_device(x.astype(dtype))
    yt = torch.empty_like(xt)
    name = "mk_dct1_f64" if eb == 8 else "mk_dct1_f32"
    mkgpu.check(name, xt.data_ptr(), yt.data_ptr(), nrow, n, work.data_ptr(), work.numel() * 8, device.stream_ptr())
    return yt.cpu().numpy()


@pytest.mark.parametrize("n", [2, 3, 4, 17, 100, 1001, 4097, 65536, 40001, 262144])
def test_dct1_f64_matches_scipy(env, n):
    rng = np.random.default_rng(n)
    nrow = 3 if n < 100000 else 1
    x = rng.standard_normal((nrow, n))
    ref = sf.dct(x, type=1, axis=-1)
    y = gpu_dct1(env, x, np.float64)
    assert rel_err(y, ref) < 1e-13, n
    assert max_rel_err(y, ref) < 1e-12, n


@pytest.mark.parametrize("n", [5, 1000, 65536, 100001])
def test_dct1_f32_matches_scipy(env, n):
    rng = np.random.default_rng(n + 1)
    x = rng.standard_normal((2, n))
    ref = sf.dct(x, type=1, axis=-1)
    y = gpu_dct1(env, x, np.float32)
    assert rel_err(y, ref) < 2e-6, n


def test_dct1_involution_large(env):
    # DCT-I twice = 2(n-1) I, at a config-3 row length
    n = 262144
    x = np.random.default_rng(9).standard_normal((1, n))
    y = gpu_dct1(env, gpu_dct1(env, x, np.float64), np.float64)
    assert rel_err(y / (2 * (n - 1)), x) < 1e-12


def _filter(env, rows_n, dtype, halve, seed=0, wdct_from_gpu=False, geometry=None, p1=None):
    """Batched filter over segments of (rows, n); returns gpu, ref, info.
    geometry: optional list that receives each segment's (P1, P2, cw, P).
    p1: optional per-segment column lengths (mk_filt_plan_create_p1)."""
    torch, device, mkgpu = env
    rng = np.random.default_rng(seed)
    desc = []
    off = 0
    xs, ws = [], []
    for rows, n in rows_n:
        desc += [rows, n, off, n, off, n]
        off += rows * n
        xs.append(rng.standard_normal((rows, n)))
        ws.append(np.abs(rng.standard_normal((rows, n))) + 0.05)
    desc = np.asarray(desc, dtype=np.int64)
    eb = 8 if dtype == np.float64 else 4
    h = ctypes.c_void_p()
    if p1 is None:
        mkgpu.check("mk_filt_plan_create", len(rows_n), desc.ctypes.data, eb, ctypes.byref(h))
    else:
        p1a = np.asarray(p1, dtype=np.int32)
        mkgpu.check("mk_filt_plan_create_p1", len(rows_n), desc.ctypes.data, eb, p1a.ctypes.data, ctypes.byref(h))
    try:
        if wdct_from_gpu:
            wd = np.concatenate([gpu_dct1(env, w, np.float64).ravel() for w in ws])
        else:
            wd = np.concatenate([sf.dct(w, type=1, axis=-1).ravel() for w in ws])
        wdt = device.to_device(wd)
        mkgpu.check("mk_filt_plan_set_weights", h, wdt.data_ptr(), device.stream_ptr())
        mkgpu.check("mk_filt_plan_set_halve", h, int(halve))
        xt = device.to_device(np.concatenate([x.ravel() for x in xs]).astype(dtype))
        yt = torch.empty_like(xt)
        name = "mk_filt_plan_exec_f64" if eb == 8 else "mk_filt_plan_exec_f32"
        mkgpu.check(name, h, xt.data_ptr(), yt.data_ptr(), device.stream_ptr())
        out = yt.cpu().numpy()
        info = np.zeros(8, dtype=np.int64)
        mkgpu.check("mk_filt_plan_info", h, info.ctypes.data)
        if geometry is not None:
            g = np.zeros(4 * len(rows_n), dtype=np.int64)
            mkgpu.check("mk_filt_plan_geometry", h, g.ctypes.data)
            geometry.extend(tuple(int(v) for v in r) for r in g.reshape(-1, 4))
    finally:
        mkgpu.check("mk_filt_plan_destroy", h)
    refs = []
    for x, w in zip(xs, ws):
        r = sf.dct(sf.dct(x, type=1, axis=-1) * w, type=1, axis=-1)
        if halve:
            r[:, 0] *= 0.5
            r[:, -1] *= 0.5
        refs.append(r.ravel())
    return out, np.concatenate(refs), info


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_filter_batched_matches_dct_pair(env, dtype):
    segs = [(3, 17), (2, 2), (4, 1001), (2, 65536), (3, 40001), (1, 131073)]
    out, ref, info = _filter(env, segs, dtype, halve=True)
    tol = 1e-12 if dtype == np.float64 else 5e-6
    assert rel_err(out, ref) < tol
    off = 0
    for rows, n in segs:
        o = out[off:off + rows * n]
        r = ref[off:off + rows * n]
        e, em = rel_err(o, r), max_rel_err(o, r)
        print("filter", dtype.__name__, rows, n, "norm %.2e max %.2e" % (e, em))
        assert e < tol and em < tol, (rows, n, e, em)
        off += rows * n


def test_filter_weights_from_gpu_dct(env):
    out, ref, _ = _filter(env, [(2, 5000), (2, 77777)], np.float64, halve=False, seed=3, wdct_from_gpu=True)
    assert rel_err(out, ref) < 1e-12


def test_filter_is_symmetric_and_kills_constants(env):
    """<a, F b> = <F a, b> and F(const) = 0 when W[0] = 0 (noise.py:677)."""
    torch, device, mkgpu = env
    n = 30001
    rng = np.random.default_rng(4)
    w = np.abs(rng.standard_normal((1, n))) + 0.1
    w[:, 0] = 0
    desc = np.asarray([2, n, 0, n, 0, n], dtype=np.int64)
    h = ctypes.c_void_p()
    mkgpu.check("mk_filt_plan_create", 1, desc.ctypes.data, 8, ctypes.byref(h))
    try:
        wd = np.concatenate([sf.dct(w, type=1, axis=-1)] * 2).ravel()
        wdt = device.to_device(wd)
        mkgpu.check("mk_filt_plan_set_weights", h, wdt.data_ptr(), device.stream_ptr())
        mkgpu.check("mk_filt_plan_set_halve", h, 1)
        a, b = rng.standard_normal(n), rng.standard_normal(n)
        xt = device.to_device(np.stack([a, b]))
        yt = torch.empty_like(xt)
        mkgpu.check("mk_filt_plan_exec_f64", h, xt.data_ptr(), yt.data_ptr(), device.stream_ptr())
        Fa, Fb = yt.cpu().numpy()
        assert abs(np.dot(a, Fb) - np.dot(Fa, b)) < 1e-10 * np.linalg.norm(Fa) * np.linalg.norm(b)
        ct = device.to_device(np.ones((2, n)) * 3.0)
        mkgpu.check("mk_filt_plan_exec_f64", h, ct.data_ptr(), yt.data_ptr(), device.stream_ptr())
        assert np.abs(yt.cpu().numpy()).max() < 1e-9 * 3.0 * n
    finally:
        mkgpu.check("mk_filt_plan_destroy", h)


def test_filter_fp32_hot_and_general_paths_agree(env):
    """Rows of every FP32 column count P1 = 2..256 (n = 2500 .. 200001) take
    the specialised 2^11-point K2; MK_FFT_NO_HOT forces the general K2.
    Both match the FP64 reference to the FP32 tolerance and each other."""
    import os
    segs = [(2, 2500), (1, 5001), (3, 20001), (1, 99999), (2, 200001), (2, 1001)]
    out_hot, ref, _ = _filter(env, segs, np.float32, halve=True, seed=7)
    os.environ["MK_FFT_NO_HOT"] = "1"
    try:
        out_gen, ref2, _ = _filter(env, segs, np.float32, halve=True, seed=7)
    finally:
        del os.environ["MK_FFT_NO_HOT"]
    assert np.array_equal(ref, ref2)
    assert rel_err(out_hot, ref) < 5e-6
    assert rel_err(out_gen, ref) < 5e-6
    assert rel_err(out_hot, out_gen) < 5e-6


def _smooth(v):
    for d in (2, 3, 5):
        while v % d == 0:
            v //= d
    return v == 1


def test_filter_fp32_every_hot_column_length(env):
    """One row per hot column length P1 (2^a 3^b 5^c <= 256): n is chosen so
    that the smallest 5-smooth P1 with 2^11 P1 >= (3(n-1)+2)/2 is exactly that
    value, the plan's reported geometry confirms it, and each row matches the
    FP64 DCT-I pair and the power-of-two general path (MK_FFT_NO_HOT)."""
    import os
    p1s = [v for v in range(2, 257) if _smooth(v)]
    assert len(p1s) == 51
    segs = [(1, (2 * 2048 * p1 - 2) // 3 + 1) for p1 in p1s]
    geo = []
    out, ref, _ = _filter(env, segs, np.float32, halve=True, seed=11, geometry=geo)
    assert [g[0] for g in geo] == p1s
    assert all(g[1] == 2048 and g[3] == g[0] * 2048 for g in geo)
    os.environ["MK_FFT_NO_HOT"] = "1"
    try:
        out_gen, _, _ = _filter(env, segs, np.float32, halve=True, seed=11)
    finally:
        del os.environ["MK_FFT_NO_HOT"]
    off = 0
    for (rows, n), p1 in zip(segs, p1s):
        o, r, g = out[off:off + n], ref[off:off + n], out_gen[off:off + n]
        assert rel_err(o, r) < 5e-6, (p1, n, rel_err(o, r))
        assert rel_err(o, g) < 5e-6, (p1, n)
        off += n


def test_filter_fp32_longer_column_lengths_are_exact(env):
    """Rows filtered with a caller-chosen P1 above the smallest that fits
    (what the autotuner may pick) still equal the FP64 DCT-I pair; a P1 that
    is too short or not a hot length is refused."""
    p1s = [v for v in range(2, 257) if _smooth(v)]
    segs, choice = [], []
    for n in (2800, 9000, 65536, 100001, 200000, 349000):
        need = -(-((3 * (n - 1) + 2) // 2) // 2048)
        fits = [p for p in p1s if p >= need]
        for p in fits[:3] + fits[-1:]:
            segs.append((2, n))
            choice.append(p)
    geo = []
    out, ref, _ = _filter(env, segs, np.float32, halve=True, seed=13, geometry=geo, p1=choice)
    assert [g[0] for g in geo] == choice
    off = 0
    for (rows, n), p in zip(segs, choice):
        o, r = out[off:off + rows * n], ref[off:off + rows * n]
        assert rel_err(o, r) < 5e-6, (p, n, rel_err(o, r))
        off += rows * n
    torch, device, mkgpu = env
    h = ctypes.c_void_p()
    for bad in (3, 7):  # 3 is too short for n = 65536; 7 is not 5-smooth
        desc = np.asarray([1, 65536, 0, 65536, 0, 65536], dtype=np.int64)
        p = np.asarray([bad], dtype=np.int32)
        with pytest.raises(mkgpu.MkError):
            mkgpu.check("mk_filt_plan_create_p1", 1, desc.ctypes.data, 4, p.ctypes.data, ctypes.byref(h))
