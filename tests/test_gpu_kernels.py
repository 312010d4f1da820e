"""GPU parity of the pointing, projection and PCG vector kernels, called
through the C ABI and compared with the CPU oracle on the same inputs."""
import os

import numpy as np
import pytest

from oracle import core
from tests.helpers import ARCSEC, raster_pointing, rel_err

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def env():
    import torch
    from minkasi_amd import device
    from minkasi_amd.lib import mkgpu
    return torch, device, mkgpu


def _gpu_pix(env, w, nx, ny, ra, dec):
    torch, device, mkgpu = env
    from minkasi_amd.maps.wcs import CarWCS
    cw = CarWCS(w["crval"], w["cdelt"], w["crpix"])
    rat = device.to_device(ra, torch.float64)
    dect = device.to_device(dec, torch.float64)
    ip = torch.empty(rat.shape, dtype=torch.int32, device=rat.device)
    nbad = torch.zeros(1, dtype=torch.int64, device=rat.device)
    prm = cw.params10()
    mkgpu.check("mk_pix_from_radec_car", rat.data_ptr(), dect.data_ptr(), rat.numel(), prm.ctypes.data,
                nx, ny, ip.data_ptr(), nbad.data_ptr(), device.stream_ptr())
    px = torch.empty_like(rat)
    py = torch.empty_like(rat)
    mkgpu.check("mk_world2pix_car", rat.data_ptr(), dect.data_ptr(), rat.numel(), prm.ctypes.data,
                px.data_ptr(), py.data_ptr(), device.stream_ptr())
    return ip.cpu().numpy(), int(nbad.item()), px.cpu().numpy(), py.cpu().numpy()


def test_pix_matches_astropy_fixtures(env):
    z = np.load(os.path.join(ROOT, "tests/golden/pix_car_fixtures.npz"))
    for c in z["cases"]:
        lims = z[c + "_lims"]
        w = core.car_wcs(lims, float(z[c + "_pixsize"]))
        nx, ny = (int(v) for v in z[c + "_nxny"])
        ip, nbad, px, py = _gpu_pix(env, w, nx, ny, z[c + "_ra"], z[c + "_dec"])
        print(c, "max|dpx| vs astropy", np.abs(px - z[c + "_px"]).max(), np.abs(py - z[c + "_py"]).max())
        assert np.array_equal(ip, z[c + "_ipix"]), c
        assert nbad == 0


def test_pix_matches_oracle_config0_shape(env):
    """64 x 65536 raster at 2 arcsec (config 0 shape): bit-exact except at
    exact rounding ties, which are counted (near-tie audit)."""
    rng = np.random.default_rng(11)
    ra_b, dec_b, dra, ddec = raster_pointing(rng, 64, 65536)
    ra = ra_b[None, :] + dra[:, None]
    dec = dec_b[None, :] + ddec[:, None]
    lims = [ra.min(), ra.max(), dec.min(), dec.max()]
    w = core.car_wcs(lims, 2 * ARCSEC)
    nx, ny = core.car_npix(w, lims, 2)
    ref = core.pix_from_radec(w, ny, ra, dec)
    ip, nbad, px, py = _gpu_pix(env, w, nx, ny, ra, dec)
    qx, qy = core.car_world2pix(w, ra, dec)
    fx = np.abs((px - 1) - np.floor(px - 1) - 0.5)
    fy = np.abs((py - 1) - np.floor(py - 1) - 0.5)
    near = (fx < 1e-9) | (fy < 1e-9)
    mism = ip != ref
    print("near ties", int(near.sum()), "mismatches", int(mism.sum()),
          "max|dpx| gpu-cpu", np.abs(px - qx).max(), np.abs(py - qy).max())
    assert not np.any(mism & ~near)
    assert nbad == 0


def test_pix_raster_variant_equals_full(env):
    torch, device, mkgpu = env
    rng = np.random.default_rng(12)
    ra_b, dec_b, dra, ddec = raster_pointing(rng, 8, 5000)
    ra = ra_b[None, :] + dra[:, None]
    dec = dec_b[None, :] + ddec[:, None]
    lims = [ra.min(), ra.max(), dec.min(), dec.max()]
    w = core.car_wcs(lims, 2 * ARCSEC)
    nx, ny = core.car_npix(w, lims, 2)
    ip_full, _, _, _ = _gpu_pix(env, w, nx, ny, ra, dec)
    from minkasi_amd.maps.wcs import CarWCS
    prm = CarWCS(w["crval"], w["cdelt"], w["crpix"]).params10()
    t = [device.to_device(a, torch.float64) for a in (ra_b, dec_b, dra, ddec)]
    ip = torch.empty((8, 5000), dtype=torch.int32, device=t[0].device)
    mkgpu.check("mk_pix_from_raster_car", t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), t[3].data_ptr(),
                8, 5000, prm.ctypes.data, nx, ny, ip.data_ptr(), None, device.stream_ptr())
    assert np.array_equal(ip.cpu().numpy(), ip_full)


def test_pix_counts_out_of_range(env):
    rng = np.random.default_rng(13)
    ra_b, dec_b, dra, ddec = raster_pointing(rng, 4, 2000)
    ra = ra_b[None, :] + dra[:, None]
    dec = dec_b[None, :] + ddec[:, None]
    lims = [ra.min(), ra.max(), dec.min(), dec.max()]
    w = core.car_wcs(lims, 2 * ARCSEC)
    nx, ny = core.car_npix(w, lims, 2)
    _, nbad, _, _ = _gpu_pix(env, w, nx // 2, ny, ra, dec)
    assert nbad > 0


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_map2tod_bit_exact(env, dtype):
    torch, device, mkgpu = env
    tdt = getattr(torch, dtype)
    rng = np.random.default_rng(3)
    npix = 256 * 256
    for ndet, n in [(64, 65536), (3, 1001), (1, 3)]:
        ipix = rng.integers(0, npix, size=(ndet, n)).astype(np.int32)
        m = rng.standard_normal(npix).astype(dtype)
        d0 = rng.standard_normal((ndet, n)).astype(dtype)
        ipt = device.to_device(ipix)
        mt = device.to_device(m)
        for do_add in (0, 1):
            dt = device.to_device(d0.copy())
            name = "mk_map2tod_f32" if dtype == "float32" else "mk_map2tod_f64"
            mkgpu.check(name, dt.data_ptr(), mt.data_ptr(), ndet * n, ipt.data_ptr(), do_add, device.stream_ptr())
            expect = (d0 + m[ipix]) if do_add else m[ipix]  # numpy float32 add is IEEE round-to-nearest too
            assert np.array_equal(dt.cpu().numpy(), expect.astype(dtype)), (ndet, n, do_add)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_map2tod_misaligned_views_bit_exact(env, dtype):
    """ipix / dat views at any element offset (one TOD inside a batched
    arena) gather exactly; equal and unequal misalignments of the two."""
    torch, device, mkgpu = env
    rng = np.random.default_rng(4)
    npix, n = 300 * 257, 10007
    ipix = rng.integers(0, npix, size=n + 8).astype(np.int32)
    m = rng.standard_normal(npix).astype(dtype)
    d0 = rng.standard_normal(n + 8).astype(dtype)
    ipt, mt = device.to_device(ipix), device.to_device(m)
    name = "mk_map2tod_f32" if dtype == "float32" else "mk_map2tod_f64"
    for oi, od in [(1, 1), (3, 0), (0, 2), (2, 3), (5, 5)]:
        for do_add in (0, 1):
            dt = device.to_device(d0.copy())
            cnt = n - 1
            mkgpu.check(name, dt[od:].data_ptr(), mt.data_ptr(), cnt, ipt[oi:].data_ptr(), do_add, device.stream_ptr())
            got = dt.cpu().numpy()
            expect = d0.copy()
            g = m[ipix[oi:oi + cnt]]
            expect[od:od + cnt] = (d0[od:od + cnt] + g) if do_add else g
            assert np.array_equal(got, expect.astype(dtype)), (oi, od, do_add)


def _plan(env, ipix_t, nx, ny, tx=32, ty=32, item=4096):
    torch, device, mkgpu = env
    import ctypes
    h = ctypes.c_void_p()
    mkgpu.check("mk_t2m_plan_create", ipix_t.data_ptr(), ipix_t.numel(), nx, ny, tx, ty, item, ctypes.byref(h),
                device.stream_ptr())
    return h


@pytest.mark.parametrize("dtype,tile", [("float32", (64, 16)), ("float64", (64, 16)), ("float32", (32, 32)),
                                        ("float64", (32, 32)), ("float32", (32, 16))])
def test_tod2map_plan_matches_oracle_and_is_deterministic(env, dtype, tile):
    """Tiles of 1024 pixels (the most a plan takes: a lane word keeps a
    10-bit local pixel) in two shapes, and the 32 x 16 the autotuner picks on
    the benchmark workload."""
    torch, device, mkgpu = env
    rng = np.random.default_rng(4)
    ra_b, dec_b, dra, ddec = raster_pointing(rng, 64, 65536)
    ra = ra_b[None, :] + dra[:, None]
    dec = dec_b[None, :] + ddec[:, None]
    lims = [ra.min(), ra.max(), dec.min(), dec.max()]
    w = core.car_wcs(lims, 2 * ARCSEC)
    nx, ny = core.car_npix(w, lims, 2)
    ipix = core.pix_from_radec(w, ny, ra, dec)
    tod = rng.standard_normal(ipix.shape)
    ref = np.zeros(nx * ny)
    core.tod2map(ref, tod, ipix)
    ipt = device.to_device(ipix)
    h = _plan(env, ipt, nx, ny, *tile)
    try:
        td = device.to_device(tod.astype(dtype))
        name = "mk_t2m_plan_exec_f32" if dtype == "float32" else "mk_t2m_plan_exec_f64"
        outs = []
        for _ in range(2):
            mt = torch.full((nx * ny,), 7.0, dtype=getattr(torch, dtype), device=td.device)
            mkgpu.check(name, h, mt.data_ptr(), td.data_ptr(), 0, device.stream_ptr())
            outs.append(mt.cpu().numpy())
        assert np.array_equal(outs[0], outs[1])  # bitwise reproducible
        tol = 1e-5 if dtype == "float32" else 1e-12
        assert rel_err(outs[0], ref) < tol
        assert np.abs(outs[0] - ref).max() <= tol * np.abs(ref).max() * 10
        # accumulate adds onto the existing map
        mt = torch.from_numpy(outs[0].copy()).to(td.device)
        mkgpu.check(name, h, mt.data_ptr(), td.data_ptr(), 1, device.stream_ptr())
        assert rel_err(mt.cpu().numpy(), 2 * ref) < tol
        # hits: exact integers
        ones = torch.ones_like(td)
        mkgpu.check(name, h, mt.data_ptr(), ones.data_ptr(), 0, device.stream_ptr())
        hits = np.bincount(ipix.ravel(), minlength=nx * ny)
        assert np.array_equal(mt.cpu().numpy(), hits.astype(dtype))
    finally:
        mkgpu.check("mk_t2m_plan_destroy", h)


@pytest.mark.parametrize("k", [1, 2, 3, 7, 16, 63, 64, 100])
def test_tod2map_plan_run_coded_groups(env, k):
    """Samples alternate between two map tiles in blocks of k, so each tile's
    stable-sorted entries form runs of k consecutive samples: groups close at
    the fifth run (4k entries when 4k < 64), at 64 entries, or with runs that
    cross group boundaries (k >= 64).  FP64 sums equal np.bincount to
    summation-order rounding, FP32 sums within 1e-5, both bitwise
    repeatable; every case also has a partial last group."""
    torch, device, mkgpu = env
    rng = np.random.default_rng(100 + k)
    nx, ny, tile = 64, 64, 16
    n = 3 * 4096 + 77
    blk = (np.arange(n) // k) % 2
    # tile (0, 0) or tile (1, 2), a random pixel inside it
    x = np.where(blk == 0, 0, 16) + rng.integers(0, tile, n)
    y = np.where(blk == 0, 0, 32) + rng.integers(0, tile, n)
    ipix = (x * ny + y).astype(np.int32)
    tod = rng.standard_normal(n)
    ref = np.bincount(ipix, tod, minlength=nx * ny)
    ipt = device.to_device(ipix)
    for item in (64, 4096):
        h = _plan(env, ipt, nx, ny, tile, tile, item)
        try:
            for dtype, tol in (("float64", 1e-12), ("float32", 1e-5)):
                td = device.to_device(tod.astype(dtype))
                name = "mk_t2m_plan_exec_f64" if dtype == "float64" else "mk_t2m_plan_exec_f32"
                outs = []
                for _ in range(2):
                    mt = torch.full((nx * ny,), 3.0, dtype=getattr(torch, dtype), device=td.device)
                    mkgpu.check(name, h, mt.data_ptr(), td.data_ptr(), 0, device.stream_ptr())
                    outs.append(mt.cpu().numpy())
                assert np.array_equal(outs[0], outs[1])
                assert np.abs(outs[0] - ref).max() <= tol * np.abs(ref).max(), (k, item, dtype)
        finally:
            mkgpu.check("mk_t2m_plan_destroy", h)


def test_tod2map_plan_edge_cases(env):
    torch, device, mkgpu = env
    rng = np.random.default_rng(6)
    # map not a multiple of the tile, skewed hits (many repeats of one pixel),
    # item boundaries inside a tile
    nx, ny = 37, 29
    ipix = rng.integers(0, nx * ny, size=(3, 777)).astype(np.int32)
    ipix[1, :500] = 5
    tod = rng.standard_normal(ipix.shape)
    ref = np.bincount(ipix.ravel(), tod.ravel(), minlength=nx * ny)
    ipt = device.to_device(ipix)
    tdev = device.to_device(tod)
    for tx, ty, item in [(8, 8, 64), (16, 4, 128), (64, 16, 4096), (32, 32, 192), (4, 4, 64)]:
        h = _plan(env, ipt, nx, ny, tx, ty, item)
        try:
            mt = torch.zeros(nx * ny, dtype=torch.float64, device=ipt.device)
            mkgpu.check("mk_t2m_plan_exec_f64", h, mt.data_ptr(), tdev.data_ptr(), 0, device.stream_ptr())
            assert np.allclose(mt.cpu().numpy(), ref, rtol=1e-12, atol=1e-12), (tx, ty, item)
        finally:
            mkgpu.check("mk_t2m_plan_destroy", h)
    # out of range ipix is refused at plan creation
    bad = ipix.copy()
    bad[0, 3] = nx * ny
    import ctypes
    h = ctypes.c_void_p()
    badt = device.to_device(bad)
    st = mkgpu.fn("mk_t2m_plan_create")(badt.data_ptr(), bad.size, nx, ny, 8, 8, 64, ctypes.byref(h),
                                         device.stream_ptr())
    assert st == mkgpu.MK_ERR_RANGE
    # a tile over 1024 pixels is refused
    st = mkgpu.fn("mk_t2m_plan_create")(ipt.data_ptr(), ipix.size, nx, ny, 64, 32, 4096, ctypes.byref(h),
                                         device.stream_ptr())
    assert st == mkgpu.MK_ERR_ARG
    # empty input: map cleared (accumulate=0) or untouched (accumulate=1)
    empty = torch.zeros(0, dtype=torch.int32, device=ipt.device)
    h = _plan(env, empty, nx, ny, 8, 8, 64)
    try:
        mt = torch.ones(nx * ny, dtype=torch.float32, device=ipt.device)
        dummy = torch.zeros(1, dtype=torch.float32, device=ipt.device)
        mkgpu.check("mk_t2m_plan_exec_f32", h, mt.data_ptr(), dummy.data_ptr(), 1, device.stream_ptr())
        assert float(mt.sum()) == nx * ny
        mkgpu.check("mk_t2m_plan_exec_f32", h, mt.data_ptr(), dummy.data_ptr(), 0, device.stream_ptr())
        assert float(mt.abs().sum()) == 0.0
    finally:
        mkgpu.check("mk_t2m_plan_destroy", h)


def test_tod2map_atomic_matches(env):
    torch, device, mkgpu = env
    rng = np.random.default_rng(7)
    ipix = rng.integers(0, 5000, size=(16, 4000)).astype(np.int32)
    tod = rng.standard_normal(ipix.shape)
    ref = np.bincount(ipix.ravel(), tod.ravel(), minlength=5000)
    mt = torch.zeros(5000, dtype=torch.float64, device=device.device())
    td, it = device.to_device(tod), device.to_device(ipix)  # keep both alive across the call
    mkgpu.check("mk_tod2map_atomic_f64", mt.data_ptr(), td.data_ptr(), ipix.size, it.data_ptr(),
                device.stream_ptr())
    assert np.allclose(mt.cpu().numpy(), ref, rtol=1e-12, atol=1e-12)


def test_pcg_vector_ops(env):
    torch, device, mkgpu = env
    from minkasi_amd.device import scratch
    rng = np.random.default_rng(8)
    n = 1_000_003
    S = scratch()
    f = lambda a: device.to_device(a.astype(np.float32))
    x, r, p, Ap, w = (rng.standard_normal(n).astype(np.float32) for _ in range(5))
    w = np.abs(w)
    z = (w * r).astype(np.float32)
    xt, rt, pt, Apt, wt, zt = f(x), f(r), f(p), f(Ap), f(w), f(z)
    sp = device.stream_ptr()
    sc = S.sc
    sc.zero_()
    mkgpu.check("mk_dot_f32", xt.data_ptr(), rt.data_ptr(), n, sc.data_ptr(), S.work.data_ptr(), sp)
    assert abs(sc[mkgpu.SC_DOT].item() - np.dot(x.astype(np.float64), r)) < 1e-9 * np.sqrt(n)
    mkgpu.check("mk_pcg_init", rt.data_ptr(), zt.data_ptr(), pt.data_ptr(), n, sc.data_ptr(), S.work.data_ptr(), sp)
    zr = np.dot(r.astype(np.float64), z)
    assert abs(sc[mkgpu.SC_ZR].item() - zr) < 1e-9 * abs(zr)
    assert np.array_equal(pt.cpu().numpy(), z)
    p = z.copy()
    mkgpu.check("mk_pcg_alpha", pt.data_ptr(), Apt.data_ptr(), n, sc.data_ptr(), S.work.data_ptr(), sp)
    pAp = np.dot(p.astype(np.float64), Ap)
    if pAp > 0:
        alpha = zr / pAp
        assert abs(sc[mkgpu.SC_ALPHA].item() - alpha) < 1e-9 * abs(alpha)
        mkgpu.check("mk_pcg_xrz_jacobi", xt.data_ptr(), rt.data_ptr(), zt.data_ptr(), pt.data_ptr(), Apt.data_ptr(),
                    wt.data_ptr(), n, sc.data_ptr(), S.work.data_ptr(), sp)
        xn = (x.astype(np.float64) + alpha * p).astype(np.float32)
        rn = (r.astype(np.float64) - alpha * Ap.astype(np.float64)).astype(np.float32)
        zn = (w * rn).astype(np.float32)
        assert np.abs(xt.cpu().numpy() - xn).max() <= 1e-6 * np.abs(xn).max()
        assert np.abs(rt.cpu().numpy() - rn).max() <= 1e-6 * np.abs(rn).max()
        zr2 = np.dot(rn.astype(np.float64), zn)
        assert abs(sc[mkgpu.SC_ZR].item() - zr2) < 1e-6 * abs(zr2)
        assert abs(sc[mkgpu.SC_BETA].item() - zr2 / zr) < 1e-6 * abs(zr2 / zr)
    else:
        assert sc[mkgpu.SC_FLAG].item() != 0
    # zero curvature is flagged, not divided
    sc.zero_()
    sc[mkgpu.SC_ZR] = 1.0
    zeros = torch.zeros_like(pt)
    mkgpu.check("mk_pcg_alpha", zeros.data_ptr(), Apt.data_ptr(), n, sc.data_ptr(), S.work.data_ptr(), sp)
    assert sc[mkgpu.SC_FLAG].item() != 0 and sc[mkgpu.SC_ALPHA].item() == 0.0
    # determinism of the reduction
    vals = []
    for _ in range(3):
        mkgpu.check("mk_dot_f32", xt.data_ptr(), rt.data_ptr(), n, sc.data_ptr(), S.work.data_ptr(), sp)
        vals.append(sc[mkgpu.SC_DOT].item())
    assert vals[0] == vals[1] == vals[2]


def _tiled_index_np(ipix, ny, tx, ty):
    x, y = ipix // ny, ipix % ny
    nty = -(-ny // ty)
    return ((x // tx) * nty + y // ty) * (tx * ty) + (x % tx) * ty + y % ty


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tiled_gather_layout_is_bit_exact(env, dtype):
    """mk_pix_tile_remap and mk_map_to_tiles_* against the layout formula;
    map2tod from the tiled copy equals map.ravel()[ipix] bit for bit, for
    ragged map sizes and every tile shape the batched operator uses."""
    torch, device, mkgpu = env
    rng = np.random.default_rng(8)
    sfx = "f64" if dtype == np.float64 else "f32"
    for nx, ny in ((37, 53), (64, 64), (1, 5), (300, 7)):
        m = rng.standard_normal((nx, ny)).astype(dtype)
        ipix = rng.integers(0, nx * ny, size=4 * 1001).astype(np.int32)
        ipix[:3] = [0, nx * ny - 1, ny - 1]
        d_m, d_ip = device.to_device(m), device.to_device(ipix)
        for tx, ty in ((8, 4), (16, 2), (4, 8), (32, 1), (3, 5)):
            size = int(mkgpu.lib.mk_map_tiled_size(nx, ny, tx, ty))
            assert size == (-(-nx // tx) * tx) * (-(-ny // ty) * ty)
            tiled = torch.full((size,), 7.0, dtype=d_m.dtype, device=d_m.device)
            mkgpu.check("mk_map_to_tiles_" + sfx, d_m.data_ptr(), nx, ny, tx, ty, tiled.data_ptr(),
                        device.stream_ptr())
            want = np.zeros(size, dtype)
            xs, ys = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
            want[_tiled_index_np((xs * ny + ys).ravel(), ny, tx, ty)] = m.ravel()
            assert np.array_equal(tiled.cpu().numpy(), want), (nx, ny, tx, ty)
            it = torch.empty_like(d_ip)
            nbad = torch.zeros(1, dtype=torch.int64, device=d_ip.device)
            mkgpu.check("mk_pix_tile_remap", d_ip.data_ptr(), ipix.size, nx, ny, tx, ty, it.data_ptr(),
                        nbad.data_ptr(), device.stream_ptr())
            assert int(nbad.item()) == 0
            assert np.array_equal(it.cpu().numpy(), _tiled_index_np(ipix.astype(np.int64), ny, tx, ty))
            out = torch.empty(ipix.size, dtype=d_m.dtype, device=d_m.device)
            mkgpu.check("mk_map2tod_" + sfx, out.data_ptr(), tiled.data_ptr(), ipix.size, it.data_ptr(), 0,
                        device.stream_ptr())
            assert np.array_equal(out.cpu().numpy(), m.ravel()[ipix])
    # indices outside the map are counted (and written as 0)
    bad = np.array([0, -1, 12, 12 * 13, 5], dtype=np.int32)
    d_b = device.to_device(bad)
    it = torch.empty_like(d_b)
    nbad = torch.zeros(1, dtype=torch.int64, device=d_b.device)
    mkgpu.check("mk_pix_tile_remap", d_b.data_ptr(), 5, 12, 13, 8, 4, it.data_ptr(), nbad.data_ptr(),
                device.stream_ptr())
    assert int(nbad.item()) == 2
    assert it.cpu().numpy()[1] == 0 and it.cpu().numpy()[3] == 0


def test_batched_gather_equals_rowmajor_gather(env):
    """The batched operator's map2tod stage (tiled copy, remapped indices)
    reproduces map.ravel()[ipix] exactly for every gather tile."""
    torch, device, mkgpu = env
    from minkasi_amd.mapmaking.batch import GATHER_TILES
    from minkasi_amd.workloads import build_problem
    P = build_problem(1, ntod=2, verbose=False)
    op = P.op
    n = op.nsamp
    x = torch.randn(op.nx, op.ny, device=op.buf_a.device)
    want = x.reshape(-1)[op.ipix[:n].long()]
    for t in GATHER_TILES:
        op._set_gather_tile(t)
        a = torch.empty(n, dtype=torch.float32, device=x.device)
        op._map2tod(x, a)
        assert torch.equal(a, want), t


def _decode_pix_code(hdr, payload, n, ny):
    """numpy decoder of the coded pointing (block header {x0 << 16 | y0,
    offset in 256-byte units | type << 30}; types 1, 2, 3)."""
    hdr = hdr.reshape(-1, 2).view(np.uint32)
    out = np.empty(n, dtype=np.int64)
    for b in range((n + 255) // 256):
        x0, y0 = int(hdr[b, 0] >> 16), int(hdr[b, 0] & 0xFFFF)
        typ, off = int(hdr[b, 1] >> 30), int(hdr[b, 1] & 0x3FFFFFFF) * 256
        m = min(256, n - 256 * b)
        if typ == 1:
            dx = payload[off:off + 256].view(np.int8)[:m].astype(np.int64)
            x, y = x0 + dx, np.full(m, y0)
        elif typ == 2:
            d = payload[off:off + 512].view(np.int8).reshape(256, 2)[:m].astype(np.int64)
            x, y = x0 + d[:, 0], y0 + d[:, 1]
        else:
            w = payload[off:off + 1024].view(np.uint32)[:m].astype(np.int64)
            x, y = w >> 16, w & 0xFFFF
        out[256 * b:256 * b + m] = x * ny + y
    return out


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_coded_pointing_decodes_and_gathers_bit_exact(dtype):
    """Coded pointing (1, 2 or 4 bytes per sample in 256-sample blocks): the
    stream decodes to the original indices, and the coded gather from the
    tiled map equals map2tod on the original indices bit for bit, for every
    gather tile, with and without do_add.  The input mixes raster rows (type
    1), blocks that step in dec (type 2), blocks that jump across the map
    (type 3) and a last block of n % 256 = 237 samples (n % 4 = 1)."""
    import torch
    from minkasi_amd import device
    from minkasi_amd.lib import mkgpu
    from minkasi_amd.lib.plans import PixCode
    from minkasi_amd.mapmaking.batch import GATHER_TILES
    rng = np.random.default_rng(21)
    nx, ny = 1027, 771
    rows, n_row = 7, 40001
    x = (np.arange(n_row) * 0.3).astype(np.int64) % 600 + 100
    parts = []
    for r in range(rows):
        y = np.full(n_row, 50 + 13 * r)
        y[21000:] += 10                                    # a dec step inside a block (no x wrap there)
        parts.append(x * ny + y)
    jump = rng.integers(0, nx * ny, 3 * 256)                # blocks of scattered pixels
    ipix = np.concatenate(parts + [jump]).astype(np.int32)
    ipix = np.concatenate([ipix, ipix[:(237 - ipix.size % 256) % 256]])
    n = ipix.size
    assert n % 256 == 237 and n % 4 == 1
    ip_d = torch.from_numpy(ipix).cuda()
    code = PixCode(ip_d, nx, ny)
    assert code.blocks_by_type[1] > 0 and code.blocks_by_type[2] > 0 and code.blocks_by_type[3] > 0
    print("coded pointing: %.3f bytes per sample, blocks by type %s" % (code.nbytes / n, code.blocks_by_type))
    assert code.nbytes < 1.5 * n
    dec = _decode_pix_code(code.hdr.cpu().numpy(), code.payload.cpu().numpy(), n, ny)
    assert np.array_equal(dec, ipix)
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    m = torch.from_numpy(rng.standard_normal((nx, ny)).astype(dtype)).cuda()
    ref = torch.empty(n, dtype=tdt, device="cuda")
    mkgpu.check("mk_map2tod_f32" if dtype == np.float32 else "mk_map2tod_f64", ref.data_ptr(), m.data_ptr(), n,
                ip_d.data_ptr(), 0, device.stream_ptr())
    base = torch.from_numpy(rng.standard_normal(n).astype(dtype)).cuda()
    ref_add = base + ref
    for tx, ty in GATHER_TILES:
        tiled = torch.zeros(int(mkgpu.lib.mk_map_tiled_size(nx, ny, tx, ty)), dtype=tdt, device="cuda")
        mkgpu.check("mk_map_to_tiles_f32" if dtype == np.float32 else "mk_map_to_tiles_f64", m.data_ptr(), nx, ny,
                    tx, ty, tiled.data_ptr(), device.stream_ptr())
        got = torch.full((n,), float("nan"), dtype=tdt, device="cuda")
        code.gather(got, tiled, (tx, ty))
        assert torch.equal(got, ref), (tx, ty)
        acc = base.clone()
        code.gather(acc, tiled, (tx, ty), do_add=True)
        assert torch.equal(acc, ref_add), (tx, ty)


def test_coded_pointing_rejects_bad_indices():
    import torch
    from minkasi_amd.lib.plans import PixCode
    ipix = np.arange(1000, dtype=np.int32)
    ipix[777] = 64 * 64
    with pytest.raises(ValueError):
        PixCode(torch.from_numpy(ipix).cuda(), 64, 64)
