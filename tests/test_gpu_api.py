"""End-to-end parity of the drop-in API (SkyMap, Tod, TodVec, make_hits,
make_rhs, TodVec.dot, run_pcg) against the FP64 oracle on the same seeded
inputs.  Noise models share the oracle's singular basis (u_in) so the two
operators are the same matrix."""
import numpy as np
import pytest

from oracle import core, noise as onoise, solve
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu
ARCSEC = np.pi / 180 / 3600


def build(ntod=3, ndet=32, n=20001, pix=2.0, dtype=np.float32, fwhm=10.0, seed=0):
    import torch
    import minkasi_amd as mk
    from minkasi_amd.synth import ScanSpec, make_tod_info
    spec = ScanSpec(center_jitter=30.0, gain_sigma=0.05)
    infos = [make_tod_info(70 + seed, i, ndet, n + 37 * i, spec, dtype=dtype) for i in range(ntod)]
    todvec = mk.TodVec()
    otods = []
    for info in infos:
        onz = onoise.NoiseSmoothedSVD(info["dat_calib"].astype(np.float64), fwhm=fwhm)
        info["onoise"] = onz
    for info in infos:
        t = mk.Tod({k: v for k, v in info.items() if k != "onoise"})
        todvec.add_tod(t)
    lims = todvec.lims()
    m = mk.SkyMap(lims, pix * ARCSEC, dtype=torch.float32 if dtype == np.float32 else torch.float64)
    for tod, info in zip(todvec.tods, infos):
        tod.set_noise(mk.NoiseSmoothedSVD, fwhm=fwhm, u_in=info["onoise"].v.T.copy())
        ipix = m.get_pix(tod)
        otods.append(solve.OTod(info["dat_calib"], ipix.cpu().numpy(), info["onoise"]))
    # oracle pixelization agrees bit for bit
    w = core.car_wcs(lims, pix * ARCSEC)
    assert core.car_npix(w, lims, 2) == (m.nx, m.ny)
    for info, ot in zip(infos, otods):
        assert np.array_equal(core.pix_from_radec(w, m.ny, info["dx"], info["dy"]), ot.ipix)
    return todvec, m, otods


def test_hits_rhs_and_dot_match_oracle():
    import torch
    import minkasi_amd as mk
    todvec, m, otods = build()
    shape = (m.nx, m.ny)
    hits = mk.make_hits(todvec, m)
    assert np.array_equal(hits.map.cpu().numpy(), solve.make_hits(otods, shape))
    ms = mk.Mapset()
    ms.add_map(m)
    rhs = ms.copy()
    todvec.make_rhs(rhs)
    b_ref = solve.make_rhs(otods, shape)
    e_rhs = rel_err(rhs.maps[0].map.cpu().numpy(), b_ref)
    x = np.random.default_rng(1).standard_normal(shape)
    xs = ms.copy()
    xs.maps[0].assign(x.astype(np.float32))
    out = todvec.dot(xs)                      # batched path
    ref = solve.todvec_dot(otods, x.astype(np.float32).astype(np.float64))
    e_dot = rel_err(out.maps[0].map.cpu().numpy(), ref)
    out2, _ = todvec.dot(xs, report_times=True)  # per-TOD path
    e_dot2 = rel_err(out2.maps[0].map.cpu().numpy(), ref)
    print("rhs", e_rhs, "dot batched", e_dot, "dot per-tod", e_dot2)
    assert e_rhs < 1e-5 and e_dot < 1e-5 and e_dot2 < 1e-5
    # bitwise reproducible A.x
    out3 = todvec.dot(xs)
    assert torch.equal(out.maps[0].map, out3.maps[0].map)


def test_graph_replay_matches_eager():
    import torch
    todvec, m, otods = build(ntod=2, ndet=16, n=8001)
    import minkasi_amd as mk
    ms = mk.Mapset()
    ms.add_map(m)
    op = todvec.batched(ms)
    x = torch.randn((m.nx, m.ny), device=m.map.device)
    eager = torch.zeros_like(x)
    op.apply(x, eager, accumulate=False)
    out = torch.zeros_like(x)
    op.capture(x, out, accumulate=False)
    for k in range(3):
        out.fill_(123.0)
        op.apply(x, out, accumulate=False)
        torch.cuda.synchronize()
        assert torch.equal(out, eager), k
        x2 = x * 2  # eager work between replays
        torch.cuda.synchronize()


def test_skymap_tod2map_variants_agree():
    """The reference's per-method wrappers (tod2map_omp / simple / cached /
    everyone / old / atomic) all add P^T d into the map: bit-equal through
    the deterministic plan, the atomic one to FP32 rounding; against the
    oracle's tod2map."""
    import torch
    todvec, m, otods = build(ntod=2, ndet=8, n=3001)
    tod, ot = todvec.tods[0], otods[0]
    d = np.random.default_rng(2).standard_normal(ot.dat.shape).astype(np.float32)
    ref = np.zeros(m.nx * m.ny)
    core.tod2map(ref.reshape(1, -1), d.astype(np.float64), ot.ipix)
    outs = {}
    for name in ("tod2map_omp", "tod2map_simple", "tod2map_cached", "tod2map_everyone", "tod2map_atomic"):
        m.clear()
        getattr(m, name)(tod, d)
        outs[name] = m.map.cpu().numpy().ravel().copy()
    m.clear()
    m.tod2map_old(tod, d)
    outs["tod2map_old"] = m.map.cpu().numpy().ravel().copy()
    for name, o in outs.items():
        assert rel_err(o, ref) < 1e-6, name
        if name != "tod2map_atomic":
            assert np.array_equal(o, outs["tod2map_omp"]), name


def test_tod_noise_wrappers():
    """set_noise_smoothed_svd is set_noise(NoiseSmoothedSVD) (with func, of
    the data minus the model); apply_noise_white_masked is mask * dat *
    mywt per detector."""
    import torch
    import minkasi_amd as mk
    rng = np.random.default_rng(6)
    ndet, n = 6, 2001
    dat = np.cumsum(rng.standard_normal((ndet, n)), axis=1)
    a, b = mk.Tod({"dat_calib": dat.copy()}), mk.Tod({"dat_calib": dat.copy()})
    a.set_noise_smoothed_svd(fwhm=20)
    b.set_noise(mk.NoiseSmoothedSVD, fwhm=20)
    x = rng.standard_normal((ndet, n))
    assert rel_err(a.apply_noise(x).cpu().numpy(), b.apply_noise(x).cpu().numpy()) < 1e-12
    model = np.outer(np.ones(ndet), np.linspace(0, 1, n))
    c, e = mk.Tod({"dat_calib": dat.copy()}), mk.Tod({"dat_calib": dat - model})
    c.set_noise_smoothed_svd(fwhm=20, func=lambda pars, tod: pars * model, pars=1.0)
    e.set_noise(mk.NoiseSmoothedSVD, fwhm=20)
    assert rel_err(c.apply_noise(x).cpu().numpy(), e.apply_noise(x).cpu().numpy()) < 1e-12
    mask = rng.random((ndet, n)) > 0.1
    w = mk.Tod({"dat_calib": dat.copy(), "mask": mask})
    w.set_noise_white_masked()
    w.info["mywt"] = np.arange(1, ndet + 1, dtype=float)
    got = w.apply_noise_white_masked().cpu().numpy()
    assert np.array_equal(got, mask * dat * w.info["mywt"][:, None])


def test_fits_tod_through_device_path():
    """A MUSTANG TOD read from FITS (an astropy-written fixture) pixelizes on
    the device exactly as the oracle does, and its hit map counts every
    unmasked sample."""
    import os
    import torch
    import minkasi_amd as mk
    from minkasi_amd.tods.io import read_tod_from_fits
    fname = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mustang_tod.fits")
    info = read_tod_from_fits(fname, branch=180.0)
    tv = mk.TodVec()
    tv.add_tod(mk.Tod(info))
    lims = tv.lims()
    m = mk.SkyMap(lims, 10 * ARCSEC, dtype=torch.float32)
    ipix = m.get_pix(tv.tods[0]).cpu().numpy()
    w = core.car_wcs(lims, 10 * ARCSEC)
    assert np.array_equal(core.pix_from_radec(w, m.ny, info["dx"], info["dy"]), ipix)
    hits = mk.make_hits(tv, m)
    assert float(hits.map.sum()) == float(info["mask"].sum())
    want = np.bincount(ipix.ravel(), weights=info["mask"].ravel().astype(float), minlength=m.nx * m.ny)
    assert np.array_equal(hits.map.cpu().numpy().ravel(), want.astype(np.float32))


def test_per_tod_path_after_batched_operator():
    """Once the batched operator has made every TOD's ipix a view into its
    index arena (offsets that are not multiples of 4 here), the per-TOD
    path still works: TodVec.dot over FP64 copies of the TODs, which runs
    map2tod TOD by TOD, matches the oracle's A.x."""
    import torch
    import minkasi_amd as mk
    todvec, m, otods = build(ntod=3, ndet=7, n=4001)
    ms = mk.Mapset()
    ms.add_map(m)
    op = todvec.batched(ms)
    assert any(o % 4 for o in op.offs)
    tv = mk.TodVec()
    for tod in todvec.tods:
        ip = tod.info["ipix"]
        assert ip.data_ptr() != 0 and ip.is_contiguous()
        d = tod.get_data()
        d = d.to(torch.float64) if torch.is_tensor(d) else np.asarray(d, dtype=np.float64)
        tv.add_tod(mk.Tod({"dat_calib": d, "ipix": ip}))
    for t64, ot in zip(tv.tods, otods):
        t64.set_noise(mk.NoiseSmoothedSVD, fwhm=10.0, u_in=ot.noise.v.T.copy(), keep_fp64=True)
    m64 = m.copy()
    m64.map = torch.randn(m.nx, m.ny, dtype=torch.float64, device=m.map.device)
    x = mk.Mapset()
    x.add_map(m64)
    out = tv.dot(x)
    ref = solve.todvec_dot(otods, m64.map.cpu().numpy())
    assert rel_err(out.maps[0].map.cpu().numpy(), ref) < 1e-10


def test_short_memory_operator_matches_default():
    """The operator as built when device memory is short (filter in groups of
    TODs sharing one work buffer, map2tod gathered from the row-major map)
    gives the default operator's A.x and rhs bit for bit, eagerly and from a
    captured graph."""
    import torch
    import minkasi_amd as mk
    from minkasi_amd.mapmaking.batch import BatchedDot
    todvec, m, otods = build(ntod=5, ndet=16, n=8001)
    op = BatchedDot(todvec, m, tiled_gather=True)
    work = [t.noise.filt32.work_bytes for t in todvec.tods]
    small = BatchedDot(todvec, m, tiled_gather=False, filter_work_cap=2 * max(work))
    assert op.tiled and len(op.filts) == 1
    assert not small.tiled and small.code is None and len(small.filts) >= 3
    assert sorted(i for g in small.filter_groups for i in g) == list(range(5))
    assert small.filter_work.numel() <= 2 * max(work)
    x = torch.randn((m.nx, m.ny), device=m.map.device)
    want = torch.zeros_like(x)
    op.apply(x, want, accumulate=False)
    got = torch.zeros_like(x)
    small.apply(x, got, accumulate=False)
    assert torch.equal(got, want)
    r0, r1 = torch.zeros_like(x), torch.zeros_like(x)
    op.rhs(r0, accumulate=False)
    small.rhs(r1, accumulate=False)
    assert torch.equal(r0, r1)
    g = torch.zeros_like(x)
    small.capture(x, g, accumulate=False)
    for k in range(3):
        g.fill_(7.0)
        small.apply(x, g, accumulate=False)
        torch.cuda.synchronize()
        assert torch.equal(g, want), k
    rep = small.memory_report()
    assert rep["tiled_gather_ipix"] == 0 and rep["filter_groups"] == len(small.filts)


def test_run_pcg_matches_oracle():
    import torch
    import minkasi_amd as mk
    todvec, m, otods = build(ntod=2, ndet=24, n=16001, seed=3)
    shape = (m.nx, m.ny)
    ms = mk.Mapset()
    ms.add_map(m)
    rhs = ms.copy()
    todvec.make_rhs(rhs)
    hits = mk.make_hits(todvec, m)
    pre = ms.copy()
    h = hits.map
    pre.maps[0].map.copy_(torch.where(h > 0, 1.0 / torch.where(h > 0, h, torch.ones_like(h)), torch.zeros_like(h)))
    x0 = rhs.copy()
    x0.clear()
    hist = []
    x = mk.run_pcg(rhs, x0, todvec, pre, maxiter=20, verbose=False, history=hist)
    b_ref = solve.make_rhs(otods, shape)
    pre_ref = solve.jacobi_from_hits(solve.make_hits(otods, shape))
    xr, hr = solve.run_pcg(b_ref, np.zeros(shape), lambda v: solve.todvec_dot(otods, v), pre_ref, maxiter=20)
    xg = x.maps[0].map.cpu().numpy()
    e = rel_err(xg, xr)
    zr = [h_["zr"] for h_ in hist]
    zrr = [h_["zr"] for h_ in hr]
    print("pcg 20 it: map rel err", e, "zr rel err last", abs(zr[-1] - zrr[-1]) / abs(zrr[-1]))
    assert e < 1e-3
    assert np.allclose(zr[:5], zrr[:5], rtol=1e-3)
    # the generic (Mapset-level) loop, forced by a mask of ones, runs the same
    # update order with host scalars and gives the same iterates
    ones = ms.copy()
    ones.maps[0].map.fill_(1.0)
    hist2 = []
    x2 = mk.run_pcg(rhs, x0, todvec, pre, maxiter=20, verbose=False, mask=ones, history=hist2)
    e2 = rel_err(x2.maps[0].map.cpu().numpy(), xg)
    zr2 = [h_["zr"] for h_ in hist2]
    print("generic vs device path: map rel err", e2)
    assert len(zr2) == len(zr) == 20
    assert e2 < 1e-4
    assert np.allclose(zr2, zr, rtol=1e-4)


def test_fp64_config0_path_matches_oracle():
    """FP64 TOD + FP64 map (config 0 precision): per-TOD path, FP64 noise."""
    import torch
    import minkasi_amd as mk
    todvec, m, otods = build(ntod=1, ndet=16, n=8193, dtype=np.float64, seed=5)
    shape = (m.nx, m.ny)
    ms = mk.Mapset()
    ms.add_map(m)
    x = np.random.default_rng(2).standard_normal(shape)
    xs = ms.copy()
    xs.maps[0].assign(x)
    out = todvec.dot(xs)
    ref = solve.todvec_dot(otods, x)
    e = rel_err(out.maps[0].map.cpu().numpy(), ref)
    print("fp64 dot rel err", e)
    assert e < 1e-10
