"""Extra Mapset components on the device against the FP64 oracle: cut
samples (Cuts, CutsCompact, tsModel) and polarised maps (PolMap), from the
C-ABI kernels up to run_pcg with a mixed Mapset."""
import numpy as np
import pytest

from oracle import components as oc
from oracle import core, solve
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    from minkasi_amd import device
    from minkasi_amd.lib import mkgpu
    return torch, device, mkgpu


# ---------------------------------------------------------------------- cuts
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_cut_kernels_exact(env, dtype):
    torch, device, mkgpu = env
    rng = np.random.default_rng(0)
    sfx = "f64" if dtype == np.float64 else "f32"
    for ndet, n, ncut in ((1, 10, 0), (3, 1000, 1), (7, 5000, 3000), (64, 20001, 100000)):
        dat = rng.standard_normal((ndet, n)).astype(dtype)
        imap = rng.choice(ndet * n, ncut, replace=False).astype(np.int64)  # unsorted, unique
        vec0 = rng.standard_normal(ncut).astype(dtype)
        d_dat, d_imap = device.to_device(dat), device.to_device(imap)
        for do_add in (0, 1):
            v = device.to_device(vec0.copy())
            mkgpu.check("mk_tod2cuts_" + sfx, v.data_ptr(), d_dat.data_ptr(), dat.size, d_imap.data_ptr(), ncut,
                        do_add, None, device.stream_ptr())
            want = (vec0 + dat.ravel()[imap]) if do_add else dat.ravel()[imap]
            assert np.array_equal(v.cpu().numpy(), want.astype(dtype)), (ndet, n, ncut, do_add)
            dd = d_dat.clone()
            d_vec = device.to_device(vec0)  # held until the kernel has run
            mkgpu.check("mk_cuts2tod_" + sfx, dd.data_ptr(), d_vec.data_ptr(), dat.size,
                        d_imap.data_ptr(), ncut, do_add, None, device.stream_ptr())
            w = dat.ravel().copy()
            w[imap] = (w[imap] + vec0) if do_add else vec0
            assert np.array_equal(dd.cpu().numpy().ravel(), w), (ndet, n, ncut, do_add)
    # out-of-range indices are skipped and counted
    dat = rng.standard_normal((2, 50)).astype(dtype)
    imap = np.array([0, -1, 99, 100, 7], dtype=np.int64)
    v = torch.zeros(5, dtype=torch.float64 if dtype == np.float64 else torch.float32, device=device.device())
    nbad = torch.zeros(1, dtype=torch.int64, device=device.device())
    d_dat, d_imap = device.to_device(dat), device.to_device(imap)
    mkgpu.check("mk_tod2cuts_" + sfx, v.data_ptr(), d_dat.data_ptr(), 100, d_imap.data_ptr(), 5, 0,
                nbad.data_ptr(), device.stream_ptr())
    assert int(nbad.item()) == 2
    assert np.array_equal(v.cpu().numpy(), np.array([dat.flat[0], 0, dat.flat[99], 0, dat.flat[7]], dtype))


class _FakeTod:
    def __init__(self, info):
        self.info = info

    def get_ndet(self):
        return self.info["dat_calib"].shape[0]

    def get_ndata(self):
        return self.info["dat_calib"].shape[1]

    def get_data_dims(self):
        return tuple(self.info["dat_calib"].shape)

    def get_data(self, dat=None):
        return self.info["dat_calib"] if dat is None else dat


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_cuts_compact_api(env, dtype):
    torch, device, mkgpu = env
    from minkasi_amd.tods.cuts import CutsCompact
    rng = np.random.default_rng(1)
    ndet, n = 9, 4000
    dat = rng.standard_normal((ndet, n)).astype(dtype)
    tod = _FakeTod({"dat_calib": device.to_device(dat)})
    mask = rng.random((ndet, n)) > 0.05
    mask[4] = True
    c = CutsCompact(tod)
    assert c.map.dtype == (torch.float32 if dtype == np.float32 else torch.float64)
    c.cuts_from_array(mask)
    c.get_imap()
    assert np.array_equal(c.imap.cpu().numpy(), np.flatnonzero(~mask))
    c.tod2map(tod)
    assert np.array_equal(c.map.cpu().numpy(), dat.ravel()[~mask.ravel()])
    c.tod2map(tod, do_add=True)
    assert np.array_equal(c.map.cpu().numpy(), (2 * dat.ravel()[~mask.ravel()]).astype(dtype))
    out = torch.zeros((ndet, n), dtype=c.map.dtype, device=device.device())
    c.map2tod(tod, out, do_add=False)
    want = np.zeros(ndet * n, dtype)
    want[~mask.ravel()] = c.map.cpu().numpy()
    assert np.array_equal(out.cpu().numpy().ravel(), want)
    hostd = np.ones((ndet, n), dtype)
    c.map2tod(tod, hostd, do_add=True)  # numpy data are updated in place too
    assert np.array_equal(hostd.ravel(), (1 + want * (~mask.ravel())).astype(dtype))
    # vector operations against numpy
    a = c.copy()
    b = c.copy()
    b.map.copy_(torch.as_tensor(rng.standard_normal(len(c.imap)), device=device.device()))
    am, bm = a.map.cpu().numpy().astype(np.float64), b.map.cpu().numpy().astype(np.float64)
    assert abs(a.dot(b) - float(am @ bm)) <= 1e-12 * np.abs(am * bm).sum()
    a.axpy(b, -0.75)
    tol = 0 if dtype == np.float64 else 1e-6
    assert np.allclose(a.map.cpu().numpy(), am - 0.75 * bm, rtol=tol, atol=1e-12 + tol)
    prod = a * b
    assert np.array_equal(prod.map.cpu().numpy(), a.map.cpu().numpy() * b.map.cpu().numpy())
    Ax = b.copy()
    c.apply_prior(a, Ax)
    assert np.allclose(Ax.map.cpu().numpy(), bm.astype(dtype) + c.map.cpu().numpy() * a.map.cpu().numpy(),
                       rtol=1e-6, atol=1e-6)
    with pytest.raises(ValueError):
        c.tod2map(None, np.zeros((ndet, n + 1), dtype))


def test_cuts_compact_segments_per_detector_and_merge(env):
    torch, device, mkgpu = env
    from minkasi_amd.tods.cuts import CutsCompact
    ndet, n = 4, 500
    dat = np.arange(ndet * n, dtype=np.float64).reshape(ndet, n)
    tod = _FakeTod({"dat_calib": device.to_device(dat)})
    c = CutsCompact(tod)
    for det, a, b in ((0, 100, 150), (3, 10, 20), (3, 15, 30), (1, 490, 600), (2, 700, 800)):
        c.add_cut(det, a, b)
    assert c.istart == [[100], [490], [], [10, 15]] and c.istop == [[150], [500], [], [20, 30]]
    c.get_imap()
    assert np.array_equal(c.imap.cpu().numpy(), oc.cuts_imap(c.istart, c.istop, n))
    c.tod2map(tod)
    assert np.array_equal(c.map.cpu().numpy(), dat.ravel()[oc.cuts_imap(c.istart, c.istop, n)])
    with pytest.raises(ValueError):  # samples 15..19 of detector 3 appear twice
        c.map2tod(tod, torch.zeros((ndet, n), dtype=torch.float64, device=device.device()))
    c.merge_cuts()
    ma, mb = oc.merge_segments([[100], [490], [], [10, 15]], [[150], [500], [], [20, 30]], n)
    assert c.istart == ma and c.istop == mb
    c.get_imap()
    out = torch.zeros((ndet, n), dtype=torch.float64, device=device.device())
    c.tod2map(tod)
    c.map2tod(tod, out, do_add=False)
    want = np.zeros(ndet * n)
    idx = oc.cuts_imap(ma, mb, n)
    want[idx] = dat.ravel()[idx]
    assert np.array_equal(out.cpu().numpy().ravel(), want)


def test_cuts_class_bad_samples(env):
    torch, device, mkgpu = env
    from minkasi_amd.tods.cuts import Cuts
    rng = np.random.default_rng(2)
    dat = rng.standard_normal((5, 300))
    bad = rng.random((5, 300)) < 0.1
    tod = _FakeTod({"dat_calib": device.to_device(dat), "bad_samples": bad})
    c = Cuts(tod)
    assert c.nsamp == bad.sum()
    c.tod2map(tod)
    assert np.array_equal(c.map.cpu().numpy(), dat[bad])
    out = device.to_device(np.ones((5, 300)))
    c.map2tod(tod, out)  # do_add=True overwrites, as the reference
    want = np.ones((5, 300))
    want[bad] = dat[bad]
    assert np.array_equal(out.cpu().numpy(), want)
    c2 = Cuts(tod, do_add=False)
    c2.map.copy_(c.map)
    c2.map2tod(tod, out)
    want[bad] += dat[bad]
    assert np.array_equal(out.cpu().numpy(), want)
    assert abs(c.dot(c2) - float(dat[bad] @ dat[bad])) < 1e-12 * float(dat[bad] @ dat[bad])


def _cut_problem(seed=0):
    from tests.test_gpu_api import build
    import minkasi_amd as mk
    from minkasi_amd.mapmaking.timestream import tod_key
    todvec, m, otods = build(ntod=3, ndet=16, n=6001, dtype=np.float64, seed=seed)
    rng = np.random.default_rng(seed + 10)
    cuts = mk.tsModel(todvec, mk.CutsCompact)
    imaps = []
    for tod, ot in zip(todvec.tods, otods):
        mask = rng.random(ot.dat.shape) > 0.02
        cc = cuts.data[tod_key(tod)]
        cc.cuts_from_array(mask)
        cc.get_imap()
        imaps.append(np.flatnonzero(~mask).astype(np.int64))
    ms = mk.Mapset()
    ms.add_map(m)
    ms.add_map(cuts)
    return todvec, m, otods, ms, imaps


def _oracle_dot(otods, imaps, x_map, x_cuts):
    out_m = np.zeros(x_map.shape)
    out_c = []
    for ot, imap, xc in zip(otods, imaps, x_cuts):
        tmp = np.zeros(ot.dat.shape)
        core.map2tod(tmp, np.ascontiguousarray(x_map).ravel(), ot.ipix, do_add=True)
        core.cuts2tod(tmp, np.ascontiguousarray(xc), imap, do_add=True)
        tmp = np.ascontiguousarray(ot.noise.apply_noise(tmp))
        core.tod2map(out_m.reshape(-1), tmp, ot.ipix)
        oc_ = np.zeros(len(imap))
        core.tod2cuts(oc_, tmp, imap, do_add=True)
        out_c.append(oc_)
    return out_m, out_c


def test_tsmodel_cuts_in_mapset_dot_and_pcg(env):
    torch, device, mkgpu = env
    import minkasi_amd as mk
    todvec, m, otods, ms, imaps = _cut_problem()
    rng = np.random.default_rng(3)
    x = ms.copy()
    xm = rng.standard_normal((m.nx, m.ny))
    x.maps[0].assign(xm)
    xcs = []
    for k in list(x.maps[1].data.keys()):
        v = rng.standard_normal(len(x.maps[1].data[k].imap))
        x.maps[1].data[k].map.copy_(torch.as_tensor(v, device=device.device()))
        xcs.append(v)
    out = todvec.dot(x)
    rm, rc = _oracle_dot(otods, imaps, xm, xcs)
    assert rel_err(out.maps[0].map.cpu().numpy(), rm) < 1e-10
    for k, r in zip(list(out.maps[1].data.keys()), rc):
        assert rel_err(out.maps[1].data[k].map.cpu().numpy(), r) < 1e-10
    # x.A x through Mapset.dot equals the oracle's
    xAx = x.dot(out)
    ref = float(np.sum(xm * rm)) + sum(float(a @ b) for a, b in zip(xcs, rc))
    assert abs(xAx - ref) < 1e-9 * abs(ref)

    # PCG over the mixed Mapset (generic path) against the oracle PCG on the
    # concatenated vector; Jacobi on the map, identity on the cuts
    rhs = ms.copy()
    todvec.make_rhs(rhs)
    hits = mk.make_hits(todvec, m)
    pre = ms.copy()
    h = hits.map.cpu().numpy()
    pm = np.where(h > 0, 1.0 / np.where(h > 0, h, 1), 0.0)
    pre.maps[0].assign(pm)
    for k in pre.maps[1].data:
        pre.maps[1].data[k].map.fill_(1.0)
    x0 = rhs.copy()
    x0.clear()
    xs = mk.run_pcg(rhs, x0, todvec, pre, maxiter=12, verbose=False)

    sizes = [m.nx * m.ny] + [len(i) for i in imaps]
    offs = np.cumsum([0] + sizes)

    def split(v):
        return v[:offs[1]].reshape(m.nx, m.ny), [v[offs[i]:offs[i + 1]] for i in range(1, len(sizes))]

    def A(v):
        a, b = split(v)
        om, ocs = _oracle_dot(otods, imaps, a, b)
        return np.concatenate([om.ravel()] + ocs)

    b_m = solve.make_rhs(otods, (m.nx, m.ny))
    b_c = []
    for ot, imap in zip(otods, imaps):
        t = np.zeros(len(imap))
        core.tod2cuts(t, np.ascontiguousarray(ot.noise.apply_noise(ot.dat)), imap, do_add=True)
        b_c.append(t)
    b = np.concatenate([b_m.ravel()] + b_c)
    got_b = np.concatenate([rhs.maps[0].map.cpu().numpy().ravel()] +
                           [rhs.maps[1].data[k].map.cpu().numpy() for k in list(rhs.maps[1].data)])
    assert rel_err(got_b, b) < 1e-10
    P = np.concatenate([pm.ravel()] + [np.ones(s) for s in sizes[1:]])
    xr, _ = solve.run_pcg(b, np.zeros_like(b), A, P, maxiter=12)
    got = np.concatenate([xs.maps[0].map.cpu().numpy().ravel()] +
                         [xs.maps[1].data[k].map.cpu().numpy() for k in list(xs.maps[1].data)])
    assert rel_err(got, xr) < 1e-8


def test_prior_from_skymap(env):
    torch, device, mkgpu = env
    from tests.test_gpu_api import build
    todvec, m, otods = build(ntod=1, ndet=8, n=3001, dtype=np.float64)
    rng = np.random.default_rng(4)
    g = np.zeros((m.nx, m.ny))
    g[m.nx // 3: m.nx // 2, :] = rng.uniform(0.5, 2.0, (m.nx // 2 - m.nx // 3, m.ny))
    m.assign(g)
    prior = todvec.tods[0].prior_from_skymap(m)
    vals = g.ravel()[otods[0].ipix.ravel()]
    idx = np.flatnonzero(vals != 0)
    assert np.array_equal(prior.imap.cpu().numpy(), idx)
    assert rel_err(prior.map.cpu().numpy(), 1.0 / vals[idx] ** 2) < 1e-15


# ---------------------------------------------------------------------- polarisation
def _pol_inputs(rng, ndet, n, npix, dtype):
    ipix = rng.integers(0, npix, (ndet, n)).astype(np.int32)
    tg = rng.uniform(-np.pi, 2 * np.pi, (ndet, n))
    dat = rng.standard_normal((ndet, n))
    return ipix, tg, dat


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_polmap_kernels_match_oracle(env, dtype):
    torch, device, mkgpu = env
    from minkasi_amd.mapmaking.map2tod import polmap2tod
    from minkasi_amd.mapmaking.tod2map import tod2polmap
    rng = np.random.default_rng(5)
    ndet, n, npix = 12, 9001, 513
    ipix, tg, dat = _pol_inputs(rng, ndet, n, npix, dtype)
    tg_c = tg.astype(dtype).astype(np.float64)   # the oracle sees the same angles the device does
    dat_c = dat.astype(dtype).astype(np.float64)
    tol = 1e-12 if dtype == np.float64 else 2e-6
    d_ipix, d_tg = device.to_device(ipix), device.to_device(tg.astype(dtype))
    for poltag, npol in (("QU", 2), ("IQU", 3)):
        mp = rng.standard_normal(npix * npol)
        mp_c = mp.astype(dtype).astype(np.float64)
        ref = dat_c.copy()
        core.polmap2tod(ref, mp_c, tg_c, ipix, npol, do_add=True)
        got = device.to_device(dat.astype(dtype))
        polmap2tod(got, device.to_device(mp.astype(dtype)), poltag, d_tg, d_ipix, do_add=True)
        assert rel_err(got.cpu().numpy(), ref) < tol, poltag
    for poltag, mode in (("QU", 1), ("IQU", 2), ("QU_PRECON", 3), ("IQU_PRECON", 4)):
        nc = core.POL_NCOMP[mode]
        ref = np.zeros(npix * nc)
        core.tod2polmap(ref, dat_c, tg_c, ipix, mode)
        got = torch.zeros(npix * nc, dtype=d_tg.dtype, device=device.device())
        tod2polmap(got, device.to_device(dat.astype(dtype)), poltag, d_tg, d_ipix)
        assert rel_err(got.cpu().numpy(), ref) < (1e-12 if dtype == np.float64 else 1e-5), poltag
    with pytest.raises(ValueError):
        bad = ipix.copy()
        bad[3, 5] = npix
        polmap2tod(device.to_device(dat.astype(dtype)), device.to_device(np.zeros(npix * 3, dtype)), "IQU", d_tg,
                   device.to_device(bad))


@pytest.mark.parametrize("poltag", ["QU_PRECON", "IQU_PRECON"])
def test_polmap_invert_and_precon_mul(env, poltag):
    torch, device, mkgpu = env
    import minkasi_amd as mk
    rng = np.random.default_rng(6)
    k = 2 if poltag == "QU_PRECON" else 3
    nc = 3 if k == 2 else 6
    npx = 4000
    q, _ = np.linalg.qr(rng.standard_normal((npx, k, k)))
    lam = rng.uniform(0.5, 50.0, (npx, k))
    lam[:40, -1] = 1e-12 * lam[:40, 0]     # near-singular: pinv drops that direction
    lam[40:60] = 0.0                       # empty pixels
    B = np.einsum("nij,nj,nkj->nik", q, lam, q)
    if k == 2:
        packed = np.stack([B[:, 0, 0], B[:, 1, 1], B[:, 0, 1]], 1)
    else:
        packed = np.stack([B[:, r, c] for r, c in [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]], 1)
    packed = packed.reshape(80, 50, nc)
    lims = [1.0, 1.001, 0.5, 0.501]
    pm = mk.PolMap(lims, 1e-5, poltag=poltag, nx=80, ny=50, dtype=torch.float64)
    pm.assign(packed)
    pm.invert(1e-6)
    ref = oc.polmap_invert(packed, poltag, 1e-6)
    scale = np.abs(ref).max()
    assert np.max(np.abs(pm.map.cpu().numpy() - ref)) < 1e-9 * scale
    sp = "QU" if k == 2 else "IQU"
    xm = mk.PolMap(lims, 1e-5, poltag=sp, nx=80, ny=50, dtype=torch.float64)
    xv = rng.standard_normal((80, 50, k))
    xm.assign(xv)
    out = pm * xm
    assert out.poltag == sp
    assert rel_err(out.map.cpu().numpy(), oc.precon_mul(pm.map.cpu().numpy(), xv, poltag)) < 1e-13
    with pytest.raises(ValueError):
        xm * xm.copy()  # an IQU map is not a preconditioner


def test_polmap_pcg_matches_oracle(env):
    torch, device, mkgpu = env
    import minkasi_amd as mk
    from tests.test_gpu_api import build
    todvec, m, otods = build(ntod=3, ndet=16, n=6001, dtype=np.float64, seed=2)
    rng = np.random.default_rng(7)
    tgs = []
    for tod, ot in zip(todvec.tods, otods):
        tg = rng.uniform(0, np.pi, ot.dat.shape[0])[:, None] + np.linspace(0, 0.5, ot.dat.shape[1])[None, :]
        tod.info["twogamma_saved"] = tg
        tgs.append(np.ascontiguousarray(tg))
    pm = mk.PolMap(todvec.lims(), m.pixsize, poltag="IQU", dtype=torch.float64)
    assert (pm.nx, pm.ny) == (m.nx, m.ny)
    npix = pm.nx * pm.ny
    hits = mk.make_hits(todvec, pm)
    assert hits.poltag == "IQU_PRECON"
    href = np.zeros(npix * 6)
    for ot, tg in zip(otods, tgs):
        core.tod2polmap(href, np.ones(ot.dat.shape), tg, ot.ipix, 4)
    assert rel_err(hits.map.cpu().numpy().ravel(), href) < 1e-12
    hits.invert(1e-6)
    pre_ref = oc.polmap_invert(href.reshape(pm.nx, pm.ny, 6), "IQU_PRECON", 1e-6)
    assert np.max(np.abs(hits.map.cpu().numpy() - pre_ref)) < 1e-9 * np.abs(pre_ref).max()

    ms = mk.Mapset()
    ms.add_map(pm)
    rhs = ms.copy()
    todvec.make_rhs(rhs)
    b = np.zeros(npix * 3)
    for ot, tg in zip(otods, tgs):
        core.tod2polmap(b, np.ascontiguousarray(ot.noise.apply_noise(ot.dat)), tg, ot.ipix, 2)
    assert rel_err(rhs.maps[0].map.cpu().numpy().ravel(), b) < 1e-10

    def A(v):
        out = np.zeros(npix * 3)
        for ot, tg in zip(otods, tgs):
            t = np.zeros(ot.dat.shape)
            core.polmap2tod(t, v, tg, ot.ipix, 3, do_add=True)
            core.tod2polmap(out, np.ascontiguousarray(ot.noise.apply_noise(t)), tg, ot.ipix, 2)
        return out

    pre = mk.Mapset()
    pre.add_map(hits)
    x0 = rhs.copy()
    x0.clear()
    xs = mk.run_pcg(rhs, x0, todvec, pre, maxiter=10, verbose=False)
    M = lambda r: oc.precon_mul(pre_ref, r.reshape(pm.nx, pm.ny, 3), "IQU_PRECON").ravel()
    xr, _ = solve.run_pcg(b, np.zeros_like(b), A, M, maxiter=10)
    assert rel_err(xs.maps[0].map.cpu().numpy().ravel(), xr) < 1e-8


def test_batched_map_plus_cuts_matches_per_tod_path(env):
    """FP32 SkyMap + tsModel(CutsCompact): the batched flat operator
    (BatchedCutsDot) gives the per-TOD loop's A x, rhs and PCG iterates."""
    torch, device, mkgpu = env
    import minkasi_amd as mk
    from minkasi_amd.mapmaking import pcg
    from minkasi_amd.mapmaking.timestream import tod_key
    from tests.test_gpu_api import build
    todvec, m, otods = build(ntod=3, ndet=32, n=20001, dtype=np.float32, seed=5)
    rng = np.random.default_rng(11)
    cuts = mk.tsModel(todvec, mk.CutsCompact)
    for tod in todvec.tods:
        mask = np.ones(tod.get_data_dims(), bool)
        for _ in range(20):  # segments of 50-400 samples on random detectors
            d, t0 = rng.integers(0, mask.shape[0]), rng.integers(0, mask.shape[1] - 400)
            mask[d, t0:t0 + rng.integers(50, 400)] = False
        cuts.data[tod_key(tod)].cuts_from_array(mask)
        cuts.data[tod_key(tod)].get_imap()
    ms = mk.Mapset()
    ms.add_map(m)
    ms.add_map(cuts)
    bc = todvec.batched_cuts(ms)
    assert bc is not None and bc.ncut > 0
    x = ms.copy()
    x.maps[0].assign(rng.standard_normal((m.nx, m.ny)).astype(np.float32))
    for k in x.maps[1].data:
        c = x.maps[1].data[k]
        c.map.copy_(torch.as_tensor(rng.standard_normal(c.map.numel()), device=c.map.device))
    fast = todvec.dot(x)
    slow, _ = todvec.dot(x, report_times=True)  # the reference's per-TOD loop
    pack = lambda s: bc.pack(s).cpu().numpy()
    assert rel_err(pack(fast), pack(slow)) < 1e-5
    rf = ms.copy()
    todvec.make_rhs(rf)
    rs = ms.copy()
    for tod in todvec.tods:
        dat_filt = tod.apply_noise()
        for mm in rs.maps:
            mm.tod2map(tod, dat_filt)
    assert rel_err(pack(rf), pack(rs)) < 1e-5
    hits = mk.make_hits(todvec, m)
    pre = ms.copy()
    h = hits.map.cpu().numpy()
    pre.maps[0].assign(np.where(h > 0, 1.0 / np.where(h > 0, h, 1), 0.0).astype(np.float32))
    for k in pre.maps[1].data:
        pre.maps[1].data[k].map.fill_(1.0)
    x0 = rf.copy()
    x0.clear()
    xf = mk.run_pcg(rf, x0, todvec, pre, maxiter=10, verbose=False)
    xs = pcg._run_pcg_generic(rf, x0, todvec, None, pre, None, 10, "map", [-1], 0, ".fits", None, None, False)
    assert rel_err(pack(xf), pack(xs)) < 1e-3


def test_cross_product_kernel(env):
    """mk_cross_f64: A.B^T with a long inner axis against numpy."""
    torch, device, mkgpu = env
    from minkasi_amd.tods.processing import _cross64
    rng = np.random.default_rng(40)
    for M, N, n in [(7, 3, 20001), (70, 65, 9000), (1, 1, 5)]:
        A, B = rng.standard_normal((M, n)), rng.standard_normal((N, n))
        got = _cross64(device.to_device(A), device.to_device(B)).cpu().numpy()
        assert rel_err(got, A @ B.T) < 1e-13, (M, N, n)


@pytest.mark.parametrize("host", [False, True])
def test_gapfill_eig_matches_oracle(env, host):
    """gapfill_eig (device products, host eig) against the FP64 oracle on a
    rank-2 signal with white noise and short cuts; insert_cuts adds the fill
    into the data at the cut samples."""
    torch, device, mkgpu = env
    from minkasi_amd.tods.cuts import CutsCompact
    from minkasi_amd.tods.processing import gapfill_eig
    rng = np.random.default_rng(30)
    ndet, n = 24, 2000
    modes = np.cumsum(rng.standard_normal((2, n)), axis=1)
    modes *= 30.0 / np.abs(modes).max()
    d = rng.standard_normal((ndet, 2)) @ modes + rng.standard_normal((ndet, n))
    mask = np.ones(d.shape, dtype=bool)
    mask[3, 100:104] = False
    mask[7, 400:403] = False
    mask[0, 1990:1995] = False
    tod = _FakeTod({"dat_calib": device.to_device(d)})
    cuts = CutsCompact(tod)
    cuts.cuts_from_array(mask)
    cuts.get_imap()
    imap = np.flatnonzero(~mask)
    ref = oc.gapfill_eig(d, imap, thresh=5.0, niter_eig=3, niter_inner=3)
    dat = d.copy() if host else device.to_device(d)
    out = gapfill_eig(dat, cuts, thresh=5.0, niter_eig=3, niter_inner=3, insert_cuts=True)
    assert out is not cuts and out.map.dtype == cuts.map.dtype
    assert rel_err(out.map.cpu().numpy(), ref) < 1e-9
    assert np.array_equal(cuts.map.cpu().numpy(), np.zeros(len(imap)))  # the input cuts are untouched
    got = dat if host else dat.cpu().numpy()
    want = d.copy()
    want.flat[imap] += out.map.cpu().numpy()
    assert rel_err(got, want) < 1e-12


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_polmap_deterministic_binning_matches_oracle(env, dtype):
    """Polarised binning through the tod2map plan (one pass per component,
    FP64 sums): equal to the oracle's serial sums (oracle/oracle.c, pinned
    against the reference's own minkasi.c) to summation-order rounding for
    every mode, bitwise repeatable, and PolMap.tod2map uses it."""
    torch, device, mkgpu = env
    import minkasi_amd as mk
    from minkasi_amd.lib.plans import T2MPlan
    rng = np.random.default_rng(6)
    nx, ny = 31, 27
    npix = nx * ny
    ndet, n = 10, 7001
    ipix, tg, dat = _pol_inputs(rng, ndet, n, npix, dtype)
    tg_c = tg.astype(dtype).astype(np.float64)
    dat_c = dat.astype(dtype).astype(np.float64)
    d_ipix = device.to_device(ipix)
    d_tg, d_dat = device.to_device(tg.astype(dtype)), device.to_device(dat.astype(dtype))
    plan = T2MPlan(d_ipix.reshape(-1), nx, ny, tile=(16, 8))
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    for mode in (1, 2, 3, 4):
        nc = core.POL_NCOMP[mode]
        ref = np.zeros(npix * nc)
        core.tod2polmap(ref, dat_c, tg_c, ipix, mode)
        got = torch.zeros(npix * nc, dtype=tdt, device=device.device())
        plan.exec_pol(got, d_dat.reshape(-1), d_tg.reshape(-1), mode, False)
        again = torch.full_like(got, 9.0)
        plan.exec_pol(again, d_dat.reshape(-1), d_tg.reshape(-1), mode, False)
        assert torch.equal(got, again), mode
        e = np.abs(got.cpu().numpy() - ref).max() / np.abs(ref).max()
        # FP32: the angles' FP32 sin/cos and one rounding of each product; FP64: order only
        assert e < (1e-13 if dtype == np.float64 else 2e-6), (mode, e)
        acc = torch.ones_like(got)
        plan.exec_pol(acc, d_dat.reshape(-1), d_tg.reshape(-1), mode, True)
        assert np.abs(acc.cpu().numpy() - 1.0 - got.cpu().numpy()).max() <= 1e-6 * np.abs(ref).max()
    # PolMap.tod2map: the deterministic path by default, the atomic one on request
    lims = [1.0, 1.001, 0.5, 0.501]
    pm = mk.PolMap(lims, 1e-5, poltag="IQU", nx=nx, ny=ny, dtype=tdt)
    tod = mk.Tod({"dat_calib": dat.astype(dtype), "twogamma_saved": tg.astype(dtype), "dx": np.zeros((ndet, n)),
                  "dy": np.zeros((ndet, n))})
    tod.info[pm.tag] = d_ipix
    pm.tod2map(tod, tod.get_data())
    ref = np.zeros(npix * 3)
    core.tod2polmap(ref, dat_c, tg_c, ipix, 2)
    first = pm.map.clone()
    pm.clear()
    pm.tod2map(tod, tod.get_data())
    assert torch.equal(pm.map, first)
    assert np.abs(first.cpu().numpy().ravel() - ref).max() / np.abs(ref).max() < (1e-13 if dtype == np.float64
                                                                                   else 2e-6)
