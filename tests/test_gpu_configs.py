"""The HIP path at the workloads of the five benchmark configurations.

- Configuration 0 (1 TOD of 64 x 65536, FP64, 256^2 map at 2 arcsec, fwhm 10,
  20 Jacobi PCG iterations, the call order of minkasi's
  examples/minkasi_mpi_example.py:35-112) at full size against the FP64
  oracle's solve (minkasi/mapmaking/pcg.py:101-177): map and r.z sequence.
- Configurations 1, 2 and 4 at full size (7.5e8, 8.2e9 and 1.9e9 samples):
  properties that hold at any size -- <y, Ax> = <Ay, x>, A (constant on hit
  pixels) = 0, <x, Ax> > 0, bitwise repeatability, graph replay = eager.
  Configuration 2 also checks its grouped filter (the memory plan that
  shares one work buffer between groups of TODs) against each TOD's own
  single-segment filter, bit for bit; configuration 4 solves to a true
  residual of 1e-6 checked with the all-FP64 operator.
- The extreme TOD shapes of those distributions (512 x 200000, 64 x 10000,
  and an n - 1 that is prime) through the batched operator against the
  FP64 oracle, apply_noise norm-wise and max-elementwise.
Configuration 3 runs in test_gpu_precision.py."""
import numpy as np
import pytest

from oracle import core, noise as onoise, solve
from tests.helpers import max_rel_err, rel_err

pytestmark = pytest.mark.gpu
ARCSEC = np.pi / 180 / 3600


@pytest.fixture(autouse=True)
def _release_device_memory():
    """Each workload here fills a large part of the device: what the previous
    test left in torch's cache goes back before the next one builds."""
    import gc
    import torch
    gc.collect()
    torch.cuda.empty_cache()
    yield
    gc.collect()
    torch.cuda.empty_cache()


def _hit_random(P, seed):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    hit = (P.hits.map > 0).to(torch.float32)
    return torch.randn(P.hits.map.shape, device="cuda", generator=g) * hit


def _dot(a, b):
    import torch
    return float(torch.dot(a.reshape(-1).to(torch.float64), b.reshape(-1).to(torch.float64)).item())


def _operator_properties(P, label):
    import torch
    op = P.op
    x, y = _hit_random(P, 1), _hit_random(P, 2)
    ax, ay = torch.zeros_like(x), torch.zeros_like(x)
    op.apply(x, ax, accumulate=False)
    op.apply(y, ay, accumulate=False)
    nx_, nax = float(torch.linalg.vector_norm(x)), float(torch.linalg.vector_norm(ax))
    asym = abs(_dot(y, ax) - _dot(ay, x)) / (float(torch.linalg.vector_norm(y)) * nax)
    xax = _dot(x, ax)
    u = (P.hits.map > 0).to(torch.float32)
    au = torch.zeros_like(u)
    op.apply(u, au, accumulate=False)
    null = (float(torch.linalg.vector_norm(au)) / float(torch.linalg.vector_norm(u))) / (nax / nx_)
    again = torch.zeros_like(x)
    op.apply(x, again, accumulate=False)
    gx, gout = x.clone(), torch.zeros_like(x)
    op.capture(gx, gout, accumulate=False)
    replays = []
    for k in range(3):
        gout.fill_(7.0)
        op.apply(gx, gout, accumulate=False)
        torch.cuda.synchronize()
        replays.append(bool(torch.equal(gout, ax)))
    op.graph = None
    print("%s: asymmetry %.2e, null-mode ratio %.2e, <x,Ax>/|x||Ax| %.3e, repeat %s, graph %s" % (
        label, asym, null, xax / (nx_ * nax), torch.equal(again, ax), replays))
    assert asym < 1e-5
    assert null < 1e-4
    assert xax > 0
    assert torch.equal(again, ax)
    assert all(replays)


def test_config0_fp64_solve_matches_oracle():
    """Configuration 0 at full size: FP64 TOD and map, per-TOD path, 20
    Jacobi PCG iterations; map and r.z within 1e-10 of the FP64 oracle."""
    import torch
    import minkasi_amd as mk
    from minkasi_amd.synth import CONFIGS, ScanSpec, make_tod_info
    c = CONFIGS[0]
    info = make_tod_info(0, 0, 64, 65536, ScanSpec(**c["spec"]), dtype=np.float64)
    todvec = mk.TodVec()
    todvec.add_tod(mk.Tod(info))
    lims = todvec.lims()
    m = mk.SkyMap(lims, c["pix"] * ARCSEC, nx=256, ny=256, dtype=torch.float64)
    tod = todvec.tods[0]
    ipix = m.get_pix(tod).cpu().numpy()
    assert ipix.min() >= 0 and ipix.max() < 256 * 256
    onz = onoise.NoiseSmoothedSVD(info["dat_calib"], fwhm=c["fwhm"])
    tod.set_noise(mk.NoiseSmoothedSVD, fwhm=c["fwhm"], u_in=onz.v.T.copy())
    ot = solve.OTod(info["dat_calib"], ipix, onz)
    shape = (256, 256)
    hits = mk.make_hits(todvec, m)
    assert np.array_equal(hits.map.cpu().numpy(), solve.make_hits([ot], shape))
    ms = mk.Mapset()
    ms.add_map(m)
    rhs = ms.copy()
    rhs.clear()
    todvec.make_rhs(rhs)
    b_ref = solve.make_rhs([ot], shape)
    assert rel_err(rhs.maps[0].map.cpu().numpy(), b_ref) < 1e-10
    pre = ms.copy()
    h = hits.map
    pre.maps[0].map.copy_(torch.where(h > 0, 1.0 / torch.where(h > 0, h, torch.ones_like(h)), torch.zeros_like(h)))
    x0 = rhs.copy()
    x0.clear()
    hist = []
    x = mk.run_pcg(rhs, x0, todvec, pre, maxiter=c["iters"], verbose=False, history=hist)
    xr, hr = solve.run_pcg(b_ref, np.zeros(shape), lambda v: solve.todvec_dot([ot], v),
                           solve.jacobi_from_hits(solve.make_hits([ot], shape)), maxiter=c["iters"])
    e_map = rel_err(x.maps[0].map.cpu().numpy(), xr)
    zr, zrr = np.array([h_["zr"] for h_ in hist]), np.array([h_["zr"] for h_ in hr])
    e_zr = float(np.max(np.abs(zr - zrr) / np.abs(zrr)))
    print("config 0: 20 iterations, map rel err %.3e, max r.z rel err %.3e, zr %.3e -> %.3e" % (
        e_map, e_zr, zrr[0], zrr[-1]))
    assert len(zr) == len(zrr) == c["iters"]
    assert e_map < 1e-10
    assert e_zr < 1e-10


def test_config1_full_operator_properties():
    from minkasi_amd.workloads import build_problem
    P = build_problem(1)
    assert len(P.shapes) == 32 and all(s[0] == 215 for s in P.shapes)
    _operator_properties(P, "config 1 (%.3g samples)" % P.nsamp)


@pytest.mark.timeout(1200)
def test_config2_full_operator_and_grouped_filter():
    """Configuration 2 (256 TODs, 8.2e9 samples) with the memory plan it
    needs on one GPU: operator properties, and each sampled TOD's segment of
    the grouped filter equal to that TOD's own single-segment filter."""
    import torch
    from minkasi_amd.workloads import build_problem
    P = build_problem(2, verbose=True)
    op = P.op
    print("config 2: %d TODs, %.3g samples, %d filter groups, tiled gather %s" % (
        len(P.shapes), P.nsamp, len(op.filts), op.tiled))
    assert len(op.filts) > 1
    _operator_properties(P, "config 2")
    n = op.nsamp
    # tod2map over all 8.2e9 samples against torch's index_add in FP64: the
    # plan's build runs one wave per group of 64 samples, more groups than
    # one dispatch can hold at a wave each
    a = op.buf_a[:n]
    a.normal_(generator=torch.Generator(device="cuda").manual_seed(3))
    ipix = op.ipix[:n]
    ref = torch.zeros(P.map.nx * P.map.ny, dtype=torch.float64, device="cuda")
    for c0 in range(0, n, 1 << 26):
        c1 = min(n, c0 + (1 << 26))
        ref.index_add_(0, ipix[c0:c1].long(), a[c0:c1].double())
    got = torch.zeros(P.map.nx * P.map.ny, dtype=torch.float32, device="cuda")
    op.t2m.exec(got, a, False, acc64=False)
    d = got.double() - ref
    e, em = float(d.norm() / ref.norm()), float(d.abs().max() / ref.abs().max())
    print("config 2 tod2map against index_add: rel %.2e, max %.2e" % (e, em))
    assert e < 2e-6 and em < 1e-5
    del ref, got, d
    b = op.buf_b[:n]
    for tod, o, (ndet, nn) in zip(op.tods, op.offs, op.sizes):
        b[o:o + ndet * nn].copy_(tod.get_data().reshape(-1))
    op._filter(b)
    groups = op.filter_groups
    picks = sorted({groups[0][0], groups[len(groups) // 2][-1], groups[-1][-1],
                    max(range(len(op.sizes)), key=lambda i: op.sizes[i][0] * op.sizes[i][1])})
    for k in picks:
        tod = op.tods[k]
        ndet, nn = op.sizes[k]
        o = op.offs[k]
        own = torch.empty((ndet, nn), dtype=torch.float32, device="cuda")
        tod.noise.filt32.exec(tod.get_data(), own)
        same = torch.equal(own.reshape(-1), b[o:o + ndet * nn])
        tod.noise.filt32.release_work()
        print("TOD %d (%d x %d): grouped filter == own filter: %s" % (k, ndet, nn, same))
        assert same


def test_config4_full_operator_and_true_residual_solve():
    """Configuration 4 (64 TODs): operator properties, then Jacobi PCG to a
    true ||b - Ax||/||b|| < 1e-6 (FP64 right-hand side, FP64-binned check),
    confirmed by the all-FP64 operator."""
    import torch
    from minkasi_amd.workloads import build_problem, solve_to_tol
    P = build_problem(4, keep_fp64=True)
    _operator_properties(P, "config 4")
    acc = P.op.accurate()
    b64 = torch.zeros((P.map.nx, P.map.ny), dtype=torch.float64, device="cuda")
    acc.rhs(b64)
    r = solve_to_tol(P, None, 1e-6, 300, check=P.op.f64_view(), b64=b64, verify=acc)
    print("config 4 Jacobi to true 1e-6:", {k: v for k, v in r.items() if k not in ("x", "history")})
    assert r["reached"] and r["true_rel_res"] < 1e-6 and r["true_rel_res_fp64"] < 1e-6


def _is_prime(k):
    if k < 2:
        return False
    f = 2
    while f * f <= k:
        if k % f == 0:
            return False
        f += 1
    return True


@pytest.mark.timeout(900)
def test_extreme_shapes_batched_vs_oracle():
    """The largest and smallest TODs of the 64-512 x 10000-200000
    distributions and a prime n - 1, in one batched operator on a 2048^2
    map at 1 arcsec: apply_noise per TOD and A.x against the FP64 oracle."""
    import torch
    import minkasi_amd as mk
    from minkasi_amd.synth import CONFIGS, ScanSpec, make_tod_info
    n_prime = 100004
    assert _is_prime(n_prime - 1)
    shapes = [(512, 200000), (64, 10000), (96, n_prime)]
    c = CONFIGS[4]
    spec = ScanSpec(**c["spec"])
    todvec = mk.TodVec()
    infos = [make_tod_info(4, 900 + i, nd, n, spec, dtype=np.float32) for i, (nd, n) in enumerate(shapes)]
    for info in infos:
        todvec.add_tod(mk.Tod(info))
    m = mk.SkyMap(todvec.lims(), c["pix"] * ARCSEC, nx=2048, ny=2048)
    otods = []
    for tod, info in zip(todvec.tods, infos):
        d64 = info["dat_calib"].astype(np.float64)
        onz = onoise.NoiseSmoothedSVD(d64, fwhm=c["fwhm"])
        tod.set_noise(mk.NoiseSmoothedSVD, fwhm=c["fwhm"], u_in=onz.v.T.copy())
        otods.append(solve.OTod(d64, m.get_pix(tod).cpu().numpy(), onz))
        got = tod.noise.apply_noise(tod.get_data()).cpu().numpy()
        want = onz.apply_noise(d64)
        e, em = rel_err(got, want), max_rel_err(got, want)
        print("apply_noise %s: norm %.2e max %.2e" % (tod.get_data_dims(), e, em))
        assert e < 1e-5 and em < 1e-5
        tod.noise.filt32.release_work()
    ms = mk.Mapset()
    ms.add_map(m)
    assert todvec.batched(ms) is not None
    x = np.random.default_rng(7).standard_normal((2048, 2048)).astype(np.float32)
    xs = ms.copy()
    xs.maps[0].assign(x)
    out = todvec.dot(xs).maps[0].map.cpu().numpy()
    ref = solve.todvec_dot(otods, x.astype(np.float64))
    hit = solve.make_hits(otods, (2048, 2048)) > 0
    e, em = rel_err(out, ref), max_rel_err(out[hit], ref[hit])
    print("batched A.x over the extreme shapes: norm %.2e max %.2e" % (e, em))
    assert e < 1e-5 and em < 1e-4
