"""Solving to a true residual: the FP64-binned tod2map, the all-FP64
operator (AccurateDot), residual replacement and the checked stop of the
device PCG, against the FP64 oracle; plus the regressions for a cached
batched operator whose noise models change, for pixel-index validation and
for a breakdown in a quiet run."""
import numpy as np
import pytest

from oracle import solve
from tests.helpers import max_rel_err, rel_err
from tests.test_gpu_api import build

pytestmark = pytest.mark.gpu


def _fp64_build(**kw):
    """build() with FP64 filters kept by every noise model."""
    import minkasi_amd as mk
    todvec, m, otods = build(**kw)
    for tod, ot in zip(todvec.tods, otods):
        tod.set_noise(mk.NoiseSmoothedSVD, fwhm=kw.get("fwhm", 10.0), u_in=ot.noise.v.T.copy(), keep_fp64=True)
    return todvec, m, otods


@pytest.mark.parametrize("tile", [(32, 32), (16, 64)])
def test_t2m_fp64_sums_match_bincount(tile):
    """FP32 samples binned with FP64 sums: equal to the FP64 bincount of the
    same samples to summation-order rounding (1e-13), into an FP64 map and
    (rounded once) into an FP32 map; bitwise repeatable."""
    import torch
    from minkasi_amd.lib.plans import T2MPlan
    rng = np.random.default_rng(11)
    nx, ny, n = 300, 260, 3_000_001
    x = np.cumsum(rng.integers(-1, 2, n)) % nx
    y = (np.arange(n) // 7) % ny
    ipix = (x * ny + y).astype(np.int32)
    dat = (rng.standard_normal(n) * 1e3 + 5e4).astype(np.float32)   # large cancelling offsets
    dat[::2] *= -1
    ref = np.bincount(ipix, weights=dat.astype(np.float64), minlength=nx * ny)
    ip_d = torch.from_numpy(ipix).cuda()
    d_d = torch.from_numpy(dat).cuda()
    plan = T2MPlan(ip_d, nx, ny, tile=tile)
    m64 = torch.zeros(nx * ny, dtype=torch.float64, device="cuda")
    plan.exec(m64, d_d, False)
    got = m64.cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(dat).max() * 64
    m64b = torch.zeros_like(m64)
    plan.exec(m64b, d_d, False)
    assert torch.equal(m64, m64b)
    m32 = torch.zeros(nx * ny, dtype=torch.float32, device="cuda")
    plan.exec(m32, d_d, False, acc64=True)
    assert np.array_equal(m32.cpu().numpy(), ref.astype(np.float32))
    # the FP32-sum kernel is measurably worse on the same input
    m32s = torch.zeros_like(m32)
    plan.exec(m32s, d_d, False, acc64=False)
    e32 = np.abs(m32s.cpu().numpy() - ref).max()
    e64 = np.abs(m32.cpu().numpy() - ref).max()
    print("tod2map max error: FP32 sums %.3e, FP64 sums %.3e" % (e32, e64))
    assert e64 <= e32


def test_accurate_operator_matches_fp64_oracle():
    """AccurateDot (every stage FP64) against the FP64 oracle's A.x and rhs
    to 1e-10, norm-wise and max-elementwise; the FP64-binned view of the FP32
    operator within 1e-6 of it."""
    import torch
    todvec, m, otods = _fp64_build(ntod=3, ndet=24, n=12001, seed=4)
    shape = (m.nx, m.ny)
    import minkasi_amd as mk
    ms = mk.Mapset()
    ms.add_map(m)
    op = todvec.batched(ms)
    acc = op.accurate()
    x = np.random.default_rng(5).standard_normal(shape)
    x64 = torch.from_numpy(x).cuda()
    out = torch.zeros_like(x64)
    acc.apply(x64, out)
    ref = solve.todvec_dot(otods, x)
    e, emax = rel_err(out.cpu().numpy(), ref), max_rel_err(out.cpu().numpy(), ref)
    b64 = torch.zeros_like(x64)
    acc.rhs(b64)
    bref = solve.make_rhs(otods, shape)
    eb = rel_err(b64.cpu().numpy(), bref)
    binned = torch.zeros_like(x64)
    op.f64_view().apply(x64, binned)
    eb32 = rel_err(binned.cpu().numpy(), ref)
    print("AccurateDot A.x %.3e (max %.3e), rhs %.3e; FP64-binned FP32 operator %.3e" % (e, emax, eb, eb32))
    assert e < 1e-10 and emax < 1e-10 and eb < 1e-10
    assert eb32 < 1e-6


def test_checked_solve_reaches_true_residual():
    """Jacobi PCG with the FP64-binned check and the FP64 right-hand side:
    the iterate accumulates in FP64, the residual is replaced once, the stop
    is accepted only on the check operator's true residual, and the
    all-FP64 operator confirms ||b - A x||/||b|| < tol for the final iterate.
    The recursive-residual stop of the same solve is reported beside it."""
    import torch
    import minkasi_amd as mk
    from minkasi_amd.mapmaking.pcg import DevicePCG, make_check
    todvec, m, otods = _fp64_build(ntod=4, ndet=32, n=20001, seed=6)
    ms = mk.Mapset()
    ms.add_map(m)
    op = todvec.batched(ms)
    rhs = ms.copy()
    todvec.make_rhs(rhs)
    hits = mk.make_hits(todvec, m)
    pre = ms.copy()
    h = hits.map
    pre.maps[0].map.copy_(torch.where(h > 0, 1.0 / torch.where(h > 0, h, torch.ones_like(h)), torch.zeros_like(h)))
    chk, b64 = make_check(op, "fp64")
    tol = 1e-6
    solver = DevicePCG(op, rhs.maps[0].map, pre, check=chk, b64=b64)
    x = torch.zeros_like(rhs.maps[0].map)
    solver.init(x)
    it = 0
    while it < 400:
        solver.step()
        it += 1
        if solver.converged(tol):
            break
    assert it < 400
    assert solver.replacements >= 1 and solver.checks >= 1
    t_chk = solver.last_true
    t64 = solver.true_rel_res64(op.accurate())
    print("checked solve: %d iterations, %d replacements, %d checks, true residual %.3e (check) %.3e (FP64)" % (
        it, solver.replacements, solver.checks, t_chk, t64))
    assert t_chk < tol and t64 < tol
    assert abs(t64 - t_chk) < 0.1 * tol
    # the FP32 map holds the FP64 iterate's rounding
    assert torch.equal(x, solver.x64.to(torch.float32))
    # the drop-in call path: run_pcg(check="fp64") stops on the same criterion
    x0 = rhs.copy()
    x0.clear()
    hist = []
    xr = mk.run_pcg(rhs, x0, todvec, pre, maxiter=400, tol=tol, check="fp64", verbose=False, history=hist)
    assert hist[-1]["true_rel_res"] < tol
    assert any(h_.get("replaced") for h_ in hist)
    e = rel_err(xr.maps[0].map.cpu().numpy(), x.cpu().numpy())
    assert e < 1e-5, e


def test_cached_operator_follows_new_noise_model():
    """TodVec caches its batched operator; re-estimating a TOD's noise
    (Tod.set_noise, as the reference's multi-pass examples do between
    solves) must rebuild it: the next dot equals the oracle with the new
    noise, not the old one."""
    import torch
    import minkasi_amd as mk
    from oracle import noise as onoise
    todvec, m, otods = build(ntod=2, ndet=16, n=9001, seed=8)
    shape = (m.nx, m.ny)
    ms = mk.Mapset()
    ms.add_map(m)
    x = np.random.default_rng(3).standard_normal(shape).astype(np.float32)
    xs = ms.copy()
    xs.maps[0].assign(x)
    out1 = todvec.dot(xs).maps[0].map.cpu().numpy()
    op1 = todvec._batch
    assert rel_err(out1, solve.todvec_dot(otods, x.astype(np.float64))) < 1e-5
    # new noise on the second TOD: a different smoothing width
    tod = todvec.tods[1]
    d64 = otods[1].dat
    onz = onoise.NoiseSmoothedSVD(d64, fwhm=40.0)
    tod.set_noise(mk.NoiseSmoothedSVD, fwhm=40.0, u_in=onz.v.T.copy())
    otods[1] = solve.OTod(otods[1].dat, otods[1].ipix, onz)
    out2 = todvec.dot(xs).maps[0].map.cpu().numpy()
    assert todvec._batch is not op1
    ref2 = solve.todvec_dot(otods, x.astype(np.float64))
    e2 = rel_err(out2, ref2)
    print("dot after set_noise vs oracle with the new noise: %.3e" % e2)
    assert e2 < 1e-5
    assert rel_err(out2, out1) > 1e-3


def test_saved_pixelization_outside_map_raises():
    """A saved pixelization with an index outside the map raises before any
    projection reads it (the reference indexes unchecked)."""
    import torch
    import minkasi_amd as mk
    todvec, m, otods = build(ntod=1, ndet=8, n=5001, seed=9)
    tod = todvec.tods[0]
    ip = tod.info[m.tag].clone()
    ip.view(-1)[1234] = m.nx * m.ny + 5
    tod.info[m.tag] = ip
    dat = torch.zeros(tod.get_data_dims(), dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        m.map2tod(tod, dat)
    m.tod2map_method = "atomic"
    with pytest.raises(ValueError):
        m.tod2map(tod, dat)
    # an in-place change to an already checked tensor is seen (version counter)
    ip2 = otods[0].ipix.copy()
    tod.info[m.tag] = torch.from_numpy(ip2.reshape(tod.get_data_dims())).cuda()
    m.map2tod(tod, dat)       # valid: passes
    tod.info[m.tag].view(-1)[7] = -3
    with pytest.raises(ValueError):
        m.map2tod(tod, dat)


def test_quiet_run_raises_on_breakdown():
    """A breakdown (p.Ap <= 0) in a run that reads no per-iteration scalars
    still raises PCGBreakdown at the end (forced from the host with an
    operator of negative curvature)."""
    import torch
    import minkasi_amd as mk
    from minkasi_amd.mapmaking.pcg import PCGBreakdown
    todvec, m, otods = build(ntod=1, ndet=8, n=5001, seed=10)
    ms = mk.Mapset()
    ms.add_map(m)
    real = todvec.batched(ms)

    class Negative:
        nx, ny = real.nx, real.ny

        def null_mask(self):
            return None

        def apply(self, x, out, accumulate=False):
            out.copy_(x)
            out.neg_()
            return out

    neg = Negative()
    todvec.batched = lambda mapset, **kw: neg
    b = ms.copy()
    b.maps[0].map.fill_(1.0)
    x0 = ms.copy()
    x0.clear()
    with pytest.raises(PCGBreakdown):
        mk.run_pcg(b, x0, todvec, None, maxiter=3, verbose=False)
