"""GPU rotation GEMMs, smoothing and NoiseSmoothedSVD against the FP64 oracle."""
import numpy as np
import pytest

from oracle import noise as onoise
from tests.helpers import max_rel_err, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    from minkasi_amd import device
    from minkasi_amd.lib import mkgpu
    return torch, device, mkgpu


def _orth(rng, k):
    q, _ = np.linalg.qr(rng.standard_normal((k, k)))
    return q


@pytest.mark.parametrize("ndet,n", [(64, 5000), (215, 9001), (512, 1000), (1024, 700), (7, 33)])
def test_split_bf16_rotation_accuracy(env, ndet, n):
    torch, device, mkgpu = env
    from minkasi_amd.lib.plans import RotPlan, SplitMatrix
    rng = np.random.default_rng(ndet + n)
    v = _orth(rng, ndet)
    d = (rng.standard_normal((ndet, n)) + 3 * np.cumsum(rng.standard_normal(n))[None, :] / np.sqrt(n)).astype(np.float32)
    dt = device.to_device(d)
    ref = v @ d.astype(np.float64)
    for mode, kind, tol in ((0, "bf16", 2e-5), (1, "bf16", 1e-6), (2, "f16", 1e-6)):
        plan = RotPlan([(SplitMatrix(v, kind=kind), n, 0, n, 0, n)], mode=mode)
        out = torch.empty_like(dt)
        plan.exec(dt, out)
        e, em = rel_err(out.cpu().numpy(), ref), max_rel_err(out.cpu().numpy(), ref)
        print("split rotation mode", mode, ndet, n, "rel", e, "max", em)
        assert e < tol and em < 10 * tol


def test_f16_rotation_range(env):
    """Mode 2 scales each column of B and the matrix by powers of two before
    the fp16 split: columns of 1e-30 and 1e30, zero columns, a column with
    one huge entry in the first K step or a later one, columns that grow
    over every step, a column with an empty first step and a matrix with
    entries far above 1 (a prewhitened v) all keep FP32-level accuracy
    column by column, and an infinity passes through to its column only."""
    torch, device, mkgpu = env
    from minkasi_amd.lib.plans import RotPlan, SplitMatrix
    rng = np.random.default_rng(5)
    ndet, n = 200, 700
    v = _orth(rng, ndet) * (1.0 / rng.uniform(1e-3, 1e-2, ndet))[None, :]  # entries up to ~1e3
    d = rng.standard_normal((ndet, n))
    colscale = 10.0 ** rng.uniform(-30, 30, n)
    colscale[:5] = [1e-30, 1e30, 0.0, 1.0, 1e-38]
    colscale[300:341] = 1.0
    d = d * colscale[None, :]
    d[7, 300] = 3e6  # one entry far above the rest of its column
    d[150, 301] = 3e6  # ... in a later K step (the column is rescaled)
    d[:, 302:340] *= (10.0 ** (np.arange(ndet) / 20.0))[:, None]  # rising over every step
    d[:40, 340] = 0.0  # nothing in the first step
    d = d.astype(np.float32)
    ref = v @ d.astype(np.float64)
    sm = SplitMatrix(v, kind="f16")
    assert np.max(np.abs(v)) * 2.0 ** sm.aexp < 2 ** 14
    plan = RotPlan([(sm, n, 0, n, 0, n)])
    assert plan.mode == 2
    out = torch.empty(ndet * n, dtype=torch.float32, device=device.device())
    plan.exec(device.to_device(d), out)
    out = out.cpu().numpy().reshape(ndet, n).astype(np.float64)
    assert np.all(out[:, 2] == 0)
    cols = [j for j in range(n) if j != 2]
    err = np.array([np.linalg.norm(out[:, j] - ref[:, j]) / np.linalg.norm(ref[:, j]) for j in cols])
    print("fp16-pair rotation over 60 decades of column scale: worst column rel err %.2e" % err.max())
    assert err.max() < 2e-6
    d[11, 40] = np.inf
    o = torch.empty(ndet * n, dtype=torch.float32, device=device.device())
    plan.exec(device.to_device(d), o)
    o = o.cpu().numpy().reshape(ndet, n)
    assert not np.all(np.isfinite(o[:, 40]))
    assert np.all(np.isfinite(np.delete(o, 40, axis=1)))


def test_rotation_batched_segments(env):
    torch, device, mkgpu = env
    from minkasi_amd.lib.plans import RotPlan, SplitMatrix
    rng = np.random.default_rng(1)
    # one plan mixing every full-height class width and the general kernel (ndet > 512)
    shapes = [(64, 1000), (200, 3001), (130, 129), (600, 257), (33, 77), (448, 500)]
    vs, ds, segs, off = [], [], [], 0
    for ndet, n in shapes:
        v = _orth(rng, ndet)
        d = rng.standard_normal((ndet, n)).astype(np.float32)
        sm = SplitMatrix(v)
        segs.append((sm, n, off, n, off, n))
        vs.append(v)
        ds.append(d)
        off += ndet * n
    plan = RotPlan(segs)
    b = device.to_device(np.concatenate([d.ravel() for d in ds]))
    c = torch.zeros_like(b)
    plan.exec(b, c)
    c = c.cpu().numpy()
    off = 0
    for v, d in zip(vs, ds):
        k = d.size
        assert rel_err(c[off:off + k].reshape(d.shape), v @ d.astype(np.float64)) < 2e-5
        off += k


def test_fp64_gemm_and_gram(env):
    torch, device, mkgpu = env
    from minkasi_amd.mapmaking.noise import _gemm64, _gram
    rng = np.random.default_rng(2)
    for dt in (np.float32, np.float64):
        d = rng.standard_normal((77, 20001)).astype(dt)
        G = _gram(device.to_device(d)).cpu().numpy()
        ref = d.astype(np.float64) @ d.astype(np.float64).T
        assert rel_err(G, ref) < 1e-13
        A = rng.standard_normal((77, 77))
        C = _gemm64(device.to_device(A), device.to_device(d)).cpu().numpy()
        assert rel_err(C, A @ d.astype(np.float64)) < 1e-13


@pytest.mark.parametrize("n,fwhm", [(5000, 10.0), (65536, 50.0), (60, 50.0), (3, 10.0)])
def test_smoothing_matches_reference_formula(env, n, fwhm):
    torch, device, mkgpu = env
    from minkasi_amd.tools.smooth import smooth_rows
    rng = np.random.default_rng(n)
    P = np.abs(rng.standard_normal((3, n))) * np.logspace(0, 3, n)[None, ::-1] + 0.01
    ref = onoise.smooth_many_vecs(P, fwhm)
    out = smooth_rows(device.to_device(P), fwhm).cpu().numpy()
    assert max_rel_err(out, ref) < 1e-11
    # square + invert + zero bin 0, as NoiseSmoothedSVD uses it
    Y = rng.standard_normal((2, n))
    W = smooth_rows(device.to_device(Y), fwhm, square=True, invert=True).cpu().numpy()
    refW = onoise.smooth_many_vecs(Y ** 2, fwhm)
    refW[:, 1:] = 1.0 / refW[:, 1:]
    refW[:, 0] = 0
    assert max_rel_err(W, refW) < 1e-10


def _tod(rng, ndet, n, distinct=False):
    t = np.arange(n)
    cm = np.cumsum(rng.standard_normal(n)) * 0.006
    sig = 0.1 * (1 + np.arange(ndet) / ndet * 3) if distinct else np.full(ndet, 0.1)
    gains = 1 + 0.05 * rng.standard_normal(ndet)
    return gains[:, None] * cm[None, :] + sig[:, None] * rng.standard_normal((ndet, n)) + 0.01 * np.sin(t / 50.0)


def test_noise_apply_parity_same_basis(env):
    """Same singular basis on both sides (u_in): FP64 path to 1e-10, FP32 path
    (split-BF16 rotations + FP32 filter) to 1e-5 against the FP64 oracle."""
    torch, device, mkgpu = env
    from minkasi_amd.mapmaking.noise import NoiseSmoothedSVD
    rng = np.random.default_rng(5)
    ndet, n = 48, 30001
    d = _tod(rng, ndet, n)
    ref = onoise.NoiseSmoothedSVD(d, fwhm=10.0)
    u = ref.v.T.copy()
    g64 = NoiseSmoothedSVD(device.to_device(d), fwhm=10.0, u_in=u)
    assert max_rel_err(np.asarray(g64.mywt), ref.mywt) < 1e-6  # mywt kept in FP32
    x = rng.standard_normal((ndet, n))
    e64 = rel_err(g64.apply_noise(device.to_device(x)).cpu().numpy(), ref.apply_noise(x))
    g32 = NoiseSmoothedSVD(device.to_device(d.astype(np.float32)), fwhm=10.0, u_in=u)
    xs = [x, d, np.cumsum(rng.standard_normal((ndet, n)), axis=1) * 0.01]
    o64 = g64.apply_noise(device.to_device(x)).cpu().numpy()
    emax64 = max_rel_err(o64, ref.apply_noise(x))
    errs, emax = [], []
    for xx in xs:
        o = g32.apply_noise(device.to_device(xx.astype(np.float32))).cpu().numpy()
        r = ref.apply_noise(xx.astype(np.float32).astype(np.float64))
        errs.append(rel_err(o, r))
        emax.append(max_rel_err(o, r))
    print("apply_noise rel err fp64", e64, emax64, "fp32 [white, tod, random-walk] norm", errs, "max", emax)
    assert e64 < 1e-10 and emax64 < 1e-10
    # the apply_noise tolerance, norm-wise and max-elementwise
    assert max(errs) < 1e-5 and max(emax) < 1e-5


def test_noise_init_eigh_matches_svd_when_modes_distinct(env):
    torch, device, mkgpu = env
    from minkasi_amd.mapmaking.noise import NoiseSmoothedSVD
    rng = np.random.default_rng(6)
    ndet, n = 16, 20000
    d = _tod(rng, ndet, n, distinct=True)
    ref = onoise.NoiseSmoothedSVD(d, fwhm=50.0)
    g = NoiseSmoothedSVD(device.to_device(d), fwhm=50.0)
    # same basis up to signs
    assert np.allclose(np.abs(np.sum(g.v * ref.v, axis=1)), 1.0, atol=1e-8)
    x = rng.standard_normal((ndet, n))
    e = rel_err(g.apply_noise(device.to_device(x)).cpu().numpy(), ref.apply_noise(x))
    assert e < 1e-8
    w = g.get_det_weights()
    assert rel_err(w, ref.get_det_weights()) < 1e-5


def test_noise_prewhiten_matches_oracle(env):
    torch, device, mkgpu = env
    from minkasi_amd.mapmaking.noise import NoiseSmoothedSVD
    rng = np.random.default_rng(7)
    ndet, n = 12, 5001
    d = _tod(rng, ndet, n, distinct=True)
    ref = onoise.NoiseSmoothedSVD(d, fwhm=10.0, prewhiten=True)
    g = NoiseSmoothedSVD(device.to_device(d), fwhm=10.0, prewhiten=True)
    assert np.allclose(g.noisevec, ref.noisevec, rtol=1e-12)
    x = rng.standard_normal((ndet, n))
    assert rel_err(g.apply_noise(device.to_device(x)).cpu().numpy(), ref.apply_noise(x)) < 1e-8


@pytest.mark.parametrize("fwhm", [10.0, 50.0])
def test_device_smoothing_width_closed_form(env, fwhm):
    """The device smoothing (mk_smooth_rows_f64) against the same closed form
    as the oracle: a single-bin spectrum comes back as the Gaussian of the
    stated FWHM with peak 1/(sqrt(2 pi) sigma)."""
    from minkasi_amd.tools.smooth import smooth_many_vecs
    from tests.test_oracle_noise import _single_bin_closed_form
    _single_bin_closed_form(smooth_many_vecs, fwhm)
