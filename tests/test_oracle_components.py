"""Pins the oracle for the extra Mapset components (cut samples and polarised
maps): oracle/components.py and the or_tod2cuts / or_cuts2tod /
or_polmap2tod / or_tod2polmap functions of oracle/oracle.c."""
import ctypes

import numpy as np
import pytest

from oracle import components as oc
from oracle import core


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _runs_brute(v):
    out, i, n = [], 0, len(v)
    while i < n:
        if not v[i]:
            j = i
            while j < n and not v[j]:
                j += 1
            out.append((i, j))
            i = j
        else:
            i += 1
    return out


def test_segs_from_vec_matches_brute_force():
    rng = np.random.default_rng(0)
    for n in (1, 2, 7, 100, 513):
        for p in (0.05, 0.5, 0.95):
            v = rng.random(n) > p
            nseg, a, b = oc.segs_from_vec(v)
            assert list(zip(a, b)) == _runs_brute(v)
            assert nseg == len(a)
    assert oc.segs_from_vec(np.ones(9, bool)) == (0, [], [])
    assert oc.segs_from_vec(np.zeros(4, bool)) == (1, [0], [4])


def test_imap_from_mask_is_flatnonzero_and_merge_is_union():
    rng = np.random.default_rng(1)
    ndet, imax = 6, 300
    mask = rng.random((ndet, imax)) > 0.1
    mask[2] = True  # a detector without cuts
    segs = [oc.segs_from_vec(mask[d]) for d in range(ndet)]
    imap = oc.cuts_imap([s[1] for s in segs], [s[2] for s in segs], imax)
    assert np.array_equal(imap, np.flatnonzero(~mask))
    # overlapping and touching intervals: merge gives the set union
    a = [[5, 10, 30, 40, 100], [0], []]
    b = [[12, 20, 35, 41, 120], [300], []]
    ma, mb = oc.merge_segments(a, b, imax)
    for d in range(3):
        want = set()
        for s, e in zip(a[d], b[d]):
            want |= set(range(s, e))
        got = set()
        for s, e in zip(ma[d], mb[d]):
            got |= set(range(s, e))
        assert got == want
        assert all(e < s2 for e, s2 in zip(mb[d][:-1], ma[d][1:]))  # disjoint, not touching
    assert ma[0] == [5, 30, 40, 100] and mb[0] == [20, 35, 41, 120]
    assert oc.merge_segments(ma, mb, imax) == (ma, mb)


def _ref_or_none():
    return core.ref_lib()


@pytest.mark.parametrize("do_add", [0, 1])
def test_cut_gather_scatter_closed_form_and_reference(do_add):
    rng = np.random.default_rng(2)
    ndet, n = 5, 1000
    dat = rng.standard_normal((ndet, n))
    imap = np.sort(rng.choice(ndet * n, 700, replace=False)).astype(np.int64)
    vec0 = rng.standard_normal(len(imap))
    vec = vec0.copy()
    core.tod2cuts(vec, dat, imap, do_add)
    assert np.array_equal(vec, (vec0 if do_add else 0) + dat.ravel()[imap])
    d2 = dat.copy()
    core.cuts2tod(d2, vec0, imap, do_add)
    want = dat.ravel().copy()
    want[imap] = (want[imap] if do_add else 0) + vec0
    assert np.array_equal(d2.ravel(), want)
    L = _ref_or_none()
    if L is not None:
        vr = vec0.copy()
        L.tod2cuts(_p(vr), _p(dat), _p(imap), len(imap), do_add)
        assert np.array_equal(vr, vec)
        dr = dat.copy()
        L.cuts2tod(_p(dr), _p(vec0), _p(imap), len(imap), do_add)
        assert np.array_equal(dr, d2)


def _pol_problem(seed=3, ndet=4, n=500, npix=37):
    rng = np.random.default_rng(seed)
    ipix = rng.integers(0, npix, (ndet, n)).astype(np.int32)
    tg = rng.uniform(-np.pi, 2 * np.pi, (ndet, n))
    dat = rng.standard_normal((ndet, n))
    return rng, ipix, tg, dat, npix


@pytest.mark.parametrize("npol", [2, 3])
def test_polmap2tod_closed_form_and_reference(npol):
    rng, ipix, tg, dat, npix = _pol_problem()
    m = rng.standard_normal(npix * npol)
    out = np.zeros_like(dat)
    core.polmap2tod(out, m, tg, ipix, npol, do_add=False)
    mm = m.reshape(npix, npol)
    c, s = np.cos(tg), np.sin(tg)
    want = (mm[ipix, 0] + mm[ipix, 1] * c + mm[ipix, 2] * s) if npol == 3 else (mm[ipix, 0] * c + mm[ipix, 1] * s)
    assert np.allclose(out, want, rtol=1e-14, atol=1e-14)
    # a dropped term or swapped cos/sin fails the closed form
    if npol == 3:
        assert not np.allclose(out, mm[ipix, 0] + mm[ipix, 1] * s + mm[ipix, 2] * c, atol=1e-6)
    o2 = dat.copy()
    core.polmap2tod(o2, m, tg, ipix, npol, do_add=True)
    assert np.allclose(o2, dat + out, rtol=1e-14, atol=1e-14)
    L = _ref_or_none()
    if L is not None:
        f = L.map2tod_iqu_omp if npol == 3 else L.map2tod_qu_omp
        r = dat.copy()
        f(_p(r), _p(m), _p(tg), ipix.shape[0], ipix.shape[1], _p(ipix), 1)
        assert np.array_equal(r, o2)


@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_tod2polmap_bincount_and_reference(mode):
    rng, ipix, tg, dat, npix = _pol_problem(seed=4)
    nc = core.POL_NCOMP[mode]
    m = np.zeros(npix * nc)
    core.tod2polmap(m, dat, tg, ipix, mode)
    c, s, d = np.cos(tg).ravel(), np.sin(tg).ravel(), dat.ravel()
    w = {1: [d * c, d * s], 2: [d, d * c, d * s], 3: [d * c * c, d * s * s, d * s * c],
         4: [d, d * c, d * s, d * c * c, d * s * c, d * s * s]}[mode]
    want = np.stack([np.bincount(ipix.ravel(), wk, minlength=npix) for wk in w], axis=1)
    assert np.allclose(m.reshape(npix, nc), want, rtol=1e-12, atol=1e-12)
    L = _ref_or_none()
    if L is not None:
        f = {1: L.tod2map_qu_simple, 2: L.tod2map_iqu_simple, 3: L.tod2map_qu_precon_simple,
             4: L.tod2map_iqu_precon_simple}[mode]
        r = np.zeros_like(m)
        f(_p(r), _p(dat), _p(tg), ipix.shape[0], ipix.shape[1], _p(ipix))
        assert np.array_equal(r, m)


def _spd_blocks(rng, n, k, cond=10.0):
    q, _ = np.linalg.qr(rng.standard_normal((n, k, k)))
    lam = rng.uniform(1.0, cond, (n, k))
    return np.einsum("nij,nj,nkj->nik", q, lam, q)


def _pack(blocks, poltag):
    n = blocks.shape[0]
    if poltag == "QU_PRECON":
        return np.stack([blocks[:, 0, 0], blocks[:, 1, 1], blocks[:, 0, 1]], 1).reshape(n, 1, 3)
    idx = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    return np.stack([blocks[:, r, c] for r, c in idx], 1).reshape(n, 1, 6)


@pytest.mark.parametrize("poltag,k", [("QU_PRECON", 2), ("IQU_PRECON", 3)])
def test_invert_is_the_inverse_on_well_conditioned_blocks(poltag, k):
    rng = np.random.default_rng(5)
    B = _spd_blocks(rng, 50, k)
    inv = oc.polmap_invert(_pack(B, poltag), poltag, 1e-6)
    assert np.allclose(inv, _pack(np.linalg.inv(B), poltag), rtol=1e-10, atol=1e-12)


def test_invert_pseudo_inverse_identities_and_zeroing_rule():
    rng = np.random.default_rng(6)
    # rank-1 QU blocks: the Moore-Penrose identities hold for the result
    v = rng.standard_normal((20, 2))
    B = np.einsum("ni,nj->nij", v, v)
    inv = oc.polmap_invert(_pack(B, "QU_PRECON"), "QU_PRECON", 1e-6).reshape(20, 3)
    P = np.zeros((20, 2, 2))
    P[:, 0, 0], P[:, 1, 1], P[:, 0, 1], P[:, 1, 0] = inv[:, 0], inv[:, 1], inv[:, 2], inv[:, 2]
    assert np.allclose(B @ P @ B, B, atol=1e-10)
    assert np.allclose(P @ B @ P, P, atol=1e-10)
    assert np.allclose(B @ P, np.transpose(B @ P, (0, 2, 1)), atol=1e-10)
    assert np.allclose(P @ B, np.transpose(P @ B, (0, 2, 1)), atol=1e-10)
    # IQU: pixel 0 has a tiny determinant (zeroed, 1/I kept), pixel 1 is
    # empty (all zero), pixel 2 is well conditioned
    good = _spd_blocks(rng, 1, 3)[0]
    thin = np.diag([4.0, 1e-9, 1e-9])
    blocks = np.stack([thin, np.zeros((3, 3)), good])
    out = oc.polmap_invert(_pack(blocks, "IQU_PRECON"), "IQU_PRECON", 1e-6).reshape(3, 6)
    assert np.allclose(out[0], [0.25, 0, 0, 0, 0, 0])
    assert np.array_equal(out[1], np.zeros(6))
    assert np.allclose(out[2], _pack(np.linalg.inv(good)[None], "IQU_PRECON").ravel(), rtol=1e-10)


@pytest.mark.parametrize("poltag,k", [("QU_PRECON", 2), ("IQU_PRECON", 3)])
def test_precon_mul_is_blockwise_matvec(poltag, k):
    rng = np.random.default_rng(7)
    B = rng.standard_normal((30, k, k))
    B = B + np.transpose(B, (0, 2, 1))
    x = rng.standard_normal((30, 1, k))
    out = oc.precon_mul(_pack(B, poltag), x, poltag)
    assert np.allclose(out.reshape(30, k), np.einsum("nij,nj->ni", B, x.reshape(30, k)), rtol=1e-13, atol=1e-13)


def _lowrank(rng, ndet=12, n=600, k=2, amp=None):
    modes = np.cumsum(rng.standard_normal((k, n)), axis=1)
    if amp is not None:
        modes *= amp / np.abs(modes).max()
    mix = rng.standard_normal((ndet, k))
    return mix @ modes


def test_gapfill_eig_recovers_low_rank_signal():
    """Rank-2 signal (amplitude 30) plus unit white noise over 24 detectors,
    with short cuts: the fill recovers the signal at the cut samples to the
    noise level, better after more iterations than after one."""
    rng = np.random.default_rng(30)
    sig = _lowrank(rng, ndet=24, n=2000, amp=30.0)
    d = sig + rng.standard_normal(sig.shape)
    mask = np.ones(d.shape, dtype=bool)
    mask[3, 100:104] = False
    mask[7, 400:403] = False
    mask[0, 1990:1995] = False
    imap = np.flatnonzero(~mask)
    truth = sig.flat[imap]
    errs = [np.abs(oc.gapfill_eig(d, imap, niter_eig=ne, niter_inner=ni) - truth).max()
            for ne, ni in [(1, 1), (3, 3), (10, 10)]]
    assert errs[0] > 2 * errs[1] and max(errs[1:]) < 1.0 < np.abs(truth).max() / 10, errs


def test_gapfill_eig_inner_step_is_least_squares():
    """One eigen step with one inner step: the fill equals the cut samples of
    the least-squares fit of the (zero-filled) data by the selected modes."""
    rng = np.random.default_rng(31)
    d = _lowrank(rng, k=3) + 0.01 * rng.standard_normal((12, 600))
    imap = np.arange(5 * 600 + 50, 5 * 600 + 90)
    fill = oc.gapfill_eig(d, imap, thresh=2.0, niter_eig=1, niter_inner=1)
    tmp = d.copy()
    tmp.flat[imap] = 0.0
    w, v = np.linalg.eigh(tmp @ tmp.T)
    uu = v[:, w > 4.0 * np.median(w)].T @ tmp
    coef, *_ = np.linalg.lstsq(uu.T, tmp.T, rcond=None)  # tmp^T ~ uu^T coef
    assert np.allclose(fill, (coef.T @ uu).flat[imap], rtol=1e-9, atol=1e-9 * np.abs(tmp).max())


def test_gapfill_eig_without_modes_fills_zero():
    rng = np.random.default_rng(32)
    d = rng.standard_normal((6, 100))
    imap = np.arange(10, 30)
    assert np.array_equal(oc.gapfill_eig(d, imap, thresh=1e6), np.zeros(20))
