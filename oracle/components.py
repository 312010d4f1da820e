"""Oracle: the extra Mapset components, cut samples and polarised maps, in
FP64 (test infrastructure).

Functions, what they follow, and what pins them (tests/test_oracle_components.py):
  segs_from_vec     -- minkasi/tods/cuts.py:19-57 (False runs of a
                       True-padded vector).  Pinned by a brute-force scan
                       of runs on random masks, and the all-True case.
  cuts_imap         -- cuts.py:330-347 (det * imax + sample, det-major,
                       segments in stored order).  Pinned by
                       np.flatnonzero(~mask) for segments from a mask.
  merge_segments    -- cuts.py:311-328 (union of [istart, istop) per
                       detector).  Pinned by the set union of the intervals
                       and by idempotence.
  polmap_invert     -- maps/polmap.py:300-384 (np.linalg.pinv with rcond =
                       thresh; IQU_PRECON blocks under thresh * max det
                       zeroed except 1/I).  Pinned by np.linalg.inv on
                       well-conditioned blocks, the four Moore-Penrose
                       identities on rank-deficient ones, and a hand-built
                       map whose zeroed pixels are known.
  precon_mul        -- polmap.py:624-684 (the block times the Stokes vector
                       of each pixel).  Pinned by an explicit per-pixel
                       3x3 (2x2) matrix-vector product.
  gapfill_eig       -- minkasi/tods/processing.py:357-421 (eigenmodes of
                       the gap-filled detector covariance, least-squares fit
                       of the data by them, the fit replacing the cut
                       samples).  Pinned by exactly low-rank data, whose cut
                       samples the iteration recovers (error shrinking with
                       the iteration count), by an explicit least-squares
                       solve of one inner step, and by the no-mode case
                       (threshold above every eigenvalue: zero fill).
"""
import numpy as np


def segs_from_vec(vec, pad=True):
    """(nseg, istart, istop) of the False runs of vec (cuts.py:19-57)."""
    if pad:
        vv = np.ones(len(vec) + 2, dtype=bool)
        vv[1:-1] = vec
    else:
        vv = np.asarray(vec, dtype=bool)
    if vv.all():
        return 0, [], []
    edges = np.flatnonzero(vv[1:] != vv[:-1])
    return len(edges) // 2, list(edges[0::2]), list(edges[1::2])


def cuts_imap(istart, istop, imax):
    """Flat indices of every cut sample; istart/istop are per-detector lists."""
    parts = [d * imax + np.arange(a, b, dtype=np.int64)
             for d in range(len(istart)) for a, b in zip(istart[d], istop[d])]
    return np.concatenate(parts) if parts else np.zeros(0, dtype=np.int64)


def merge_segments(istart, istop, imax):
    """Per detector, the union of its [istart, istop) intervals as sorted,
    disjoint, non-touching runs."""
    out_a, out_b = [], []
    for a_list, b_list in zip(istart, istop):
        if len(a_list) < 2:
            out_a.append(list(a_list))
            out_b.append(list(b_list))
            continue
        tmp = np.ones(imax + 2, dtype=bool)
        for a, b in zip(a_list, b_list):
            tmp[a + 1:b + 1] = False
        _, s, e = segs_from_vec(tmp, pad=False)
        out_a.append(s)
        out_b.append(e)
    return out_a, out_b


def _blocks(map_, poltag):
    n = map_.shape[0] * map_.shape[1]
    m = map_.reshape(n, -1)
    if poltag == "QU_PRECON":
        t = np.zeros((n, 2, 2))
        t[:, 0, 0], t[:, 1, 1], t[:, 0, 1], t[:, 1, 0] = m[:, 0], m[:, 1], m[:, 2], m[:, 2]
    else:
        t = np.zeros((n, 3, 3))
        t[:, 0, 0] = m[:, 0]
        t[:, 0, 1] = t[:, 1, 0] = m[:, 1]
        t[:, 0, 2] = t[:, 2, 0] = m[:, 2]
        t[:, 1, 1] = m[:, 3]
        t[:, 1, 2] = t[:, 2, 1] = m[:, 4]
        t[:, 2, 2] = m[:, 5]
    return t


def polmap_invert(map_, poltag, thresh=1e-6):
    """Returns the inverted (nx, ny, ncomp) map (polmap.py:300-384)."""
    map_ = np.asarray(map_, dtype=np.float64)
    nx, ny = map_.shape[:2]
    t = _blocks(map_, poltag)
    out = map_.copy()
    if poltag == "QU_PRECON":
        p = np.linalg.pinv(t, thresh)
        out[:, :, 0] = p[:, 0, 0].reshape(nx, ny)
        out[:, :, 1] = p[:, 1, 1].reshape(nx, ny)
        out[:, :, 2] = p[:, 0, 1].reshape(nx, ny)
        return out
    dets = np.linalg.det(t)
    bad = dets < thresh * dets.max()
    pos = t[:, 0, 0] > 0
    keep_i = t[bad & pos, 0, 0]
    p = np.linalg.pinv(t, thresh)
    p[bad] = 0
    p[bad & pos, 0, 0] = 1.0 / keep_i
    for k, (r, c) in enumerate([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
        out[:, :, k] = p[:, r, c].reshape(nx, ny)
    return out


def precon_mul(pre, x, poltag):
    """The preconditioner block of each pixel times its Stokes vector."""
    pre = np.asarray(pre, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    out = np.empty_like(x)
    if poltag == "QU_PRECON":
        out[..., 0] = pre[..., 0] * x[..., 0] + pre[..., 2] * x[..., 1]
        out[..., 1] = pre[..., 2] * x[..., 0] + pre[..., 1] * x[..., 1]
        return out
    a = pre
    out[..., 0] = a[..., 0] * x[..., 0] + a[..., 1] * x[..., 1] + a[..., 2] * x[..., 2]
    out[..., 1] = a[..., 1] * x[..., 0] + a[..., 3] * x[..., 1] + a[..., 4] * x[..., 2]
    out[..., 2] = a[..., 2] * x[..., 0] + a[..., 4] * x[..., 1] + a[..., 5] * x[..., 2]
    return out


def gapfill_eig(dat, imap, thresh=5.0, niter_eig=3, niter_inner=3):
    """Fill values of the cut samples imap (flat indices into dat, ndet x n)
    after gapfill_eig's iterations; the steps of processing.py:395-419 with
    the cut vector as a flat array (map2tod(do_add=False) sets
    tmp.flat[imap] = fill, tod2map(do_add=False) reads fill = pred.flat[imap])."""
    dat = np.asarray(dat, dtype=np.float64)
    fill = np.zeros(len(imap))
    for _ in range(niter_eig):
        tmp = dat.copy()
        tmp.flat[imap] = fill
        mycov = tmp @ tmp.T
        ee, vv = np.linalg.eig(mycov)
        mask = ee > thresh * thresh * np.median(ee)
        vv = np.real(vv[:, mask])
        uu = vv.T @ tmp
        lhs_inv = np.linalg.inv(uu @ uu.T) if mask.any() else np.zeros((0, 0))
        for __ in range(niter_inner):
            rhs = tmp @ uu.T
            fitp = lhs_inv @ rhs.T
            pred = fitp.T @ uu
            fill = pred.flat[imap].copy()
            tmp.flat[imap] = fill
    return fill

