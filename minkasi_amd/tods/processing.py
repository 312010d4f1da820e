"""TOD preprocessing of minkasi's canonical pipeline
(examples/minkasi_mpi_example.py:42-52): truncation to an FFT-friendly
length, DCT-I downsampling and the common-mode plus polynomial-drift fit.

truncate_tod follows minkasi/tods/processing.py:648-681.  The TOD keeps
ndata_new = max(find_good_fft_lens(ndata - 1, primes)) + 1 samples (nothing
happens when that is not shorter), and every array of the info dict with an
axis of length ndata is cut to its first ndata_new entries along the first
such axis.  Device tensors are cut on the device (a contiguous copy, as
np.take makes on the host).

find_good_fft_lens restates minkasi/tools/py_fft.py:300-395: the products of
the given primes enumerated in log2 space.  A product is kept when its log2
is at most log2(n + 0.5) + 1e-3; every partial product without the last
prime's factor must stay below log2(n + 0.5).  The 1e-3 slack means a
product slightly above n can be returned, in which case truncate_tod leaves
the TOD as it is, like the reference.
"""
import math

import numpy as np
import torch

from .. import device as dev
from ..lib import mkgpu


def _products(lim, logs, start, out):
    # products of the remaining primes times 2**start, log2 <= lim (+ slack on the last)
    step = logs[0]
    if len(logs) == 1:
        for v in np.arange(start, lim + 1e-3, step):
            out.append(v)
        return
    for v in np.arange(start, lim, step):
        _products(lim, logs[1:], v, out)


def find_good_fft_lens(n, primes=(2, 3, 5, 7)):
    """Sorted array of the products of `primes` up to about n."""
    lim = math.log2(n + 0.5)
    vals = []
    _products(lim, np.log2(np.asarray(primes, dtype=float)), 0.0, vals)
    return np.sort(np.asarray(np.round(2.0 ** np.asarray(vals)), dtype=np.int64))


def good_length(ndata, primes=(2, 3, 5, 7, 11)):
    """The TOD length truncate_tod keeps for ndata samples."""
    new = int(find_good_fft_lens(ndata - 1, primes).max()) + 1
    return new if new < ndata else ndata


def truncate_tod(tod_info, primes=(2, 3, 5, 7, 11)):
    """Cut the TOD in place to a good FFT length (processing.py:648-681).
    Requires 'dat_calib' (2-D); returns the new length (or None without it)."""
    if "dat_calib" not in tod_info:
        return None
    dat = tod_info["dat_calib"]
    if len(dat.shape) != 2:
        raise ValueError("dat_calib not 2d")
    ndata = int(dat.shape[1])
    new = good_length(ndata, primes)
    if new >= ndata:
        return ndata
    for key in list(tod_info.keys()):
        a = tod_info[key]
        shape = getattr(a, "shape", None)
        if shape is None or len(shape) == 0:
            continue
        axes = [i for i, s in enumerate(shape) if s == ndata]
        if not axes:
            continue
        if hasattr(a, "narrow"):  # torch tensor, any device
            tod_info[key] = a.narrow(axes[0], 0, new).contiguous()
        else:
            tod_info[key] = np.take(a, indices=range(0, new), axis=axes[0])
    return new


def downsample_tod(tod_info, fac=10):
    """DCT-I downsampling of every array whose last axis has ndata samples
    (processing.py:621-645, via downsample_array_r2r).  As in the reference,
    dt is multiplied by 10 whatever fac is (processing.py:637)."""
    from ..tools.array_ops import downsample_array_r2r
    ndata = tod_info["dat_calib"].shape[1]
    tod_info["dt"] *= 10
    for key in list(tod_info.keys()):
        a = tod_info[key]
        shape = getattr(a, "shape", None)
        if shape is None or len(shape) == 0 or shape[-1] != ndata:
            continue
        tod_info[key] = downsample_array_r2r(a, fac)


def fit_cm_plus_poly(dat, ord=2, cm_ord=1, niter=1, medsub=False, full_out=False):
    """Fit the detector-median common mode times Legendre polynomials of order
    < cm_ord plus per-detector Legendre drifts of order <= ord, and return
    the data with the drifts (not the common mode) removed
    (processing.py:510-579).

    Device tensors (ndet x ndata, FP32 or FP64) are fitted on the GPU and
    keep their dtype; numpy input is fitted in FP64 and returned as FP64
    numpy arrays, as the reference's float64 arithmetic returns them.  The
    median (mk_median_cols_*), the design, the normal-equation sums
    (mk_design_dots_*) and the subtraction (mk_design_apply_*) run on the
    device; the (ord + 1 + cm_ord)^2 inverse is taken on the host in FP64,
    as np.linalg.inv does in _linfit_2mat (processing.py:28-39).
    full_out: also return pred2 (the fitted common mode, ndet x ndata) and cm."""
    if medsub:
        raise NotImplementedError("medsub=True (a per-detector median over ndata samples) is not provided")
    if ord < 0 or cm_ord < 1:
        raise ValueError("need ord >= 0 and cm_ord >= 1 (legvander of a negative degree)")
    host = not isinstance(dat, torch.Tensor)
    if host:
        d = torch.from_numpy(np.ascontiguousarray(dat, dtype=np.float64)).to(dev.device())
    else:
        d = dat.to(dev.device())
        if d.dtype not in (torch.float32, torch.float64):
            d = d.to(torch.float64)
        d = d.contiguous()
    if d.dim() != 2:
        raise ValueError("dat must be ndet x ndata")
    ndet, n = d.shape
    if n < 2 or ndet < 1 or ndet > 1024:
        raise ValueError("need ndata >= 2 and 1 <= ndet <= 1024 (got %d x %d)" % (ndet, n))
    sfx = "f64" if d.dtype == torch.float64 else "f32"
    ncol = ord + 1 + cm_ord
    if ncol > 8:
        raise ValueError("ord + 1 + cm_ord must be at most 8")
    f64 = dict(dtype=torch.float64, device=d.device)
    sp = dev.stream_ptr()
    dd = d.clone()
    cm = torch.zeros(n, dtype=d.dtype, device=d.device)
    pred2 = torch.zeros(n, dtype=d.dtype, device=d.device)  # the reference's initial value (niter = 0)
    X = torch.empty((n, ncol), **f64)
    Xt = torch.empty((ncol, n), **f64)
    lhs = torch.empty((ncol, ncol), **f64)
    rhs = torch.empty((ncol, ndet), **f64)
    wb = max(int(mkgpu.lib.mk_design_dots_workspace_bytes(n, ncol, ndet)),
             int(mkgpu.lib.mk_design_dots_workspace_bytes(n, ncol, ncol)))
    work = torch.empty(wb // 8 + 32, **f64)
    for _ in range(niter):
        mkgpu.check("mk_median_cols_" + sfx, dd.data_ptr(), ndet, n, n, cm.data_ptr(), sp)
        mkgpu.check("mk_cm_poly_design_" + sfx, cm.data_ptr(), n, ord, cm_ord, X.data_ptr(), Xt.data_ptr(), sp)
        mkgpu.check("mk_design_dots_f64", X.data_ptr(), n, ncol, Xt.data_ptr(), ncol, n, lhs.data_ptr(),
                    work.data_ptr(), work.numel() * 8, sp)
        mkgpu.check("mk_design_dots_" + sfx, X.data_ptr(), n, ncol, d.data_ptr(), ndet, n, rhs.data_ptr(),
                    work.data_ptr(), work.numel() * 8, sp)
        fitp = np.dot(np.linalg.inv(lhs.cpu().numpy()), rhs.cpu().numpy())
        coef = torch.from_numpy(np.ascontiguousarray(fitp)).to(d.device)
        mkgpu.check("mk_design_apply_" + sfx, d.data_ptr(), ndet, n, n, X.data_ptr(), ncol, 0, ord + 1,
                    coef.data_ptr(), dd.data_ptr(), n, sp)
        if full_out:
            pred2 = torch.empty_like(d)
            coef_cm = coef[ord + 1:]  # rows of a row-major matrix: contiguous
            mkgpu.check("mk_design_apply_" + sfx, 0, ndet, n, n, X.data_ptr(), ncol, ord + 1, cm_ord,
                        coef_cm.data_ptr(), pred2.data_ptr(), n, sp)
    if host:
        dd, pred2, cm = dd.cpu().numpy(), pred2.cpu().numpy(), cm.cpu().numpy()
    if full_out:
        return dd, pred2, cm
    return dd


def gapfill_lin(dat):
    """Linear interpolation over NaN runs (tods/processing.py:286-354), in
    place on a host array ndet x nsamp.  Returns (dat, cuts) with cuts[det]
    the [last good sample before, first good sample after] pair of each run.
    As in the reference, a run that starts at sample 0 takes its left edge
    from the last sample (numpy's index -1) and is left unfilled, a run that
    reaches the last sample raises IndexError, and a detector that is NaN
    throughout raises ValueError."""
    dat = np.asarray(dat)
    ndet, nsamp = dat.shape
    samps = np.arange(nsamp)
    full_cuts = {}
    for i in range(ndet):
        bad = np.flatnonzero(np.isnan(dat[i]))
        if len(bad) == 0:
            continue
        if len(bad) == nsamp:
            raise ValueError("Error: detector non-responsive. Please cut before passing to gapfill_lin.")
        brk = np.flatnonzero(np.diff(bad) > 1)
        starts = np.concatenate([[0], brk + 1])
        ends = np.concatenate([brk, [len(bad) - 1]])
        cur = []
        for s0, e0 in zip(starts, ends):
            lo, hi = samps[bad[s0] - 1], samps[bad[e0] + 1]
            lo_v, hi_v = dat[i, bad[s0] - 1], dat[i, bad[e0] + 1]
            cur.append([lo, hi])
            miss = np.arange(lo + 1, hi)
            dat[i, miss] = np.interp(miss, [lo, hi], [lo_v, hi_v])
        full_cuts[i] = cur
    return dat, full_cuts



def _cross64(A, B):
    """A . B^T on the device in FP64 (A: M x n, B: N x n, contiguous)."""
    M, n = A.shape
    N = B.shape[0]
    C = torch.empty((M, N), dtype=torch.float64, device=A.device)
    nb = int(mkgpu.lib.mk_cross_workspace_bytes(M, N, n))
    work = torch.empty(nb // 8 + 32, dtype=torch.float64, device=A.device)
    mkgpu.check("mk_cross_f64", M, N, n, dev.ptr(A, torch.float64, "A"), n, dev.ptr(B, torch.float64, "B"), n,
                C.data_ptr(), work.data_ptr(), work.numel() * 8, dev.stream_ptr())
    return C


def gapfill_eig(dat, cuts, tod=None, thresh=5.0, niter_eig=3, niter_inner=3, insert_cuts=False):
    """Fits the cut samples of dat (ndet x n) from its dominant detector
    eigenmodes (tods/processing.py:357-421), returning a copy of the
    CutsCompact `cuts` holding the fitted values.

    niter_eig times: the data with the current fill at the cut samples gives
    the detector covariance tmp.tmp^T (device, FP64); its eigenvectors with
    eigenvalues above thresh^2 times the median (numpy eig on the host, as
    the reference) span the modes uu = vv^T tmp; niter_inner times the data
    are fitted by those modes and the fit replaces the cut samples.  The
    long products (tmp.tmp^T, vv^T tmp, uu.uu^T, tmp.uu^T, fit.uu) are
    device FP64 kernels; the eigen-decomposition and the neig x neig inverse
    are host numpy.  With insert_cuts the fill is added into dat at the cut
    samples (in place; dat must then be a device FP64 tensor or a numpy
    array)."""
    from ..mapmaking.noise import _gemm64
    host = not isinstance(dat, torch.Tensor)
    d = dev.to_device(np.ascontiguousarray(dat), torch.float64) if host else dat.to(torch.float64).contiguous()
    if d.dim() != 2:
        raise ValueError("dat must be ndet x n")
    cuts_cur = cuts.copy()
    if len(cuts_cur.imap) == 0 and int(np.sum(cuts_cur.nseg)) > 0:
        cuts_cur.get_imap()
    out_dtype = cuts_cur.map.dtype
    cuts_cur.map = cuts_cur.map.to(torch.float64)
    cuts_cur.clear()
    for _ in range(niter_eig):
        tmp = d.clone()
        cuts_cur.map2tod(tod, tmp, do_add=False)
        mycov = _cross64(tmp, tmp).cpu().numpy()
        ee, vv = np.linalg.eig(mycov)
        mask = ee > thresh * thresh * np.median(ee)
        neig = int(np.sum(mask))
        print("working with " + repr(neig) + " eigenvectors.")
        if neig == 0:  # the fit of no modes is zero
            cuts_cur.clear()
            cuts_cur.map2tod(tod, tmp, do_add=False)
            continue
        vv = np.real(vv[:, mask])
        uu = _gemm64(torch.from_numpy(np.ascontiguousarray(vv.T)).to(d.device), tmp)
        lhs_inv = np.linalg.inv(_cross64(uu, uu).cpu().numpy())
        for __ in range(niter_inner):
            rhs = _cross64(tmp, uu).cpu().numpy()                       # ndet x neig
            fitp = lhs_inv @ rhs.T                                       # neig x ndet
            pred = _gemm64(torch.from_numpy(np.ascontiguousarray(fitp.T)).to(d.device), uu)
            cuts_cur.tod2map(tod, pred, do_add=False)
            cuts_cur.map2tod(tod, tmp, do_add=False)
    if insert_cuts:
        if host:
            cuts_cur.map2tod(tod, d)
            np.copyto(dat, d.cpu().numpy().reshape(np.shape(dat)).astype(np.asarray(dat).dtype, copy=False))
        else:
            if dat.dtype != torch.float64 or not dat.is_contiguous():
                raise TypeError("insert_cuts updates dat in place: pass a contiguous float64 device tensor")
            cuts_cur.map2tod(tod, dat)
    cuts_cur.map = cuts_cur.map.to(out_dtype)
    return cuts_cur
