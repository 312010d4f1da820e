"""Oracle: Tod.dot, TodVec.dot, make_rhs, make_hits and run_pcg in FP64
(test infrastructure).

Follows minkasi/tods/core.py:582-691 (mapset2tod zero-then-add, apply_noise,
tod2map), :967-1047 (TODs in list order), minkasi/mapmaking/tod2map.py:
343-400 (hits) and minkasi/mapmaking/pcg.py:101-177 (update order).
Pins (tests/test_oracle_solve.py): on tiny problems A is assembled column by
column from todvec_dot and checked symmetric PSD with the constant-on-hit-
pixels null mode; run_pcg equals a dense-matrix PCG with the same
preconditioner; with the identity preconditioner it is plain CG; r.z >= 0.
"""
import numpy as np

from . import core


class OTod:
    def __init__(self, dat, ipix, noise):
        self.dat = np.asarray(dat, dtype=np.float64)
        self.ipix = np.ascontiguousarray(ipix, dtype=np.int32)
        self.noise = noise


def tod_dot(tod, m, out):
    """out += P^T N^-1 P m for one TOD (core.py:649-691)."""
    tmp = np.zeros(tod.dat.shape)
    core.map2tod(tmp, np.ascontiguousarray(m).ravel(), tod.ipix, do_add=True)
    tmp = tod.noise.apply_noise(tmp)
    core.tod2map(out.reshape(-1), tmp, tod.ipix)
    return out


def todvec_dot(tods, m):
    out = np.zeros(m.shape)
    for tod in tods:
        tod_dot(tod, m, out)
    return out


def make_rhs(tods, shape):
    b = np.zeros(shape)
    for tod in tods:
        core.tod2map(b.reshape(-1), tod.noise.apply_noise(tod.dat), tod.ipix)
    return b


def make_hits(tods, shape):
    h = np.zeros(shape)
    for tod in tods:
        core.tod2map(h.reshape(-1), np.ones(tod.dat.shape), tod.ipix)
    return h


def jacobi_from_hits(hits):
    """examples/minkasi_mpi_example.py:99-105: 1/hits where hits > 0."""
    pre = np.zeros_like(hits)
    ii = hits > 0
    pre[ii] = 1.0 / hits[ii]
    return pre


def run_pcg(b, x0, apply_A, precon=None, maxiter=25, tol=None):
    """pcg.py:101-177 with numpy maps; precon is an elementwise weight map or
    a callable.  Returns (x, history)."""
    def M(r):
        if precon is None:
            return r.copy()
        if callable(precon):
            return precon(r)
        return precon * r

    Ax = apply_A(x0)
    r = b.copy() - Ax
    z = M(r)
    p = z.copy()
    zr = float(np.sum(r * z))
    x = x0.copy()
    bb = float(np.sum(b * b))
    hist = []
    for i in range(maxiter):
        Ap = apply_A(p)
        pAp = float(np.sum(p * Ap))
        alpha = zr / pAp
        x = x + alpha * p
        r = r - alpha * Ap
        z = M(r)
        zr_new = float(np.sum(r * z))
        beta = zr_new / zr
        p = z + beta * p
        zr = zr_new
        rel = np.sqrt(float(np.sum(r * r)) / bb)
        hist.append(dict(it=i, zr=zr, alpha=alpha, beta=beta, rel_res=rel))
        if tol is not None and rel < tol:
            break
    return x, hist


def dense_A(tods, shape):
    """A assembled column by column (tiny maps only)."""
    npix = shape[0] * shape[1]
    A = np.zeros((npix, npix))
    e = np.zeros(shape)
    for j in range(npix):
        e.flat[j] = 1.0
        A[:, j] = todvec_dot(tods, e).ravel()
        e.flat[j] = 0.0
    return A
