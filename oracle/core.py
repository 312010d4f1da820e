"""Oracle: projections, CAR WCS and pixelization in FP64 (test infrastructure).

Functions and what pins them (tests/test_oracle.py):
  car_wcs / car_npix      -- astropy fixtures (nx, ny per case).  Follows
                             minkasi/maps/utils.py:47-66, skymap.py:206-234.
  pix_from_radec          -- astropy fixtures (bit-exact ipix, raw pixel
                             coordinates to a few ulp).  skymap.py:457-490.
  map2tod                 -- closed form m.ravel()[ipix] and the reference's
                             own minkasi.c (oracle/_ref).  minkasi.c:116-144.
  tod2map                 -- np.bincount and the reference's tod2map_simple
                             (oracle/_ref).  minkasi.c:50-55, 97-112.
  tod2cuts / cuts2tod     -- closed forms (fancy indexing) and the reference's
                             own tod2cuts / cuts2tod (oracle/_ref).
                             minkasi.c:260-284.
  polmap2tod / tod2polmap -- closed forms (cos/sin-weighted gathers and
                             np.bincount) and the reference's own
                             map2tod_{iqu,qu}_omp, tod2map_*_simple
                             (oracle/_ref).  minkasi.c:146-257.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise OSError("oracle/liboracle.so missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I64 = ctypes.c_int64
        I = ctypes.c_int
        L.or_map2tod.argtypes = [P, P, I64, I64, P, I, I]
        L.or_tod2map_simple.argtypes = [P, P, I64, I64, P]
        L.or_tod2map_omp.argtypes = [P, P, I64, I64, P, I64, I]
        L.or_car_euler.argtypes = [ctypes.c_double, ctypes.c_double, P]
        L.or_car_euler.restype = I
        L.or_car_world2pix.argtypes = [P, P, I64, P, P, P, P]
        L.or_pix_from_radec.argtypes = [P, P, I64, P, P, ctypes.c_int32, P, I]
        L.or_dct1_direct.argtypes = [P, I64, I64, P]
        L.or_tod2cuts.argtypes = [P, P, P, I64, I]
        L.or_cuts2tod.argtypes = [P, P, P, I64, I]
        L.or_polmap2tod.argtypes = [P, P, P, I64, P, I, I]
        L.or_tod2polmap.argtypes = [P, P, P, I64, P, I]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own minkasi.c, compiled from /root/reference into
    oracle/_ref by oracle/Makefile.  None when it was not built."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libminkasi_ref.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I = ctypes.c_int
        L.map2tod_omp.argtypes = [P, P, I, I, P, I]
        L.map2tod_simple.argtypes = [P, P, I, I, P, I]
        L.tod2map_simple.argtypes = [P, P, I, I, P]
        L.tod2map_omp.argtypes = [P, P, I, I, P, I]
        L.set_nthread.argtypes = [I]
        L.tod2cuts.argtypes = [P, P, P, I, I]
        L.cuts2tod.argtypes = [P, P, P, I, I]
        L.map2tod_iqu_omp.argtypes = [P, P, P, I, I, P, I]
        L.map2tod_qu_omp.argtypes = [P, P, P, I, I, P, I]
        L.tod2map_qu_simple.argtypes = [P, P, P, I, I, P]
        L.tod2map_iqu_simple.argtypes = [P, P, P, I, I, P]
        L.tod2map_qu_precon_simple.argtypes = [P, P, P, I, I, P]
        L.tod2map_iqu_precon_simple.argtypes = [P, P, P, I, I, P]
        _ref = L
    return _ref


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def nthreads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


# --------------------------------------------------------------------------
# WCS (minkasi/maps/utils.py:47-66, proj='CAR', ref_equ=False)

def car_wcs(lims, pixsize, cosdec=None):
    """Returns dict(crval, cdelt, crpix) in degrees exactly as get_wcs sets
    them: crval = lims[1], lims[2] (rad -> deg), cdelt = [-pix, pix]."""
    cosdec = 1.0  # CAR forces cosdec = 1 (utils.py:53)
    crval = np.array([lims[1] * 180 / np.pi, lims[2] * 180 / np.pi])
    cdelt = np.array([-pixsize / cosdec * 180 / np.pi, pixsize * 180 / np.pi])
    crpix = np.array([1.0, 1.0])
    return dict(crval=crval, cdelt=cdelt, crpix=crpix)


def car_euler(w):
    eul = np.zeros(6)
    st = lib().or_car_euler(float(w["crval"][0]), float(w["crval"][1]), _p(eul))
    if st != 0:
        raise ValueError("invalid CAR WCS (status %d)" % st)
    return eul


def _lin(w):
    return np.array([w["cdelt"][0], w["cdelt"][1], w["crpix"][0], w["crpix"][1]])


def car_world2pix(w, ra, dec):
    """1-based pixel coordinates of (ra, dec) in radians (wcs_world2pix(.., 1))."""
    ra = np.ascontiguousarray(ra, dtype=np.float64)
    dec = np.ascontiguousarray(dec, dtype=np.float64)
    px = np.empty_like(ra)
    py = np.empty_like(ra)
    eul = car_euler(w)
    lin_ = _lin(w)
    lib().or_car_world2pix(_p(ra), _p(dec), ra.size, _p(eul), _p(lin_), _p(px), _p(py))
    return px, py


def car_npix(w, lims, pad=2, nx=None, ny=None):
    """SkyMap.get_npix (skymap.py:206-234)."""
    corners = np.zeros([4, 2])
    corners[0, :] = [lims[0], lims[2]]
    corners[1, :] = [lims[0], lims[3]]
    corners[2, :] = [lims[1], lims[2]]
    corners[3, :] = [lims[1], lims[3]]
    px, py = car_world2pix(w, corners[:, 0].copy(), corners[:, 1].copy())
    px = np.round(px)
    py = np.round(py)
    if nx is None:
        nx = px.max() + pad
    if ny is None:
        ny = py.max() + pad
    return int(nx), int(ny)


def pix_from_radec(w, ny, ra, dec):
    """SkyMap.pix_from_radec (skymap.py:457-490) -> int32 (ndet, ndata)."""
    ra = np.ascontiguousarray(ra, dtype=np.float64)
    dec = np.ascontiguousarray(dec, dtype=np.float64)
    ipix = np.empty(ra.shape, dtype=np.int32)
    eul = car_euler(w)
    lin_ = _lin(w)
    lib().or_pix_from_radec(_p(ra), _p(dec), ra.size, _p(eul), _p(lin_), int(ny), _p(ipix), nthreads())
    return ipix


# --------------------------------------------------------------------------
# Projections (minkasi.c:116-144, 50-55, 97-112)

def map2tod(dat, map_, ipix, do_add=True, nthread=None):
    assert dat.dtype == np.float64 and map_.dtype == np.float64
    assert ipix.dtype == np.int32 and dat.flags.c_contiguous
    ndet, ndata = dat.shape
    m = np.ascontiguousarray(map_)
    lib().or_map2tod(_p(dat), _p(m), ndet, ndata, _p(ipix), int(bool(do_add)), nthread or nthreads())
    return dat


def tod2map(map_, dat, ipix, nthread=None):
    """Adds tod into map in place (deterministic private-map reduction)."""
    assert map_.dtype == np.float64 and map_.flags.c_contiguous
    dat = np.ascontiguousarray(dat, dtype=np.float64)
    ndet, ndata = dat.shape
    lib().or_tod2map_omp(_p(map_), _p(dat), ndet, ndata, _p(ipix), map_.size, nthread or nthreads())
    return map_


def tod2map_simple(map_, dat, ipix):
    dat = np.ascontiguousarray(dat, dtype=np.float64)
    ndet, ndata = dat.shape
    lib().or_tod2map_simple(_p(map_), _p(dat), ndet, ndata, _p(ipix))
    return map_


def dct1_direct(x):
    """DCT-I by definition (O(n^2)), rows of a 2-D FP64 array."""
    x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64)
    out = np.empty_like(x)
    lib().or_dct1_direct(_p(x), x.shape[0], x.shape[1], _p(out))
    return out


# --------------------------------------------------------------------------
# Cut samples and polarised projections (minkasi.c:146-284)

POL_NCOMP = {1: 2, 2: 3, 3: 3, 4: 6}


def tod2cuts(vec, dat, imap, do_add=True):
    """vec (+)= dat.ravel()[imap], serially (minkasi.c:260-268)."""
    assert vec.dtype == np.float64 and dat.dtype == np.float64 and imap.dtype == np.int64
    assert dat.flags.c_contiguous and vec.flags.c_contiguous
    lib().or_tod2cuts(_p(vec), _p(dat), _p(imap), len(imap), int(bool(do_add)))


def cuts2tod(dat, vec, imap, do_add=True):
    """dat.ravel()[imap] (+)= vec, serially (minkasi.c:271-284)."""
    assert vec.dtype == np.float64 and dat.dtype == np.float64 and imap.dtype == np.int64
    assert dat.flags.c_contiguous and vec.flags.c_contiguous
    lib().or_cuts2tod(_p(dat), _p(vec), _p(imap), len(imap), int(bool(do_add)))


def polmap2tod(dat, map_, twogamma, ipix, npol, do_add=True):
    """npol 3 (IQU) or 2 (QU), map_ pixel-major (npix * npol)."""
    assert dat.dtype == map_.dtype == twogamma.dtype == np.float64 and ipix.dtype == np.int32
    lib().or_polmap2tod(_p(dat), _p(map_), _p(twogamma), dat.size, _p(ipix), int(npol), int(bool(do_add)))


def tod2polmap(map_, dat, twogamma, ipix, mode):
    """mode 1 QU, 2 IQU, 3 QU_PRECON, 4 IQU_PRECON; map_ += serially."""
    assert dat.dtype == map_.dtype == twogamma.dtype == np.float64 and ipix.dtype == np.int32
    lib().or_tod2polmap(_p(map_), _p(dat), _p(twogamma), dat.size, _p(ipix), int(mode))
