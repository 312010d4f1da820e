"""map2tod wrapper (minkasi/mapmaking/map2tod.py:19-54) over mk_map2tod_*,
and the one-off validation of pixel indices the projections rely on."""
import weakref

import torch

from .. import device as dev
from ..lib import mkgpu

# (id of an index tensor, npix) -> (weak reference, data pointer, version) of
# the index tensors already found inside [0, npix)
_PIX_OK = {}


def check_pix(ipix, npix):
    """Raises ValueError when any index of the int32 device tensor ipix lies
    outside [0, npix).  The reference's C kernels index the map unchecked
    (minkasi/lib/minkasi.c:97-144); here a foreign or stale pixelization is
    an error instead of an out-of-bounds access.  Each (tensor, npix) pair is
    checked once: a later call with the same, unmodified tensor returns at
    once (the tensor's version counter changes on in-place writes)."""
    key = (id(ipix), int(npix))
    ent = _PIX_OK.get(key)
    if ent is not None and ent[0]() is ipix and ent[1] == ipix.data_ptr() and ent[2] == ipix._version:
        return
    n = ipix.numel()
    if n:
        nbad = torch.zeros(1, dtype=torch.int64, device=ipix.device)
        mkgpu.check("mk_pix_count_bad", dev.ptr(ipix, torch.int32, "ipix"), n, int(npix), nbad.data_ptr(),
                    dev.stream_ptr())
        nb = int(nbad.item())
        if nb:
            raise ValueError("%d of %d pixel indices fall outside a map of %d pixels" % (nb, n, npix))
    _PIX_OK[key] = (weakref.ref(ipix, lambda _r, k=key: _PIX_OK.pop(k, None)), ipix.data_ptr(), ipix._version)


def map2tod(dat, map, ipix, do_add=False, do_omp=True, check=True):
    """Unroll a map into a TOD in place: dat (+)= map.flat[ipix].
    check=False skips the index validation (callers that validated ipix
    against this map's size themselves)."""
    if dat.dtype != map.dtype:
        raise TypeError("dat (%s) and map (%s) must share a dtype" % (dat.dtype, map.dtype))
    name = {torch.float32: "mk_map2tod_f32", torch.float64: "mk_map2tod_f64"}[dat.dtype]
    if dat.numel() != ipix.numel():
        raise ValueError("dat has %d samples, ipix %d" % (dat.numel(), ipix.numel()))
    if check:
        check_pix(ipix, map.numel())
    mkgpu.check(name, dev.ptr(dat, name="dat"), dev.ptr(map, name="map"), dat.numel(),
                dev.ptr(ipix, torch.int32, "ipix"), int(bool(do_add)), dev.stream_ptr())


def polmap2tod(dat, map, poltag, twogamma, ipix, do_add=False, do_omp=True, npix=None):
    """Unroll a QU or IQU map into a TOD in place (mapmaking/map2tod.py:57-120):
    dat (+)= [I +] Q cos 2g + U sin 2g.  map is (nx, ny, npol); npix defaults
    to nx * ny.  Samples whose pixel falls outside the map raise."""
    npol = {"QU": 2, "IQU": 3}.get(poltag)
    if npol is None:
        raise ValueError("unrecognized poltag " + repr(poltag) + " in polmap2tod.")
    if dat.dtype != map.dtype or twogamma.dtype != map.dtype:
        raise TypeError("dat, map and twogamma must share a dtype")
    if dat.numel() != ipix.numel() or twogamma.numel() != dat.numel():
        raise ValueError("dat, twogamma and ipix differ in size")
    if npix is None:
        npix = map.numel() // npol
    if map.numel() < npix * npol:
        raise ValueError("map holds fewer than npix * npol values")
    name = {torch.float32: "mk_polmap2tod_f32", torch.float64: "mk_polmap2tod_f64"}[dat.dtype]
    nbad = torch.zeros(1, dtype=torch.int64, device=dat.device)
    mkgpu.check(name, dev.ptr(dat, name="dat"), dev.ptr(map, name="map"), dev.ptr(twogamma, name="twogamma"),
                dat.numel(), dev.ptr(ipix, torch.int32, "ipix"), npix, npol, int(bool(do_add)), nbad.data_ptr(),
                dev.stream_ptr())
    nb = int(nbad.item())
    if nb:
        raise ValueError("%d samples fall outside the map in polmap2tod" % nb)
