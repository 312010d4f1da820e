"""tod2map wrappers and make_hits (minkasi/mapmaking/tod2map.py:29-161,
343-400).  The deterministic plan lives on SkyMap.tod2map; these functions
keep the reference's free-function forms."""
import torch

from .. import device as dev
from ..lib import mkgpu
from ..lib.plans import T2MPlan


def tod2map_omp(map, dat, ipix, atomic=False):
    """map += tod2map(dat).  atomic=True: global-atomic scatter (order
    nondeterministic, like the reference); otherwise a one-shot deterministic
    plan (build cost included: use SkyMap.tod2map to reuse a plan)."""
    if ipix.dtype != torch.int32:
        print("Warning - ipix is not int32 in tod2map_omp.  this is likely to produce garbage results.")
    if atomic:
        from .map2tod import check_pix
        if dat.numel() != ipix.numel():
            raise ValueError("dat has %d samples, ipix %d" % (dat.numel(), ipix.numel()))
        check_pix(ipix, map.numel())
        name = {torch.float32: "mk_tod2map_atomic_f32", torch.float64: "mk_tod2map_atomic_f64"}[dat.dtype]
        mkgpu.check(name, dev.ptr(map, name="map"), dev.ptr(dat, name="dat"), dat.numel(),
                    dev.ptr(ipix, torch.int32, "ipix"), dev.stream_ptr())
        return
    nx, ny = map.shape if map.dim() == 2 else (map.numel(), 1)
    T2MPlan(ipix, nx, ny).exec(map, dat, True)


def tod2map_simple(map, dat, ipix):
    tod2map_omp(map, dat, ipix)


def tod2map_cached(map, dat, ipix):
    tod2map_omp(map, dat, ipix)


_POL_MODE = {"QU": mkgpu.POL_QU, "IQU": mkgpu.POL_IQU, "QU_PRECON": mkgpu.POL_QU_PRECON,
             "IQU_PRECON": mkgpu.POL_IQU_PRECON}
_POL_NCOMP = {"QU": 2, "IQU": 3, "QU_PRECON": 3, "IQU_PRECON": 6}


def tod2polmap(map, dat, poltag, twogamma, ipix, npix=None):
    """Bin a TOD into a polarised map (mapmaking/tod2map.py:164-215): map +=
    the per-sample terms of poltag (QU, IQU, QU_PRECON or IQU_PRECON), with
    device atomics.  Samples whose pixel falls outside the map raise."""
    mode = _POL_MODE.get(poltag)
    if mode is None:
        raise ValueError("unrecognized poltag " + repr(poltag) + " in tod2polmap.")
    nc = _POL_NCOMP[poltag]
    if dat.dtype != map.dtype or twogamma.dtype != map.dtype:
        raise TypeError("dat, map and twogamma must share a dtype")
    if dat.numel() != ipix.numel() or twogamma.numel() != dat.numel():
        raise ValueError("dat, twogamma and ipix differ in size")
    if npix is None:
        npix = map.numel() // nc
    if map.numel() < npix * nc:
        raise ValueError("map holds fewer than npix * ncomp values")
    name = {torch.float32: "mk_tod2polmap_atomic_f32", torch.float64: "mk_tod2polmap_atomic_f64"}[dat.dtype]
    nbad = torch.zeros(1, dtype=torch.int64, device=dat.device)
    mkgpu.check(name, dev.ptr(map, name="map"), dev.ptr(dat, name="dat"), dev.ptr(twogamma, name="twogamma"),
                dat.numel(), dev.ptr(ipix, torch.int32, "ipix"), npix, mode, nbad.data_ptr(), dev.stream_ptr())
    nb = int(nbad.item())
    if nb:
        raise ValueError("%d samples fall outside the map in tod2polmap" % nb)


def make_hits(todvec, map, do_weights=False):
    """Hit (or weight) map (tod2map.py:343-400).  For a polarised map the
    hits are the PRECON blocks (poltag + '_PRECON')."""
    import numpy as np
    hits = map.copy()
    if getattr(map, "npol", 1) > 1:
        hits.set_polstate(map.poltag + "_PRECON")
    hits.clear()
    if not do_weights and not any("mask" in t.info for t in todvec.tods):
        from ..maps.mapset import Mapset
        ms = Mapset()
        ms.maps.append(hits)
        ms.nmap = 1
        b = todvec.batched(ms)
        if b is not None:
            b.hits(hits.map, accumulate=True)
            return hits
    for tod in todvec.tods:
        shape = tod.get_data_dims()
        if do_weights:
            weights = tod.get_det_weights()
            if weights is None:
                print("error in making weight map.  Detector weights requested, but do not appear to be present.  Do you have a noise model?")
                tmp = torch.ones(shape, dtype=hits.map.dtype, device=hits.map.device)
            else:
                w = torch.as_tensor(np.asarray(weights), dtype=hits.map.dtype, device=hits.map.device)
                tmp = w[:, None].expand(shape).contiguous()
        else:
            tmp = torch.ones(shape, dtype=hits.map.dtype, device=hits.map.device)
        if "mask" in tod.info:
            tmp = tmp * dev.to_device(tod.info["mask"], hits.map.dtype)
        hits.tod2map(tod, tmp)
    return hits
