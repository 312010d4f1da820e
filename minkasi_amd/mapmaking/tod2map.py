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


def make_hits(todvec, map, do_weights=False):
    """Hit (or weight) map (tod2map.py:343-400)."""
    import numpy as np
    hits = map.copy()
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
