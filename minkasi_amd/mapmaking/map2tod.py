"""map2tod wrapper (minkasi/mapmaking/map2tod.py:19-54) over mk_map2tod_*."""
import torch

from .. import device as dev
from ..lib import mkgpu


def map2tod(dat, map, ipix, do_add=False, do_omp=True):
    """Unroll a map into a TOD in place: dat (+)= map.flat[ipix]."""
    if dat.dtype != map.dtype:
        raise TypeError("dat (%s) and map (%s) must share a dtype" % (dat.dtype, map.dtype))
    name = {torch.float32: "mk_map2tod_f32", torch.float64: "mk_map2tod_f64"}[dat.dtype]
    if dat.numel() != ipix.numel():
        raise ValueError("dat has %d samples, ipix %d" % (dat.numel(), ipix.numel()))
    mkgpu.check(name, dev.ptr(dat, name="dat"), dev.ptr(map, name="map"), dat.numel(),
                dev.ptr(ipix, torch.int32, "ipix"), int(bool(do_add)), dev.stream_ptr())
