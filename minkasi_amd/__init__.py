"""minkasi_amd -- MI355X-native mapmaking inner loop with minkasi's API.

Public names mirror the reference package (minkasi/__init__.py,
minkasi/minkasi_all.py): Tod, TodVec, SkyMap, Mapset, NoiseSmoothedSVD,
run_pcg, run_pcg_wprior, make_hits.  Device work runs in libmkgpu (HIP,
gfx950); importing the compute modules raises if that library is missing.
"""
__version__ = "0.1.0"


def __getattr__(name):
    # lazy imports keep `import minkasi_amd` cheap and CPU-safe
    if name in ("SkyMap",):
        from .maps.skymap import SkyMap
        return SkyMap
    if name in ("Mapset",):
        from .maps.mapset import Mapset
        return Mapset
    if name in ("Tod", "TodVec"):
        from .tods import core
        return getattr(core, name)
    if name in ("NoiseSmoothedSVD",):
        from .mapmaking.noise import NoiseSmoothedSVD
        return NoiseSmoothedSVD
    if name in ("run_pcg", "run_pcg_wprior", "null_precon", "HasPrior"):
        from .mapmaking import pcg
        return getattr(pcg, name)
    if name in ("make_hits",):
        from .mapmaking.tod2map import make_hits
        return make_hits
    raise AttributeError(name)
