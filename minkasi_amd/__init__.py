"""minkasi_amd -- MI355X-native mapmaking inner loop with minkasi's API.

Public names mirror the reference package (minkasi/__init__.py,
minkasi/minkasi_all.py): Tod, TodVec, SkyMap, PolMap, Mapset,
NoiseSmoothedSVD, run_pcg, run_pcg_wprior, make_hits, Cuts, CutsCompact,
tsModel, tsMultiModel, truncate_tod, downsample_tod, fit_cm_plus_poly,
gapfill_lin, read_tod_from_fits (and the other TOD readers).
Device work runs in libmkgpu (HIP, gfx950); importing the compute modules
raises if that library is missing.
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
    if name in ("PolMap",):
        from .maps.polmap import PolMap
        return PolMap
    if name in ("Cuts", "CutsCompact", "segs_from_vec"):
        from .tods import cuts
        return getattr(cuts, name)
    if name in ("tsModel", "tsMultiModel"):
        from .mapmaking import timestream
        return getattr(timestream, name)
    if name in ("myrank", "nproc", "comm", "have_mpi", "MPI", "barrier", "report_mpi", "set_nthread", "get_nthread"):
        from . import parallel
        return getattr(parallel, name)
    if name in ("truncate_tod", "downsample_tod", "fit_cm_plus_poly", "find_good_fft_lens", "gapfill_lin",
                "gapfill_eig"):
        from .tods import processing
        return getattr(processing, name)
    if name in ("read_tod_from_fits", "read_tod_from_fits_cbass", "read_tod_from_fits_NIKA2", "read_octave_struct",
                "todvec_from_files_octave", "cut_blacklist", "fits_open"):
        from .tods import io
        return getattr(io, name)
    raise AttributeError(name)
