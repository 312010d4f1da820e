"""Every public name in one flat namespace, as minkasi/minkasi_all.py gives
legacy scripts (`import minkasi_amd.minkasi_all as minkasi`)."""
from .parallel import *  # noqa: F401,F403
from .maps.skymap import SkyMap  # noqa: F401
from .maps.polmap import PolMap  # noqa: F401
from .maps.mapset import Mapset  # noqa: F401
from .maps.wcs import get_wcs  # noqa: F401
from .mapmaking.noise import NoiseSmoothedSVD  # noqa: F401
from .mapmaking.pcg import run_pcg, run_pcg_wprior  # noqa: F401
from .mapmaking.map2tod import map2tod, polmap2tod  # noqa: F401
from .mapmaking.tod2map import make_hits, tod2map_cached, tod2map_omp, tod2map_simple, tod2polmap  # noqa: F401
from .mapmaking.timestream import tsModel, tsMultiModel  # noqa: F401
from .tods.core import Tod, TodVec  # noqa: F401
from .tods.cuts import Cuts, CutsCompact, segs_from_vec  # noqa: F401
from .tods.processing import (downsample_tod, find_good_fft_lens, fit_cm_plus_poly, gapfill_eig,  # noqa: F401
                              gapfill_lin, truncate_tod)
from .tods.io import (cut_blacklist, read_octave_struct, read_tod_from_fits, read_tod_from_fits_cbass,  # noqa: F401
                      read_tod_from_fits_NIKA2, todvec_from_files_octave)
from .tools.array_ops import downsample_array_r2r, downsample_vec_r2r  # noqa: F401
from .tools.fft import fft_r2r, read_wisdom, write_wisdom  # noqa: F401
