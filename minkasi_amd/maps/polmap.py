"""PolMap: a CAR map with several Stokes components per pixel
(minkasi/maps/polmap.py:47-724).

Geometry, WCS and pixelization are SkyMap's (the reference repeats the same
code, polmap.py:88-195 and :386-462).  With npol > 1 the map is an
(nx, ny, npol) device tensor, the reference's layout.  Its projections run
in libmkgpu:
  map2tod   mk_polmap2tod_*            (QU, IQU)
  tod2map   mk_t2m_plan_exec_pol_*     (QU, IQU, QU_PRECON, IQU_PRECON; deterministic)
  invert    mk_polmap_invert_*         (per-pixel pseudo-inverse)
  __mul__   mk_polmap_precon_mul_*     (preconditioner block times Stokes vector)
Each sample's polarisation angle comes from tod.info['twogamma_saved'], as
in the reference.  It is copied to the device once, in the map's precision,
and kept in tod.info.  With npol == 1 ('I') a PolMap behaves as a SkyMap.

Differences from the reference:
  * tod2map with npol > 1 bins each component through the TOD's
    deterministic tod2map plan (FP64 sums, fixed order); the reference sums
    serially, in sample order.  tod2map_method = "atomic" keeps the device
    atomics (order not fixed).
  * invert does not print the determinant range.
"""
import copy

import torch

from .. import device as dev
from ..lib import mkgpu
from .skymap import SkyMap


def _poltag2pols(poltag):
    """polmap.py:32-44."""
    return {"I": ["I"], "IQU": ["I", "Q", "U"], "QU": ["Q", "U"],
            "IQU_PRECON": ["I", "Q", "U", "QQ", "QU", "UU"], "QU_PRECON": ["QQ", "UU", "QU"]}.get(poltag)


_MODE = {"QU": mkgpu.POL_QU, "IQU": mkgpu.POL_IQU, "QU_PRECON": mkgpu.POL_QU_PRECON,
         "IQU_PRECON": mkgpu.POL_IQU_PRECON}


def _sfx(t):
    return {torch.float32: "f32", torch.float64: "f64"}[t.dtype]


class PolMap(SkyMap):
    def __init__(self, lims, pixsize, poltag="I", proj="CAR", pad=2, primes=None, cosdec=None, nx=None, ny=None,
                 mywcs=None, tag="ipix", purge_pixellization=False, ref_equ=False, dtype=torch.float32):
        pols = _poltag2pols(poltag)
        if pols is None:
            raise ValueError("Unrecognized polarization state " + poltag + " in PolMap.__init__")
        super().__init__(lims, pixsize, proj=proj, pad=pad, primes=primes, cosdec=cosdec, nx=nx, ny=ny,
                         mywcs=mywcs, tag=tag, purge_pixellization=purge_pixellization, ref_equ=ref_equ, dtype=dtype)
        self._set_pols(poltag, pols)

    def _set_pols(self, poltag, pols):
        self.npol = len(pols)
        self.poltag = poltag
        self.pols = pols
        shape = (self.nx, self.ny, self.npol) if self.npol > 1 else (self.nx, self.ny)
        self.map = torch.zeros(shape, dtype=self.map.dtype, device=self.map.device)

    def set_polstate(self, poltag):
        """New polarisation state and a zero map (polmap.py:273-298)."""
        pols = _poltag2pols(poltag)
        if pols is None:
            print("Unrecognized polarization state " + poltag + " in PolMap.set_polstate.")
            return
        self._set_pols(poltag, pols)

    def copy(self):
        return copy.deepcopy(self)

    def assign(self, arr):
        assert arr.shape[0] == self.nx
        assert arr.shape[1] == self.ny
        if self.npol > 1:
            assert arr.shape[2] == self.npol
        t = arr if isinstance(arr, torch.Tensor) else torch.as_tensor(arr)
        self.map.copy_(t.to(self.map.device, self.map.dtype))

    # ------------------------------------------------------------------ projections
    def _twogamma(self, tod):
        key = "_twogamma_" + _sfx(self.map)
        tg = tod.info.get(key)
        if tg is None:
            if "twogamma_saved" not in tod.info:
                raise KeyError("a polarised map needs tod.info['twogamma_saved']")
            tg = dev.to_device(tod.info["twogamma_saved"], self.map.dtype)
            tod.info[key] = tg
        return tg

    def _npix(self):
        return self.nx * self.ny

    def map2tod(self, tod, dat, do_add=True, do_omp=True):
        """polmap.py:464-502."""
        if self.npol == 1:
            return super().map2tod(tod, dat, do_add, do_omp)
        from ..mapmaking.map2tod import polmap2tod
        polmap2tod(dat, self.map, self.poltag, self._twogamma(tod), self.get_pix(tod), do_add, do_omp,
                   npix=self._npix())

    def tod2map(self, tod, dat=None, do_add=True, do_omp=True):
        """polmap.py:504-545."""
        if self.npol == 1:
            return super().tod2map(tod, dat, do_add, do_omp)
        from ..mapmaking.tod2map import tod2polmap
        if dat is None:
            dat = tod.get_data()
        if not do_add:
            self.clear()
        d = dev.to_device(dat, self.map.dtype)
        ipix = self.get_pix(tod)
        if self.tod2map_method == "atomic":
            tod2polmap(self.map, d, self.poltag, self._twogamma(tod), ipix, npix=self._npix())
        else:
            # deterministic: the TOD's cached tod2map plan, one pass per component
            self._t2m_plan(tod, ipix).exec_pol(self.map, d, self._twogamma(tod), _MODE[self.poltag], True)
        if self.purge_pixellization:
            tod.clear_saved_pix(self.tag)

    # ------------------------------------------------------------------ vector ops
    def invert(self, thresh=1e-6):
        """Per-pixel pseudo-inverse of the preconditioner blocks, or 1/x where
        x != 0 for a one-component map (polmap.py:300-384)."""
        if self.npol == 1:
            return super().invert()
        if self.poltag not in ("QU_PRECON", "IQU_PRECON"):
            return  # the reference leaves other states untouched
        work = torch.empty(1, dtype=torch.int64, device=self.map.device)
        mkgpu.check("mk_polmap_invert_" + _sfx(self.map), self.map.data_ptr(), self._npix(), _MODE[self.poltag],
                    float(thresh), work.data_ptr(), 8, dev.stream_ptr())
        torch.cuda.current_stream().synchronize()  # work must outlive the kernels

    def __mul__(self, map):
        """polmap.py:624-684: elementwise for one component, else this
        PRECON map's blocks times map's Stokes vectors."""
        if self.npol == 1:
            return super().__mul__(map)
        if map.poltag + "_PRECON" != self.poltag:
            raise ValueError("Tag from input map does not match this map")
        new_map = map.copy()
        x = map.map.to(self.map.dtype).contiguous()
        out = new_map.map if new_map.map.dtype == self.map.dtype else torch.empty_like(x)
        mkgpu.check("mk_polmap_precon_mul_" + _sfx(self.map), out.data_ptr(), self.map.data_ptr(), x.data_ptr(),
                    self._npix(), _MODE[self.poltag], dev.stream_ptr())
        if out is not new_map.map:
            new_map.map.copy_(out)
        return new_map

    # ------------------------------------------------------------------ output
    def write(self, fname="map.fits"):
        """polmap.py:585-622: one FITS file per component, named
        <head>_<pol>.fits, each holding that component transposed."""
        if self.npol == 1:
            return super().write(fname)
        ind = fname.rfind(".")
        if ind > 0 and fname[ind + 1:] == "fits":
            head, tail = fname[:ind], fname[ind:]
        else:
            head, tail = fname, ".fits"
        for i in range(self.npol):
            self._write_fits(head + "_" + self.pols[i] + tail, self.map[:, :, i])
