"""SkyMap with the reference's interface (minkasi/maps/skymap.py:38-825),
the map held as a device tensor.

Pixelization (pix_from_radec/get_pix), map2tod and tod2map run in libmkgpu.
Vector operations (copy, axpy, dot, __mul__, invert, clear) run on the
device; dot accumulates in FP64.  tod2map uses a deterministic plan cached
next to the TOD's pixelization, so repeated calls are bitwise identical
(the reference's tod2map_omp sums thread maps in arrival order).
"""
import copy
import os
import struct

import numpy as np
import torch

from .. import device as dev
from ..lib import mkgpu
from ..lib.plans import T2MPlan
from .wcs import CarWCS, get_wcs


def find_good_fft_lens(n, primes=(2, 3, 5, 7)):
    """All integers <= n whose prime factors are in `primes`, sorted
    (minkasi/tools/py_fft.py:364-395)."""
    vals = [1]
    for p in primes:
        for v in list(vals):
            x = v * p
            while x <= n:
                vals.append(x)
                x *= p
    return np.asarray(sorted(set(vals)), dtype=int)


class SkyMap:
    def __init__(self, lims, pixsize, proj="CAR", pad=2, square=False, multiple=False, primes=None, cosdec=None,
                 nx=None, ny=None, mywcs=None, tag="ipix", purge_pixellization=False, ref_equ=False,
                 dtype=torch.float32):
        if mywcs is None:
            assert pixsize != 0
            self.wcs = get_wcs(lims, pixsize, proj, cosdec, ref_equ)
        else:
            self.wcs = mywcs
            pixsize = mywcs.cdelt[1] * np.pi / 180
        self.lims = lims
        nx, ny = self.get_npix(pad, nx=nx, ny=ny)
        if primes is not None:
            # the reference reads self.nx before it exists here (skymap.py:150);
            # this applies the intent: round each side up to a primes-smooth length
            lens = find_good_fft_lens(2 * (nx + ny), primes)
            nx = int(lens[lens >= nx].min())
            ny = int(lens[lens >= ny].min())
            self.primes = list(primes)
        else:
            self.primes = None
        if square:
            if nx != ny:
                nmax = max(nx, ny)
                ratio_x = nmax / nx
                ratio_y = nmax / ny
                self.lims[0] = self.lims[1] - ratio_x * (self.lims[1] - self.lims[0])
                self.lims[3] = self.lims[2] + ratio_y * (self.lims[3] - self.lims[2])
                nx, ny = self.get_npix(pad)
        if multiple:
            assert type(multiple) == int
            xmax = 2 * np.ceil(nx / multiple)
            xdiff = self.lims[1] - self.lims[0]
            self.lims[0] = self.lims[1] - xdiff * xmax / nx
            ymax = 2 * np.ceil(ny / multiple)
            ydiff = self.lims[3] - self.lims[2]
            self.lims[3] = self.lims[2] + ydiff * ymax / ny
            nx, ny = self.get_npix(pad)
        self.nx = int(nx)
        self.ny = int(ny)
        self.caches = None
        self.cosdec = cosdec
        self.tod2map_method = None
        self.tag = tag
        self.purge_pixellization = purge_pixellization
        self.pixsize = pixsize
        self.proj = proj
        self.pad = pad
        self.map = torch.zeros((self.nx, self.ny), dtype=dtype, device=dev.device())

    # ------------------------------------------------------------------ geometry
    def get_npix(self, pad, nx=None, ny=None):
        """skymap.py:206-234."""
        corners = np.zeros([4, 2])
        corners[0, :] = [self.lims[0], self.lims[2]]
        corners[1, :] = [self.lims[0], self.lims[3]]
        corners[2, :] = [self.lims[1], self.lims[2]]
        corners[3, :] = [self.lims[1], self.lims[3]]
        px, py = self.wcs.world2pix(corners[:, 0], corners[:, 1])
        px, py = np.round(px), np.round(py)
        if min(px.min(), py.min()) < -0.5:
            print("corners seem to have gone negative in SkyMap projection.  not good, you may want to check this.")
        if nx is None:
            nx = px.max() + pad
        if ny is None:
            ny = py.max() + pad
        return int(nx), int(ny)

    def get_caches(self):
        """The deterministic tod2map needs no per-thread caches; kept so
        Mapset.get_caches / TodVec.dot(cache_maps=True) work unchanged."""
        self.caches = None

    def clear_caches(self):
        self.caches = None

    # The reference's per-method tod2map wrappers (skymap.py:310-397, 611-640).
    # Its cached / everyone variants reduce per-thread map copies; here every
    # one of them accumulates into self.map through the deterministic plan,
    # except tod2map_atomic, which uses the global-atomic scatter.
    def tod2map_atomic(self, tod, dat):
        """skymap.py:310-323 (atomic adds, order-dependent in FP32)."""
        from ..mapmaking.tod2map import tod2map_omp
        tod2map_omp(self.map, dev.to_device(dat, self.map.dtype), self.get_pix(tod), atomic=True)

    def tod2map_omp(self, tod, dat):
        """skymap.py:325-338."""
        self._t2m_plan(tod, self.get_pix(tod)).exec(self.map, dev.to_device(dat, self.map.dtype), True)

    def tod2map_simple(self, tod, dat):
        """skymap.py:340-353."""
        self.tod2map_omp(tod, dat)

    def tod2map_cached(self, tod, dat):
        """skymap.py:379-397 (no per-thread caches are needed)."""
        self.tod2map_omp(tod, dat)

    def tod2map_everyone(self, tod, dat):
        """skymap.py:355-377: the reference needs tod.info[tag + '_edges'],
        the per-thread pixel ranges of its cached reduction; the plan needs
        none, so they are not required."""
        self.tod2map_omp(tod, dat)

    def tod2map_old(self, tod, dat=None, do_add=True, do_omp=True):
        """skymap.py:611-640."""
        self.tod2map(tod, dat, do_add, do_omp)

    def set_tod2map(self, method=None, todvec=None):
        """skymap.py:254-308: stored, and honoured: 'atomic' selects the
        non-deterministic global-atomic scatter, anything else the plan."""
        self.tod2map_method = method

    # ------------------------------------------------------------------ vector ops
    def copy(self):
        return copy.deepcopy(self)

    def __deepcopy__(self, memo):
        new = self.__class__.__new__(self.__class__)
        for k, v in self.__dict__.items():
            if k == "map":
                new.map = v.clone()
            else:
                new.__dict__[k] = copy.deepcopy(v, memo)
        return new

    def clear(self):
        self.map.zero_()

    def axpy(self, map, a):
        """self.map += a * map.map (skymap.py:428-440)."""
        sfx = {torch.float32: "f32", torch.float64: "f64"}[self.map.dtype]
        x = map.map.to(self.map.dtype).contiguous()
        mkgpu.check("mk_axpy_" + sfx, self.map.data_ptr(), x.data_ptr(), float(a), self.map.numel(),
                    dev.stream_ptr())

    def assign(self, arr):
        assert arr.shape[0] == self.nx
        assert arr.shape[1] == self.ny
        t = arr if isinstance(arr, torch.Tensor) else torch.from_numpy(np.asarray(arr))
        self.map.copy_(t.to(self.map.device, self.map.dtype))

    def dot(self, map):
        """float(sum(a*b)) with FP64 accumulation (skymap.py:651-666)."""
        S = dev.scratch()
        if self.map.dtype == torch.float32 and map.map.dtype == torch.float32:
            mkgpu.check("mk_dot_f32", self.map.data_ptr(), map.map.data_ptr(), self.map.numel(), S.sc.data_ptr(),
                        S.work.data_ptr(), dev.stream_ptr())
        else:
            a = self.map.to(torch.float64).contiguous()
            b = map.map.to(torch.float64).contiguous()
            mkgpu.check("mk_dot_f64", a.data_ptr(), b.data_ptr(), a.numel(), S.sc.data_ptr(), S.work.data_ptr(),
                        dev.stream_ptr())
        return float(S.sc[mkgpu.SC_DOT].item())

    def __mul__(self, map):
        """skymap.py:724-740: elementwise product in a copy of `map`."""
        new_map = map.copy()
        sfx = {torch.float32: "f32", torch.float64: "f64"}[new_map.map.dtype]
        a = self.map.to(new_map.map.dtype).contiguous()
        mkgpu.check("mk_mul_" + sfx, new_map.map.data_ptr(), a.data_ptr(), map.map.data_ptr(),
                    self.map.numel(), dev.stream_ptr())
        return new_map

    def invert(self):
        """skymap.py:785-790."""
        mask = torch.abs(self.map) > 0
        self.map[mask] = 1.0 / self.map[mask]

    def mpi_reduce(self, chunksize=1e5):
        """One GPU holds every TOD: nothing to reduce (skymap.py:742-781)."""
        return None

    def r_th_maps(self):
        xvec = np.arange(self.nx)
        xvec = xvec - xvec.mean()
        yvec = np.arange(self.ny)
        yvec = yvec - yvec.mean()
        ymat, xmat = np.meshgrid(yvec, xvec)
        return np.sqrt(xmat ** 2 + ymat ** 2), np.arctan2(xmat, ymat)

    # ------------------------------------------------------------------ pointing
    def pix_from_radec(self, ra, dec):
        """skymap.py:457-490 on the device.  Raises if any sample falls off
        the map (the reference has no bounds check and would corrupt
        memory in tod2map)."""
        rat = dev.to_device(ra, torch.float64)
        dect = dev.to_device(dec, torch.float64)
        if rat.shape != dect.shape:
            raise ValueError("ra and dec shapes differ")
        ipix = torch.empty(rat.shape, dtype=torch.int32, device=rat.device)
        nbad = torch.zeros(1, dtype=torch.int64, device=rat.device)
        prm = self.wcs.params10()
        mkgpu.check("mk_pix_from_radec_car", rat.data_ptr(), dect.data_ptr(), rat.numel(), prm.ctypes.data,
                    self.nx, self.ny, ipix.data_ptr(), nbad.data_ptr(), dev.stream_ptr())
        nb = int(nbad.item())
        if nb:
            raise ValueError("%d samples fall outside the %d x %d map" % (nb, self.nx, self.ny))
        return ipix

    def pix_from_raster(self, ra_b, dec_b, dra, ddec):
        """Pixelization of ra[d,t] = ra_b[t] + dra[d], dec[d,t] = dec_b[t] +
        ddec[d] without materialising the FP64 pointing (same per-sample
        arithmetic as pix_from_radec)."""
        t = [dev.to_device(a, torch.float64) for a in (ra_b, dec_b, dra, ddec)]
        ndet, n = t[2].numel(), t[0].numel()
        ipix = torch.empty((ndet, n), dtype=torch.int32, device=t[0].device)
        nbad = torch.zeros(1, dtype=torch.int64, device=t[0].device)
        prm = self.wcs.params10()
        mkgpu.check("mk_pix_from_raster_car", t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), t[3].data_ptr(),
                    ndet, n, prm.ctypes.data, self.nx, self.ny, ipix.data_ptr(), nbad.data_ptr(), dev.stream_ptr())
        nb = int(nbad.item())
        if nb:
            raise ValueError("%d samples fall outside the %d x %d map" % (nb, self.nx, self.ny))
        return ipix

    def get_pix(self, tod, savepix=True):
        """skymap.py:492-540 (cached in tod.info[self.tag])."""
        if self.tag is not None:
            ipix = tod.get_saved_pix(self.tag)
            if ipix is not None:
                if not isinstance(ipix, torch.Tensor):
                    ipix = dev.to_device(np.asarray(ipix, dtype=np.int32))
                    tod.info[self.tag] = ipix
                # a saved pixelization may belong to another map (the default
                # tag is shared) or be user-supplied: checked once per tensor
                from ..mapmaking.map2tod import check_pix
                check_pix(ipix, self.nx * self.ny)
                return ipix
        if "dx" not in tod.info and "ra_b" in tod.info:
            ipix = self.pix_from_raster(tod.info["ra_b"], tod.info["dec_b"], tod.info["dra"], tod.info["ddec"])
        else:
            ra, dec = tod.get_radec()
            ipix = self.pix_from_radec(ra, dec)
        if savepix and self.tag is not None:
            tod.save_pixellization(self.tag, ipix)
        return ipix

    def _t2m_plan(self, tod, ipix):
        key = "_t2m_" + str(self.tag)
        plan = tod.info.get(key)
        if plan is None or plan.nx != self.nx or plan.ny != self.ny or plan.nsamp != ipix.numel():
            plan = T2MPlan(ipix, self.nx, self.ny)
            tod.info[key] = plan
        return plan

    def map2tod(self, tod, dat, do_add=True, do_omp=True):
        """skymap.py:542-567 -> map2tod (mapmaking/map2tod.py:19-54)."""
        from ..mapmaking.map2tod import map2tod
        ipix = self.get_pix(tod)
        map2tod(dat, self.map, ipix, do_add, do_omp)

    def tod2map(self, tod, dat=None, do_add=True, do_omp=True):
        """skymap.py:569-609."""
        if dat is None:
            dat = tod.get_data()
        if not do_add:
            self.clear()
        ipix = self.get_pix(tod)
        dat = dev.to_device(dat, self.map.dtype)
        if self.tod2map_method == "atomic":
            from ..mapmaking.tod2map import tod2map_omp
            tod2map_omp(self.map, dat, ipix, atomic=True)
        else:
            self._t2m_plan(tod, ipix).exec(self.map, dat, True)
        if self.purge_pixellization:
            tod.clear_saved_pix(self.tag)

    # ------------------------------------------------------------------ output
    def write(self, fname="map.fits"):
        """skymap.py:705-722: FITS primary HDU of map.T with the WCS cards."""
        self._write_fits(fname, self.map)

    def _write_fits(self, fname, plane):
        write_map_fits(fname, plane.to(torch.float64).cpu().numpy(), self.wcs)

    def plot(self, plot_info=None):
        raise NotImplementedError("plotting is outside the mapmaking inner loop")


def write_map_fits(fname, m, wcs):
    """A FITS primary HDU holding the (nx, ny) map m transposed, as FP64
    (BITPIX -64, big-endian), with the WCS cards (skymap.py:705-722 writes
    fits.PrimaryHDU(map.T, header=wcs.to_header()))."""
    data = np.asarray(m, dtype=np.float64).T.copy()
    cards = [("SIMPLE", True), ("BITPIX", -64), ("NAXIS", 2), ("NAXIS1", data.shape[1]),
             ("NAXIS2", data.shape[0])] + wcs.to_header()
    fname = os.path.abspath(fname)
    os.makedirs(os.path.dirname(fname), exist_ok=True)
    with open(fname, "wb") as f:
        f.write(_fits_header(cards))
        raw = data.astype(">f8").tobytes()
        f.write(raw)
        pad = (-len(raw)) % 2880
        f.write(b"\0" * pad)


def _fits_header(cards):
    out = []
    for k, v in cards:
        if isinstance(v, bool):
            val = "T" if v else "F"
            s = "%-8s= %20s" % (k, val)
        elif isinstance(v, int):
            s = "%-8s= %20d" % (k, v)
        elif isinstance(v, float):
            s = "%-8s= %20s" % (k, repr(float(v)).upper() if "e" in repr(float(v)) else repr(float(v)))
        else:
            s = "%-8s= '%-8s'" % (k, v)
        out.append(s.ljust(80)[:80])
    out.append("END".ljust(80))
    blob = "".join(out).encode("ascii")
    return blob + b" " * ((-len(blob)) % 2880)


def read_fits_map(fname):
    """Reads back what SkyMap.write wrote: (map (nx, ny), header dict)."""
    with open(fname, "rb") as f:
        raw = f.read()
    hdr = {}
    pos = 0
    while True:
        card = raw[pos:pos + 80].decode("ascii")
        pos += 80
        if card.startswith("END"):
            break
        if "=" in card[:10]:
            k, v = card[:8].strip(), card[10:].strip()
            hdr[k] = v.strip("'").strip()
    pos = ((pos + 2879) // 2880) * 2880
    n1, n2 = int(hdr["NAXIS1"]), int(hdr["NAXIS2"])
    data = np.frombuffer(raw[pos:pos + 8 * n1 * n2], dtype=">f8").reshape(n2, n1)
    return data.T.astype(np.float64), hdr
