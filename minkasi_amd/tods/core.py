"""Tod and TodVec with the reference's interface (minkasi/tods/core.py:30-1047).

TOD arrays in ``info`` may be numpy arrays or device tensors; projections,
noise and dot products run on the device.  ``TodVec.dot`` sends every TOD
through one batched operator (mapmaking/batch.py: one launch per stage over
all TODs) when the mapset is a single FP32 SkyMap and every TOD has a
NoiseSmoothedSVD; otherwise it follows the reference's per-TOD loop
(core.py:1012-1020) with device kernels.
"""
import copy
import time

import numpy as np
import torch

from .. import device as dev
from ..mapmaking.noise import NoiseSmoothedSVD, WithDetWeights


def _torch_dtype(a):
    if isinstance(a, torch.Tensor):
        return a.dtype
    return {np.dtype(np.float32): torch.float32}.get(np.asarray(a).dtype, torch.float64)


class Tod:
    def __init__(self, info):
        self.info = info.copy()
        self.jumps = None
        self.cuts = None
        self.noise = None
        self.noise_delayed = False
        self.noise_args = ()
        self.noise_kwargs = {}

    # ---------------------------------------------------------------- shape
    def lims(self):
        if "dx" not in self.info and "ra_b" in self.info:
            # separable raster: min/max of a sum are the sums of min/max
            # (rounding is monotone), so this equals the materialised result
            i = self.info
            return [float(np.min(i["ra_b"]) + np.min(i["dra"])), float(np.max(i["ra_b"]) + np.max(i["dra"])),
                    float(np.min(i["dec_b"]) + np.min(i["ddec"])), float(np.max(i["dec_b"]) + np.max(i["ddec"]))]
        dx, dy = self.info["dx"], self.info["dy"]
        f = (lambda a: float(a.min()), lambda a: float(a.max()))
        return [f[0](dx), f[1](dx), f[0](dy), f[1](dy)]

    def get_ndet(self):
        return int(self.info["dat_calib"].shape[0])

    def get_ndata(self):
        return int(self.info["dat_calib"].shape[1])

    def get_nsamp(self):
        return self.get_ndet() * self.get_ndata()

    def get_data_dims(self):
        return (self.get_ndet(), self.get_ndata())

    # ---------------------------------------------------------------- pixelization
    def get_saved_pix(self, tag="ipix"):
        return self.info.get(tag, None)

    def clear_saved_pix(self, tag="ipix"):
        if tag in self.info.keys():
            del self.info[tag]

    def save_pixellization(self, tag, ipix):
        if tag in self.info.keys():
            print("warning - overwriting key ", tag, " in tod.save_pixellization.")
        self.info[tag] = ipix

    def set_tag(self, tag):
        if "tag" in self.info:
            print("Warning: overwriting tag")
        self.info["tag"] = tag

    def set_pix(self, map):
        ipix = map.get_pix(self)
        self.save_pixellization(map.tag, ipix)

    # ---------------------------------------------------------------- data
    def get_data(self, dat=None):
        if dat is None:
            dat = self.info.get("dat_calib", None)
        if dat is None:
            raise ValueError("This TOD has no stored data. Populare tod.info['dat_calib']")
        return dat

    def get_tvec(self):
        return self.info["ctime"]

    def get_radec(self):
        return self.info["dx"], self.info["dy"]

    def get_empty(self, clear=False):
        """Device tensor (ndet, ndata) in the dtype of info['dtype'] or of
        dat_calib (core.py:286-311)."""
        if "dtype" in self.info.keys():
            dtype = _torch_dtype(np.zeros(1, dtype=self.info["dtype"]))
        elif "dat_calib" in self.info.keys():
            dtype = _torch_dtype(self.info["dat_calib"])
        else:
            dtype = torch.float64
        f = torch.zeros if clear else torch.empty
        return f((self.get_ndet(), self.get_ndata()), dtype=dtype, device=dev.device())

    def set_noise_smoothed_svd(self, fwhm=50, func=None, pars=None, prewhiten=False, fit_powlaw=False):
        """The deprecated spelling of set_noise(NoiseSmoothedSVD)
        (core.py:503-515 -> _depracated.py:77-93): the model is fitted to
        dat_calib, or to dat_calib - func(pars, tod) when func is given."""
        dat = self.get_data()
        if func is not None:
            model = func(pars, self)
            dat = dat - (dev.to_device(model, dat.dtype) if isinstance(dat, torch.Tensor) else model)
        self.set_noise(NoiseSmoothedSVD, dat, fwhm=fwhm, prewhiten=prewhiten, fit_powlaw=fit_powlaw)

    def set_noise_white_masked(self):
        """core.py:411-416: white noise of unit weight per detector."""
        self.info["noise"] = "white_masked"
        self.info["mywt"] = np.ones(self.get_ndet())

    def apply_noise_white_masked(self, dat=None):
        """mask * dat * mywt per detector (core.py:418-445), on the device
        (elementwise products in the library); a new tensor of dat's dtype."""
        from ..lib import mkgpu
        d = dev.to_device(self.get_data(dat))
        if d.dtype not in (torch.float32, torch.float64):
            d = d.to(torch.float64)
        d = d.contiguous()
        sfx = "f32" if d.dtype == torch.float32 else "f64"
        w = dev.to_device(np.asarray(self.info["mywt"], dtype=np.float64), d.dtype).reshape(-1, 1)
        w = w.expand(d.shape).contiguous()
        m = dev.to_device(self.info["mask"], d.dtype).expand(d.shape).contiguous()
        out = torch.empty_like(d)
        mkgpu.check("mk_mul_" + sfx, out.data_ptr(), m.data_ptr(), d.data_ptr(), d.numel(), dev.stream_ptr())
        mkgpu.check("mk_mul_" + sfx, out.data_ptr(), out.data_ptr(), w.data_ptr(), d.numel(), dev.stream_ptr())
        return out

    def set_jumps(self, jumps):
        """core.py:693-703."""
        self.jumps = jumps

    def cut_detectors(self, isgood):
        """Keeps only the detectors where isgood is True (core.py:705-728):
        every numpy array or device tensor in info whose first axis has one
        entry per detector is sliced (a copy), and the per-detector jump and
        cut lists drop the removed detectors.  Arrays with another first
        axis are left alone, where the reference would index them by
        detector number too.  The noise model is not rebuilt."""
        isgood = np.asarray(isgood.cpu() if isinstance(isgood, torch.Tensor) else isgood).astype(bool)
        ndet = self.get_ndet()
        if isgood.shape != (ndet,):
            raise ValueError("isgood has %s entries, the TOD %d detectors" % (isgood.shape, ndet))
        keep = np.flatnonzero(isgood)
        bad = np.flatnonzero(~isgood)[::-1]  # deleted in reverse, as the reference
        for key in list(self.info.keys()):
            v = self.info[key]
            if isinstance(v, np.ndarray) and v.ndim >= 1 and v.shape[0] == ndet:
                self.info[key] = np.take(v, keep, axis=0).copy()
            elif isinstance(v, torch.Tensor) and v.dim() >= 1 and v.shape[0] == ndet:
                self.info[key] = v.index_select(0, torch.from_numpy(keep).to(v.device)).contiguous()
        if self.jumps is not None:
            for i in bad:
                del self.jumps[i]
        if self.cuts is not None:
            for i in bad:
                del self.cuts[i]

    def set_apix(self):
        """info['apix'] (core.py:102-116): the detectors' elevation offsets
        from a quadratic fit to the mean elevation track, scaled to +-1."""
        elev = self.info["elev"]
        elev = elev.detach().cpu().numpy() if isinstance(elev, torch.Tensor) else np.asarray(elev)
        track_mean = np.mean(elev, axis=0)
        x = np.arange(track_mean.shape[0]) / track_mean.shape[0]
        a = np.polyfit(x, track_mean, 2)
        track = a[2] + a[1] * x + a[0] * x ** 2
        delev = elev - track[None, :]
        self.info["apix"] = delev / np.max(np.abs(delev))

    def copy(self, copy_info=False):
        if copy_info:
            myinfo = self.info.copy()
            for key in myinfo.keys():
                try:
                    myinfo[key] = self.info[key].clone() if isinstance(self.info[key], torch.Tensor) \
                        else self.info[key].copy()
                except Exception:
                    pass
            tod = Tod(myinfo)
        else:
            tod = Tod(self.info)
        if self.jumps is not None:
            tod.jumps = list(self.jumps)
        if self.cuts is not None:
            tod.cuts = list(self.cuts)
        tod.noise = self.noise
        tod.noise_delayed = self.noise_delayed
        tod.noise_args, tod.noise_kwargs = self.noise_args, self.noise_kwargs
        if hasattr(self, "noise_modelclass"):
            tod.noise_modelclass = self.noise_modelclass
        return tod

    # ---------------------------------------------------------------- noise
    def set_noise(self, modelclass=NoiseSmoothedSVD, dat=None, delayed=False, *args, **kwargs):
        if delayed:
            self.noise_args = copy.deepcopy(args)
            self.noise_kwargs = copy.deepcopy(kwargs)
            self.noise_delayed = True
            self.noise_modelclass = modelclass
        else:
            self.noise_delayed = False
            dat = self.get_data(dat)
            self.noise = modelclass(dat, *args, **kwargs)

    def get_det_weights(self):
        if self.noise is None:
            print("noise model not set in get_det_weights.")
            return None
        if isinstance(self.noise, WithDetWeights):
            return self.noise.get_det_weights()
        print("noise model does not support detector weights in get_det_weights.")
        return None

    def apply_noise(self, dat=None):
        """core.py:517-552.  The reference silently falls back to a hard-wired
        model when the noise object raises; here the error propagates."""
        dat = self.get_data(dat)
        if self.noise_delayed:
            self.noise = self.noise_modelclass(dat, *(self.noise_args), **(self.noise_kwargs))
            self.noise_delayed = False
        if self.noise is None:
            raise RuntimeError("tod.noise is not set; call set_noise first")
        return self.noise.apply_noise(dat)

    # ---------------------------------------------------------------- projections
    def mapset2tod(self, mapset, dat=None):
        if dat is None:
            dat = self.get_empty(True)
        for map in mapset.maps:
            map.map2tod(self, dat)
        return dat

    def tod2mapset(self, mapset, dat=None):
        if dat is None:
            dat = self.get_data()
        for map in mapset.maps:
            map.tod2map(self, dat)

    def dot(self, mapset, mapset_out, times=False, mask=None):
        """core.py:649-691 (mask applied in place to the input, as the reference)."""
        t1 = time.time()
        if mask is not None:
            for i, imap in enumerate(mapset.maps):
                imap.map *= mask.maps[i].map
        tmp = self.mapset2tod(mapset)
        t2 = time.time()
        tmp = self.apply_noise(tmp)
        t3 = time.time()
        self.tod2mapset(mapset_out, tmp)
        if mask is not None:
            for i, imap in enumerate(mapset_out.maps):
                imap.map *= mask.maps[i].map
        t4 = time.time()
        if times:
            return np.asarray([t2 - t1, t3 - t2, t4 - t3])

    def iter_init(self):
        return

    def iter_finalize(self):
        return

    def timestream_chisq(self, dat=None):
        dat = self.get_data(dat)
        d = dev.to_device(dat)
        filt = self.apply_noise(d)
        return float((filt.to(torch.float64) * d.to(torch.float64)).sum().item())

    def prior_from_skymap(self, skymap):
        """A CutsCompact prior on the samples where skymap is non-zero, with
        weight 1/value^2 (core.py:751-782)."""
        from .cuts import CutsCompact
        tmp = self.get_empty(True).to(skymap.map.dtype)
        skymap.map2tod(self, tmp)
        mask = tmp == 0
        prior = CutsCompact(self, dtype=skymap.map.dtype)
        prior.cuts_from_array(mask)
        prior.get_imap()
        prior.tod2map(self, tmp)
        if prior.map is None:
            raise ValueError("Prior map is None, something went wrong")
        prior.map = 1.0 / prior.map ** 2
        return prior


class TodVec:
    def __init__(self):
        self.tods = []
        self.ntod = 0
        self._batch = None

    def add_tod(self, tod, copy_info=False):
        self.tods.append(tod.copy(copy_info))
        self.tods[-1].set_tag(self.ntod)
        self.ntod = self.ntod + 1
        self._batch = None
        self._batch_cuts = None

    def lims(self):
        if self.ntod == 0:
            return None
        xmin, xmax, ymin, ymax = self.tods[0].lims()
        for i in range(1, self.ntod):
            x1, x2, y1, y2 = self.tods[i].lims()
            xmin = min(x1, xmin)
            xmax = max(x2, xmax)
            ymin = min(y1, ymin)
            ymax = max(y2, ymax)
        return [xmin, xmax, ymin, ymax]

    def get_nsamp(self, reduce=True):
        tot = 0
        for tod in self.tods:
            tot = tot + tod.get_nsamp()
        return tot

    def set_pix(self, map):
        for tod in self.tods:
            tod.set_pix(map)

    # ---------------------------------------------------------------- A.x
    def batched(self, mapset, **kw):
        """The batched operator for this TodVec and mapset geometry (built once,
        cached).  None when the inputs do not fit the batched path."""
        from ..mapmaking.batch import BatchedDot
        if not BatchedDot.eligible(self, mapset):
            return None
        b = self._batch
        if b is None or not b.matches(self, mapset):
            # a stale operator (other noise models, pixelization or map size)
            # is released before the new one takes its device memory
            self._batch = b = None
            self._batch_cuts = None
            torch.cuda.empty_cache()
            b = BatchedDot(self, mapset.maps[0], **kw)
            self._batch = b
        return b

    def batched_cuts(self, mapset):
        """The batched operator for a Mapset of one FP32 SkyMap plus a tsModel
        of CutsCompact (built once per cut layout, cached), or None."""
        from ..mapmaking.batch import BatchedCutsDot
        if not BatchedCutsDot.eligible(self, mapset):
            return None
        bc = getattr(self, "_batch_cuts", None)
        if bc is None or not bc.matches(self, mapset):
            one = type(mapset)()
            one.maps, one.nmap = [mapset.maps[0]], 1
            bc = BatchedCutsDot(self.batched(one), mapset.maps[1])
            self._batch_cuts = bc
        return bc

    def dot_cached(self, mapset, mapset_out=None, mask=None):
        return self.dot(mapset, mapset_out, mask=mask)

    def dot(self, mapset, mapset_out=None, report_times=False, cache_maps=False, mask=None):
        """core.py:967-1026."""
        if mapset_out is None:
            mapset_out = mapset.copy()
            mapset_out.clear()
        b = None if report_times else self.batched(mapset)
        if b is not None:
            if mask is not None:
                mapset.maps[0].map *= mask.maps[0].map
            b.apply(mapset.maps[0].map, mapset_out.maps[0].map, accumulate=True)
            if mask is not None:
                mapset_out.maps[0].map *= mask.maps[0].map
            return mapset_out
        bc = None if (report_times or mask is not None) else self.batched_cuts(mapset)
        if bc is not None:
            y = bc.pack(mapset_out)
            bc.apply(bc.pack(mapset), y, accumulate=True)
            bc.unpack(y, mapset_out)
            return mapset_out
        times = np.zeros(self.ntod)
        tot_times = np.zeros(3)
        for i in range(self.ntod):
            tod = self.tods[i]
            t1 = time.time()
            tod.iter_init()
            mytimes = tod.dot(mapset, mapset_out, True, mask=mask)
            tod.iter_finalize()
            t2 = time.time()
            tot_times = tot_times + mytimes
            times[i] = t2 - t1
        if report_times:
            return mapset_out, times
        return mapset_out

    def make_rhs(self, mapset, do_clear=False):
        """core.py:1028-1047."""
        if do_clear:
            mapset.clear()
        b = self.batched(mapset)
        if b is not None:
            b.rhs(mapset.maps[0].map, accumulate=True)
            return
        bc = self.batched_cuts(mapset)
        if bc is not None:
            y = bc.pack(mapset)
            bc.rhs(y, accumulate=True)
            bc.unpack(y, mapset)
            return
        for tod in self.tods:
            dat_filt = tod.apply_noise()
            for map in mapset.maps:
                map.tod2map(tod, dat_filt)
