"""Cut samples as solvable parameters: Cuts (minkasi/tods/cuts.py:60-198),
CutsCompact (cuts.py:201-498) and segs_from_vec (cuts.py:19-57).

The segment lists (nseg, istart, istop) stay on the host, as in the
reference.  imap (int64 flat det * ndata + sample indices) and map (one value
per cut sample) live on the device.  The gather and scatter run in libmkgpu
(mk_tod2cuts_*, mk_cuts2tod_*), and the vector operations use the map
kernels (mk_dot_*, mk_axpy_*, mk_mul_*, mk_addmul_*).

Differences from the reference:
  * map takes the dtype of the TOD data (FP32 or FP64; the reference's is
    FP64), or dtype= when given.
  * Every detector has its own segment lists.  The reference initialises them
    as [[]] * ndet (cuts.py:239-240), one list shared by every detector, so
    add_cut on one detector also appends to the others' lists and get_imap
    then reads the wrong segments.
  * axpy updates map in place (the reference rebinds a new array), so a
    copy(deep=False) sees the update.
  * map2tod needs imap without repeated samples (overlapping segments):
    call merge_cuts() before get_imap().  The reference's serial loop would
    add the repeats one after another.
"""
import numpy as np
import torch

from .. import device as dev
from ..lib import mkgpu


def segs_from_vec(vec, pad=True):
    """(nseg, istart, istop) of the runs where vec is False (cuts.py:19-57):
    istart is the first sample of a run and istop one past its last."""
    v = vec.detach().cpu().numpy() if isinstance(vec, torch.Tensor) else np.asarray(vec)
    if pad:
        vv = np.ones(len(v) + 2, dtype=bool)
        vv[1:-1] = v
    else:
        vv = v.astype(bool)
    if vv.all():
        return 0, [], []
    inds = np.flatnonzero(vv[1:] != vv[:-1])
    assert len(inds) % 2 == 0
    return len(inds) // 2, [int(i) for i in inds[0::2]], [int(i) for i in inds[1::2]]


def _sfx(t):
    if t.dtype == torch.float32:
        return "f32"
    if t.dtype == torch.float64:
        return "f64"
    raise TypeError("cut vectors are float32 or float64, got %s" % t.dtype)


class Cuts:
    """Bad samples of a TOD as free parameters (cuts.py:60-198): inds are the
    flat indices of tod.info['bad_samples'], map one value per bad sample.
    As in the reference, map2tod with do_add=True (the default) overwrites the
    bad samples with map, and do_add=False adds map to them (cuts.py:148-152)."""

    def __init__(self, tod, do_add=True, dtype=None):
        if isinstance(tod, Cuts):
            self.map = tod.map.clone()
            self.inds = tod.inds.clone()
            self.nsamp = tod.nsamp
            self.do_add = tod.do_add
            self._n = tod._n
            return
        try:
            bad = tod.info["bad_samples"]
            dims = tod.get_data_dims()
        except AttributeError as exc:
            raise ValueError("tod must either be a Tod or Cuts object") from exc
        b = bad.detach().cpu().numpy() if isinstance(bad, torch.Tensor) else np.asarray(bad)
        inds = np.ravel_multi_index(np.where(b), dims).astype(np.int64)
        self.nsamp = len(inds)
        self._n = int(np.prod(dims))
        self.inds = torch.from_numpy(inds).to(dev.device())
        if dtype is None:
            d = tod.info.get("dat_calib")
            dtype = d.dtype if isinstance(d, torch.Tensor) and d.dtype in (torch.float32, torch.float64) \
                else torch.float64
        self.map = torch.zeros(self.nsamp, dtype=dtype, device=dev.device())
        self.do_add = do_add

    def clear(self):
        self.map.zero_()

    def axpy(self, cuts, a):
        """self.map += a * cuts.map (cuts.py:120-132)."""
        if self.nsamp:
            x = cuts.map.to(self.map.dtype)
            mkgpu.check("mk_axpy_" + _sfx(self.map), self.map.data_ptr(), x.data_ptr(), float(a), self.nsamp,
                        dev.stream_ptr())

    def map2tod(self, tod, dat=None):
        """Bad samples of dat := map (do_add=True) or += map (cuts.py:134-153);
        dat is a contiguous device tensor of the map's dtype, updated in place."""
        if dat is None:
            dat = tod.get_data()
        if not isinstance(dat, torch.Tensor) or dat.dtype != self.map.dtype or not dat.is_contiguous():
            raise TypeError("Cuts.map2tod updates its data in place: pass a contiguous %s device tensor"
                            % self.map.dtype)
        if dat.numel() != self._n:
            raise ValueError("data has %d samples, the cuts were built for %d" % (dat.numel(), self._n))
        if self.nsamp:
            mkgpu.check("mk_cuts2tod_" + _sfx(self.map), dat.data_ptr(), self.map.data_ptr(), dat.numel(),
                        self.inds.data_ptr(), self.nsamp, 0 if self.do_add else 1, None, dev.stream_ptr())

    def tod2map(self, tod, dat=None):
        """map := the data at the bad samples (cuts.py:155-170)."""
        if dat is None:
            dat = tod.get_data()
        d = dev.to_device(dat, self.map.dtype)
        if d.numel() != self._n:
            raise ValueError("data has %d samples, the cuts were built for %d" % (d.numel(), self._n))
        if self.nsamp:
            mkgpu.check("mk_tod2cuts_" + _sfx(self.map), self.map.data_ptr(), d.data_ptr(), d.numel(),
                        self.inds.data_ptr(), self.nsamp, 0, None, dev.stream_ptr())

    def dot(self, cuts):
        """sum(self.map * cuts.map), FP64 accumulation (cuts.py:172-187)."""
        if self.nsamp == 0:
            return 0.0
        S = dev.scratch()
        a, b = self.map, cuts.map.to(self.map.dtype)
        mkgpu.check("mk_dot_" + _sfx(a), a.data_ptr(), b.data_ptr(), a.numel(), S.sc.data_ptr(), S.work.data_ptr(),
                    dev.stream_ptr())
        return float(S.sc[mkgpu.SC_DOT].item())

    def copy(self):
        return Cuts(self)


class CutsCompact:
    def __init__(self, tod, dtype=None):
        if isinstance(tod, CutsCompact):
            self.ndet = tod.ndet
            self.nseg = tod.nseg
            self.imax = tod.imax
            self.istart = tod.istart
            self.istop = tod.istop
            if dtype is None:
                dtype = tod.map.dtype
        else:
            ndet = tod.get_ndet()
            self.ndet = ndet
            self.nseg = np.zeros(ndet, dtype="int")
            self.imax = tod.get_ndata()
            self.istart = [[] for _ in range(ndet)]
            self.istop = [[] for _ in range(ndet)]
            if dtype is None:
                d = tod.info.get("dat_calib")
                dtype = getattr(d, "dtype", None)
                if not isinstance(dtype, torch.dtype):
                    dtype = torch.float32 if np.dtype(dtype or np.float64) == np.float32 else torch.float64
                if dtype not in (torch.float32, torch.float64):
                    dtype = torch.float64
        self._unique = True
        self.imap = torch.zeros(0, dtype=torch.int64, device=dev.device())
        self.map = torch.zeros(0, dtype=dtype, device=dev.device())

    def copy(self, deep=True):
        c = CutsCompact(self)
        c._unique = self._unique
        if deep:
            c.imap = self.imap.clone()
            c.map = self.map.clone()
        else:
            c.imap = self.imap
            c.map = self.map
        return c

    # ------------------------------------------------------------ segments
    def add_cut(self, det, istart, istop):
        """Cut [istart, istop) of detector det (cuts.py:271-294)."""
        if istart >= self.imax:
            return
        if istop > self.imax:
            istop = self.imax
        self.nseg[det] = self.nseg[det] + 1
        self.istart[det].append(istart)
        self.istop[det].append(istop)

    def cuts_from_array(self, cutmat):
        """Segments from an (ndet, imax) mask, False where cut (cuts.py:296-309)."""
        m = cutmat.detach().cpu().numpy() if isinstance(cutmat, torch.Tensor) else np.asarray(cutmat)
        for det in range(m.shape[0]):
            nseg, istart, istop = segs_from_vec(m[det, :])
            self.nseg[det] = nseg
            self.istart[det] = istart
            self.istop[det] = istop

    def merge_cuts(self):
        """Merge overlapping or touching segments per detector (cuts.py:312-328)."""
        tmp = np.ones(self.imax + 2, dtype="bool")
        for det in range(self.ndet):
            if self.nseg[det] < 2:
                continue
            tmp[:] = True
            for i in range(self.nseg[det]):
                tmp[(self.istart[det][i] + 1):(self.istop[det][i] + 1)] = False
            nseg, istart, istop = segs_from_vec(tmp, pad=False)
            self.nseg[det] = nseg
            self.istart[det] = istart
            self.istop[det] = istop

    def get_imap(self):
        """Builds imap from the segments (detector-major, segments in stored
        order) and a zero map (cuts.py:330-349)."""
        starts, lens = [], []
        for det in range(self.ndet):
            for i in range(self.nseg[det]):
                starts.append(det * self.imax + self.istart[det][i])
                lens.append(self.istop[det][i] - self.istart[det][i])
        starts = np.asarray(starts, dtype=np.int64)
        lens = np.maximum(np.asarray(lens, dtype=np.int64), 0)
        ncut = int(lens.sum())
        print("ncut is " + repr(ncut))
        if ncut:
            first = np.repeat(np.cumsum(lens) - lens, lens)
            imap = np.repeat(starts, lens) + (np.arange(ncut, dtype=np.int64) - first)
            order = np.argsort(starts[lens > 0], kind="stable")
            s, e = starts[lens > 0][order], (starts + lens)[lens > 0][order]
            self._unique = bool(np.all(s[1:] >= e[:-1]))
        else:
            imap = np.zeros(0, dtype=np.int64)
            self._unique = True
        self.imap = torch.from_numpy(imap).to(dev.device())
        self.map = torch.zeros(ncut, dtype=self.map.dtype, device=dev.device())

    # ------------------------------------------------------------ projections
    def _tod(self, tod, mat):
        if mat is None:
            if tod is None:
                raise ValueError("Both tod and mat cannot be none")
            mat = tod.get_data()
        return mat

    def _check_size(self, n):
        if n != self.ndet * self.imax:
            raise ValueError("data has %d samples, the cuts were built for %d x %d" % (n, self.ndet, self.imax))

    def tod2map(self, tod, mat=None, do_add=True, do_omp=False):
        """map (+)= data at the cut samples (cuts.py:351-383)."""
        mat = self._tod(tod, mat)
        if len(self.imap) == 0:
            return
        d = dev.to_device(mat, self.map.dtype)
        self._check_size(d.numel())
        name = "mk_tod2cuts_" + _sfx(self.map)
        mkgpu.check(name, self.map.data_ptr(), d.data_ptr(), d.numel(), self.imap.data_ptr(), len(self.imap),
                    int(bool(do_add)), None, dev.stream_ptr())

    def map2tod(self, tod, mat=None, do_add=True, do_omp=False):
        """Data at the cut samples (+)= map (cuts.py:385-417).  mat is updated
        in place (a device tensor of the map's dtype, or a numpy array)."""
        mat = self._tod(tod, mat)
        if len(self.imap) == 0:
            return
        if not self._unique:
            raise ValueError("cut segments overlap: call merge_cuts() and get_imap() before map2tod")
        host = not isinstance(mat, torch.Tensor)
        if host:
            d = dev.to_device(np.ascontiguousarray(mat), self.map.dtype)
        else:
            if mat.dtype != self.map.dtype or not mat.is_cuda or not mat.is_contiguous():
                raise TypeError("map2tod updates its data in place: pass a contiguous %s device tensor"
                                % self.map.dtype)
            d = mat
        self._check_size(d.numel())
        name = "mk_cuts2tod_" + _sfx(self.map)
        mkgpu.check(name, d.data_ptr(), self.map.data_ptr(), d.numel(), self.imap.data_ptr(), len(self.imap),
                    int(bool(do_add)), None, dev.stream_ptr())
        if host:
            np.copyto(mat, d.cpu().numpy().reshape(np.shape(mat)).astype(np.asarray(mat).dtype, copy=False))

    # ------------------------------------------------------------ vector ops
    def clear(self):
        self.map.zero_()

    def dot(self, other=None):
        """sum(self.map * other.map) with FP64 accumulation (cuts.py:426-443)."""
        if other is None:
            other = self
        if self.map.numel() != other.map.numel():
            raise ValueError("cut vectors differ in length")
        if self.map.numel() == 0:
            return 0.0
        S = dev.scratch()
        a, b = self.map, other.map
        if a.dtype != b.dtype:
            a, b = a.to(torch.float64), b.to(torch.float64)
        mkgpu.check("mk_dot_" + _sfx(a), a.data_ptr(), b.data_ptr(), a.numel(), S.sc.data_ptr(), S.work.data_ptr(),
                    dev.stream_ptr())
        return float(S.sc[mkgpu.SC_DOT].item())

    def axpy(self, common, a):
        """self.map += a * common.map (cuts.py:445-457), in place."""
        x = common.map.to(self.map.dtype)
        if self.map.numel():
            mkgpu.check("mk_axpy_" + _sfx(self.map), self.map.data_ptr(), x.data_ptr(), float(a), self.map.numel(),
                        dev.stream_ptr())

    def apply_prior(self, x, Ax):
        """Ax.map += self.map * x.map (cuts.py:459-474)."""
        if self.map is None:
            return
        if Ax.map.numel():
            xm = x.map.to(Ax.map.dtype)
            pm = self.map.to(Ax.map.dtype)
            mkgpu.check("mk_addmul_" + _sfx(Ax.map), Ax.map.data_ptr(), pm.data_ptr(), xm.data_ptr(), Ax.map.numel(),
                        dev.stream_ptr())

    def write(self, fname=None):
        """Not implemented in the reference either (cuts.py:476-481)."""
        pass

    def __mul__(self, to_mul):
        """Elementwise product of the cut vectors, in a copy (cuts.py:483-495)."""
        tt = self.copy()
        b = to_mul.map.to(self.map.dtype)
        if tt.map.numel():
            mkgpu.check("mk_mul_" + _sfx(tt.map), tt.map.data_ptr(), self.map.data_ptr(), b.data_ptr(),
                        tt.map.numel(), dev.stream_ptr())
        return tt
