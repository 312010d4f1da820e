"""The batched operator A.x = P^T N^-1 P x over every TOD of a TodVec.

The reference loops over TODs and, per TOD, calls map2tod, apply_noise and
tod2map one after another (minkasi/tods/core.py:649-691, :1012-1020).  Here
all TODs live in one arena (concatenated row-major ndet x n blocks) and each
stage is ONE launch over the whole arena:

  1. map2tod      flat gather over the arena from a tiled copy of the map
                  (mk_map_to_tiles_f32 + mk_map2tod_f32 with remapped indices)
  2. v.d          split-BF16 MFMA, one segment per TOD       (mk_rot_plan_exec)
  3. C W C        four-step FFT filter, one segment per TOD  (mk_filt_plan_exec_f32)
                  (edge halving folded in)
  4. vT.(.)       split-BF16 MFMA                            (mk_rot_plan_exec)
  5. tod2map      deterministic plan over the arena          (mk_t2m_plan_exec_f32)

The sequence can be captured once in a HIP graph and replayed.
"""
import os
import time

import numpy as np
import torch

from .. import device as dev
from ..lib import mkgpu
from ..lib.plans import DeflPlan, FiltPlan, PixCode, RotPlan, T2MPlan
from .map2tod import map2tod
from .noise import NoiseSmoothedSVD


# map2tod gather tiles (tx, ty), 32 FP32 pixels = one 128-byte line; the
# first is the default, the autotuner times all of them
GATHER_TILES = ((8, 4), (16, 2), (4, 8), (32, 1))

# Device memory left free once the operator is built (the solver's map
# vectors, the caller's temporaries).
MEM_RESERVE = 6 << 30
# Work buffer below which the filter would rather give up the tiled gather's
# index copy than be split into more groups.
MIN_FILTER_WORK = 8 << 30


def filter_groups(work, cap):
    """Consecutive TODs grouped so that each group's filter work buffer
    (sum of its TODs' work bytes) stays within cap, with the groups about
    equal.  Returns lists of TOD indices."""
    if not work:
        return []
    total, top = sum(work), max(work)
    if cap < top:
        raise MemoryError("the filter needs %.2f GB of work memory for its largest TOD, %.2f GB are left"
                          % (top / 1e9, cap / 1e9))
    if total <= cap:
        return [list(range(len(work)))]
    ngroups = -(-total // cap)
    target = min(cap, max(top, -(-total // ngroups)))
    groups, cur, acc = [], [], 0
    for i, w in enumerate(work):
        if cur and acc + w > target:
            groups.append(cur)
            cur, acc = [], 0
        cur.append(i)
        acc += w
    groups.append(cur)
    if max(sum(work[i] for i in g) for g in groups) > cap:  # target overshoot: plain greedy fill to cap
        groups, cur, acc = [], [], 0
        for i, w in enumerate(work):
            if cur and acc + w > cap:
                groups.append(cur)
                cur, acc = [], 0
            cur.append(i)
            acc += w
        groups.append(cur)
    return groups


def _hot_p1(plan):
    """The FFT column length of a one-segment FP32 filter plan when it is on
    the hot path (rows of 2^11 points, mixed-radix columns), else 0."""
    g = plan.geometry()[0]
    return int(g[0]) if int(g[1]) == 2048 and not os.environ.get("MK_FFT_NO_HOT") else 0


class BatchedDot:
    @staticmethod
    def eligible(todvec, mapset):
        from ..maps.skymap import SkyMap
        if len(mapset.maps) != 1 or not isinstance(mapset.maps[0], SkyMap):
            return False
        if getattr(mapset.maps[0], "npol", 1) != 1:
            return False
        if mapset.maps[0].map.dtype != torch.float32 or todvec.ntod == 0:
            return False
        for tod in todvec.tods:
            if tod.noise_delayed or not isinstance(tod.noise, NoiseSmoothedSVD):
                return False
            if tod.noise.dtype != torch.float32 and tod.noise.filt32 is None:
                return False
        return True

    def __init__(self, todvec, skymap, tile=(32, 32), item_samples=None, tiled_gather=None, filter_work_cap=None):
        """tiled_gather: gather map2tod from a tiled copy of the map through a
        second index array (4 bytes per sample), or from the row-major map
        (False); None decides from the free device memory.
        filter_work_cap: largest filter work buffer in bytes (None: what the
        free device memory allows).  TODs whose filter work exceeds it in
        total are filtered in groups that share one buffer."""
        self.tods = list(todvec.tods)
        self.nx, self.ny, self.tag = skymap.nx, skymap.ny, skymap.tag
        d = dev.device()
        offs, sizes = [], []
        off = 0
        for tod in self.tods:
            ndet, n = tod.noise.ndet, tod.noise.n
            if (ndet, n) != tod.get_data_dims():
                raise ValueError("noise model shape differs from TOD data")
            offs.append(off)
            sizes.append((ndet, n))
            off += ndet * n
        self.offs, self.sizes, self.nsamp = offs, sizes, off
        # ipix arena; the TODs' cached pixelizations become views into it
        self.ipix = torch.empty(max(off, 4), dtype=torch.int32, device=d)
        for tod, o, (ndet, n) in zip(self.tods, offs, sizes):
            ip = skymap.get_pix(tod)
            self.ipix[o:o + ndet * n].copy_(ip.reshape(-1))
            tod.info[self.tag] = self.ipix[o:o + ndet * n].view(ndet, n)
        # what the operator was built from: the TODs' pixelizations (views into
        # the arena) and noise models; matches() rebuilds when either changes
        self.ipix_views = [tod.info[self.tag] for tod in self.tods]
        self.noises = [tod.noise for tod in self.tods]
        # the two FP32 work arenas are halves of one block, which the FP64
        # operator of the true-residual check (AccurateDot) reuses as one FP64
        # arena; each half starts 256-byte aligned
        half = -(-max(off, 4) // 64) * 64
        self._bufs = torch.empty(2 * half, dtype=torch.float32, device=d)
        self.buf_a = self._bufs[:half]
        self.buf_b = self._bufs[half:]
        # the per-TOD index blocks the arena replaced go back to the device,
        # where the library's own allocations can use them
        torch.cuda.empty_cache()
        from .. import tune
        tuner = tune.active()
        # the tod2map plan first: its build needs sort buffers over 2^30-sample
        # segments and a 4-byte sorted entry per sample on top of the 2.3 bytes
        # per sample it keeps
        if tuner is not None and off > 0 and tile == (32, 32) and item_samples is None:
            self.t2m = tuner.t2m_plan(self.ipix[:off], self.nx, self.ny, dat=self.buf_a[:off])
        else:
            self.t2m = T2MPlan(self.ipix[:off], self.nx, self.ny, tile=tile, item_samples=item_samples)
        work = [int(tod.noise.filt32.work_bytes) for tod in self.tods]
        # the tiled gather reads coded pointing (about 1 byte per sample; its
        # exact size from one sizing pass) and a tiled copy of the map
        units, code_bytes = (PixCode.sizes(self.ipix[:off], self.nx, self.ny) if off and tiled_gather is not False
                             else (None, 0))
        self.tiled, cap = self._memory_plan(work, off, tiled_gather, filter_work_cap, index_bytes=code_bytes)
        self.code = None
        self.map_t = None
        self.gather_tile = None
        if self.tiled and off:
            self.code = PixCode(self.ipix[:off], self.nx, self.ny, units=units)
        del units
        if self.tiled and tuner is not None and off > 0:
            sig = tune.ipix_signature(self.ipix[:off], self.nx, self.ny)
            x = torch.randn(self.nx, self.ny, device=d)

            def time_tile(t):
                self._set_gather_tile(tuple(t))
                return self._time(lambda: self._map2tod(x, self.buf_a[:off]))

            best = tuner.choose("gather_tile", sig, [list(t) for t in GATHER_TILES], time_tile)
            self._set_gather_tile(tuple(best))
        elif self.tiled:
            self._set_gather_tile(GATHER_TILES[0])
        # filter plans over groups of TODs (one group unless memory is short),
        # with the same FFT column lengths as the noise models whose spectra
        # they borrow, sharing one work buffer
        self.filts = []
        self.filter_groups = filter_groups(work, cap)
        for grp in self.filter_groups:
            segs = [(sizes[i][0], sizes[i][1], offs[i], sizes[i][1], offs[i], sizes[i][1]) for i in grp]
            fp = FiltPlan(segs, 4, p1=[_hot_p1(self.tods[i].noise.filt32) for i in grp])
            for k, i in enumerate(grp):
                fp.borrow(k, self.tods[i].noise.filt32, 0)
            fp.set_halve(True)
            self.filts.append(fp)
        wmax = max((sum(work[i] for i in g) for g in self.filter_groups), default=0)
        self.filter_work = torch.empty(max(wmax, 16), dtype=torch.uint8, device=d)
        for fp in self.filts:
            fp.set_work(self.filter_work)
        self.rot_v = RotPlan([(tod.noise.v_split, n, o, n, o, n) for tod, o, (ndet, n) in zip(self.tods, offs, sizes)])
        self.rot_vT = RotPlan([(tod.noise.vT_split, n, o, n, o, n)
                               for tod, o, (ndet, n) in zip(self.tods, offs, sizes)])
        # deflation of dominant detector modes ahead of the first rotation, for
        # the TODs whose noise model asks for it
        dsegs = [(ndet, n, o, tod.noise.v[:tod.noise.defl_k].T)
                 for tod, o, (ndet, n) in zip(self.tods, offs, sizes) if getattr(tod.noise, "defl_k", 0)]
        self.defl = DeflPlan(dsegs) if dsegs else None
        for tod in self.tods:  # per-TOD filter work buffers are not needed on the batched path
            tod.noise.filt32.release_work()
        self.graph = None
        self._gx = self._gout = None
        # tod2map sums of A.x: FP32 (3.7 ms at 1.9e9 samples) unless set;
        # into an FP64 map (f64_view) and for the right-hand side they are
        # FP64 (6.1 ms); the PCG iteration needs no more than FP32 sums
        # because its residual is replaced from the FP64-binned check
        self.t2m_acc64 = False

    @staticmethod
    def _memory_plan(work, nsamp, tiled_gather, filter_work_cap, free=None, index_bytes=None):
        """(tiled gather?, filter work cap in bytes) from the device memory
        free once the tod2map plan is built: less a reserve, it goes to the
        tiled gather's index structure (index_bytes, at most 4 bytes per
        sample) and the filter work buffer, and the tiled gather is given up
        first when the filter would otherwise get less than MIN_FILTER_WORK
        (or its largest TOD's work)."""
        if not work:
            return bool(tiled_gather is not False), 0
        if free is None:
            free = int(torch.cuda.mem_get_info()[0])
        if index_bytes is None:
            index_bytes = 4 * nsamp
        left = free - MEM_RESERVE
        total, top = sum(work), max(work)
        if tiled_gather is None:
            tiled_gather = left - index_bytes >= max(top, min(total, MIN_FILTER_WORK))
        if tiled_gather:
            left -= index_bytes
        if left < top:
            raise MemoryError(
                "batched operator over %.3g samples: %.1f GB of device memory free after the tod2map plan, not "
                "enough for %sa filter work buffer of %.1f GB and a %.1f GB reserve" % (
                    nsamp, free / 1e9, "tiled gather indices (%.1f GB), " % (index_bytes / 1e9) if tiled_gather else "",
                    top / 1e9, MEM_RESERVE / 1e9))
        cap = left if filter_work_cap is None else min(left, int(filter_work_cap))
        return bool(tiled_gather), cap

    def _filter(self, b):
        """b = C W C b in place, group by group."""
        for fp in self.filts:
            fp.exec(b, b)

    def memory_report(self):
        """Device bytes held by the operator's own buffers and plans."""
        n = self.nsamp
        t = self.t2m.info()
        return {"work_tods": 8 * n, "ipix": 4 * n, "tiled_gather_ipix": self.code.nbytes if self.code else 0,
                "tod2map_plan": t["plan_bytes"] + t["partial_bytes"], "filter_work": int(self.filter_work.numel()),
                "filter_groups": len(self.filts)}

    def _set_gather_tile(self, tile):
        """Size the tiled map buffer for tx x ty map tiles (powers of two);
        the coded pointing does not depend on the tile."""
        if tile == self.gather_tile:
            return
        tx, ty = int(tile[0]), int(tile[1])
        size = int(mkgpu.lib.mk_map_tiled_size(self.nx, self.ny, tx, ty))
        if self.map_t is None or self.map_t.numel() != size:
            self.map_t = torch.zeros(size, dtype=torch.float32, device=self.ipix.device)
        self.gather_tile = (tx, ty)
        self.graph = None

    def _rot_v(self, a, b):
        """b = v.a per TOD (a is overwritten by its deflated copy when some
        TOD's modes are deflated)."""
        if self.defl is not None:
            self.defl.exec(a, a)
            self.rot_v.exec(a, b)
            self.defl.addback(b)
        else:
            self.rot_v.exec(a, b)

    def _map2tod(self, x, a):
        """a = x.flat[ipix] over the arena, gathered from the tiled copy (or
        from the row-major map when memory is short)."""
        if not self.nsamp:
            return
        if not self.tiled:
            # the arena's indices were checked against nx * ny at construction
            map2tod(a, x.reshape(-1), self.ipix[:self.nsamp], do_add=False, check=False)
            return
        tx, ty = self.gather_tile
        mkgpu.check("mk_map_to_tiles_f32", dev.ptr(x, torch.float32, "map"), self.nx, self.ny, tx, ty,
                    self.map_t.data_ptr(), dev.stream_ptr())
        self.code.gather(a, self.map_t, self.gather_tile)

    @staticmethod
    def _time(fn, reps=3):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    def matches(self, todvec, mapset):
        """Same TODs, map size, pixelizations and noise models.  A TOD whose
        noise was re-estimated (Tod.set_noise, as the reference's multi-pass
        examples do between solves) or whose pixelization was replaced makes
        the cached operator stale."""
        m = mapset.maps[0]
        if not (len(todvec.tods) == len(self.tods) and all(a is b for a, b in zip(todvec.tods, self.tods))
                and m.nx == self.nx and m.ny == self.ny and m.tag == self.tag):
            return False
        for tod, ip, nz in zip(self.tods, self.ipix_views, self.noises):
            if tod.noise_delayed or tod.noise is not nz or tod.info.get(self.tag) is not ip:
                return False
        return True

    # ------------------------------------------------------------------
    def stages(self, x, out, accumulate):
        n = self.nsamp
        a, b = self.buf_a[:n], self.buf_b[:n]
        self._map2tod(x, a)
        self._rot_v(a, b)
        self._filter(b)
        self.rot_vT.exec(b, a)
        self.t2m.exec(out.reshape(-1), a, accumulate, acc64=self.t2m_acc64)

    def apply(self, x, out, accumulate=True):
        """out (+)= A x for device maps x, out (nx, ny) FP32."""
        if self.graph is not None and x is self._gx and out is self._gout and accumulate == self._gacc:
            self.graph.replay()
            return out
        self.stages(x, out, accumulate)
        return out

    def capture(self, x, out, accumulate=False, warmup=1):
        """Captures apply(x, out) into a HIP graph replayed by later calls
        with the same tensors."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self.stages(x, out, accumulate)
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.stages(x, out, accumulate)
        self.graph, self._gx, self._gout, self._gacc = g, x, out, accumulate
        return g

    def rhs(self, out, accumulate=True):
        """out (+)= sum over TODs of P^T N^-1 d (TodVec.make_rhs,
        minkasi/tods/core.py:1028-1047) with the batched kernels."""
        n = self.nsamp
        a, b = self.buf_a[:n], self.buf_b[:n]
        for tod, o, (ndet, nn) in zip(self.tods, self.offs, self.sizes):
            d = tod.get_data()
            d = d if isinstance(d, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(d))
            a[o:o + ndet * nn].copy_(d.reshape(-1).to(a.device, torch.float32))
        self._rot_v(a, b)
        self._filter(b)
        self.rot_vT.exec(b, a)
        self.t2m.exec(out.reshape(-1), a, accumulate, acc64=True)
        return out

    def hits(self, out, accumulate=True):
        """Hit counts through the same deterministic plan (make_hits)."""
        n = self.nsamp
        self.buf_a[:n].fill_(1.0)
        self.t2m.exec(out.reshape(-1), self.buf_a[:n], accumulate)
        return out

    def null_mask(self):
        """Hit pixels (bool, nx x ny): the support of A's null vector, the
        constant on hit pixels (computed once with the tod2map plan)."""
        m = getattr(self, "_null", None)
        if m is None:
            h = torch.zeros((self.nx, self.ny), dtype=torch.float32, device=self.buf_a.device)
            self.hits(h, accumulate=False)
            m = self._null = h > 0
        return m

    def accurate(self):
        """The all-FP64 operator over these TODs (AccurateDot, built once)."""
        a = getattr(self, "_accurate", None)
        if a is None:
            a = self._accurate = AccurateDot(self)
        return a

    def f64_view(self):
        """This operator with FP64 binning into FP64 maps (apply(x64, out64),
        rhs(out64)): the FP32 stages see fl32(x64), tod2map sums in FP64."""
        v = getattr(self, "_f64", None)
        if v is None:
            v = self._f64 = _Binned64(self)
        return v

    def stage_times(self, x, out, reps=3):
        """Per-stage device time (ms) with events on the launch stream."""
        n = self.nsamp
        a, b = self.buf_a[:n], self.buf_b[:n]
        steps = [
            ("map2tod", lambda: self._map2tod(x, a)),
            ("rot_v", lambda: self._rot_v(a, b)),
            ("filter", lambda: self._filter(b)),
            ("rot_vT", lambda: self.rot_vT.exec(b, a)),
            ("tod2map", lambda: self.t2m.exec(out.reshape(-1), a, False, acc64=self.t2m_acc64)),
        ]
        res = {}
        for name, fn in steps:
            fn()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            e1.synchronize()
            res[name] = e0.elapsed_time(e1) / reps
        return res


class _Binned64:
    """BatchedDot.f64_view: the FP32 operator evaluated from an FP64 map
    into an FP64 map, its tod2map summing in FP64."""

    def __init__(self, base):
        self.base = base
        self.nx, self.ny = base.nx, base.ny
        self.x32 = torch.empty((base.nx, base.ny), dtype=torch.float32, device=base.buf_a.device)

    def matches(self, base):
        return base is self.base

    def apply(self, x64, out64):
        self.x32.copy_(x64)
        self.base.stages(self.x32, out64, False)
        return out64

    def rhs(self, out64):
        return self.base.rhs(out64, accumulate=False)


class BatchedCutsDot:
    """The batched operator for a Mapset of one FP32 SkyMap plus a tsModel of
    CutsCompact (the reference's cuts example, examples/minkasi_cuts_example.py),
    on flat vectors [map (nx*ny) | every TOD's cut values, in TOD order].

    A [x; c] = [P^T N^-1 (P x + C c); C^T N^-1 (P x + C c)] with C the cut
    scatter (mk_cuts2tod_f32 over one arena-wide int64 index, after the
    map2tod stage) and C^T the gather (mk_tod2cuts_f32, after the second
    rotation), so every stage stays one launch over all TODs."""

    @staticmethod
    def eligible(todvec, mapset):
        from ..tods.cuts import CutsCompact
        from .timestream import tsModel, tod_key
        if len(mapset.maps) != 2 or not isinstance(mapset.maps[1], tsModel):
            return False
        one = type(mapset)()
        one.maps, one.nmap = [mapset.maps[0]], 1
        if not BatchedDot.eligible(todvec, one):
            return False
        cm = mapset.maps[1]
        for tod in todvec.tods:
            c = cm.data.get(tod_key(tod))
            if not isinstance(c, CutsCompact) or c.map.dtype != torch.float32 or not c._unique:
                return False
            if c.ndet * c.imax != tod.get_nsamp():
                return False
        return len(cm.data) == todvec.ntod

    def __init__(self, base, cuts_model):
        from .timestream import tod_key
        self.base = base
        self.keys = [tod_key(t) for t in base.tods]
        self.npix = base.nx * base.ny
        parts, self.counts = [], []
        for key, o in zip(self.keys, base.offs):
            c = cuts_model.data[key]
            parts.append(c.imap + o)
            self.counts.append(int(c.imap.numel()))
        d = base.buf_a.device
        self.cimap = torch.cat(parts) if parts else torch.zeros(0, dtype=torch.int64, device=d)
        self.ncut = int(self.cimap.numel())
        self.n = self.npix + self.ncut
        self.cut_ids = [id(cuts_model.data[k].imap) for k in self.keys]

    def matches(self, todvec, mapset):
        """Same TODs, map size and cut samples (a Mapset copy holds copies of
        the cut indices: those are compared by value)."""
        if len(mapset.maps) != 2 or not hasattr(mapset.maps[1], "data") or not self.base.matches(todvec, mapset):
            return False
        cm = mapset.maps[1]
        for j, k in enumerate(self.keys):
            c = cm.data.get(k)
            if c is None or int(c.imap.numel()) != self.counts[j]:
                return False
            if id(c.imap) != self.cut_ids[j] and not torch.equal(c.imap, self._imap_of(j)):
                return False
        return True

    def null_mask(self):
        """[hit pixels | no cut samples]: A's null vector is the constant on
        hit pixels and zero on the cut samples."""
        m = getattr(self, "_null", None)
        if m is None:
            m = torch.zeros(self.n, dtype=torch.bool, device=self.base.buf_a.device)
            m[:self.npix] = self.base.null_mask().reshape(-1)
            self._null = m
        return m

    def _imap_of(self, j):
        o = sum(self.counts[:j])
        return self.cimap[o:o + self.counts[j]] - self.base.offs[j]

    # ---------------------------------------------------------------- packing
    def pack(self, mapset, out=None):
        """Flat FP32 vector [map | cuts] of a Mapset of this structure."""
        if out is None:
            out = torch.empty(self.n, dtype=torch.float32, device=self.base.buf_a.device)
        out[:self.npix].copy_(mapset.maps[0].map.reshape(-1))
        o = self.npix
        for k, c in zip(self.keys, self.counts):
            if c:
                out[o:o + c].copy_(mapset.maps[1].data[k].map)
            o += c
        return out

    def unpack(self, flat, mapset):
        mapset.maps[0].map.copy_(flat[:self.npix].view_as(mapset.maps[0].map))
        o = self.npix
        for k, c in zip(self.keys, self.counts):
            if c:
                mapset.maps[1].data[k].map.copy_(flat[o:o + c])
            o += c
        return mapset

    # ---------------------------------------------------------------- operator
    def _cuts_in(self, c, a):
        if self.ncut:
            mkgpu.check("mk_cuts2tod_f32", a.data_ptr(), c.data_ptr(), a.numel(), self.cimap.data_ptr(), self.ncut, 1,
                        None, dev.stream_ptr())

    def _cuts_out(self, a, c, accumulate):
        if self.ncut:
            mkgpu.check("mk_tod2cuts_f32", c.data_ptr(), a.data_ptr(), a.numel(), self.cimap.data_ptr(), self.ncut,
                        int(bool(accumulate)), None, dev.stream_ptr())

    def apply(self, x, out, accumulate=True):
        """out (+)= A x on flat vectors (a captured graph is replayed for the
        tensors it was captured with)."""
        g = getattr(self, "graph", None)
        if g is not None and x is self._gx and out is self._gout and accumulate == self._gacc:
            g.replay()
            return out
        return self.stages(x, out, accumulate)

    def capture(self, x, out, accumulate=False, warmup=1):
        """Captures apply(x, out) into a HIP graph (as BatchedDot.capture)."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self.stages(x, out, accumulate)
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.stages(x, out, accumulate)
        self.graph, self._gx, self._gout, self._gacc = g, x, out, accumulate
        return g

    def stages(self, x, out, accumulate):
        B = self.base
        n = B.nsamp
        a, b = B.buf_a[:n], B.buf_b[:n]
        xm = x[:self.npix].view(B.nx, B.ny)
        B._map2tod(xm, a)
        self._cuts_in(x[self.npix:], a)
        B._rot_v(a, b)
        B._filter(b)
        B.rot_vT.exec(b, a)
        B.t2m.exec(out[:self.npix], a, accumulate, acc64=B.t2m_acc64)
        self._cuts_out(a, out[self.npix:], accumulate)
        return out

    def rhs(self, out, accumulate=True):
        """out (+)= [P^T N^-1 d; C^T N^-1 d] (TodVec.make_rhs)."""
        B = self.base
        n = B.nsamp
        B.rhs(out[:self.npix].view(B.nx, B.ny), accumulate)  # leaves N^-1 d in buf_a
        self._cuts_out(B.buf_a[:n], out[self.npix:], accumulate)
        return out


class AccurateDot:
    """FP64 A.x and right-hand side over the TODs of a BatchedDot: the
    operator the true residual ||b - A x||/||b|| is measured with, and the
    one the device PCG replaces its recursive residual from.

    The reference evaluates all of A in FP64 (minkasi/tods/core.py:649-691
    on numpy maps, noise.py:680-714).  The batched FP32 operator differs from
    it by a few 1e-7 relative at the solution, as much as the residual a
    solve to 1e-6 is asked to reach, so the check runs every stage in FP64:

      map2tod   FP64 gather over the arena           (mk_map2tod_f64)
      v.d       FP64 GEMM per TOD                    (mk_gemm_nn_f64)
      C W C     FP64 filter, batched plans borrowing the noise models' FP64
                spectra, TODs in groups sharing one work buffer
      vT.(.)    FP64 GEMM per TOD
      tod2map   FP64 through the batched deterministic plan

    Every TOD's noise model must keep its FP64 filter (keep_fp64=True).  The
    FP64 arena aliases the FP32 operator's two work arenas (the two operators
    never run at the same time on the launch stream); a second FP64 arena is
    this object's own (8 bytes per sample)."""

    def __init__(self, base, work_cap=8 << 30):
        missing = [k for k, tod in enumerate(base.tods) if getattr(tod.noise, "filt64", None) is None]
        if missing:
            raise ValueError("AccurateDot needs noise models built with keep_fp64=True (TODs %s lack the FP64 "
                             "filter)" % missing[:8])
        self.base = base
        self.nx, self.ny = base.nx, base.ny
        n = base.nsamp
        d = base.buf_a.device
        self.a64 = base._bufs.view(torch.float64)
        if self.a64.numel() < n:
            raise RuntimeError("FP64 view of the work arenas is shorter than the arena")
        self.b64 = torch.empty(max(self.a64.numel(), 4), dtype=torch.float64, device=d)
        self.v64 = [tod.noise._v64 for tod in base.tods]
        self.vT64 = [tod.noise._vT64 for tod in base.tods]
        work = [int(tod.noise.filt64.work_bytes) for tod in base.tods]
        for tod in base.tods:
            tod.noise.filt64.release_work()
        cap = max(max(work, default=0), min(int(work_cap), sum(work)))
        self.filter_groups = filter_groups(work, cap)
        self.filts = []
        for grp in self.filter_groups:
            segs = [(base.sizes[i][0], base.sizes[i][1], base.offs[i], base.sizes[i][1], base.offs[i],
                     base.sizes[i][1]) for i in grp]
            fp = FiltPlan(segs, 8)
            for k, i in enumerate(grp):
                fp.borrow(k, base.tods[i].noise.filt64, 0)
            fp.set_halve(True)
            self.filts.append(fp)
        wmax = max((sum(work[i] for i in g) for g in self.filter_groups), default=0)
        self.filter_work = torch.empty(max(wmax, 16), dtype=torch.uint8, device=d)
        for fp in self.filts:
            fp.set_work(self.filter_work)

    def matches(self, base):
        return base is self.base

    def _rotate(self, mats, src, dst):
        sp = dev.stream_ptr()
        for m, o, (ndet, n) in zip(mats, self.base.offs, self.base.sizes):
            mkgpu.check("mk_gemm_nn_f64", ndet, n, ndet, m.data_ptr(), ndet, src[o:o + ndet * n].data_ptr(), 0, n,
                        dst[o:o + ndet * n].data_ptr(), n, sp)

    def _noise(self, a, b):
        """b = v.a ; b = C W C b ; a = vT.b (per TOD, FP64)."""
        self._rotate(self.v64, a, b)
        for fp in self.filts:
            fp.exec(b, b)
        self._rotate(self.vT64, b, a)

    def apply(self, x64, out64):
        """out64 = A x64 for FP64 device maps (nx, ny)."""
        B = self.base
        n = B.nsamp
        a, b = self.a64[:n], self.b64[:n]
        if x64.dtype != torch.float64 or out64.dtype != torch.float64:
            raise TypeError("AccurateDot works on FP64 maps")
        mkgpu.check("mk_map2tod_f64", a.data_ptr(), dev.ptr(x64, torch.float64, "map"), n, B.ipix.data_ptr(), 0,
                    dev.stream_ptr())
        self._noise(a, b)
        B.t2m.exec(out64.reshape(-1), a, False)
        return out64

    def rhs(self, out64):
        """out64 = sum over TODs of P^T N^-1 d in FP64 (TodVec.make_rhs,
        minkasi/tods/core.py:1028-1047)."""
        B = self.base
        n = B.nsamp
        a, b = self.a64[:n], self.b64[:n]
        for tod, o, (ndet, nn) in zip(B.tods, B.offs, B.sizes):
            dd = tod.get_data()
            dd = dd if isinstance(dd, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(dd))
            a[o:o + ndet * nn].copy_(dd.reshape(-1).to(a.device, torch.float64))
        self._noise(a, b)
        B.t2m.exec(out64.reshape(-1), a, False)
        return out64
