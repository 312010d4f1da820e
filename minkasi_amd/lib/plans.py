"""Owning Python handles for the library's plan objects (tod2map, filter,
rotation).  Each handle keeps the tensors its device descriptors point at
alive and destroys the plan when collected."""
import ctypes
import os

import numpy as np
import torch

from . import mkgpu
from .. import device as dev


class _Handle:
    _destroy = None

    def __init__(self):
        self.h = ctypes.c_void_p()
        self._keep = []

    def close(self):
        if self.h and self.h.value and self._destroy:
            try:
                mkgpu.check(self._destroy, self.h)
            except Exception:
                pass
        self.h = ctypes.c_void_p()

    def __del__(self):
        self.close()


# FP32 samples are binned with FP64 sums unless MK_T2M_ACC32 is set (the
# FP32-sum kernel is kept for comparison).
ACC64_DEFAULT = not os.environ.get("MK_T2M_ACC32")


class T2MPlan(_Handle):
    """Deterministic tod2map plan over a (batch) ipix array."""

    _destroy = "mk_t2m_plan_destroy"

    def __init__(self, ipix, nx, ny, tile=(32, 32), item_samples=None):
        super().__init__()
        n = int(ipix.numel())
        if item_samples is None:
            # enough items to fill 256 CUs several times, large enough that
            # the partial-tile traffic stays a few percent of the samples
            item_samples = int(min(32768, max(1024, (n // (256 * 8)) // 64 * 64)))
        self.nx, self.ny, self.tile, self.item_samples = int(nx), int(ny), tuple(tile), item_samples
        self.nsamp = n
        mkgpu.check("mk_t2m_plan_create", dev.ptr(ipix, torch.int32, "ipix") if n else 0, n, self.nx, self.ny,
                    int(tile[0]), int(tile[1]), int(item_samples), ctypes.byref(self.h), dev.stream_ptr())

    def info(self):
        out = np.zeros(8, dtype=np.int64)
        mkgpu.check("mk_t2m_plan_info", self.h, out.ctypes.data)
        keys = ["nsamp", "nitems", "ntiles_touched", "tile_x", "tile_y", "plan_bytes", "partial_bytes", "max_item"]
        return dict(zip(keys, (int(v) for v in out)))

    def exec(self, map_t, dat, accumulate, acc64=None):
        """map_t (+)= tod2map(dat) in the plan's fixed order.

        FP32 samples are summed in FP64 (acc64, the default when
        MK_T2M_ACC32 is unset) into an FP32 or FP64 map; acc64=False keeps
        the FP32 sums of mk_t2m_plan_exec_f32.  FP64 samples need an FP64
        map and are summed in FP64."""
        if map_t.numel() != self.nx * self.ny:
            raise ValueError("map size %d does not match plan %dx%d" % (map_t.numel(), self.nx, self.ny))
        if dat.numel() != self.nsamp:
            raise ValueError("TOD has %d samples, plan expects %d" % (dat.numel(), self.nsamp))
        dptr = dev.ptr(dat, name="dat") if self.nsamp else dev.ptr(map_t)
        mptr = dev.ptr(map_t, name="map")
        acc = int(bool(accumulate))
        if dat.dtype == torch.float64:
            if map_t.dtype != torch.float64:
                raise TypeError("FP64 samples need an FP64 map (got %s)" % map_t.dtype)
            mkgpu.check("mk_t2m_plan_exec_f64", self.h, mptr, dptr, acc, dev.stream_ptr())
            return
        if dat.dtype != torch.float32 or map_t.dtype not in (torch.float32, torch.float64):
            raise TypeError("unsupported tod2map dtypes (map %s, TOD %s)" % (map_t.dtype, dat.dtype))
        if acc64 is None:
            acc64 = ACC64_DEFAULT
        if map_t.dtype == torch.float64 or acc64:
            mkgpu.check("mk_t2m_plan_exec_f32_acc64", self.h, mptr, int(map_t.dtype == torch.float64), dptr, acc,
                        dev.stream_ptr())
        else:
            mkgpu.check("mk_t2m_plan_exec_f32", self.h, mptr, dptr, acc, dev.stream_ptr())


    def exec_pol(self, map_t, dat, twogamma, mode, accumulate):
        """Polarised map (nx*ny x ncomp) (+)= the per-sample terms of mode
        (mkgpu.POL_*) through this plan: deterministic, FP64 sums."""
        if dat.numel() != self.nsamp or twogamma.numel() != self.nsamp:
            raise ValueError("TOD / twogamma sizes differ from the plan's %d samples" % self.nsamp)
        if dat.dtype != map_t.dtype or twogamma.dtype != map_t.dtype:
            raise TypeError("map, TOD and twogamma must share a dtype")
        ncomp = {mkgpu.POL_QU: 2, mkgpu.POL_IQU: 3, mkgpu.POL_QU_PRECON: 3, mkgpu.POL_IQU_PRECON: 6}[mode]
        if map_t.numel() != self.nx * self.ny * ncomp:
            raise ValueError("map holds %d values, the plan's %dx%d map with %d components needs %d" % (
                map_t.numel(), self.nx, self.ny, ncomp, self.nx * self.ny * ncomp))
        name = {torch.float32: "mk_t2m_plan_exec_pol_f32", torch.float64: "mk_t2m_plan_exec_pol_f64"}[dat.dtype]
        dptr = dev.ptr(dat, name="dat") if self.nsamp else dev.ptr(map_t)
        gptr = dev.ptr(twogamma, name="twogamma") if self.nsamp else dev.ptr(map_t)
        mkgpu.check(name, self.h, dev.ptr(map_t, name="map"), dptr, gptr, int(mode), int(bool(accumulate)),
                    dev.stream_ptr())


class FiltPlan(_Handle):
    """Batched C W C filter (noise filter core)."""

    _destroy = "mk_filt_plan_destroy"

    def __init__(self, segs, elem_bytes=4, p1=None):
        """segs: list of (rows, n, in_off, in_stride, out_off, out_stride).
        p1: per segment the FFT column length (0 = the smallest that fits);
        None asks the active autotuner (minkasi_amd.tune), if any."""
        super().__init__()
        self.segs = [tuple(int(v) for v in s) for s in segs]
        self.elem_bytes = elem_bytes
        desc = np.asarray(self.segs, dtype=np.int64).reshape(-1)
        if desc.size == 0:
            desc = np.zeros(6, dtype=np.int64)
        if p1 is None and elem_bytes == 4:
            from .. import tune
            t = tune.active()
            if t is not None:
                p1 = [t.fft_p1(s[1]) for s in self.segs]
        if p1 is None:
            mkgpu.check("mk_filt_plan_create", len(self.segs), desc.ctypes.data, int(elem_bytes),
                        ctypes.byref(self.h))
        else:
            p = np.asarray(list(p1) + [0], dtype=np.int32)
            if len(p) != len(self.segs) + 1:
                raise ValueError("p1 needs one entry per segment")
            mkgpu.check("mk_filt_plan_create_p1", len(self.segs), desc.ctypes.data, int(elem_bytes), p.ctypes.data,
                        ctypes.byref(self.h))

    def set_weights(self, wdct):
        mkgpu.check("mk_filt_plan_set_weights", self.h, dev.ptr(wdct, torch.float64, "wdct"), dev.stream_ptr())

    def set_halve(self, halve):
        mkgpu.check("mk_filt_plan_set_halve", self.h, int(bool(halve)))

    def release_work(self):
        mkgpu.check("mk_filt_plan_release_work", self.h, dev.stream_ptr())
        self._work = None

    @property
    def work_bytes(self):
        out = ctypes.c_int64()
        mkgpu.check("mk_filt_plan_work_bytes", self.h, ctypes.byref(out))
        return out.value

    def set_work(self, buf):
        """Use the device tensor buf (at least work_bytes bytes) as the work
        buffer; plans run one after another on one stream may share one."""
        nbytes = buf.numel() * buf.element_size()
        mkgpu.check("mk_filt_plan_set_work", self.h, ctypes.c_void_p(buf.data_ptr()), int(nbytes),
                    dev.stream_ptr())
        self._work = buf

    def borrow(self, seg, src, src_seg=0):
        mkgpu.check("mk_filt_plan_borrow_spectrum", self.h, int(seg), src.h, int(src_seg))
        self._keep.append(src)

    def exec(self, x, out):
        if self.elem_bytes == 4:
            mkgpu.check("mk_filt_plan_exec_f32", self.h, dev.ptr(x, torch.float32, "in"),
                        dev.ptr(out, torch.float32, "out"), dev.stream_ptr())
        else:
            mkgpu.check("mk_filt_plan_exec_f64", self.h, dev.ptr(x, torch.float64, "in"),
                        dev.ptr(out, torch.float64, "out"), dev.stream_ptr())

    def info(self):
        out = np.zeros(8, dtype=np.int64)
        mkgpu.check("mk_filt_plan_info", self.h, out.ctypes.data)
        keys = ["nseg", "items13", "items2", "work_bytes", "spec_bytes", "elem_bytes", "lds13", "lds2"]
        return dict(zip(keys, (int(v) for v in out)))

    def geometry(self):
        out = np.zeros(4 * max(1, len(self.segs)), dtype=np.int64)
        mkgpu.check("mk_filt_plan_geometry", self.h, out.ctypes.data)
        return out.reshape(-1, 4)[: len(self.segs)]


# Rotation arithmetic: "f16" (mode 2, three fp16-pair products) unless
# MK_ROT_KIND=bf16 asks for the split-BF16 modes.
ROT_KIND = os.environ.get("MK_ROT_KIND", "f16")


class SplitMatrix:
    """An FP64 matrix cut into planes for the MFMA rotation, padded to the
    kernel's tiles.  kind "f16": hi/lo fp16 planes of 2^aexp * a, with aexp
    chosen so that every entry lies below 2^14 (mode 2); kind "bf16":
    hi/mid/lo bf16 planes (modes 0 and 1)."""

    def __init__(self, a, kind=None):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.kind = kind or ROT_KIND
        if self.kind not in ("f16", "bf16"):
            raise ValueError("kind must be 'f16' or 'bf16', not %r" % (self.kind,))
        M, K = a.shape
        mp, ld = ctypes.c_int64(), ctypes.c_int64()
        mkgpu.fn("mk_split_bf16_padded_dims")(M, K, ctypes.byref(mp), ctypes.byref(ld))
        self.M, self.K, self.M_pad, self.lda = M, K, mp.value, ld.value
        d = dev.device()
        a_dev = torch.from_numpy(a).to(d)
        size = self.M_pad * self.lda
        if self.kind == "f16":
            amax = float(np.max(np.abs(a))) if a.size else 0.0
            if not np.isfinite(amax):
                raise ValueError("rotation matrix has non-finite entries")
            self.aexp = 14 - int(np.frexp(amax)[1]) if amax > 0 else 0   # amax < 2^(14 - aexp)
            self.hi = torch.empty(size, dtype=torch.float16, device=d)
            self.lo = torch.empty(size, dtype=torch.float16, device=d)
            mkgpu.check("mk_split_f16_matrix", a_dev.data_ptr(), M, K, K, self.aexp, self.hi.data_ptr(),
                        self.lo.data_ptr(), dev.stream_ptr())
        else:
            self.aexp = 0
            self.hi = torch.empty(size, dtype=torch.bfloat16, device=d)
            self.mid = torch.empty(size, dtype=torch.bfloat16, device=d)
            self.lo = torch.empty(size, dtype=torch.bfloat16, device=d)
            mkgpu.check("mk_split_bf16_matrix", a_dev.data_ptr(), M, K, K, self.hi.data_ptr(), self.mid.data_ptr(),
                        self.lo.data_ptr(), dev.stream_ptr())
        self._src = a_dev  # kept until the split kernel has run on the stream

    def desc_planes(self):
        """Fields 3-5 of the plan's segment descriptor."""
        if self.kind == "f16":
            return [self.hi.data_ptr(), self.lo.data_ptr(), self.aexp]
        return [self.hi.data_ptr(), self.mid.data_ptr(), self.lo.data_ptr()]


class RotPlan(_Handle):
    """Batched MFMA rotation C = A.B over segments."""

    _destroy = "mk_rot_plan_destroy"

    def __init__(self, segs, mode=None):
        """segs: list of (SplitMatrix, N, b_off, b_ld, c_off, c_ld).
        mode 0: 3 bf16 products; mode 1: 6 bf16 products; mode 2: 3 fp16
        products (modes 1 and 2 are about FP32-accurate).  The default is 2
        for fp16 planes and 1 for bf16 planes; every segment's planes must be
        of the mode's kind."""
        super().__init__()
        kinds = {sm.kind for sm, *_ in segs} or {ROT_KIND}
        if len(kinds) != 1:
            raise ValueError("segments mix fp16 and bf16 planes")
        kind = kinds.pop()
        if mode is None:
            mode = 2 if kind == "f16" else 1
        if (int(mode) == 2) != (kind == "f16"):
            raise ValueError("mode %d does not take %s planes" % (int(mode), kind))
        rows = []
        for sm, N, b_off, b_ld, c_off, c_ld in segs:
            rows.append([sm.M, sm.K, int(N)] + sm.desc_planes() + [sm.lda, int(b_off), int(b_ld), int(c_off),
                                                                  int(c_ld)])
            self._keep.append(sm)
        desc = np.asarray(rows, dtype=np.int64).reshape(-1)
        if desc.size == 0:
            desc = np.zeros(11, dtype=np.int64)
        self.mode = int(mode)
        mkgpu.check("mk_rot_plan_create", len(rows), desc.ctypes.data, self.mode, ctypes.byref(self.h))

    def exec(self, b, c):
        mkgpu.check("mk_rot_plan_exec", self.h, dev.ptr(b, torch.float32, "B"), dev.ptr(c, torch.float32, "C"),
                    dev.stream_ptr())


class DeflPlan(_Handle):
    """Deflation of dominant detector modes ahead of the FP32 rotation
    (mk_defl_plan_*): exec(x, out) writes out = x - U U^T x per segment and
    keeps c = U^T x; addback(y) adds c into the first k rows of y = v.out."""

    _destroy = "mk_defl_plan_destroy"

    def __init__(self, segs):
        """segs: list of (ndet, n, off, U) with U an ndet x k FP64 array."""
        super().__init__()
        rows, us = [], []
        for ndet, n, off, U in segs:
            U = np.ascontiguousarray(U, dtype=np.float64)
            if U.ndim != 2 or U.shape[0] != ndet:
                raise ValueError("U must be ndet x k")
            rows.append([int(ndet), int(n), int(off), int(U.shape[1])])
            us.append(U.ravel())
        self.nseg = len(rows)
        desc = np.asarray(rows, dtype=np.int64).reshape(-1)
        u = np.concatenate(us) if us else np.zeros(1)
        if desc.size == 0:
            desc = np.zeros(4, dtype=np.int64)
        mkgpu.check("mk_defl_plan_create", len(rows), desc.ctypes.data, u.ctypes.data, ctypes.byref(self.h))

    def exec(self, x, out):
        mkgpu.check("mk_defl_plan_exec", self.h, dev.ptr(x, torch.float32, "in"), dev.ptr(out, torch.float32, "out"),
                    dev.stream_ptr())

    def addback(self, y):
        mkgpu.check("mk_defl_plan_addback", self.h, dev.ptr(y, torch.float32, "y"), dev.stream_ptr())


class PixCode:
    """Coded pointing of an ipix arena for the tiled gather
    (mk_pix_code_*, mk_map2tod_coded_*): blocks of 256 samples holding their
    first pixel and per-sample offsets in 1, 2 or 4 bytes."""

    @staticmethod
    def sizes(ipix, nx, ny):
        """(units per block, total payload bytes) for ipix; raises on
        indices outside the nx x ny map."""
        n = int(ipix.numel())
        nb = int(mkgpu.lib.mk_pix_code_blocks(n))
        units = torch.empty(max(nb, 1), dtype=torch.int32, device=ipix.device)
        nbad = torch.zeros(1, dtype=torch.int64, device=ipix.device)
        if n:
            mkgpu.check("mk_pix_code_size", dev.ptr(ipix, torch.int32, "ipix"), n, int(nx), int(ny),
                        units.data_ptr(), nbad.data_ptr(), dev.stream_ptr())
        nb_bad = int(nbad.item())
        if nb_bad:
            raise ValueError("%d samples have pixel indices outside the %dx%d map" % (nb_bad, nx, ny))
        u = units[:nb].cpu().numpy().astype(np.int64)
        return units[:nb], int(u.sum()) * 256 + 8 * nb

    def __init__(self, ipix, nx, ny, units=None):
        n = int(ipix.numel())
        d = ipix.device
        if units is None:
            units, _ = PixCode.sizes(ipix, nx, ny)
        nb = int(units.numel())
        u = units.cpu().numpy().astype(np.int64)
        off = np.zeros(max(nb, 1), dtype=np.int64)
        if nb > 1:
            off[1:nb] = np.cumsum(u)[:-1]
        total = int(u.sum())
        if total >= 1 << 30:
            raise ValueError("pointing of %d samples too large for one coded stream" % n)
        self.nsamp, self.nx, self.ny, self.nblocks = n, int(nx), int(ny), nb
        self.hdr = torch.empty(2 * max(nb, 1), dtype=torch.int32, device=d)
        self.payload = torch.empty(max(total, 1) * 256, dtype=torch.uint8, device=d)
        if n:
            off_d = torch.from_numpy(off[:nb].astype(np.int32)).to(d)
            mkgpu.check("mk_pix_code_write", dev.ptr(ipix, torch.int32, "ipix"), n, self.ny, off_d.data_ptr(),
                        units.data_ptr(), self.hdr.data_ptr(), self.payload.data_ptr(), dev.stream_ptr())
            torch.cuda.current_stream().synchronize()  # off_d is freed on return
        self.nbytes = self.hdr.numel() * 4 + self.payload.numel()
        self.blocks_by_type = {t: int((u == k).sum()) for t, k in ((1, 1), (2, 2), (3, 4))}

    def gather(self, dat, tiled, tile, do_add=False):
        """dat (+)= tiled map gathered at the coded pixels (tile: its (tx, ty))."""
        name = {torch.float32: "mk_map2tod_coded_f32", torch.float64: "mk_map2tod_coded_f64"}[dat.dtype]
        if dat.numel() != self.nsamp or tiled.dtype != dat.dtype:
            raise ValueError("coded gather: %d samples of %s for a code of %d" % (dat.numel(), dat.dtype,
                                                                                 self.nsamp))
        mkgpu.check(name, dev.ptr(dat, name="dat"), tiled.data_ptr(), self.nsamp, self.hdr.data_ptr(),
                    self.payload.data_ptr(), self.ny, int(tile[0]), int(tile[1]), int(bool(do_add)),
                    dev.stream_ptr())
