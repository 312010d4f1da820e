"""Autotuner for the batched operator's plan choices, with an on-disk cache.

It takes the role FFTW wisdom plays in the reference (minkasi/tools/fft.py:
347-384): transform plans chosen by measurement, remembered across runs.
Three choices are tuned on the device:

  * FFT column length of the noise filter.  The filter's real FFT has
    L = 2^11 * P1 points with P1 a 5-smooth number in [2, 256].  By default
    the smallest P1 with L >= 3(n-1)+1 is taken.  The tuner times the
    filter once per hot P1 (64M-point batches), stores each P1's cost per
    point, and then picks for each row length n the P1 that minimises
    P1 * cost(P1) among the lengths that fit.  A longer transform is still
    exact, because the convolution only needs L >= 3(n-1)+1.
  * map2tod gather tile: the batched operator gathers from a copy of the map
    laid out in tx x ty pixel tiles; the tile shape is timed on the arena.
  * tod2map plan tile and item size.  Candidate plans are built on the
    arena's own pixelization and their exec is timed.  The winner is
    remembered for that pixelization (map size, sample count and a checksum
    of a strided sample of ipix).

The cache is a JSON file at $MK_TUNE_CACHE or ~/.cache/minkasi_amd/tune.json,
keyed by GPU architecture, CU count and library ABI version.  Tuning is opt-in:
tune.enable() (bench.py does this unless --no-tune) makes later FiltPlans
and BatchedDots use it.  Plans built before it keep their default choices,
and BatchedDot copies each TOD's column length from that TOD's noise model.
"""
import json
import math
import os
import time

import numpy as np

_active = None

T2M_TILES = ((32, 32), (32, 16), (64, 16), (16, 32), (16, 16))
HOT_P2 = 2048


def cache_path():
    return os.environ.get("MK_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "minkasi_amd",
                                                           "tune.json")


def active():
    """The enabled Tuner, or None."""
    return _active


def enable(cache=None, retune=False, verbose=False):
    """Loads (or measures) the tuning data for this device and makes it the
    active tuner.  Returns it."""
    global _active
    t = Tuner(cache or cache_path(), verbose=verbose)
    if retune or not t.has_fft_costs():
        t.measure_fft_costs()
        t.save()
    _active = t
    return t


def disable():
    global _active
    _active = None


def p1_minimum(n):
    """Smallest column count need with 2^11 * need >= the filter's minimum
    complex length for rows of n samples (0 when the row is too short for
    the 2^11-point hot rows)."""
    m = n - 1
    p_min = (3 * m + 2) // 2
    if p_min <= HOT_P2:
        return 0
    return -(-p_min // HOT_P2)


def choose_p1(n, hot, cost):
    """The hot column length minimising P1 * cost(P1) among those that fit
    rows of n samples (0 = default, for rows the hot path does not take).
    cost: {P1: seconds per point}; lengths without a cost are skipped."""
    need = p1_minimum(n)
    if need == 0:
        return 0
    best, best_t = 0, math.inf
    for p in hot:
        if p < need or p not in cost:
            continue
        t = p * cost[p]
        if t < best_t * (1 - 1e-9):
            best, best_t = p, t
    if best and best == min(p for p in hot if p >= need):
        return 0  # the default already
    return best


def ipix_signature(ipix, nx, ny):
    """Key of a pixelization: map size, sample count and a checksum of 65536
    strided samples (device to host copy of 256 KB)."""
    n = int(ipix.numel())
    flat = ipix.reshape(-1)
    step = max(1, n // 65536)
    sample = flat[::step][:65536].cpu().numpy().astype(np.int64)
    chk = int((sample * (np.arange(len(sample), dtype=np.int64) % 1000003 + 1)).sum() % (1 << 61))
    return "%dx%d:%d:%x" % (nx, ny, n, chk)


class Tuner:
    def __init__(self, path, verbose=False):
        self.path = path
        self.verbose = verbose
        self.key = self._device_key()
        self.data = {}
        try:
            with open(path) as f:
                self.data = json.load(f)
        except (OSError, ValueError):
            self.data = {}
        self.entry = self.data.setdefault(self.key, {})
        self.hot = self._hot_list()
        self.fft_cost = {int(k): float(v) for k, v in self.entry.get("fft_cost", {}).items()}
        self.chosen = {}

    # -------------------------------------------------------------- keys, cache
    @staticmethod
    def _device_key():
        """Architecture, CU count and library ABI (the marketing name is not
        used: it can be empty, e.g. under a profiler)."""
        import torch
        from .lib import mkgpu
        if torch.cuda.is_available():
            pr = torch.cuda.get_device_properties(0)
            name = "%s|cu%d" % (getattr(pr, "gcnArchName", "") or pr.name, pr.multi_processor_count)
        else:
            name = "cpu"
        return "%s|abi%d" % (name, int(mkgpu.lib.mk_version()))

    @staticmethod
    def _hot_list():
        from .lib import mkgpu
        n = int(mkgpu.lib.mk_filt_hot_p1(None, 0))
        out = np.zeros(n, dtype=np.int32)
        mkgpu.lib.mk_filt_hot_p1(out.ctypes.data, n)
        return [int(v) for v in out]

    def has_fft_costs(self):
        return all(p in self.fft_cost for p in self.hot)

    def save(self):
        self.entry["fft_cost"] = {str(k): v for k, v in sorted(self.fft_cost.items())}
        try:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            tmp = self.path + ".tmp%d" % os.getpid()
            with open(tmp, "w") as f:
                json.dump(self.data, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError as e:  # read-only home: keep the results in memory
            if self.verbose:
                print("tune: cache not written (%s)" % e)

    # -------------------------------------------------------------- FFT
    def measure_fft_costs(self, points=1 << 26, reps=5):
        """Seconds per transform point of the FP32 filter for every hot P1
        (the first length is timed twice so the clocks have settled)."""
        import torch
        from . import device as dev
        from .lib.plans import FiltPlan
        if os.environ.get("MK_FFT_NO_HOT"):
            return  # no hot path: nothing to choose
        d = dev.device()
        for p in [self.hot[0]] + list(self.hot):
            n = (2 * HOT_P2 * p - 2) // 3 + 1          # 2^11 p >= the minimum for n
            if p1_minimum(n) == 0:
                n = (2 * HOT_P2 * 2 - 2) // 3 + 1      # shortest rows on the hot path
            rows = max(8, points // (p * HOT_P2))
            plan = FiltPlan([(rows, n, 0, n, 0, n)], 4, p1=[p if p >= p1_minimum(n) else 0])
            w = torch.ones((rows, n), dtype=torch.float64, device=d)
            plan.set_weights(w)
            del w
            x = torch.ones(rows * n, dtype=torch.float32, device=d)
            plan.exec(x, x)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                plan.exec(x, x)
            e1.record()
            e1.synchronize()
            sec = e0.elapsed_time(e1) / 1e3 / reps
            self.fft_cost[p] = sec / (rows * p * HOT_P2)
            plan.close()
            del x
            if self.verbose:
                print("tune: fft P1=%3d  %.3f ns/point" % (p, self.fft_cost[p] * 1e9))

    def fft_p1(self, n):
        """Column length for rows of n samples (0 = the default)."""
        p = choose_p1(int(n), self.hot, self.fft_cost)
        if p:
            self.chosen["fft_p1_changed"] = self.chosen.get("fft_p1_changed", 0) + 1
        return p

    # -------------------------------------------------------------- tod2map
    def t2m_plan(self, ipix, nx, ny, dat=None, reps=3):
        """A T2MPlan for ipix with the fastest tile / item size for this
        pixelization (timed, or from the cache).  dat: a scratch FP32 buffer
        of ipix.numel() values to time with (overwritten)."""
        import torch
        from .lib.plans import T2MPlan
        sig = ipix_signature(ipix, nx, ny)
        t2m = self.entry.setdefault("t2m", {})
        if sig in t2m:
            c = t2m[sig]
            self.chosen["t2m"] = c
            return T2MPlan(ipix, nx, ny, tile=tuple(c["tile"]), item_samples=c["item_samples"])
        if dat is None:
            dat = torch.empty(int(ipix.numel()), dtype=torch.float32, device=ipix.device)
        dat.fill_(1.0)
        out = torch.zeros(nx * ny, dtype=torch.float32, device=ipix.device)

        # timed with the FP32 sums the PCG iteration's A.x uses
        def timed(plan):
            plan.exec(out, dat, False, acc64=False)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                plan.exec(out, dat, False, acc64=False)
            e1.record()
            e1.synchronize()
            return e0.elapsed_time(e1) / reps

        # one candidate plan alive at a time (a plan holds about 2.3 bytes
        # per sample, its build several times that); the winner is built
        # again at the end
        best, best_ms, results = None, math.inf, []
        for tile in T2M_TILES:
            plan = T2MPlan(ipix, nx, ny, tile=tile)
            ms = timed(plan)
            results.append((tile, plan.item_samples, ms))
            if ms < best_ms:
                best, best_ms = (tuple(tile), plan.item_samples), ms
            plan.close()
        base = best[1]
        for items in (base // 2, base * 2):
            if items < 1024 or items % 64:
                continue
            plan = T2MPlan(ipix, nx, ny, tile=best[0], item_samples=items)
            ms = timed(plan)
            results.append((best[0], items, ms))
            if ms < best_ms:
                best, best_ms = (best[0], items), ms
            plan.close()
        c = {"tile": list(best[0]), "item_samples": int(best[1]), "ms": best_ms,
             "tried": [[list(t), int(i), m] for t, i, m in results], "when": time.strftime("%Y-%m-%d %H:%M:%S")}
        t2m[sig] = c
        self.chosen["t2m"] = c
        self.save()
        if self.verbose:
            for t, i, m in results:
                print("tune: t2m tile %s items %d  %.3f ms" % (t, i, m))
        return T2MPlan(ipix, nx, ny, tile=best[0], item_samples=best[1])

    # -------------------------------------------------------------- generic
    def choose(self, kind, sig, candidates, time_fn):
        """The fastest of candidates for (kind, sig): from the cache, or by
        calling time_fn(candidate) -> milliseconds for each and caching the
        winner.  Candidates are JSON-serialisable (lists are kept as lists)."""
        table = self.entry.setdefault(kind, {})
        if sig in table:
            c = table[sig]
            self.chosen[kind] = c
            return c["choice"]
        tried = []
        for cand in candidates:
            tried.append([cand, float(time_fn(cand))])
        best = min(tried, key=lambda t: t[1])
        c = {"choice": best[0], "ms": best[1], "tried": tried, "when": time.strftime("%Y-%m-%d %H:%M:%S")}
        table[sig] = c
        self.chosen[kind] = c
        self.save()
        if self.verbose:
            for cand, ms in tried:
                print("tune: %s %s  %.3f ms" % (kind, cand, ms))
        return best[0]

    def report(self):
        """What the tuner chose, for the benchmark line."""
        out = {"cache": self.path, "device": self.key}
        if "t2m" in self.chosen:
            out["t2m_tile"] = self.chosen["t2m"]["tile"]
            out["t2m_item_samples"] = self.chosen["t2m"]["item_samples"]
        out["fft_p1_changed"] = self.chosen.get("fft_p1_changed", 0)
        if "gather_tile" in self.chosen:
            out["gather_tile"] = self.chosen["gather_tile"]["choice"]
        return out
