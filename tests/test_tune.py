"""Autotuner host logic (minkasi_amd/tune.py): the column-length rule
against the library's own default geometry, the cost-based choice and the
on-disk cache.  No GPU needed."""
import json

import numpy as np
import pytest

from minkasi_amd import tune


def _hot():
    return tune.Tuner._hot_list()


def test_hot_list_is_the_5_smooth_numbers_up_to_256():
    want = [p for p in range(2, 257) if all(f in (2, 3, 5) for f in _factors(p))]
    assert _hot() == want


def _factors(p):
    out, d = [], 2
    while p > 1:
        while p % d == 0:
            out.append(d)
            p //= d
        d += 1
    return out


def test_p1_minimum_matches_the_convolution_length_rule():
    # 2^11 * need >= (3(n-1)+2)//2 > 2^11 * (need-1): the rule the library's geometry uses
    for n in (2, 2731, 2732, 10000, 65536, 200000, 349525, 349526):
        need = tune.p1_minimum(n)
        pmin = (3 * (n - 1) + 2) // 2
        if need == 0:
            assert pmin <= 2048
        else:
            assert 2048 * need >= pmin > 2048 * (need - 1)


def test_choose_p1_uses_costs_and_keeps_the_default_when_it_is_cheapest():
    hot = _hot()
    flat = {p: 1.0 for p in hot}
    for n in (5000, 65536, 200000):
        assert tune.choose_p1(n, hot, flat) == 0            # equal per-point cost: the smallest length
    n = 200000                                               # need = 147 -> default 150
    assert tune.p1_minimum(n) == 147
    cost = dict(flat)
    cost[150] = 1.2                                          # 150 * 1.2 = 180 > 160 * 1.0
    assert tune.choose_p1(n, hot, cost) == 160
    cost[160] = 1.2
    cost[162] = 1.0
    assert tune.choose_p1(n, hot, cost) == 162
    assert tune.choose_p1(1000, hot, cost) == 0              # rows below the hot path
    assert tune.choose_p1(n, hot, {}) == 0                   # no data: default


def test_cache_round_trip(tmp_path, monkeypatch):
    path = str(tmp_path / "sub" / "tune.json")
    t = tune.Tuner(path)
    assert not t.has_fft_costs()
    t.fft_cost = {p: 1e-10 * (1 + (p % 7) / 10) for p in t.hot}
    t.entry.setdefault("t2m", {})["64x64:100:abc"] = {"tile": [32, 64], "item_samples": 2048, "ms": 1.0}
    t.save()
    with open(path) as f:
        d = json.load(f)
    assert t.key in d and len(d[t.key]["fft_cost"]) == len(t.hot)
    t2 = tune.Tuner(path)
    assert t2.has_fft_costs()
    assert all(abs(t2.fft_cost[p] - t.fft_cost[p]) < 1e-25 for p in t.hot)
    assert t2.entry["t2m"]["64x64:100:abc"]["tile"] == [32, 64]
    monkeypatch.setenv("MK_TUNE_CACHE", path)
    assert tune.cache_path() == path


def test_flat_namespace_and_single_rank_names():
    """minkasi_amd.minkasi_all gives the names legacy scripts import, and the
    MPI bootstrap names describe one process (rank 0 of 1)."""
    import minkasi_amd as mk
    import minkasi_amd.minkasi_all as flat
    for name in ("Tod", "TodVec", "SkyMap", "PolMap", "Mapset", "NoiseSmoothedSVD", "run_pcg", "run_pcg_wprior",
                 "make_hits", "tsModel", "CutsCompact", "truncate_tod", "downsample_tod", "fit_cm_plus_poly",
                 "fft_r2r", "myrank", "nproc", "barrier", "comm", "read_tod_from_fits", "read_tod_from_fits_cbass",
                 "read_tod_from_fits_NIKA2", "gapfill_lin", "todvec_from_files_octave", "cut_blacklist"):
        assert hasattr(flat, name), name
    for name in ("read_tod_from_fits", "read_tod_from_fits_cbass", "gapfill_lin", "cut_blacklist", "Tod", "run_pcg"):
        assert getattr(mk, name) is getattr(flat, name), name
    assert (mk.myrank, mk.nproc, mk.have_mpi, mk.comm) == (0, 1, False, None)
    assert mk.get_nthread() >= 1 and mk.set_nthread(4) is None
    mk.barrier()  # no device work queued, no GPU needed


def test_t2m_tile_candidates_fit_the_plan():
    """Every tod2map tile the autotuner may pick is one the plan accepts: at
    most 1024 pixels (a lane word keeps a 10-bit local pixel) and tile_y a
    multiple of 4 (the per-tile sum reads 4 pixels of one tile row)."""
    from minkasi_amd import tune
    for tx, ty in tune.T2M_TILES:
        assert 4 <= tx * ty <= 1024 and ty % 4 == 0, (tx, ty)
