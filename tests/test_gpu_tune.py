"""The autotuner on the device: it measures every hot FFT column length,
caches the result, and a batched operator built with its choices applies
the same A as the default plans."""
import time

import numpy as np
import pytest

from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def test_autotuner_measures_caches_and_keeps_the_operator(tmp_path, monkeypatch):
    import torch
    from minkasi_amd import tune
    from minkasi_amd.workloads import build_problem
    monkeypatch.setenv("MK_TUNE_CACHE", str(tmp_path / "tune.json"))
    t = tune.enable()
    try:
        assert t.has_fft_costs() and all(c > 0 and np.isfinite(c) for c in t.fft_cost.values())
        t0 = time.time()
        t2 = tune.enable()  # from the cache: no measurement
        assert time.time() - t0 < 5 and t2.fft_cost == t.fft_cost
        for n in (3000, 40000, 160000):
            p = t2.fft_p1(n)
            assert p == 0 or (p in t2.hot and p >= tune.p1_minimum(n))
        P = build_problem(1, ntod=3, verbose=False)
        assert "t2m" in t2.chosen
        tile = tuple(t2.chosen["t2m"]["tile"])
    finally:
        tune.disable()
    x = torch.randn(P.map.nx, P.map.ny, device=P.rhs.maps[0].map.device)
    y_tuned = torch.zeros_like(x)
    P.op.apply(x, y_tuned, accumulate=False)
    Q = build_problem(1, ntod=3, verbose=False)  # default plans
    y_def = torch.zeros_like(x)
    Q.op.apply(x, y_def, accumulate=False)
    assert rel_err(y_tuned.cpu().numpy(), y_def.cpu().numpy()) < 2e-6, tile
    # a second build with the tuner finds the tod2map choice in the cache
    t3 = tune.enable()
    try:
        build_problem(1, ntod=3, verbose=False)
        assert t3.chosen["t2m"]["tile"] == list(tile)
    finally:
        tune.disable()
