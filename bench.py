"""Benchmark: batched PCG mapmaking on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 4]

Workload (default): 64 synthetic TODs, detector counts uniform in 64-512 and
lengths uniform in 10000-200000 samples, FP32, on a 2048 x 2048 map at
1 arcsec, NoiseSmoothedSVD noise (fwhm 50 bins), hits-count Jacobi PCG.
A step is one PCG iteration over all TODs: the batched A.p (map2tod, split-
BF16 MFMA rotation, FFT noise filter, rotation back, deterministic tod2map,
captured in one HIP graph) plus the fused FP64-scalar vector updates.

Prints ONE JSON line: value = PCG iterations per second (whole job), plus
the reference's own throughput definition (Msamp/s = sum ndet*n / seconds
per iteration, minkasi/mapmaking/pcg.py:191-198), time to ||r||/||b|| < tol
for the solvers, a roofline entry for the dominant kernel and a CPU baseline
measured with the FP64 oracle on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

METRIC = "time-to-‖r‖/‖b‖<1e-6 & PCG iters/s, learned vs Jacobi PCG, 64 TODs"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=4)
    ap.add_argument("--ntod", type=int, default=None)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--maxiter", type=int, default=600, help="cap for the solve-to-tolerance runs")
    ap.add_argument("--no-solve", action="store_true", help="skip the time-to-tolerance runs")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cuts", type=float, default=0.0, metavar="FRACTION",
                    help="also solve for cut samples: this fraction of the samples, in 200-sample segments on random "
                         "detectors, as a tsModel of CutsCompact next to the map (one flat PCG vector)")
    ap.add_argument("--no-tune", action="store_true",
                    help="default plans instead of the autotuner's (FFT column lengths, tod2map tiles)")
    ap.add_argument("--truncate", action="store_true",
                    help="cut every TOD so that n-1 is {2,3,5,7,11}-smooth first (truncate_tod)")
    ap.add_argument("--profile", action="store_true", help="short run for rocprofv3: setup + steps only")
    ap.add_argument("--weights", default=os.path.join(HERE, "minkasi_amd", "learned", "weights_c4.npz"),
                    help="trained learned-preconditioner weights")
    ap.add_argument("--chain-weights", default=os.path.join(HERE, "minkasi_amd", "learned", "weights_c4_chain.npz"),
                    help="weights of the 4-layer chain network, solved and reported beside --weights ('' to skip)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (counter calibration) and exit")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_bytes(P):
    """Algorithmic HBM bytes per launch of each stage (FP32 TOD, int32 ipix)."""
    S = P.nsamp
    npix = P.map.nx * P.map.ny
    return {
        "map2tod": 8 * S + 4 * npix,      # ipix in, TOD out (+ map once)
        "rot_v": 8 * S,                    # TOD in, rotated TOD out
        "filter": 12 * S,                  # rotated TOD in, mywt in, filtered out
        "rot_vT": 8 * S,
        "tod2map": 8 * S + 4 * npix,      # ipix in, TOD in (+ map once)
    }


def _largest_prime_factor(k):
    f, big = 2, 1
    while f * f <= k:
        while k % f == 0:
            big, k = f, k // f
        f += 1
    return max(big, k) if k > 1 else big


def cpu_baseline(P, budget_s=25.0, max_tods=12):
    """FP64 oracle (OpenMP projections, numpy GEMM, scipy DCT-I) timed on
    TODs of the workload drawn at random (fixed seed) until the budget is
    spent: the per-sample A.x cost of that sample, scaled to PCG
    iterations/s of the full job.  The DCT-I cost depends on how n - 1
    factors (scipy falls back to Bluestein for large primes), so the sample
    spans several lengths rather than one median TOD."""
    import torch
    from oracle import noise as onoise, solve
    # the box gives one GPU 16 host cores (OMP_NUM_THREADS is set to that
    # share there); a larger affinity mask here uses at most that many
    cores = min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "16") or 16), 16)
    os.environ["OMP_NUM_THREADS"] = str(cores)
    import scipy.fft
    order = np.random.default_rng(0).permutation(len(P.shapes))
    x = np.random.default_rng(0).standard_normal((P.map.nx, P.map.ny))
    t_start = time.time()
    picked, t_ax, s_ax, t_init = [], 0.0, 0, 0.0
    with scipy.fft.set_workers(cores):
        for k in order:
            if picked and (time.time() - t_start > budget_s or len(picked) >= max_tods):
                break
            ndet, n = P.shapes[k]
            tod = P.todvec.tods[k]
            d = tod.get_data()
            d = d.cpu().numpy().astype(np.float64) if isinstance(d, torch.Tensor) else np.asarray(d, np.float64)
            ipix = tod.info[P.map.tag].cpu().numpy().reshape(ndet, n)
            t0 = time.time()
            nz = onoise.NoiseSmoothedSVD(d, fwhm=50.0)
            t_init += time.time() - t0
            otod = solve.OTod(d, ipix, nz)
            t0 = time.time()
            solve.todvec_dot([otod], x)
            t_ax += time.time() - t0
            s_ax += ndet * n
            picked.append((ndet, n, _largest_prime_factor(n - 1)))
            del d, otod, nz
    per_sample = t_ax / s_ax
    return {
        "value": 1.0 / (per_sample * P.nsamp),
        "unit": "PCG iter/s",
        "cores": cores,
        "kind": "port",
        "sample": "FP64 oracle A.x on %d of the %d TODs drawn at random (seed 0; ndet x n [largest prime of n-1]: %s), "
                  "%.3g samples in %.2f s, scaled to the job's %.3g samples; noise init %.1f s untimed; %d threads, "
                  "the box's host-core share for one GPU; FFTW unavailable, scipy DCT-I" % (
                      len(picked), len(P.shapes), ", ".join("%dx%d [%d]" % p for p in picked), s_ax, t_ax,
                      P.nsamp, t_init, cores),
        "msamp_per_s": s_ax / t_ax / 1e6,
    }


TRAFFIC_FILE = os.path.join(HERE, "profiles", "pmc_traffic_pcg_64tod_2048.txt")


def measured_traffic(stage):
    """HBM bytes per launch of a stage from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, copy-calibrated), or None."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            parts = line.rstrip("\n").split("\t")
            if len(parts) == 3 and parts[0] == "stage" and parts[1] == stage:
                return int(float(parts[2]))
    return None


def copy_bandwidth(nbytes=1 << 31, reps=5):
    """Achievable HBM bandwidth on this box, GB/s of bytes read + written:
    the better of a device-to-device copy of nbytes (torch copy_) and the
    library's streaming elementwise product o = a * b (mk_mul_f32, 12 bytes
    per element), each event-timed over reps launches."""
    import torch
    from minkasi_amd import device as mdev
    from minkasi_amd.lib import mkgpu
    torch.cuda.empty_cache()
    free = int(torch.cuda.mem_get_info()[0])
    nbytes = min(nbytes, max(1 << 24, (free // 4) & ~((1 << 20) - 1)))  # three buffers within the free memory
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    c = torch.empty_like(a)

    def timed(fn):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / 1e3 / reps

    t_copy = timed(lambda: b.copy_(a))
    t_mul = timed(lambda: mkgpu.check("mk_mul_f32", c.data_ptr(), a.data_ptr(), b.data_ptr(), n, mdev.stream_ptr()))
    del a, b, c
    torch.cuda.empty_cache()
    return max(2.0 * nbytes / t_copy, 3.0 * nbytes / t_mul) / 1e9


def copy_probe(nbytes):
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": int(n * 4), "launches": 3}))


def add_cuts(P, frac, seg=200, seed=1234):
    """Cut model for the cuts variant: on every TOD, segments of `seg` samples
    on random detectors covering about `frac` of its samples.  Returns the
    flat operator, rhs, Jacobi weights [1/hits | 1], null mask and cut count."""
    import contextlib
    import numpy as np
    import torch
    from minkasi_amd.mapmaking.batch import BatchedCutsDot
    from minkasi_amd.mapmaking.timestream import tod_key, tsModel
    from minkasi_amd.tods.cuts import CutsCompact
    rng = np.random.default_rng(seed)
    cm = tsModel(P.todvec, CutsCompact)
    with contextlib.redirect_stdout(sys.stderr):  # get_imap prints the cut count, as the reference does
        for tod in P.todvec.tods:
            c = cm.data[tod_key(tod)]
            ndet, n = tod.get_data_dims()
            for _ in range(max(1, int(frac * ndet * n / seg))):
                d, t0 = int(rng.integers(ndet)), int(rng.integers(0, max(1, n - seg)))
                c.add_cut(d, t0, t0 + seg)
            c.merge_cuts()
            c.get_imap()
    op = BatchedCutsDot(P.op, cm)
    b = torch.zeros(op.n, dtype=torch.float32, device=P.op.buf_a.device)
    op.rhs(b, accumulate=False)
    npix = op.npix
    w = torch.ones(op.n, dtype=torch.float32, device=b.device)
    w[:npix] = P.precon.maps[0].map.reshape(-1)
    null = torch.zeros(op.n, dtype=torch.bool, device=b.device)
    null[:npix] = w[:npix] > 0
    return op, b, w, null, op.ncut


def main():
    args = parse()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    # stdout carries the one JSON result line; everything else (library
    # progress prints included) goes to stderr
    json_out, sys.stdout = sys.stdout, sys.stderr
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the batched solver holds every TOD on one device "
              "(sharding the TODs over several GPUs is designed but not built).", file=sys.stderr)
        sys.exit(2)
    import torch
    from minkasi_amd.mapmaking.pcg import DevicePCG
    from minkasi_amd.workloads import build_problem
    from minkasi_amd.synth import CONFIGS

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    t_setup = time.time()
    tuner, t_tune = None, 0.0
    if not args.no_tune:
        from minkasi_amd import tune
        t_tune = time.time()
        tuner = tune.enable()
        t_tune = time.time() - t_tune
        log("tune: FFT costs ready in %.1fs (cache %s)" % (t_tune, tuner.path))
    solve = not args.no_solve and not args.cuts and not args.profile
    # the solve-to-tolerance runs judge the true residual against the FP64
    # operator, which needs the noise models' FP64 filters
    P = build_problem(args.config, ntod=args.ntod, verbose=True, truncate=args.truncate, keep_fp64=solve)
    log("setup: generate %.1fs pixelize %.1fs noise %.1fs plans %.1fs rhs %.1fs; %d TODs, %.3g samples" % (
        P.t_generate, P.t_pixelize, P.t_noise, P.t_plans, P.t_rhs, len(P.shapes), P.nsamp))
    log("operator memory (GB):", json.dumps({k: (round(v / 1e9, 2) if k != "filter_groups" else v)
                                              for k, v in P.op.memory_report().items()}))
    op = P.op
    b = P.rhs.maps[0].map
    ncut = 0
    if args.cuts > 0:
        op, b, w, null, ncut = add_cuts(P, args.cuts)
        solver = DevicePCG(op, b, w, null_mask=null)
    else:
        solver = DevicePCG(op, b, P.precon)
    # one graph for A.p on the solver's own buffers
    x = torch.zeros_like(b)
    solver.init(x)
    if not args.no_graph:
        op.capture(solver.p, solver.Ap, accumulate=False)
    t_setup = time.time() - t_setup

    # ---- timed PCG steps (Jacobi)
    for _ in range(args.warmup):
        solver.step()
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(args.steps):
        solver.step()
    torch.cuda.synchronize()
    dt = time.time() - t0
    ms_per_step = 1000.0 * dt / args.steps
    iters_per_s = args.steps / dt
    rel_after = solver.rel_res()
    log("PCG: %.2f ms/iter, %.2f iter/s, %.0f Msamp/s; rel res after %d its %.3e" % (
        ms_per_step, iters_per_s, P.nsamp * iters_per_s / 1e6, args.warmup + args.steps, rel_after))

    # ---- per-stage times (events on the launch stream) and roofline
    pm = solver.p[:P.map.nx * P.map.ny].view(P.map.nx, P.map.ny)
    st = P.op.stage_times(pm, torch.empty_like(pm), reps=3)
    sb = stage_bytes(P)
    dom = max(st, key=st.get)
    # the committed traffic counts are of the default workload only
    profiled = (args.config, args.ntod, args.truncate, args.cuts) == (4, None, False, 0.0)
    achieved = sb[dom] / (st[dom] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom) if profiled else None,
            "algorithmic_bytes": int(sb[dom]), "stage_ms": {k: round(v, 3) for k, v in st.items()},
            "stage_frac": {k: round(sb[k] / (st[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) for k in st}}
    copy_gbs = copy_bandwidth()
    roof["copy_gbs"] = round(copy_gbs, 1)
    roof["frac_of_copy"] = round(achieved / copy_gbs, 4)
    roof["stage_frac_of_copy"] = {k: round(sb[k] / (st[k] * 1e-3) / 1e9 / copy_gbs, 4) for k in st}
    log("stage ms:", json.dumps(roof["stage_ms"]), "frac:", json.dumps(roof["stage_frac"]))
    log("device copy: %.0f GB/s (%.2f of the 8 TB/s peak); stage frac of copy: %s" % (
        copy_gbs, copy_gbs / HBM_PEAK_GBS, json.dumps(roof["stage_frac_of_copy"])))

    result = {
        "metric": METRIC, "value": round(iters_per_s, 4), "unit": "PCG iter/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "config%d: %d TODs, ndet %s, n %s, %dx%d map at %g arcsec, NoiseSmoothedSVD, "
                               "Jacobi PCG%s" % (args.config, len(P.shapes), CONFIGS[args.config]["ndet"],
                                                 CONFIGS[args.config]["n"], P.map.nx, P.map.ny,
                                                 CONFIGS[args.config]["pix"],
                                                 (", truncate_tod lengths" if args.truncate else "") +
                                                 (", %d cut samples (%.2g%%) solved with the map" %
                                                  (ncut, 100.0 * ncut / P.nsamp) if ncut else "")),
                   "ntod": len(P.shapes), "nsamp": int(P.nsamp), "map": [P.map.nx, P.map.ny]},
        "msamp_per_s": round(P.nsamp * iters_per_s / 1e6, 1),
        "setup_s": round(t_setup, 1),
        "roofline": roof,
    }
    free, total = torch.cuda.mem_get_info()
    result["config"]["memory"] = {"hbm_used_gb": round((total - free) / 1e9, 1), "hbm_gb": round(total / 1e9, 1),
                                  "tiled_gather": bool(P.op.tiled), "filter_groups": len(P.op.filts)}
    if tuner is not None:
        result["config"]["tuning"] = dict(tuner.report(), fft_measure_s=round(t_tune, 1))
    if args.profile:
        print(json.dumps(result), file=json_out, flush=True)
        return

    # ---- time to a true residual ||b - Ax||/||b|| < tol: Jacobi, then the
    # learned preconditioner.  b is the FP64 right-hand side (every stage in
    # FP64, setup); in the timed solve the iterate accumulates in FP64, the
    # recursive residual is replaced once from the FP64-binned operator and
    # the stop is accepted only when that operator's true residual is below
    # tol.  The all-FP64 operator then re-evaluates the final iterate
    # (true_rel_res_fp64, outside the timed region).
    if solve:
        from minkasi_amd.workloads import map_distance, solve_to_tol
        res = {}
        t_acc = time.time()
        acc = P.op.accurate()
        b64 = torch.zeros((P.map.nx, P.map.ny), dtype=torch.float64, device=dev)
        acc.rhs(b64)
        chk = P.op.f64_view()
        torch.cuda.synchronize()
        res["fp64_rhs_s"] = round(time.time() - t_acc, 2)
        jac = solve_to_tol(P, None, args.tol, args.maxiter, check=chk, b64=b64, verify=acc)
        res["jacobi"] = {k: (round(v, 3) if k == "seconds" else v) for k, v in jac.items()
                         if k not in ("x", "history")}
        log("Jacobi to tol:", json.dumps(res["jacobi"]))
        if not args.no_learned and os.path.exists(args.weights):
            from minkasi_amd.learned.precon import LearnedPrecon, load
            params, arch, spec, meta = load(args.weights)
            lp = LearnedPrecon(P.hits, params, arch, spec)
            lrn = solve_to_tol(P, lp, args.tol, args.maxiter, check=chk, b64=b64, verify=acc)
            res["learned"] = {k: (round(v, 3) if k == "seconds" else v) for k, v in lrn.items()
                              if k not in ("x", "history")}
            res["learned"].update(arch=arch, spec=list(spec), weights=os.path.relpath(args.weights, HERE))
            res["speedup_wall"] = round(jac["seconds"] / lrn["seconds"], 3)
            res["iteration_ratio"] = round(jac["iters"] / max(lrn["iters"], 1), 3)
            res["map_distance"] = map_distance(P, lrn["x"], jac["x"])
            log("learned to tol:", json.dumps(res["learned"]), "speedup %.2fx" % res["speedup_wall"])
            # the network the workload names (4 dilated 3x3 layers in series,
            # 16 channels), reported beside the shipped multilevel net
            if args.chain_weights and os.path.exists(args.chain_weights):
                params, arch, spec, meta = load(args.chain_weights)
                lc = LearnedPrecon(P.hits, params, arch, spec)
                ch = solve_to_tol(P, lc, args.tol, args.maxiter, check=chk, b64=b64, verify=acc)
                res["learned_chain"] = {k: (round(v, 3) if k == "seconds" else v) for k, v in ch.items()
                                        if k not in ("x", "history")}
                res["learned_chain"].update(arch=arch, spec=list(spec),
                                            weights=os.path.relpath(args.chain_weights, HERE),
                                            speedup_wall=round(jac["seconds"] / ch["seconds"], 3))
                log("chain to tol:", json.dumps(res["learned_chain"]))
        elif not args.no_learned:
            res["learned"] = {"error": "weights file %s missing" % args.weights}
        result["time_to_tol"] = res
    if not args.no_cpu and not ncut:
        try:
            result["cpu_baseline"] = cpu_baseline(P)
        except Exception as e:  # reported, not fatal
            result["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(result), file=json_out, flush=True)


if __name__ == "__main__":
    main()
