"""Harvest, train and evaluate the learned preconditioner on one GPU.

    python -m minkasi_amd.learned.run --seeds 32 --ntod 64 \
        --out minkasi_amd/learned/weights_c4.npz [--eval]

Training problems use the 64-TOD, 2048^2 distribution with TOD seeds and
shapes disjoint from the benchmark problem (seed_base = 10000 + 1000 * seed,
shape seed 20000 + seed).  Each is solved with plain Jacobi PCG and its
(residual, error) pairs feed the energy loss (default; the problem's own
operator supplies A z) or the crop loss.  --eval then solves the benchmark
problem (seed_base 0) to ||r||/||b|| < tol with Jacobi and with the trained
network and prints one JSON line.
"""
import argparse
import json
import time

import numpy as np
import torch


def log(*a):
    print(*a, flush=True)


def parse(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=int, default=4)
    ap.add_argument("--arch", default="levels", choices=["levels", "tent", "chain"])
    ap.add_argument("--spec", default="1,4,16,64", help="level factors (levels) or dilations (chain)")
    ap.add_argument("--loss", default="energy", choices=["energy", "crop"])
    ap.add_argument("--seeds", type=int, default=32)
    ap.add_argument("--ntod", type=int, default=64, help="TODs per training problem")
    ap.add_argument("--steps-per-seed", type=int, default=60, help="energy loss: Adam steps per problem")
    ap.add_argument("--pairs", type=int, default=5000, help="crop loss: crops in the dataset")
    ap.add_argument("--crop", type=int, default=512)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--steps", type=int, default=1500, help="crop loss: Adam steps")
    ap.add_argument("--lr", type=float, default=1e-2)
    ap.add_argument("--init-frac", type=float, default=0.01, help="initial K^T K gain relative to alpha")
    ap.add_argument("--max-iter", type=int, default=150)
    ap.add_argument("--tol-final", type=float, default=5e-8)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    ap.add_argument("--init", default=None, help="start from these weights (.npz from save or .npy params)")
    ap.add_argument("--calibrate", default="", help="K^T K scale factors tried on a held-out problem, e.g. "
                    "0.25,0.5,1,2 (empty: none)")
    ap.add_argument("--calib-ntod", type=int, default=64, help="TODs of the held-out calibration problem")
    ap.add_argument("--calibrate-levels", action="store_true",
                    help="levels/tent: then scale each level's K^T K term by coordinate search on two held-out "
                         "problems")
    ap.add_argument("--level-steps", default="0.5,2",
                    help="factor pairs of the per-level search, coarse to fine, e.g. 0.5,2;0.7,1.4")
    ap.add_argument("--eval", action="store_true")
    ap.add_argument("--eval-ntod", type=int, default=None)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--maxiter", type=int, default=400)
    return ap.parse_args(argv)


class Dataset:
    """Crops (r, e, s, m) on the device, split into training and validation."""

    def __init__(self, R, E, S, M, val_frac=0.1, seed=0):
        N = R.shape[0]
        perm = torch.from_numpy(np.random.default_rng(seed).permutation(N)).to(R.device)
        nval = max(1, int(N * val_frac)) if N > 10 else 0
        self.val = [t[perm[:nval]] for t in (R, E, S, M)]
        self.tr = [t[perm[nval:]] for t in (R, E, S, M)]
        self.N = self.tr[0].shape[0]
        self.crop = R.shape[1]

    def set_margin(self, margin):
        """Loss weights restricted to pixels at least `margin` inside each crop."""
        C = self.crop
        inner = torch.zeros((C, C), dtype=torch.float32, device=self.tr[3].device)
        if C > 2 * margin:
            inner[margin:C - margin, margin:C - margin] = 1.0
        if not hasattr(self, "_m_full"):
            self._m_full = (self.tr[3], self.val[3])
        self.tr[3] = self._m_full[0] * inner
        self.val[3] = self._m_full[1] * inner

    def alpha_fit(self):
        """Closed-form alpha for K = 0: argmin sum m (alpha s^2 r - e)^2, and its loss per crop."""
        R, E, S, M = (t.double() for t in self.tr)
        s2r = S * S * R
        a = float((M * s2r * E).sum() / (M * s2r * s2r).sum())
        loss = float((M * (a * s2r - E) ** 2).sum() / self.N)
        return a, loss


def build_dataset(config, seeds, ntod, pairs, crop, margin, rng, max_iter, tol_final, shapes_seed0=20000,
                  seed_base0=10000, weight="hits"):
    from ..synth import tod_shapes
    from ..workloads import build_problem
    from .harvest import harvest
    per_seed = int(np.ceil(pairs / seeds))
    parts = []
    t0 = time.time()
    for s in range(seeds):
        shapes = tod_shapes(config, ntod, seed=shapes_seed0 + s)
        P = build_problem(config, shapes=shapes, seed_base=seed_base0 + 1000 * s)
        (R, E, S, M, ks), rel = harvest(P, per_seed, crop, margin, rng, max_iter=max_iter, tol_final=tol_final,
                                       weight=weight)
        parts.append((R, E, S, M))
        log("harvest seed %d: best rel res %.2e after %d iterations, %d crops from iterations %d..%d (%.0f s)" % (
            s, rel[-1], len(rel) - 1, R.shape[0], ks.min(), ks.max(), time.time() - t0))
        del P
        torch.cuda.empty_cache()
    out = [torch.cat([p[i] for p in parts]) for i in range(4)]
    return Dataset(*out)


def reach(arch, spec):
    """Pixels within which K^T K u depends on u."""
    if arch == "chain":
        return 2 * sum(spec)
    return 2 * max(spec) + max(spec)  # 3x3 taps of the coarsest level, plus its block


def init_scaled(arch, spec, alpha0, init_frac, seed):
    """Random weights with K^T K about init_frac * alpha in gain."""
    from .nets import param_layout
    from .precon import init_params
    lay, _ = param_layout(arch, len(spec))
    p0 = init_params(arch, spec, seed=seed, scale=1.0, log_alpha=np.log(alpha0))
    depth = len(spec) if arch == "chain" else 1
    gain = (init_frac * alpha0) ** (1.0 / (2 * depth))
    for off, co, ci in lay:
        p0[off:off + co * ci * 9] *= gain / np.sqrt(9.0 * ci * (1 if arch == "chain" else co))
    return p0


def train(ds, arch, spec, steps, lr, batch, seed=0, init_frac=0.3, log_every=100):
    """Crop-loss Adam on the training crops; returns (params numpy, history,
    meta).  The loss ignores pixels within the reach of K^T K of a crop's
    border, where the crop lacks inputs the full map has."""
    from .train import Trainer
    ds.set_margin(reach(arch, spec))
    alpha0, base_loss = ds.alpha_fit()
    p0 = init_scaled(arch, spec, alpha0, init_frac, seed)
    B, C = batch, ds.crop
    tr = Trainer(p0, arch, spec, B, C, C, lr=lr)
    R, E, S, M = ds.tr
    bufs = [torch.empty((B, 1, C, C), dtype=torch.float32, device=R.device) for _ in range(4)]
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    t1 = time.time()
    hist = []
    for it in range(steps):
        idx = torch.randint(0, ds.N, (B,), generator=g).to(R.device)
        for src, dst in zip((R, E, S, M), bufs):
            torch.index_select(src, 0, idx, out=dst.view(B, C, C))
        if it == int(steps * 0.6) or it == int(steps * 0.85):
            tr.lr *= 0.3
        tr.step(*bufs, gscale=1.0 / B)
        if log_every and (it % log_every == 0 or it == steps - 1):
            l = tr.loss() / B
            hist.append((it, l))
            log("  step %d loss/crop %.4f (alpha-only fit %.4f) alpha %.4g  %.1f ms/step" % (
                it, l, base_loss, float(np.exp(tr.params[-1].item())), 1000 * (time.time() - t1) / (it + 1)))
    torch.cuda.synchronize()
    params = tr.params.cpu().numpy()
    vloss = val_loss(ds, params, arch, spec)
    meta = dict(loss="crop", arch=arch, spec=list(spec), crops=ds.N, crop=C, steps=steps, batch=B, lr=lr,
                init_frac=init_frac, alpha_fit=alpha0, alpha_fit_loss=base_loss, val_loss=vloss,
                train_s=time.time() - t1)
    log("  trained: validation loss/crop %.4f (alpha-only fit on training crops %.4f)" % (vloss, base_loss))
    return params, hist, meta


def energy_pairs(P, max_iter, tol_final, min_ratio=10.0):
    """Full-map (r_k, e_k) pairs of one problem, each scaled by
    1/sqrt(e_k . r_k) = 1/||e_k||_A so that M = 0 has unit loss."""
    from .harvest import run_jacobi_history
    xs, rs, rel = run_jacobi_history(P, max_iter, tol_final)
    hit = P.hits.map > 0
    nhit = float(hit.sum().item())
    xstar = xs[-1]
    pairs = []
    for k in range(len(xs) - 1):
        if rel[k] < min_ratio * rel[-1]:
            continue
        e = xstar - xs[k]
        mean = float((e.to(torch.float64) * hit).sum().item()) / nhit
        e = torch.where(hit, e - mean, torch.zeros_like(e))
        er = float((e.to(torch.float64) * rs[k].to(torch.float64)).sum().item())
        if not er > 0:
            continue
        c = 1.0 / np.sqrt(er)
        pairs.append((rs[k] * c, e * c, k))
    del xs, rs
    return pairs, rel


def train_energy(config, seeds, ntod, steps_per_seed, arch, spec, lr, seed=0, init_frac=0.01, max_iter=150,
                 tol_final=5e-8, shapes_seed0=20000, seed_base0=10000, log_every=25, warmup=50, init=None):
    """Energy-norm training on harvested full-map pairs, one problem at a
    time (its operator supplies A z).  Returns (params, history, meta)."""
    from ..synth import tod_shapes
    from ..workloads import build_problem
    from .train import Trainer
    rng = np.random.default_rng(seed)
    tr = None
    hist = []
    total = seeds * steps_per_seed
    t0 = time.time()
    npairs = 0
    for si in range(seeds):
        shapes = tod_shapes(config, ntod, seed=shapes_seed0 + si)
        P = build_problem(config, shapes=shapes, seed_base=seed_base0 + 1000 * si)
        pairs, rel = energy_pairs(P, max_iter, tol_final)
        npairs += len(pairs)
        hits = P.hits.map
        hf = hits.to(torch.float64)
        smap = torch.where(hits > 0, torch.rsqrt(torch.where(hits > 0, hf, torch.ones_like(hf))),
                           torch.zeros_like(hf)).to(torch.float32).contiguous()
        op = P.op
        if tr is None:
            # alpha of the best pure Jacobi step on the first pair: (z.r)/(z.Az), z = J r
            r0 = pairs[0][0]
            z0 = smap * smap * r0
            az0 = torch.empty_like(z0)
            op.apply(z0, az0, accumulate=False)
            alpha0 = float((z0.double() * r0.double()).sum() / (z0.double() * az0.double()).sum())
            if init is not None:
                p0 = np.array(init, dtype=np.float32, copy=True)
            else:
                p0 = init_scaled(arch, spec, alpha0, init_frac, seed)
            tr = Trainer(p0, arch, spec, 1, hits.shape[0], hits.shape[1], lr=lr)
            # energy loss of the pure Jacobi step on the first pairs (the reference level)
            jl = []
            for r_, e_, _ in pairs[:: max(1, len(pairs) // 8)]:
                zj = alpha0 * smap * smap * r_
                azj = torch.empty_like(zj)
                op.apply(zj, azj, accumulate=False)
                jl.append(float(((zj - e_).double() * (azj - r_).double()).sum()))
            jac_loss = float(np.mean(jl))
            tr.energy_buffers()
            log("  alpha of the Jacobi step: %.4g; its energy loss %.4f" % (alpha0, jac_loss))
        op.capture(tr.z, tr.az, accumulate=False)

        def apply_A(z, out):
            op.apply(z, out, accumulate=False)

        acc, nacc = 0.0, 0
        for st in range(steps_per_seed):
            it = si * steps_per_seed + st
            tr.lr = lr * min(1.0, (it + 1) / float(warmup)) * (0.3 if it >= int(total * 0.6) else 1.0) * (
                0.3 if it >= int(total * 0.85) else 1.0)
            r, e, k = pairs[int(rng.integers(len(pairs)))]
            tr.step_energy(r, e, smap, apply_A)
            if log_every and (st % log_every == log_every - 1 or st == steps_per_seed - 1):
                l = tr.loss()
                acc += l
                nacc += 1
                hist.append((it, l))
        log("  seed %d: %d pairs (best rel res %.1e after %d its); mean energy loss %.4f; alpha %.4g; %.0f s" % (
            si, len(pairs), rel[-1], len(rel) - 1, acc / max(nacc, 1), float(np.exp(tr.params[-1].item())),
            time.time() - t0))
        op.graph = None
        del P, pairs, op
        torch.cuda.empty_cache()
    params = tr.params.cpu().numpy()
    meta = dict(loss="energy", arch=arch, spec=list(spec), seeds=seeds, ntod=ntod, pairs=npairs, steps=total, lr=lr,
                init_frac=init_frac, train_s=time.time() - t0, jacobi_step_loss=jac_loss,
                final_loss=float(np.mean([h[1] for h in hist[-8:]])) if hist else None)
    return params, hist, meta


def val_loss(ds, params, arch, spec):
    """Loss per crop on the validation crops."""
    from .train import Trainer
    R, E, S, M = ds.val
    n = R.shape[0]
    if n == 0:
        return float("nan")
    C = ds.crop
    tr = Trainer(params, arch, spec, 1, C, C)
    tot = 0.0
    for i in range(n):
        tr.forward_backward(*(t[i:i + 1].reshape(1, 1, C, C).contiguous() for t in (R, E, S, M)))
        tot += tr.loss()
    return tot / n


def scale_ktk(params, arch, spec, c):
    """Parameters with K^T K scaled by c, alpha kept."""
    from .nets import param_layout
    p = np.array(params, dtype=np.float32, copy=True)
    lay, _ = param_layout(arch, len(spec))
    depth = len(spec) if arch == "chain" else 1
    f = np.float32(c ** (1.0 / (2 * depth)))
    for off, co, ci in lay:
        p[off:off + co * ci * 9] *= f
    return p


def calibrate(P, params, arch, spec, tol, maxiter, factors=(0.25, 0.5, 1.0, 2.0, 4.0)):
    """Picks the K^T K scale (relative to alpha) that needs the fewest PCG
    iterations to tol on problem P (a held-out training problem): CG is
    invariant to the scale of M, not to the balance of its two terms."""
    from ..workloads import solve_to_tol
    from .precon import LearnedPrecon
    best = None
    rows = []
    for c in factors:
        pc = scale_ktk(params, arch, spec, c)
        res = solve_to_tol(P, LearnedPrecon(P.hits, pc, arch, spec), tol, maxiter)
        its = res["iters"] if res["reached"] else maxiter + 1
        rows.append((c, its))
        log("  calibrate: K^T K scale %.3g -> %d iterations" % (c, its))
        if best is None or its < best[1]:
            best = (c, its)
    return scale_ktk(params, arch, spec, best[0]), best[0], rows


def scale_levels(params, arch, spec, factors):
    """Parameters with level l's K^T K term scaled by factors[l] (levels and
    tent: K^T K is a sum of one term per level)."""
    from .nets import param_layout
    if arch == "chain":
        raise ValueError("per-level scales need the levels or tent architecture")
    p = np.array(params, dtype=np.float32, copy=True)
    lay, _ = param_layout(arch, len(spec))
    for (off, co, ci), c in zip(lay, factors):
        p[off:off + co * ci * 9] *= np.float32(np.sqrt(c))
    return p


def calibrate_levels(Ps, params, arch, spec, tol, maxiter, steps=((0.5, 2.0),), max_sweeps=3):
    """Per-level K^T K scales by coordinate search on the held-out problems
    Ps: each level's scale is multiplied by each factor of a step in turn
    and kept only when PCG then needs strictly fewer iterations to tol,
    summed over the problems (several problems and strict improvement keep
    the four scales from fitting one problem's iteration counts); sweeps
    repeat until one changes nothing.  Returns (params, factors, rows)."""
    from ..workloads import solve_to_tol
    from .precon import LearnedPrecon

    def score(c):
        pc = scale_levels(params, arch, spec, c)
        tot = 0
        for P in Ps:
            res = solve_to_tol(P, LearnedPrecon(P.hits, pc, arch, spec), tol, maxiter)
            tot += res["iters"] if res["reached"] else maxiter + 1
        return tot

    c = [1.0] * len(spec)
    best = score(c)
    rows = [(list(c), best)]
    log("  levels %s -> %d iterations over %d problems" % (c, best, len(Ps)))
    for fs in steps:
        for _ in range(max_sweeps):
            changed = False
            for lvl in range(len(spec)):
                for f in fs:
                    t = list(c)
                    t[lvl] *= f
                    sc = score(t)
                    rows.append((t, sc))
                    if sc < best:
                        best, c, changed = sc, t, True
                        log("  levels %s -> %d iterations" % (["%.3g" % v for v in c], sc))
            if not changed:
                break
    return scale_levels(params, arch, spec, c), c, rows


def evaluate(P, params, arch, spec, tol, maxiter, jac=None):
    from ..workloads import map_distance, solve_to_tol
    from .precon import LearnedPrecon
    if jac is None:
        jac = solve_to_tol(P, None, tol, maxiter)
    lp = LearnedPrecon(P.hits, params, arch, spec)
    lrn = solve_to_tol(P, lp, tol, maxiter)
    dist = map_distance(P, lrn["x"], jac["x"])
    out = dict(learned=dict((k, v) for k, v in lrn.items() if k not in ("x", "history")),
               jacobi=dict((k, v) for k, v in jac.items() if k not in ("x", "history")),
               map_distance=dist, speedup=jac["seconds"] / lrn["seconds"],
               iter_ratio=jac["iters"] / max(lrn["iters"], 1))
    del lrn, lp
    return out, jac


def main(argv=None):
    args = parse(argv)
    from ..workloads import build_problem
    from .precon import load, save
    torch.cuda.set_device(0)
    arch = args.arch
    spec = tuple(int(d) for d in args.spec.split(","))
    init = None
    if args.init:
        if args.init.endswith(".npz"):
            init, a0, s0, _ = load(args.init)
            if (a0, s0) != (arch, spec):
                raise ValueError("--init weights are %s%s, not %s%s" % (a0, s0, arch, spec))
        else:
            init = np.load(args.init, allow_pickle=False)
    if args.seeds == 0:
        if init is None:
            raise ValueError("--seeds 0 (no training) needs --init weights")
        params, hist = np.array(init, dtype=np.float32), []
        meta = dict(load(args.init)[3]) if args.init.endswith(".npz") else dict(loss="none", arch=arch,
                                                                                 spec=list(spec))
        hist = meta.pop("history", [])
        meta["init"] = args.init
    elif args.loss == "energy":
        params, hist, meta = train_energy(args.config, args.seeds, args.ntod, args.steps_per_seed, arch, spec,
                                          args.lr, args.seed, args.init_frac, args.max_iter, args.tol_final,
                                          init=init)
    else:
        rng = np.random.default_rng(args.seed)
        ds = build_dataset(args.config, args.seeds, args.ntod, args.pairs, args.crop, 0, rng, args.max_iter,
                           args.tol_final)
        log("dataset: %d training + %d validation crops of %d^2" % (ds.N, ds.val[0].shape[0], args.crop))
        params, hist, meta = train(ds, arch, spec, args.steps, args.lr, args.batch, args.seed, args.init_frac)
        del ds
    meta.update(history=hist)
    torch.cuda.empty_cache()
    if args.calibrate or args.calibrate_levels:
        from ..synth import tod_shapes
        # held out from both the training problems and the benchmark problem
        Pc = build_problem(args.config, shapes=tod_shapes(args.config, args.calib_ntod, seed=29999),
                           seed_base=9_000_000)
        if args.calibrate:
            factors = tuple(float(c) for c in args.calibrate.split(","))
            params, cbest, rows = calibrate(Pc, params, arch, spec, args.tol, args.maxiter, factors)
            meta.update(calibration=dict(factors=factors, iterations=[r[1] for r in rows], chosen=cbest))
            log("calibrated K^T K scale: %g" % cbest)
        if args.calibrate_levels:
            # a second held-out problem (other shapes, other TOD seeds)
            Pc2 = build_problem(args.config, shapes=tod_shapes(args.config, args.calib_ntod, seed=29998),
                                seed_base=9_100_000)
            steps = tuple(tuple(float(v) for v in st.split(",")) for st in args.level_steps.split(";"))
            params, lf, rows = calibrate_levels([Pc, Pc2], params, arch, spec, args.tol, args.maxiter, steps)
            cal = dict(factors=lf, steps=steps, problems=2, tried=[[r[0], r[1]] for r in rows])
            if "level_calibration" in meta:  # a further search on top of an earlier one
                prev = meta["level_calibration"]
                cal["after"] = prev
                cal["factors_total"] = [a * b for a, b in zip(prev.get("factors_total", prev["factors"]), lf)]
            meta.update(level_calibration=cal)
            log("calibrated per-level K^T K scales: %s" % lf)
            del Pc2
        del Pc
        torch.cuda.empty_cache()
    res = dict(train=meta)
    if args.eval:
        P = build_problem(args.config, ntod=args.eval_ntod)
        ev, _ = evaluate(P, params, arch, spec, args.tol, args.maxiter)
        res.update(ev)
        meta["eval"] = dict(ntod=len(P.shapes), jacobi_iters=ev["jacobi"]["iters"],
                            learned_iters=ev["learned"]["iters"], speedup=ev["speedup"])
        log("Jacobi: %d its %.2f s; learned: %d its %.2f s; speedup %.2fx; map distance %.2e" % (
            ev["jacobi"]["iters"], ev["jacobi"]["seconds"], ev["learned"]["iters"], ev["learned"]["seconds"],
            ev["speedup"], ev["map_distance"]))
    if args.out:
        save(args.out, params, arch, spec, meta)
        log("saved", args.out)
    print(json.dumps(res, default=float))


if __name__ == "__main__":
    main()
