"""(residual, error) pairs harvested from plain Jacobi PCG runs.

For a problem with right-hand side b, Jacobi PCG from x_0 = 0 runs until its
residual falls below ``tol_final`` (or ``max_iter``); its last iterate
stands in for x*.  Iteration k then gives the pair (r_k, e_k = x* - x_k),
with A e_k = r_k; e_k is projected off the null mode of A (the constant
on hit pixels) by subtracting its mean over hit pixels.  Only iterations
whose residual is at least ``min_ratio`` times the final one are used, so
the stand-in's own error stays negligible.

Each pair contributes square crops centred on random hit pixels.  A crop
carries r, e, s = sqrt(1/hits) and a loss weight m on hit pixels at least
``margin`` pixels inside the crop (where a K^T K of that reach sees the same
inputs as on the full map), 0 elsewhere.  With weight="hits" (default) m is
the hit count, so sum m (M r - e)^2 = ||G (s r) - e / s||^2 with
G = alpha I + K^T K: the fit is measured in the Jacobi-normalised space
where every pixel counts alike; weight="unit" measures it on the map itself
(which weights low-hit edge pixels by 1/hits).  r and e are scaled by
1/sqrt(sum m e^2), so every crop weighs the same in the loss.
"""
import numpy as np
import torch

from ..mapmaking.pcg import DevicePCG, PCGBreakdown


def run_jacobi_history(P, max_iter, tol_final, graph=True):
    """Iterates, residuals and relative residuals of Jacobi PCG.  A
    breakdown (p.Ap <= 0, which FP32 rounding produces once the residual is
    at the precision floor) ends the run at the last good iterate."""
    op, b = P.op, P.rhs.maps[0].map
    solver = DevicePCG(op, b, P.precon)
    x = torch.zeros_like(b)
    solver.init(x)
    if graph:
        op.capture(solver.p, solver.Ap, accumulate=False)
    xs, rs, rel = [x.clone()], [solver.r.clone()], [1.0]
    for it in range(max_iter):
        solver.step()
        try:
            rr = solver.rel_res()
        except PCGBreakdown as e:
            print("harvest: %s at iteration %d (last rel res %.3e); stopping there" % (e, it, rel[-1]), flush=True)
            break
        xs.append(solver.x.clone())
        rs.append(solver.r.clone())
        rel.append(rr)
        if rr < tol_final or rr > 10 * min(rel):
            break
    kbest = int(np.argmin(rel))  # the most accurate iterate stands in for x*
    return xs[:kbest + 1], rs[:kbest + 1], np.array(rel[:kbest + 1])


def crops_from_history(P, xs, rs, rel, ncrop_total, crop, margin, rng, min_ratio=10.0, weight="hits"):
    hits = P.hits.map
    N0, N1 = hits.shape
    if crop > min(N0, N1):
        raise ValueError("crop %d larger than the map %s" % (crop, tuple(hits.shape)))
    hit = hits > 0
    hf = hits.to(torch.float64)
    s_full = torch.where(hit, torch.rsqrt(torch.where(hit, hf, torch.ones_like(hf))),
                         torch.zeros_like(hf)).to(torch.float32)
    xstar = xs[-1]
    usable = [k for k in range(len(xs) - 1) if rel[k] >= min_ratio * rel[-1]]
    if not usable:
        raise RuntimeError("no usable iterations (final rel res %.3g)" % rel[-1])
    hit_idx = torch.nonzero(hit.flatten()).flatten().cpu().numpy()
    nhit = float(hit.sum().item())
    ks = rng.choice(usable, size=ncrop_total, replace=True)
    ks.sort()
    d = hits.device
    R = torch.empty((ncrop_total, crop, crop), dtype=torch.float32, device=d)
    E = torch.empty_like(R)
    S = torch.empty_like(R)
    M = torch.empty_like(R)
    inner = torch.zeros((crop, crop), dtype=torch.bool, device=d)
    if crop > 2 * margin:
        inner[margin:crop - margin, margin:crop - margin] = True
    last_k, e_full = None, None
    for i, k in enumerate(ks):
        if k != last_k:
            e_full = xstar - xs[k]
            mean = float((e_full.to(torch.float64) * hit).sum().item()) / nhit
            e_full = torch.where(hit, e_full - mean, torch.zeros_like(e_full))
            last_k = k
        c = int(hit_idx[rng.integers(len(hit_idx))])
        ci, cj = c // N1, c % N1
        i0 = min(max(ci - crop // 2, 0), N0 - crop)
        j0 = min(max(cj - crop // 2, 0), N1 - crop)
        sl = (slice(i0, i0 + crop), slice(j0, j0 + crop))
        R[i] = rs[k][sl]
        E[i] = e_full[sl]
        S[i] = s_full[sl]
        mk = (hit[sl] & inner).to(torch.float32)
        M[i] = mk * hits[sl].to(torch.float32) if weight == "hits" else mk
    norm = torch.sqrt((M.to(torch.float64) * E.to(torch.float64) ** 2).sum(dim=(1, 2)))
    ok = norm > 0
    scale = torch.where(ok, 1.0 / torch.where(ok, norm, torch.ones_like(norm)), torch.zeros_like(norm))
    scale = scale.to(torch.float32)[:, None, None]
    R.mul_(scale)
    E.mul_(scale)
    keep = torch.nonzero(ok).flatten()
    return R[keep], E[keep], S[keep], M[keep], ks


def harvest(P, ncrop, crop, margin, rng, max_iter=150, tol_final=5e-8, min_ratio=10.0, weight="hits"):
    xs, rs, rel = run_jacobi_history(P, max_iter, tol_final)
    out = crops_from_history(P, xs, rs, rel, ncrop, crop, margin, rng, min_ratio, weight)
    del xs, rs
    return out, rel
