"""Builds the synthetic mapmaking problems of the benchmark configurations on
the GPU: TODs, map, pixelization, noise models, hits, Jacobi preconditioner,
right-hand side and the batched operator, following the canonical sequence
of minkasi's examples/minkasi_mpi_example.py:35-112."""
import time

import numpy as np
import torch

from . import device as dev
from .synth import ARCSEC, CONFIGS, ScanSpec, make_tod_info, tod_shapes


class Problem:
    pass


def build_problem(config_id, ntod=None, seed_base=0, shapes=None, verbose=False, truncate=False, keep_fp64=False):
    """truncate: cut every TOD so that n - 1 is {2,3,5,7,11}-smooth, as the
    reference's examples do before mapmaking (tods/processing.py:648-681).
    keep_fp64: the noise models also keep their FP64 filters, for the FP64
    operator that checks the solution's true residual."""
    from .maps.mapset import Mapset
    from .maps.skymap import SkyMap
    from .mapmaking.noise import NoiseSmoothedSVD
    from .mapmaking.tod2map import make_hits
    from .tods.core import Tod, TodVec

    c = CONFIGS[config_id]
    shapes = shapes or tod_shapes(config_id, ntod)
    spec = ScanSpec(**c["spec"])
    dtype = np.float32 if c["dtype"] == "float32" else np.float64
    P = Problem()
    P.config_id, P.shapes = config_id, shapes
    t0 = time.time()
    todvec = TodVec()
    for i, (ndet, n) in enumerate(shapes):
        info = make_tod_info(config_id, seed_base + i, ndet, n, spec, dtype=dtype, backend="torch",
                             want_radec=False)
        if truncate:
            from .tods.processing import truncate_tod
            truncate_tod(info)
        todvec.add_tod(Tod(info))
    if truncate:
        shapes = [tod.get_data_dims() for tod in todvec.tods]
        P.shapes = shapes
    torch.cuda.synchronize()
    P.t_generate = time.time() - t0
    t0 = time.time()
    lims = todvec.lims()
    N = c["npix"]
    m = SkyMap(lims, c["pix"] * ARCSEC, nx=N, ny=N,
               dtype=torch.float32 if dtype == np.float32 else torch.float64)
    for tod in todvec.tods:
        m.get_pix(tod)
    torch.cuda.synchronize()
    P.t_pixelize = time.time() - t0
    t0 = time.time()
    for k, tod in enumerate(todvec.tods):
        tod.set_noise(NoiseSmoothedSVD, fwhm=c["fwhm"], keep_fp64=True if keep_fp64 else None)
        if verbose and k % 16 == 0:
            print("noise model", k, "of", len(todvec.tods), flush=True)
    torch.cuda.synchronize()
    P.t_noise = time.time() - t0
    P.ndeflated = sum(1 for tod in todvec.tods if getattr(tod.noise, "defl_k", 0))
    if verbose:
        print("noise models: %d of %d TODs deflate dominant modes ahead of the FP32 rotation" % (
            P.ndeflated, len(todvec.tods)), flush=True)
    t0 = time.time()
    mapset = Mapset()
    mapset.add_map(m)
    op = todvec.batched(mapset)
    torch.cuda.synchronize()
    P.t_plans = time.time() - t0
    t0 = time.time()
    hits = make_hits(todvec, m)
    rhs = mapset.copy()
    rhs.clear()
    todvec.make_rhs(rhs)
    pre = mapset.copy()
    h = hits.map
    pre.maps[0].map.copy_(torch.where(h > 0, 1.0 / torch.where(h > 0, h, torch.ones_like(h)), torch.zeros_like(h)))
    torch.cuda.synchronize()
    P.t_rhs = time.time() - t0
    P.todvec, P.map, P.mapset, P.hits, P.rhs, P.precon, P.op = todvec, m, mapset, hits, rhs, pre, op
    P.nsamp = sum(a * b for a, b in shapes)
    return P


def solve_to_tol(P, precon=None, tol=1e-6, maxiter=600, graph=True, check=None, b64=None, verify=None):
    """Device PCG from x0 = 0 until ||r||/||b|| < tol.

    precon: None for the problem's hits-count Jacobi, else an object with
    apply_to(r, z) (e.g. the learned preconditioner).  check/b64: a check
    operator and the FP64 right-hand side (pcg.make_check): the iterate
    accumulates in FP64, the residual is replaced once, and the solve stops
    only when the true residual from the check operator is below tol (the
    check is inside the timed region).  verify: another operator (e.g. the
    all-FP64 AccurateDot) whose true residual of the final iterate is
    reported as true_rel_res_fp64, outside the timed region.  Returns a dict
    with iterations, wall seconds (including init), final recursive and true
    relative residuals, and the solution tensor (FP64 when checked)."""
    import time
    from .mapmaking.pcg import DevicePCG
    op, b = P.op, P.rhs.maps[0].map
    solver = DevicePCG(op, b, P.precon if precon is None else precon, check=check, b64=b64)
    x = torch.zeros_like(b)
    if graph:
        solver.init(x)
        op.capture(solver.p, solver.Ap, accumulate=False)
        x.zero_()
    torch.cuda.synchronize()
    t0 = time.time()
    solver.init(x)
    it, rel, hist, done = 0, 1.0, [], False
    while it < maxiter:
        solver.step()
        it += 1
        rel = solver.rel_res()
        hist.append(rel)
        if solver.converged(tol, rel):
            done = True
            break
    torch.cuda.synchronize()
    dt = time.time() - t0
    out = dict(iters=it, seconds=dt, rel_res=rel, reached=done)
    if check is not None:
        out.update(true_rel_res=solver.last_true if solver.last_true is not None else solver.true_rel_res64(),
                   replacements=solver.replacements, checks=solver.checks, x=solver.x64)
    else:
        out.update(true_rel_res=solver.true_rel_res(), x=x)
    if verify is not None:
        if solver.x64 is None:
            solver.x64 = x.to(torch.float64)
            solver.b64 = b64
        out["true_rel_res_fp64"] = solver.true_rel_res64(verify)
    out["history"] = hist
    return out


def map_distance(P, x, y):
    """||x - y|| / ||y|| over hit pixels after removing the null mode (the
    mean over hit pixels) from both."""
    hit = P.hits.map > 0
    xd = x.to(torch.float64)[hit]
    yd = y.to(torch.float64)[hit]
    xd = xd - xd.mean()
    yd = yd - yd.mean()
    return float(torch.linalg.vector_norm(xd - yd).item() / torch.linalg.vector_norm(yd).item())
