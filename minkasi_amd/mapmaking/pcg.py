"""Preconditioned conjugate gradient (minkasi/mapmaking/pcg.py:17-255).

Same update order as the reference (pcg.py:101-177).  When b and x0 are a
single FP32 SkyMap and the TodVec has a batched operator, the iteration
runs with device-resident FP64 scalars (alpha, beta, r.z and r.r formed on
the GPU by fused kernels); otherwise it runs the reference's Mapset-level
loop with device vector operations.

Additions (opt-in keywords): tol stops when ||r||/||b|| < tol (recursive
residual; the reference always runs maxiter iterations);
true_residual_every=k also records ||b - A x||/||b|| every k iterations;
history (a list) receives one dict per iteration; verbose=False silences
the per-iteration print; check="fp64" (or "binned64", or an operator)
accumulates the iterate in FP64, replaces the recursive residual once by
the true one and stops only when the true residual ||b - A x||/||b||,
evaluated by that operator against its own FP64 right-hand side, is below
tol (DevicePCG).
"""
import sys
import time

import numpy as np
import torch

from .. import device as dev
from ..lib import mkgpu
from ..parallel import myrank

try:
    from typing import Protocol, runtime_checkable
except ImportError:  # pragma: no cover
    from typing_extensions import Protocol, runtime_checkable


@runtime_checkable
class HasPrior(Protocol):
    def apply_prior(self, x, Ax): ...


class null_precon:
    """pcg.py:241-255."""

    def __init__(self):
        self.isnull = True

    def __add__(self, val):
        return val

    def __mul__(self, val):
        return val


class PCGBreakdown(RuntimeError):
    """Zero/negative curvature (p.Ap <= 0) or a non-finite scalar."""


def _single_f32_map(ms):
    from ..maps.skymap import SkyMap
    return (hasattr(ms, "maps") and len(ms.maps) == 1 and isinstance(ms.maps[0], SkyMap)
            and ms.maps[0].map.dtype == torch.float32)


def run_pcg_wprior(b, x0, tods, prior=None, precon=None, mask=None, maxiter=25, outroot="map", save_iters=[-1],
                   save_ind=0, save_tail=".fits", plot_iters=[], plot_info=None, plot_ind=0, tol=None,
                   true_residual_every=0, history=None, verbose=True, check=None):
    fast = (prior is None and mask is None and _single_f32_map(b) and _single_f32_map(x0)
            and tods.batched(b) is not None
            and (precon is None or _single_f32_map(precon) or hasattr(precon, "apply_to")))
    if fast:
        return _run_pcg_device(b, x0, tods, precon, maxiter, outroot, save_iters, save_ind, save_tail, tol,
                               true_residual_every, history, verbose, check=check)
    if check is not None:
        raise ValueError("check= needs the batched device path (one FP32 SkyMap, no prior or mask)")
    bc = None
    if prior is None and mask is None and hasattr(tods, "batched_cuts"):
        bc = tods.batched_cuts(b)
        if bc is not None and not (bc.matches(tods, x0) and (precon is None or (
                hasattr(precon, "maps") and len(precon.maps) == 2 and bc.matches(tods, precon)
                and precon.maps[0].map.dtype == torch.float32))):
            bc = None
    if bc is not None:
        return _run_pcg_device(b, x0, tods, precon, maxiter, outroot, save_iters, save_ind, save_tail, tol,
                               true_residual_every, history, verbose, flat=bc)
    return _run_pcg_generic(b, x0, tods, prior, precon, mask, maxiter, outroot, save_iters, save_ind, save_tail,
                            tol, history, verbose)


def run_pcg(b, x0, tods, precon=None, mask=None, maxiter=25, outroot="map", save_iters=[-1], save_ind=0,
            save_tail=".fits", plot_iters=[], plot_info=None, plot_ind=0, **kw):
    """pcg.py:203-238."""
    return run_pcg_wprior(b=b, x0=x0, tods=tods, prior=None, precon=precon, mask=mask, maxiter=maxiter,
                          outroot=outroot, save_iters=save_iters, save_ind=save_ind, save_tail=save_tail,
                          plot_iters=plot_iters, plot_info=plot_info, plot_ind=plot_ind, **kw)


def _run_pcg_generic(b, x0, tods, prior, precon, mask, maxiter, outroot, save_iters, save_ind, save_tail, tol,
                     history, verbose):
    """pcg.py:101-177 line for line over Mapset operations."""
    t1 = time.time()
    Ax = tods.dot(x0, mask=mask)
    try:
        r = b.copy()
        r.axpy(Ax, -1)
    except Exception:
        r = b - Ax
    z = precon * r if precon is not None else r.copy()
    p = z.copy()
    zr = r.dot(z)
    x = x0.copy()
    bb = b.dot(b) if tol is not None else None
    t2 = time.time()
    t3 = 0
    alpha = 0
    nsamp = tods.get_nsamp()
    tloop = time.time()
    for i in range(maxiter):
        if verbose and myrank == 0:
            if i > 0:
                print(i, zr, alpha, t2 - t1, t3 - t2, t3 - t1, nsamp / (t2 - t1) / 1e6)
            else:
                print(i, zr, t2 - t1)
            sys.stdout.flush()
        t1 = time.time()
        Ap = tods.dot(p, mask=mask)
        if prior is not None:
            prior.apply_prior(p, Ap)
        t2 = time.time()
        pAp = p.dot(Ap)
        if not (pAp > 0) or not np.isfinite(pAp):
            raise PCGBreakdown("p.Ap = %r at iteration %d" % (pAp, i))
        alpha = zr / pAp
        x_new = x.copy()
        x_new.axpy(p, alpha)
        r_new = r.copy()
        r_new.axpy(Ap, -alpha)
        z_new = precon * r_new if precon is not None else r_new.copy()
        zr_new = r_new.dot(z_new)
        beta = zr_new / zr
        p_new = z_new.copy()
        p_new.axpy(p, beta)
        p, z, r, zr, x = p_new, z_new, r_new, zr_new, x_new
        t3 = time.time()
        rel = None
        if tol is not None or history is not None:
            rr = r.dot(r)
            rel = float(np.sqrt(rr / bb)) if bb else None
        if history is not None:
            history.append(dict(it=i, zr=zr, alpha=alpha, beta=beta, rel_res=rel, t_Ap=t2 - t1, t_iter=t3 - t1))
        if i in save_iters and myrank == 0:
            x.maps[save_ind].write(outroot + "_" + str(i) + save_tail)
        if tol is not None and rel is not None and rel < tol:
            break
    tave = (time.time() - tloop) / max(1, i + 1) if maxiter else 0.0
    if verbose and maxiter:
        print("average time per iteration was ", tave, " with effective throughput ", nsamp / tave / 1e6,
              " Msamp/s")
    return x


class DevicePCG:
    """PCG over one FP32 map with device-resident FP64 scalars, in the
    update order of pcg.py:143-170.  `precon` is None, a Mapset holding the
    Jacobi weight map (fused update kernel), or an object with
    apply_to(r, z) (e.g. the learned preconditioner).

    check (optional): an operator with apply(x64, out64) that evaluates A x
    on FP64 maps more accurately than `op` (AccurateDot, or the batched
    operator's own apply_f64), and b64 the FP64 right-hand side the solution
    is judged against.  With it the iterate accumulates in FP64, the
    recursive residual is replaced once by fl32(b64 - A x64) when it first
    drops below replace_at (reliable updates: the recursion's drift from the
    FP32 operator is reset while the Krylov directions are kept), and
    converged(tol) accepts a solution only when ||b64 - A x64||/||b64|| from
    the check operator is below tol; a failed check makes the next step a
    replacement step."""

    def __init__(self, op, b_map, precon=None, null_mask="auto", check=None, b64=None, replace_at=1e-3):
        self.op = op
        self.b = b_map
        self.n = b_map.numel()
        self.S = dev.scratch()
        self.r = torch.empty_like(b_map)
        self.z = torch.empty_like(b_map)
        self.p = torch.empty_like(b_map)
        self.Ap = torch.empty_like(b_map)
        # Jacobi: a Mapset holding the weight map, or the weights as a tensor
        # shaped like b (the flat [map | cuts] vectors of BatchedCutsDot)
        self.jacobi = precon is not None and (isinstance(precon, torch.Tensor) or _single_f32_map(precon))
        self.w = (precon if isinstance(precon, torch.Tensor) else precon.maps[0].map) if self.jacobi else None
        self.precon = precon
        self.x = None
        self.bb = None
        # Null mode of A: the constant on hit pixels (zero on cut samples).
        # A.p (and r_0) are projected off it; an exact no-op in exact
        # arithmetic, it stops FP32 rounding from making the singular system
        # inconsistent, which otherwise makes the iteration diverge once
        # ||r||/||b|| reaches ~1e-6.  The hit pixels come from the operator
        # itself, so a preconditioner that is zero on some hit pixels (a
        # masked Jacobi map) does not change the projection.
        if isinstance(null_mask, str) and null_mask == "auto":
            null_mask = op.null_mask() if hasattr(op, "null_mask") else None
        self.u = None
        if null_mask is not None:
            self.u = null_mask.to(torch.float32).reshape(b_map.shape).contiguous()
            cnt = float(self.u.sum().item())
            self.inv_uu = 1.0 / cnt if cnt > 0 else 0.0
        self.check = check
        self.replace_at = float(replace_at) if replace_at else 0.0
        self.x64 = self.b64 = self.ax64 = None
        self.replacements = 0
        self.checks = 0
        self.last_true = None
        self._replace_next = False
        if check is not None:
            if b64 is None or b64.dtype != torch.float64 or b64.numel() != self.n:
                raise ValueError("a check operator needs the FP64 right-hand side b64 (%d values)" % self.n)
            self.b64 = b64.reshape(b_map.shape)
            self.x64 = torch.zeros(b_map.shape, dtype=torch.float64, device=b_map.device)
            self.ax64 = torch.zeros_like(self.x64)

    def _project(self, v):
        if self.u is not None:
            mkgpu.check("mk_project_out_f32", v.data_ptr(), self.u.data_ptr(), self.n, self.inv_uu,
                        self.S.sc.data_ptr(), self.S.work.data_ptr(), dev.stream_ptr())

    def _M(self, rv, zv):
        if self.precon is None:
            zv.copy_(rv)
        elif self.jacobi:
            mkgpu.check("mk_mul_f32", zv.data_ptr(), self.w.data_ptr(), rv.data_ptr(), self.n, dev.stream_ptr())
        else:
            self.precon.apply_to(rv, zv)

    def _true_r(self):
        """r = fl32(b64 - A x64) with the check operator; rr into the scalars."""
        sc, work, sp = self.S.sc, self.S.work.data_ptr(), dev.stream_ptr()
        self.check.apply(self.x64, self.ax64)
        mkgpu.check("mk_pcg_set_r", self.r.data_ptr(), self.b64.data_ptr(), self.ax64.data_ptr(), self.n,
                    sc.data_ptr(), work, sp)

    def init(self, x_map):
        """x <- x0 (in place), r = b - A x, z = M r, p = z, zr = r.z."""
        n, sc, work, sp = self.n, self.S.sc, self.S.work.data_ptr(), dev.stream_ptr()
        self.x = x_map
        sc.zero_()
        self.replacements = self.checks = 0
        self.last_true = None
        self._replace_next = False
        if self.check is not None:
            self.x64.copy_(x_map)
            mkgpu.check("mk_dot_f32", x_map.data_ptr(), x_map.data_ptr(), n, sc.data_ptr(), work, sp)
            if float(sc[mkgpu.SC_DOT].item()) > 0.0:
                self._true_r()
            else:
                self.ax64.zero_()
                mkgpu.check("mk_pcg_set_r", self.r.data_ptr(), self.b64.data_ptr(), self.ax64.data_ptr(), n,
                            sc.data_ptr(), work, sp)
        else:
            self.op.apply(self.x, self.Ap, accumulate=False)
            self.r.copy_(self.b)
            mkgpu.check("mk_axpy_f32", self.r.data_ptr(), self.Ap.data_ptr(), -1.0, n, sp)
        self._project(self.r)
        self._M(self.r, self.z)
        mkgpu.check("mk_pcg_init", self.r.data_ptr(), self.z.data_ptr(), self.p.data_ptr(), n, sc.data_ptr(), work,
                    sp)
        if self.check is not None:
            mkgpu.check("mk_dot_f64", self.b64.data_ptr(), self.b64.data_ptr(), n, sc.data_ptr(), work, sp)
        else:
            mkgpu.check("mk_dot_f32", self.b.data_ptr(), self.b.data_ptr(), n, sc.data_ptr(), work, sp)
        self.bb = float(sc[mkgpu.SC_DOT].item())

    def step(self):
        n, sc, work, sp = self.n, self.S.sc, self.S.work.data_ptr(), dev.stream_ptr()
        x, r, z, p, Ap = self.x, self.r, self.z, self.p, self.Ap
        replace = self.check is not None and self._replace_next
        self._replace_next = False
        self.op.apply(p, Ap, accumulate=False)
        self._project(Ap)
        mkgpu.check("mk_pcg_alpha", p.data_ptr(), Ap.data_ptr(), n, sc.data_ptr(), work, sp)
        if self.jacobi and not replace:
            if self.x64 is not None:
                mkgpu.check("mk_pcg_xrz_jacobi_x64", self.x64.data_ptr(), x.data_ptr(), r.data_ptr(), z.data_ptr(),
                            p.data_ptr(), Ap.data_ptr(), self.w.data_ptr(), n, sc.data_ptr(), work, sp)
            else:
                mkgpu.check("mk_pcg_xrz_jacobi", x.data_ptr(), r.data_ptr(), z.data_ptr(), p.data_ptr(),
                            Ap.data_ptr(), self.w.data_ptr(), n, sc.data_ptr(), work, sp)
        else:
            if self.x64 is not None:
                mkgpu.check("mk_pcg_xr_x64", self.x64.data_ptr(), x.data_ptr(), r.data_ptr(), p.data_ptr(),
                            Ap.data_ptr(), n, sc.data_ptr(), work, sp)
            else:
                mkgpu.check("mk_pcg_xr", x.data_ptr(), r.data_ptr(), p.data_ptr(), Ap.data_ptr(), n, sc.data_ptr(),
                            work, sp)
            if replace:
                self._true_r()
                self._project(r)
                self.replacements += 1
            self._M(r, z)
            mkgpu.check("mk_pcg_beta", r.data_ptr(), z.data_ptr(), n, sc.data_ptr(), work, sp)
        mkgpu.check("mk_pcg_p", p.data_ptr(), z.data_ptr(), n, sc.data_ptr(), sp)
        return replace

    def scalars(self):
        """Host copy of the scalar block (synchronises)."""
        s = self.S.sc.cpu().numpy()
        if s[mkgpu.SC_FLAG] != 0:
            raise PCGBreakdown("PCG breakdown (flag %g, pAp %r)" % (s[mkgpu.SC_FLAG], s[mkgpu.SC_PAP]))
        return s

    def rel_res(self, s=None):
        s = self.scalars() if s is None else s
        return float(np.sqrt(max(s[mkgpu.SC_RR], 0.0) / self.bb)) if self.bb > 0 else float("nan")

    def converged(self, tol, rel=None):
        """Whether the solve may stop at tol (synchronises).  Without a check
        operator: the recursive residual is below tol.  With one: it asks for
        the replacement step once the recursive residual is below
        replace_at, and when the recursive residual is below tol it computes
        the true residual with the check operator (stored in last_true) and
        accepts only if that is below tol too, else the next step replaces
        the residual."""
        rel = self.rel_res() if rel is None else rel
        if self.check is None:
            return rel < tol
        if self.replacements == 0 and rel < self.replace_at:
            self._replace_next = True
        if rel >= tol:
            return False
        self.checks += 1
        self.last_true = self.true_rel_res64()
        if self.last_true < tol:
            return True
        self._replace_next = True
        return False

    def true_rel_res64(self, check=None):
        """||b64 - A x64|| / ||b64|| with the check operator (or `check`)."""
        c = self.check if check is None else check
        out = torch.empty_like(self.x64)
        c.apply(self.x64, out)
        sc, work, sp = self.S.sc, self.S.work.data_ptr(), dev.stream_ptr()
        mkgpu.check("mk_axpy_f64", out.data_ptr(), self.b64.data_ptr(), -1.0, self.n, sp)  # A x - b
        saved = float(sc[mkgpu.SC_DOT].item())
        mkgpu.check("mk_dot_f64", out.data_ptr(), out.data_ptr(), self.n, sc.data_ptr(), work, sp)
        rr = float(sc[mkgpu.SC_DOT].item())
        sc[mkgpu.SC_DOT] = saved
        bb = self.bb
        return float(np.sqrt(max(rr, 0.0) / bb)) if bb > 0 else float("nan")

    def true_rel_res(self):
        """||b - A x|| / ||b|| with the same FP32 operator, FP64 norm (the
        check operator's FP64 residual when there is one)."""
        if self.check is not None:
            return self.true_rel_res64()
        rt = torch.empty_like(self.b)
        self.op.apply(self.x, rt, accumulate=False)
        d = self.b.to(torch.float64) - rt.to(torch.float64)
        return float(torch.linalg.vector_norm(d).item() / np.sqrt(self.bb))


def make_check(op, kind):
    """The check operator and FP64 right-hand side for DevicePCG(check=...).

    kind "fp64": the right-hand side from AccurateDot (every stage in FP64;
    needs the noise models' FP64 filters, keep_fp64=True) and, as the check
    operator, the batched operator with FP64 binning (BatchedDot.f64_view),
    which agrees with AccurateDot at the solution to about 3e-8 of ||b||
    at a tenth of its cost.  kind "binned64": that operator and its own
    right-hand side (no FP64 filters needed; b itself then carries the FP32
    stages' error, about 5e-7 of ||b||).  Or an object with apply(x64,
    out64) and rhs(out64).  Returns (check, b64)."""
    from .batch import AccurateDot
    b64 = torch.zeros((op.nx, op.ny), dtype=torch.float64, device=op.buf_a.device)
    if kind == "fp64":
        acc = op.accurate()
        acc.rhs(b64)
        return op.f64_view(), b64
    chk = op.f64_view() if kind == "binned64" else kind
    chk.rhs(b64)
    return chk, b64


def _run_pcg_device(b, x0, tods, precon, maxiter, outroot, save_iters, save_ind, save_tail, tol,
                    true_residual_every, history, verbose, flat=None, check=None):
    """Device-scalar PCG with the same update order as pcg.py:143-170.
    flat: a BatchedCutsDot, for a Mapset of a map and cut samples solved as
    one flat vector (packed here, unpacked into the returned Mapset).
    check: None, "fp64", "binned64" or an operator (make_check): the
    iterate accumulates in FP64 and tol is judged on the true residual."""
    x = x0.copy()
    if flat is None:
        op = tods.batched(b)
        chk, b64 = make_check(op, check) if check is not None else (None, None)
        solver = DevicePCG(op, b.maps[0].map, precon, check=chk, b64=b64)
        solver.init(x.maps[0].map)
    else:
        if check is not None:
            raise ValueError("check operators are for the map-only batched operator")
        w = flat.pack(precon) if precon is not None else None
        solver = DevicePCG(flat, flat.pack(b), w)
        solver.init(flat.pack(x))
    nsamp = tods.get_nsamp()
    tloop = time.time()
    it_done = 0
    for i in range(maxiter):
        t1 = time.time()
        replaced = solver.step()
        it_done = i + 1
        if verbose or tol is not None or history is not None:
            s = solver.scalars()
            t3 = time.time()
            rel = solver.rel_res(s)
            rec = dict(it=i, zr=float(s[mkgpu.SC_ZR]), alpha=float(s[mkgpu.SC_ALPHA]), beta=float(s[mkgpu.SC_BETA]),
                       rel_res=rel, t_iter=t3 - t1)
            if replaced:
                rec["replaced"] = True
            if true_residual_every and (i + 1) % true_residual_every == 0:
                rec["true_rel_res"] = solver.true_rel_res()
            stop = tol is not None and solver.converged(tol, rel)
            if solver.last_true is not None and stop:
                rec["true_rel_res"] = solver.last_true
            if history is not None:
                history.append(rec)
            if verbose and myrank == 0:
                print(i, rec["zr"], rec["alpha"], rec["t_iter"], nsamp / max(rec["t_iter"], 1e-12) / 1e6)
                sys.stdout.flush()
            if stop:
                break
        if i in save_iters and myrank == 0:
            if flat is not None:
                flat.unpack(solver.x, x)
            x.maps[save_ind].write(outroot + "_" + str(i) + save_tail)
    if flat is not None:
        flat.unpack(solver.x, x)
    # a breakdown in a quiet run (no per-iteration scalars read) still raises
    solver.scalars()
    if verbose and it_done:
        tave = (time.time() - tloop) / it_done
        print("average time per iteration was ", tave, " with effective throughput ", nsamp / tave / 1e6,
              " Msamp/s")
    return x
