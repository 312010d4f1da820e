"""Learned map-space preconditioner M = S (alpha I + K^T K) S.

S = diag(sqrt(1/hits)) (zero on unhit pixels) and K is a linear network of
3x3 convolutions with 16 channels (``nets.py``: a dilated chain, or four
parallel levels on box-restricted maps).  M is symmetric positive
semidefinite by construction and positive definite on hit pixels
(alpha > 0), so plain PCG stays valid; with K = 0 and alpha = 1 it is exactly
the hits-count Jacobi preconditioner of the reference
(examples/minkasi_mpi_example.py:99-105).

It plugs into the reference's preconditioner hook -- any object with
``__mul__(Mapset) -> Mapset`` (minkasi/mapmaking/pcg.py:110-111,157-158) --
and into the device-scalar PCG through ``apply_to(r, z)``.  Every operation
runs in the HIP library; torch only owns the buffers.
"""
import json

import numpy as np
import torch

from .. import device as dev
from ..lib import mkgpu
from .nets import CH, make_net, param_layout

DEFAULT_ARCH = ("levels", (1, 4, 16, 64))


def init_params(arch, spec, seed=0, scale=None, log_alpha=0.0):
    """Random initial weights (numpy FP32), N(0, scale^2) per weight
    (default 0.5/sqrt(9 Cin))."""
    rng = np.random.default_rng(seed)
    lay, n = param_layout(arch, len(spec))
    p = np.zeros(n, dtype=np.float32)
    for off, co, ci in lay:
        s = scale if scale is not None else 0.5 / np.sqrt(9.0 * ci)
        p[off:off + co * ci * 9] = rng.standard_normal(co * ci * 9) * s
    p[-1] = log_alpha
    return p


def unpack(params, arch, nlayer):
    """numpy flat params -> (list of (Cout, Cin, 9) arrays, log_alpha)."""
    params = np.asarray(params)
    lay, n = param_layout(arch, nlayer)
    if params.shape != (n,):
        raise ValueError("expected %d parameters, got %s" % (n, params.shape))
    ws = [params[off:off + co * ci * 9].reshape(co, ci, 9) for off, co, ci in lay]
    return ws, float(params[-1])


def save(path, params, arch, spec, meta=None):
    np.savez(path, params=np.asarray(params, dtype=np.float32), arch=np.asarray(arch),
             spec=np.asarray(spec, dtype=np.int32), meta=np.asarray(json.dumps(meta or {}, default=float)))


def load(path):
    """(params, arch, spec, meta) from a file written by save()."""
    with np.load(path, allow_pickle=False) as f:
        meta = json.loads(str(f["meta"])) if "meta" in f else {}
        return (f["params"].astype(np.float32), str(f["arch"]), tuple(int(d) for d in f["spec"]), meta)


class LearnedPrecon:
    """z = M r on device maps of shape (nx, ny), FP32."""

    def __init__(self, hits, params, arch=DEFAULT_ARCH[0], spec=DEFAULT_ARCH[1]):
        h = hits.map if hasattr(hits, "map") else hits
        if not isinstance(h, torch.Tensor) or not h.is_cuda or h.dim() != 2:
            raise TypeError("hits must be a 2-D device map")
        self.shape = tuple(h.shape)
        d = dev.device()
        hf = h.to(torch.float64)
        self.s = torch.where(hf > 0, torch.rsqrt(torch.where(hf > 0, hf, torch.ones_like(hf))),
                             torch.zeros_like(hf)).to(torch.float32).contiguous()
        self.net = make_net(arch, spec, params, 1, self.shape[0], self.shape[1])
        self.net.refresh_adjoint()
        self.u = torch.empty(self.shape, dtype=torch.float32, device=d)
        self.g0 = torch.empty(self.shape, dtype=torch.float32, device=d)
        self.n = self.net.n
        self.isnull = False

    @classmethod
    def from_file(cls, hits, path):
        params, arch, spec, _ = load(path)
        return cls(hits, params, arch, spec)

    def set_params(self, params):
        self.net.params.copy_(params if isinstance(params, torch.Tensor) else torch.from_numpy(params))
        self.net.refresh_adjoint()

    def apply_to(self, r, z):
        """z = S (alpha S r + K^T K S r); r, z contiguous FP32 (nx, ny)."""
        if tuple(r.shape) != self.shape or tuple(z.shape) != self.shape:
            raise ValueError("map shape %s differs from the preconditioner's %s" % (tuple(r.shape), self.shape))
        dev.ptr(r, torch.float32, "r")
        dev.ptr(z, torch.float32, "z")
        sp = dev.stream_ptr()
        mkgpu.check("mk_mul_f32", self.u.data_ptr(), self.s.data_ptr(), r.data_ptr(), self.n, sp)
        self.net.forward(self.u, self.g0)
        mkgpu.check("mk_precon_finish_f32", z.data_ptr(), self.s.data_ptr(), self.net.log_alpha_ptr,
                    self.u.data_ptr(), self.g0.data_ptr(), self.n, sp)
        return z

    def __mul__(self, mapset):
        """precon * r for a Mapset holding one SkyMap (pcg.py:110-111)."""
        out = mapset.copy()
        self.apply_to(mapset.maps[0].map.contiguous(), out.maps[0].map)
        return out
