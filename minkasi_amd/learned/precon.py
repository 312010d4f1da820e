"""Learned map-space preconditioner M = S (alpha I + K^T K) S.

S = diag(sqrt(1/hits)) (zero on unhit pixels) and K is a chain of linear
dilated 3x3 convolutions, 1 -> 16 -> 16 -> 16 -> 16 channels, no
activations.  M is symmetric positive semidefinite by construction and
positive definite on hit pixels (alpha > 0), so plain PCG stays valid; with
K = 0 and alpha = 1 it is exactly the hits-count Jacobi preconditioner of the
reference (examples/minkasi_mpi_example.py:99-105).

It plugs into the reference's preconditioner hook -- any object with
``__mul__(Mapset) -> Mapset`` (minkasi/mapmaking/pcg.py:110-111,157-158) --
and into the device-scalar PCG through ``apply_to(r, z)``.  Every operation
runs in the HIP library (mk_conv3x3_f32, mk_conv3x3_adjoint_weights,
mk_precon_finish_f32); torch only owns the buffers.

Parameters are one flat FP32 vector: the layer weights w_l[Cout][Cin][9] in
layer order, then log(alpha).
"""
import numpy as np
import torch

from .. import device as dev
from ..lib import mkgpu

CH = 16
NLAYER = 4
DEFAULT_DILS = (1, 3, 9, 27)


def layer_shapes(nlayer=NLAYER, ch=CH):
    shapes = [(ch, 1)]
    for _ in range(nlayer - 1):
        shapes.append((ch, ch))
    return shapes


def param_layout(nlayer=NLAYER, ch=CH):
    """[(offset, cout, cin)] per layer and the total length (incl. log alpha)."""
    out, off = [], 0
    for co, ci in layer_shapes(nlayer, ch):
        out.append((off, co, ci))
        off += co * ci * 9
    return out, off + 1


def init_params(dils=DEFAULT_DILS, seed=0, scale=None, log_alpha=0.0):
    """Random initial weights (numpy FP32): each layer scaled so K starts
    small against alpha I; the network starts close to Jacobi."""
    rng = np.random.default_rng(seed)
    lay, n = param_layout(len(dils))
    p = np.zeros(n, dtype=np.float32)
    for i, (off, co, ci) in enumerate(lay):
        s = scale if scale is not None else 0.5 / np.sqrt(9.0 * ci)
        p[off:off + co * ci * 9] = rng.standard_normal(co * ci * 9) * s
    p[-1] = log_alpha
    return p


def unpack(params, nlayer=NLAYER):
    """numpy flat params -> (list of (Cout, Cin, 9) arrays, log_alpha)."""
    params = np.asarray(params)
    lay, n = param_layout(nlayer)
    assert params.shape == (n,), params.shape
    ws = [params[off:off + co * ci * 9].reshape(co, ci, 9) for off, co, ci in lay]
    return ws, float(params[-1])


def save(path, params, dils, meta=None):
    np.savez(path, params=np.asarray(params, dtype=np.float32), dils=np.asarray(dils, dtype=np.int32),
             meta=np.asarray(repr(meta or {})))


def load(path):
    with np.load(path, allow_pickle=False) as f:
        return f["params"].astype(np.float32), tuple(int(d) for d in f["dils"])


class ConvChain:
    """Device buffers and launches for K and K^T on a batch of (H, W) images."""

    def __init__(self, params, dils, B, H, W):
        self.dils = tuple(int(d) for d in dils)
        self.nlayer = len(self.dils)
        self.lay, self.nparam = param_layout(self.nlayer)
        d = dev.device()
        if isinstance(params, torch.Tensor):
            self.params = params
        else:
            self.params = torch.from_numpy(np.asarray(params, dtype=np.float32)).to(d)
        if self.params.numel() != self.nparam or self.params.dtype != torch.float32:
            raise ValueError("expected %d FP32 parameters" % self.nparam)
        self.wt = torch.empty(self.nparam, dtype=torch.float32, device=d)  # adjoint weights, same layout
        self.B, self.H, self.W = int(B), int(H), int(W)
        self.n = self.B * self.H * self.W

    def w_ptr(self, l, adjoint=False):
        off = self.lay[l][0]
        base = self.wt if adjoint else self.params
        return base.data_ptr() + 4 * off

    @property
    def log_alpha_ptr(self):
        return self.params.data_ptr() + 4 * (self.nparam - 1)

    def refresh_adjoint(self):
        sp = dev.stream_ptr()
        for l, (off, co, ci) in enumerate(self.lay):
            mkgpu.check("mk_conv3x3_adjoint_weights", self.w_ptr(l), co, ci, self.w_ptr(l, True), sp)

    def conv(self, x, l, y, adjoint=False, accumulate=False):
        """y (+)= layer l (or its adjoint) applied to x."""
        off, co, ci = self.lay[l]
        cin, cout = (co, ci) if adjoint else (ci, co)
        mkgpu.check("mk_conv3x3_f32", x.data_ptr(), self.B, cin, self.H, self.W, self.w_ptr(l, adjoint), cout,
                    self.dils[l], y.data_ptr(), int(accumulate), dev.stream_ptr())


class LearnedPrecon:
    """z = M r on device maps of shape (nx, ny), FP32."""

    def __init__(self, hits, params, dils=DEFAULT_DILS):
        h = hits.map if hasattr(hits, "map") else hits
        if not isinstance(h, torch.Tensor) or not h.is_cuda or h.dim() != 2:
            raise TypeError("hits must be a 2-D device map")
        self.shape = tuple(h.shape)
        d = dev.device()
        hf = h.to(torch.float64)
        self.s = torch.where(hf > 0, torch.rsqrt(torch.where(hf > 0, hf, torch.ones_like(hf))),
                             torch.zeros_like(hf)).to(torch.float32).contiguous()
        self.chain = ConvChain(params, dils, 1, self.shape[0], self.shape[1])
        self.chain.refresh_adjoint()
        n = self.chain.n
        self.u = torch.empty(self.shape, dtype=torch.float32, device=d)
        self.g0 = torch.empty(self.shape, dtype=torch.float32, device=d)
        self.buf = [torch.empty((CH,) + self.shape, dtype=torch.float32, device=d) for _ in range(2)]
        self.n = n
        self.isnull = False

    @classmethod
    def from_file(cls, hits, path):
        params, dils = load(path)
        return cls(hits, params, dils)

    def set_params(self, params):
        self.chain.params.copy_(params if isinstance(params, torch.Tensor) else torch.from_numpy(params))
        self.chain.refresh_adjoint()

    def apply_to(self, r, z):
        """z = S (alpha S r + K^T K S r); r, z contiguous FP32 (nx, ny)."""
        if tuple(r.shape) != self.shape or tuple(z.shape) != self.shape:
            raise ValueError("map shape %s differs from the preconditioner's %s" % (tuple(r.shape), self.shape))
        dev.ptr(r, torch.float32, "r")
        dev.ptr(z, torch.float32, "z")
        sp = dev.stream_ptr()
        c = self.chain
        mkgpu.check("mk_mul_f32", self.u.data_ptr(), self.s.data_ptr(), r.data_ptr(), self.n, sp)
        a, b = self.buf
        c.conv(self.u, 0, a)
        for l in range(1, c.nlayer):
            c.conv(a, l, b)
            a, b = b, a
        for l in range(c.nlayer - 1, 0, -1):
            c.conv(a, l, b, adjoint=True)
            a, b = b, a
        c.conv(a, 0, self.g0, adjoint=True)
        mkgpu.check("mk_precon_finish_f32", z.data_ptr(), self.s.data_ptr(), c.log_alpha_ptr, self.u.data_ptr(),
                    self.g0.data_ptr(), self.n, sp)
        return z

    def __mul__(self, mapset):
        """precon * r for a Mapset holding one SkyMap (pcg.py:110-111)."""
        out = mapset.copy()
        self.apply_to(mapset.maps[0].map.contiguous(), out.maps[0].map)
        return out
