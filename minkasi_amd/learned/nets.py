"""The two K^T K networks of the learned preconditioner, on device buffers.

Both compute g = K^T K u for u of shape (B, 1, H, W) and back-propagate a
gradient q = dL/dt of t = alpha u + g into the weights.  Parameters are one
flat FP32 vector: the layer weights w_l[Cout][Cin][9] in layer order, then
log(alpha).

chain   K = conv_4 conv_3 conv_2 conv_1: four 3x3 convolutions in series,
        1 -> 16 -> 16 -> 16 -> 16 channels, layer l with dilation dils[l].
levels  K u = [C_0 u; C_1 R_1 u; C_2 R_2 R_1 u; C_3 R_3 R_2 R_1 u]: four
        3x3 convolutions (1 -> 16 channels) in parallel, level l on the map
        box-restricted by factors[l] (R_l restricts by factors[l] /
        factors[l-1]), so that
            K^T K = sum_l (R^l)^T C_l^T C_l R^l,
        an additive multilevel (BPX-like) operator: a learned PSD 5x5
        stencil per level, its coarse levels reaching the map-wide modes
        a chain of 3x3 taps cannot.
tent    as levels with full-weighting restriction (tent weights) in place of
        the box average, so the prolongation R^T is bilinear interpolation
        and coarse corrections come back smooth.

Every operation is a libmkgpu call (mk_conv3x3_f32, mk_conv3x3_wgrad_f32,
mk_conv3x3_adjoint_weights, mk_restrict_{box,tent}_f32, mk_prolong_{box,tent}_f32).
"""
import numpy as np
import torch

from .. import device as dev
from ..lib import mkgpu

CH = 16


def layer_shapes(arch, nlayer):
    if arch == "chain":
        return [(CH, 1)] + [(CH, CH)] * (nlayer - 1)
    if arch in ("levels", "tent"):
        return [(CH, 1)] * nlayer
    raise ValueError("unknown architecture %r" % (arch,))


def param_layout(arch, nlayer):
    """[(offset, cout, cin)] per layer and the total length (incl. log alpha)."""
    out, off = [], 0
    for co, ci in layer_shapes(arch, nlayer):
        out.append((off, co, ci))
        off += co * ci * 9
    return out, off + 1


class _Base:
    def __init__(self, arch, spec, params, B, H, W):
        self.arch, self.spec = arch, tuple(int(v) for v in spec)
        self.nlayer = len(self.spec)
        self.lay, self.nparam = param_layout(arch, self.nlayer)
        d = dev.device()
        if isinstance(params, torch.Tensor):
            self.params = params
        else:
            self.params = torch.from_numpy(np.asarray(params, dtype=np.float32)).to(d)
        if self.params.numel() != self.nparam or self.params.dtype != torch.float32:
            raise ValueError("expected %d FP32 parameters for %s%s" % (self.nparam, arch, self.spec))
        self.wt = torch.empty(self.nparam, dtype=torch.float32, device=d)  # adjoint weights, same layout
        self.B, self.H, self.W = int(B), int(H), int(W)
        self.n = self.B * self.H * self.W
        self.device = d

    def w_ptr(self, l, adjoint=False):
        return (self.wt if adjoint else self.params).data_ptr() + 4 * self.lay[l][0]

    @property
    def log_alpha_ptr(self):
        return self.params.data_ptr() + 4 * (self.nparam - 1)

    def refresh_adjoint(self):
        sp = dev.stream_ptr()
        for l, (off, co, ci) in enumerate(self.lay):
            mkgpu.check("mk_conv3x3_adjoint_weights", self.w_ptr(l), co, ci, self.w_ptr(l, True), sp)

    def _conv(self, x, l, y, H, W, dil, adjoint=False, accumulate=False):
        off, co, ci = self.lay[l]
        cin, cout = (co, ci) if adjoint else (ci, co)
        mkgpu.check("mk_conv3x3_f32", x.data_ptr(), self.B, cin, H, W, self.w_ptr(l, adjoint), cout, dil,
                    y.data_ptr(), int(accumulate), dev.stream_ptr())

    def _wgrad(self, a, co, x, ci, l, H, W, dil, grad, accumulate, work, work_bytes):
        mkgpu.check("mk_conv3x3_wgrad_f32", a.data_ptr(), co, x.data_ptr(), ci, self.B, H, W, dil,
                    grad.data_ptr() + 4 * self.lay[l][0], int(accumulate), work.data_ptr(), work_bytes,
                    dev.stream_ptr())

    def _f32(self, *shape):
        return torch.empty(shape, dtype=torch.float32, device=self.device)


class ChainNet(_Base):
    def __init__(self, params, dils, B, H, W, train=False):
        super().__init__("chain", dils, params, B, H, W)
        L = self.nlayer
        shp16 = (self.B, CH, self.H, self.W)
        self.train = train
        if train:
            self.h = [None] + [self._f32(*shp16) for _ in range(L)]  # h[1..L]
            self.g = [None] + [self._f32(*shp16) for _ in range(L - 1)]  # g[1..L-1]
        self.bufA = self._f32(*shp16)
        self.bufB = self._f32(*shp16)

    def conv(self, x, l, y, adjoint=False, accumulate=False):
        self._conv(x, l, y, self.H, self.W, self.spec[l], adjoint, accumulate)

    def workspace_bytes(self):
        return max(int(mkgpu.lib.mk_conv3x3_wgrad_workspace_bytes(self.B, CH, CH, self.H, self.W, d))
                   for d in self.spec)

    def forward(self, u, g0):
        """g0 = K^T K u (keeps the activations when built for training)."""
        L = self.nlayer
        if self.train:
            h = self.h
            self.conv(u, 0, h[1])
            for l in range(1, L):
                self.conv(h[l], l, h[l + 1])
            gin = h[L]
            for l in range(L, 1, -1):
                self.conv(gin, l - 1, self.g[l - 1], adjoint=True)
                gin = self.g[l - 1]
            self.conv(gin, 0, g0, adjoint=True)
            return
        a, b = self.bufA, self.bufB
        self.conv(u, 0, a)
        for l in range(1, L):
            self.conv(a, l, b)
            a, b = b, a
        for l in range(L - 1, 0, -1):
            self.conv(a, l, b, adjoint=True)
            a, b = b, a
        self.conv(a, 0, g0, adjoint=True)

    def backward(self, u, q, grad, work, work_bytes):
        """dL/dW from q = dL/dt (t = alpha u + K^T K u): with a_0 = q,
        a_l = conv_l(a_{l-1}), b_L = a_L, b_{l-1} = conv_l^T(b_l):
        dW_l = wgrad(g_l, a_{l-1}) + wgrad(b_l, h_{l-1})."""
        L, h = self.nlayer, self.h
        H, W = self.H, self.W
        gl = [None] + [self.g[l] for l in range(1, L)] + [h[L]]
        a_prev, a_prev_c = q, 1
        bufs = [self.bufA, self.bufB]
        for l in range(1, L + 1):
            self._wgrad(gl[l], CH, a_prev, a_prev_c, l - 1, H, W, self.spec[l - 1], grad, False, work, work_bytes)
            a_next = bufs[l % 2]
            self.conv(a_prev, l - 1, a_next)
            a_prev, a_prev_c = a_next, CH
        b = a_prev
        for l in range(L, 0, -1):
            x_prev, x_c = (h[l - 1], CH) if l > 1 else (u, 1)
            self._wgrad(b, CH, x_prev, x_c, l - 1, H, W, self.spec[l - 1], grad, True, work, work_bytes)
            if l > 1:
                b_next = self.bufA if b is self.bufB else self.bufB
                self.conv(b, l - 1, b_next, adjoint=True)
                b = b_next


class LevelsNet(_Base):
    def __init__(self, params, factors, B, H, W, train=False, arch="levels"):
        super().__init__(arch, factors, params, B, H, W)
        kind = "tent" if arch == "tent" else "box"
        self._rname, self._pname = "mk_restrict_%s_f32" % kind, "mk_prolong_%s_f32" % kind
        f = self.spec
        if f[0] != 1 or any(f[i] % f[i - 1] for i in range(1, len(f))):
            raise ValueError("level factors must start at 1 and each divide the next: %s" % (f,))
        self.steps = [1] + [f[i] // f[i - 1] for i in range(1, len(f))]
        self.dims = [(self.H, self.W)]
        for st in self.steps[1:]:
            h, w = self.dims[-1]
            self.dims.append(((h + st - 1) // st, (w + st - 1) // st))
        B = self.B
        self.u = [None] + [self._f32(B, 1, h, w) for h, w in self.dims[1:]]  # u[0] is the caller's
        self.c = [self._f32(B, CH, h, w) for h, w in self.dims]
        self.gl = [None] + [self._f32(B, 1, h, w) for h, w in self.dims[1:]]  # g[0] is the caller's
        self.train = train
        if train:
            self.ql = [None] + [self._f32(B, 1, h, w) for h, w in self.dims[1:]]
            self.cq = self._f32(B, CH, self.H, self.W)

    def workspace_bytes(self):
        return max(int(mkgpu.lib.mk_conv3x3_wgrad_workspace_bytes(self.B, CH, 1, h, w, 1)) for h, w in self.dims)

    def _restrict(self, x, l, y):
        h, w = self.dims[l - 1]
        mkgpu.check(self._rname, x.data_ptr(), self.B, h, w, self.steps[l], y.data_ptr(), dev.stream_ptr())

    def _prolong(self, y, l, x, accumulate):
        h, w = self.dims[l - 1]
        mkgpu.check(self._pname, y.data_ptr(), self.B, h, w, self.steps[l], x.data_ptr(), int(accumulate),
                    dev.stream_ptr())

    def forward(self, u, g0):
        L = self.nlayer
        us = [u] + self.u[1:]
        gs = [g0] + self.gl[1:]
        for l in range(1, L):
            self._restrict(us[l - 1], l, us[l])
        for l in range(L):
            h, w = self.dims[l]
            self._conv(us[l], l, self.c[l], h, w, 1)
            self._conv(self.c[l], l, gs[l], h, w, 1, adjoint=True)
        for l in range(L - 1, 0, -1):  # Horner: g_{l-1} += R_l^T g_l
            self._prolong(gs[l], l, gs[l - 1], accumulate=True)

    def backward(self, u, q, grad, work, work_bytes):
        """dW_l = wgrad(C_l u_l, q_l) + wgrad(C_l q_l, u_l), q_l = R^l q."""
        L = self.nlayer
        us = [u] + self.u[1:]
        qs = [q] + self.ql[1:]
        for l in range(1, L):
            self._restrict(qs[l - 1], l, qs[l])
        for l in range(L):
            h, w = self.dims[l]
            cq = self.cq.view(-1)[: self.B * CH * h * w].view(self.B, CH, h, w)
            self._conv(qs[l], l, cq, h, w, 1)
            self._wgrad(self.c[l], CH, qs[l], 1, l, h, w, 1, grad, False, work, work_bytes)
            self._wgrad(cq, CH, us[l], 1, l, h, w, 1, grad, True, work, work_bytes)


def make_net(arch, spec, params, B, H, W, train=False):
    if arch == "chain":
        return ChainNet(params, spec, B, H, W, train)
    if arch in ("levels", "tent"):
        return LevelsNet(params, spec, B, H, W, train, arch)
    raise ValueError("unknown architecture %r" % (arch,))
