"""Training of the learned preconditioner with the library's own forward and
backward kernels.

For a batch of crops (r, e, s, m) -- residual, error, sqrt(1/hits) and a
per-pixel loss weight -- one step evaluates

    u = s r,  h_l = conv_l(h_{l-1}) (h_0 = u),  g_{l-1} = conv_l^T(g_l) (g_L = h_L),
    L = sum m (s (alpha u + g_0) - e)^2

and back-propagates it exactly (the chain rule of a linear K^T K): with
q = dL/d(alpha u + g_0), a_0 = q, a_l = conv_l(a_{l-1}) and
b_L = a_L, b_{l-1} = conv_l^T(b_l),

    dL/dW_l = wgrad(g_l, a_{l-1}) + wgrad(b_l, h_{l-1}),   dL/dalpha = sum q u,

followed by an Adam step (mk_adam_f32).  Every step is 15 convolutions, 8
weight-gradient reductions, one loss reduction and one update, all in the
HIP library; no host synchronisation inside a step.

Two losses share that backward pass.  The crop loss above fits M r to e
pixel by pixel.  The energy loss L = ||M r - e||_A^2 = (z - e).(A z - r)
(z = M r, A e = r) is the A-norm error left after the preconditioned step
x + M r; per eigenmode of A it is lambda e^2 (m lambda - 1)^2, so it
penalises over- as well as under-amplification, which is what the CG
iteration count depends on.  It needs A z (one application of the mapmaking
operator) and runs on full maps; q = 2 s (A z - r).
"""
import numpy as np
import torch

from .. import device as dev
from ..lib import mkgpu
from .precon import CH, ConvChain, param_layout


class Trainer:
    def __init__(self, params, dils, B, H, W, lr=2e-3, betas=(0.9, 0.999), eps=1e-8):
        d = dev.device()
        p = torch.from_numpy(np.asarray(params, dtype=np.float32)).to(d)
        self.chain = ConvChain(p, dils, B, H, W)
        c = self.chain
        self.L = c.nlayer
        self.B, self.H, self.W = B, H, W
        shp1 = (B, 1, H, W)
        shp16 = (B, CH, H, W)
        f32 = dict(dtype=torch.float32, device=d)
        self.u = torch.empty(shp1, **f32)
        self.g0 = torch.empty(shp1, **f32)
        self.q = torch.empty(shp1, **f32)
        self.h = [None] + [torch.empty(shp16, **f32) for _ in range(self.L)]  # h[1..L]
        self.g = [None] + [torch.empty(shp16, **f32) for _ in range(self.L - 1)]  # g[1..L-1]
        self.A = torch.empty(shp16, **f32)
        self.Bf = torch.empty(shp16, **f32)
        self.grad = torch.zeros(c.nparam, **f32)
        self.m = torch.zeros(c.nparam, **f32)
        self.v = torch.zeros(c.nparam, **f32)
        self.sc = torch.zeros(4, dtype=torch.float64, device=d)
        wb = max(int(mkgpu.lib.mk_conv3x3_wgrad_workspace_bytes(B, CH, CH, H, W, dd)) for dd in c.dils)
        self.wwork = torch.empty(wb // 4 + 64, **f32)
        self.wwork_bytes = self.wwork.numel() * 4
        lb = int(mkgpu.lib.mk_precon_loss_workspace_bytes())
        self.lwork = torch.empty(lb // 8 + 8, dtype=torch.float64, device=d)
        self.lr, self.betas, self.eps = lr, betas, eps
        self.nstep = 0

    @property
    def params(self):
        return self.chain.params

    def _wgrad(self, a, co, x, ci, l, accumulate):
        off = self.chain.lay[l][0]
        mkgpu.check("mk_conv3x3_wgrad_f32", a.data_ptr(), co, x.data_ptr(), ci, self.B, self.H, self.W,
                    self.chain.dils[l], self.grad.data_ptr() + 4 * off, int(accumulate), self.wwork.data_ptr(),
                    self.wwork_bytes, dev.stream_ptr())

    def _check(self, **ts):
        for name, t in ts.items():
            dev.ptr(t, torch.float32, name)
            if t.numel() != self.chain.n:
                raise ValueError("%s has %d elements, expected %d" % (name, t.numel(), self.chain.n))

    def forward(self, r, s):
        """u = s r, h_l, g_l and g_0 = K^T K u (t = alpha u + g_0)."""
        c, L, sp = self.chain, self.L, dev.stream_ptr()
        c.refresh_adjoint()
        mkgpu.check("mk_mul_f32", self.u.data_ptr(), s.data_ptr(), r.data_ptr(), c.n, sp)
        h = self.h
        c.conv(self.u, 0, h[1])
        for l in range(1, L):
            c.conv(h[l], l, h[l + 1])
        gin = h[L]  # g_L
        for l in range(L, 1, -1):  # g_{l-1} = conv_l^T g_l
            c.conv(gin, l - 1, self.g[l - 1], adjoint=True)
            gin = self.g[l - 1]
        c.conv(gin, 0, self.g0, adjoint=True)

    def backward(self):
        """Weight gradients into self.grad from q = dL/dt in self.q."""
        c, L, h = self.chain, self.L, self.h
        # adjoint chain: a_0 = q, dW_l = wgrad(g_l, a_{l-1}), a_l = conv_l(a_{l-1})
        gl = [None] + [self.g[l] for l in range(1, L)] + [h[L]]
        a_prev, a_prev_c = self.q, 1
        bufs = [self.A, self.Bf]
        for l in range(1, L + 1):
            self._wgrad(gl[l], CH, a_prev, a_prev_c, l - 1, accumulate=False)
            a_next = bufs[l % 2]
            c.conv(a_prev, l - 1, a_next)
            a_prev, a_prev_c = a_next, CH
        # forward chain: b_L = a_L, dW_l += wgrad(b_l, h_{l-1}), b_{l-1} = conv_l^T b_l
        b = a_prev
        for l in range(L, 0, -1):
            x_prev, x_c = (h[l - 1], CH) if l > 1 else (self.u, 1)
            self._wgrad(b, CH, x_prev, x_c, l - 1, accumulate=True)
            if l > 1:
                b_next = self.A if b is self.Bf else self.Bf
                c.conv(b, l - 1, b_next, adjoint=True)
                b = b_next

    def forward_backward(self, r, e, s, m):
        """Crop loss sum m (M r - e)^2: loss (device FP64) in sc[0],
        gradients in self.grad, dL/dalpha in sc[1]."""
        self._check(r=r, e=e, s=s, m=m)
        c, sp = self.chain, dev.stream_ptr()
        self.forward(r, s)
        mkgpu.check("mk_precon_loss_f32", self.g0.data_ptr(), self.u.data_ptr(), s.data_ptr(), e.data_ptr(),
                    m.data_ptr(), c.log_alpha_ptr, c.n, self.q.data_ptr(), self.sc.data_ptr(), self.lwork.data_ptr(),
                    self.lwork.numel() * 8, sp)
        self.backward()

    def energy_forward_backward(self, r, e, s, apply_A):
        """Energy-norm loss ||M r - e||_A^2 on full maps (B = 1):
        apply_A(z, out) writes A z.  Loss in sc[0], gradients in self.grad,
        dL/dalpha in sc[1]; z = M r is left in self.z."""
        if self.B != 1:
            raise ValueError("the energy loss runs on one full map at a time (B = 1)")
        self._check(r=r, e=e, s=s)
        c, sp = self.chain, dev.stream_ptr()
        if not hasattr(self, "z"):
            self.z = torch.empty((self.H, self.W), dtype=torch.float32, device=self.u.device)
            self.az = torch.empty_like(self.z)
        self.forward(r, s)
        mkgpu.check("mk_precon_finish_f32", self.z.data_ptr(), s.data_ptr(), c.log_alpha_ptr, self.u.data_ptr(),
                    self.g0.data_ptr(), c.n, sp)
        apply_A(self.z, self.az)
        mkgpu.check("mk_precon_aloss_f32", self.z.data_ptr(), self.az.data_ptr(), r.data_ptr(), e.data_ptr(),
                    s.data_ptr(), self.u.data_ptr(), c.n, self.q.data_ptr(), self.sc.data_ptr(),
                    self.lwork.data_ptr(), self.lwork.numel() * 8, sp)
        self.backward()

    def adam(self, gscale=1.0):
        self.nstep += 1
        mkgpu.check("mk_adam_f32", self.params.data_ptr(), self.grad.data_ptr(), self.m.data_ptr(),
                    self.v.data_ptr(), self.chain.nparam, self.sc.data_ptr() + 8, self.lr, self.betas[0],
                    self.betas[1], self.eps, self.nstep, gscale, dev.stream_ptr())

    def step(self, r, e, s, m, gscale=1.0):
        self.forward_backward(r, e, s, m)
        self.adam(gscale)

    def step_energy(self, r, e, s, apply_A, gscale=1.0):
        self.energy_forward_backward(r, e, s, apply_A)
        self.adam(gscale)

    def loss(self):
        return float(self.sc[0].item())

    def dalpha(self):
        return float(self.sc[1].item())
