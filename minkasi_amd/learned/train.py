"""Training of the learned preconditioner with the library's own forward and
backward kernels.

With u = s r and t = alpha u + K^T K u, M r = s t.  The network object
(``nets.py``) computes K^T K u and back-propagates q = dL/dt into the weights
exactly (K^T K is linear in u and quadratic in each layer); dL/dalpha =
sum q u.  Two losses feed it:

crop loss    L = sum m (s t - e)^2 over crops (r, e, s, m) -- residual,
             error, sqrt(1/hits), per-pixel weight; q = 2 m s (s t - e).
energy loss  L = ||M r - e||_A^2 = (z - e).(A z - r), z = M r, A e = r, on
             full maps: the A-norm error left after the preconditioned step
             x + M r.  Per eigenmode of A it is lambda e^2 (m lambda - 1)^2,
             so it penalises over- as well as under-amplification; it needs
             A z (one application of the mapmaking operator);
             q = 2 s (A z - r).

An Adam step (mk_adam_f32) follows.  No host synchronisation inside a step.
"""
import numpy as np
import torch

from .. import device as dev
from ..lib import mkgpu
from .nets import make_net


class Trainer:
    def __init__(self, params, arch, spec, B, H, W, lr=2e-3, betas=(0.9, 0.999), eps=1e-8):
        d = dev.device()
        p = torch.from_numpy(np.asarray(params, dtype=np.float32)).to(d)
        self.net = make_net(arch, spec, p, B, H, W, train=True)
        self.B, self.H, self.W = B, H, W
        shp1 = (B, 1, H, W)
        f32 = dict(dtype=torch.float32, device=d)
        self.u = torch.empty(shp1, **f32)
        self.g0 = torch.empty(shp1, **f32)
        self.q = torch.empty(shp1, **f32)
        n = self.net.nparam
        self.grad = torch.zeros(n, **f32)
        self.m = torch.zeros(n, **f32)
        self.v = torch.zeros(n, **f32)
        self.sc = torch.zeros(4, dtype=torch.float64, device=d)
        wb = self.net.workspace_bytes()
        self.wwork = torch.empty(wb // 4 + 64, **f32)
        self.wwork_bytes = self.wwork.numel() * 4
        lb = int(mkgpu.lib.mk_precon_loss_workspace_bytes())
        self.lwork = torch.empty(lb // 8 + 8, dtype=torch.float64, device=d)
        self.lr, self.betas, self.eps = lr, betas, eps
        self.nstep = 0
        self.z = self.az = None

    @property
    def params(self):
        return self.net.params

    def _check(self, **ts):
        for name, t in ts.items():
            dev.ptr(t, torch.float32, name)
            if t.numel() != self.net.n:
                raise ValueError("%s has %d elements, expected %d" % (name, t.numel(), self.net.n))

    def forward(self, r, s):
        """u = s r and g_0 = K^T K u (activations kept for backward)."""
        self.net.refresh_adjoint()
        mkgpu.check("mk_mul_f32", self.u.data_ptr(), s.data_ptr(), r.data_ptr(), self.net.n, dev.stream_ptr())
        self.net.forward(self.u, self.g0)

    def backward(self):
        self.net.backward(self.u, self.q, self.grad, self.wwork, self.wwork_bytes)

    def forward_backward(self, r, e, s, m):
        """Crop loss: loss (device FP64) in sc[0], gradients in self.grad,
        dL/dalpha in sc[1]."""
        self._check(r=r, e=e, s=s, m=m)
        self.forward(r, s)
        mkgpu.check("mk_precon_loss_f32", self.g0.data_ptr(), self.u.data_ptr(), s.data_ptr(), e.data_ptr(),
                    m.data_ptr(), self.net.log_alpha_ptr, self.net.n, self.q.data_ptr(), self.sc.data_ptr(),
                    self.lwork.data_ptr(), self.lwork.numel() * 8, dev.stream_ptr())
        self.backward()

    def energy_buffers(self):
        if self.z is None:
            self.z = torch.empty((self.H, self.W), dtype=torch.float32, device=self.u.device)
            self.az = torch.empty_like(self.z)
        return self.z, self.az

    def energy_forward_backward(self, r, e, s, apply_A):
        """Energy loss on one full map (B = 1); apply_A(z, out) writes A z.
        Loss in sc[0], gradients in self.grad, dL/dalpha in sc[1]."""
        if self.B != 1:
            raise ValueError("the energy loss runs on one full map at a time (B = 1)")
        self._check(r=r, e=e, s=s)
        z, az = self.energy_buffers()
        sp = dev.stream_ptr()
        self.forward(r, s)
        mkgpu.check("mk_precon_finish_f32", z.data_ptr(), s.data_ptr(), self.net.log_alpha_ptr, self.u.data_ptr(),
                    self.g0.data_ptr(), self.net.n, sp)
        apply_A(z, az)
        mkgpu.check("mk_precon_aloss_f32", z.data_ptr(), az.data_ptr(), r.data_ptr(), e.data_ptr(), s.data_ptr(),
                    self.u.data_ptr(), self.net.n, self.q.data_ptr(), self.sc.data_ptr(), self.lwork.data_ptr(),
                    self.lwork.numel() * 8, sp)
        self.backward()

    def adam(self, gscale=1.0):
        self.nstep += 1
        mkgpu.check("mk_adam_f32", self.params.data_ptr(), self.grad.data_ptr(), self.m.data_ptr(),
                    self.v.data_ptr(), self.net.nparam, self.sc.data_ptr() + 8, self.lr, self.betas[0],
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
