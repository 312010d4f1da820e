"""Oracle: learned map-space preconditioner, forward and backward, FP64 numpy
(test infrastructure).

No reference counterpart: the reference's PCG takes any object with
``__mul__(Mapset)`` as its preconditioner (minkasi/mapmaking/pcg.py:110-111,
157-158; null_precon :241-255), and this is one such object:

    M r = S (alpha u + K^T K u),  u = S r,  S = diag(sqrt(1/hits)) (0 if unhit)

K is a chain of linear 3x3 convolutions with per-layer dilation; tap
d = 3*di + dj reads offset (dil*(di-1), dil*(dj-1)), zero padding; weights
w[Cout][Cin][9]; activations [B][C][H][W].

Training losses for (r, e) pairs: the crop loss L = sum m (M r - e)^2 with
per-pixel weights m, and the energy loss L = (M r - e)^T A (M r - e) with
A e = r; parameters = the conv weights and log(alpha).

LevelsNet replaces the chain by one conv per level of box-restricted maps
(restrict / prolong below): K^T K = sum_l (R^l)^T C_l^T C_l R^l.

Pins (tests/test_oracle_learned.py): conv3x3 against scipy.signal.correlate2d
with the explicitly dilated kernel; the adjoint identity <conv x, y> =
<x, conv^T y>; wgrad against brute force with unit weights; loss_and_grad
and energy_loss_and_grad against central finite differences; M symmetric and positive semidefinite;
zero weights with alpha = 1 reduce M to the hits-count Jacobi 1/hits;
restrict/prolong adjoint to each other and restrict equal to block means;
LevelsNet gradients against finite differences and symmetric PSD.
"""
import numpy as np


def conv3x3(x, w, dil=1):
    """y[b,o] = sum_{c,d} w[o,c,d] x[b,c](p + off_d), zero padding."""
    x = np.asarray(x, dtype=np.float64)
    B, C, H, W = x.shape
    O = w.shape[0]
    assert w.shape == (O, C, 9)
    pad = np.zeros((B, C, H + 2 * dil, W + 2 * dil))
    pad[:, :, dil:dil + H, dil:dil + W] = x
    y = np.zeros((B, O, H, W))
    for d in range(9):
        di, dj = d // 3, d % 3
        sh = pad[:, :, di * dil:di * dil + H, dj * dil:dj * dil + W]
        y += np.einsum("oc,bchw->bohw", w[:, :, d], sh)
    return y


def adjoint_weights(w):
    """wt[c][o][8-d] = w[o][c][d]: conv3x3(., wt) is the adjoint of conv3x3(., w)."""
    return np.ascontiguousarray(np.transpose(w, (1, 0, 2))[:, :, ::-1])


def wgrad(a, x, dil=1):
    """g[o,c,d] = sum_{b,p} a[b,o,p] x[b,c,p+off_d] = d<a, conv3x3(x, w)>/dw."""
    a = np.asarray(a, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    B, C, H, W = x.shape
    O = a.shape[1]
    pad = np.zeros((B, C, H + 2 * dil, W + 2 * dil))
    pad[:, :, dil:dil + H, dil:dil + W] = x
    g = np.zeros((O, C, 9))
    for d in range(9):
        di, dj = d // 3, d % 3
        sh = pad[:, :, di * dil:di * dil + H, dj * dil:dj * dil + W]
        g[:, :, d] = np.einsum("bohw,bchw->oc", a, sh)
    return g


class Net:
    """Weights ws[l] of shape (Cout, Cin, 9), dilations dils[l], log_alpha."""

    def __init__(self, ws, dils, log_alpha=0.0):
        self.ws = [np.asarray(w, dtype=np.float64) for w in ws]
        self.dils = list(dils)
        self.log_alpha = float(log_alpha)

    def K(self, u):
        h = u
        for w, d in zip(self.ws, self.dils):
            h = conv3x3(h, w, d)
        return h

    def KT(self, y):
        g = y
        for w, d in zip(self.ws[::-1], self.dils[::-1]):
            g = conv3x3(g, adjoint_weights(w), d)
        return g

    def apply(self, r, s):
        """M r for r, s of shape (B, 1, H, W) (or (H, W))."""
        shape = np.shape(r)
        r = np.asarray(r, dtype=np.float64).reshape((-1, 1) + shape[-2:])
        s = np.asarray(s, dtype=np.float64).reshape((-1, 1) + shape[-2:])
        u = s * r
        z = s * (np.exp(self.log_alpha) * u + self.KT(self.K(u)))
        return z.reshape(shape)

    def _forward(self, r, s):
        u = s * r
        hs = [u]
        for w, d in zip(self.ws, self.dils):
            hs.append(conv3x3(hs[-1], w, d))
        L_ = len(self.ws)
        gs = [None] * (L_ + 1)  # gs[l]: input of the adjoint of layer l (gs[L] = K u)
        gs[L_] = hs[L_]
        for l in range(L_, 0, -1):
            gs[l - 1] = conv3x3(gs[l], adjoint_weights(self.ws[l - 1]), self.dils[l - 1])
        return u, hs, gs

    def loss_and_grad(self, r, e, s, m):
        """L = sum m (M r - e)^2 and dL/d(ws), dL/d(log_alpha); arrays (B,1,H,W)."""
        alpha = np.exp(self.log_alpha)
        u, hs, gs = self._forward(r, s)
        t = alpha * u + gs[0]
        delta = s * t - e
        loss = float(np.sum(m * delta * delta))
        q = 2.0 * m * s * delta  # dL/dt
        grads, dla = self._backward(q, u, hs, gs)
        return loss, grads, dla

    def energy_loss_and_grad(self, r, e, s, A):
        """L = (M r - e)^T A (M r - e) for one (H, W) map, A a dense
        symmetric (H*W, H*W) matrix; gradients as loss_and_grad."""
        shape = np.shape(r)
        r4 = np.asarray(r, dtype=np.float64).reshape((1, 1) + shape)
        s4 = np.asarray(s, dtype=np.float64).reshape((1, 1) + shape)
        alpha = np.exp(self.log_alpha)
        u, hs, gs = self._forward(r4, s4)
        z = (s4 * (alpha * u + gs[0])).ravel()
        d = z - np.asarray(e, dtype=np.float64).ravel()
        Ad = A @ d
        loss = float(d @ Ad)
        q = (2.0 * s4.ravel() * Ad).reshape(r4.shape)
        grads, dla = self._backward(q, u, hs, gs)
        return loss, grads, dla

    def _backward(self, q, u, hs, gs):
        """Weight and log(alpha) gradients from q = dL/d(alpha u + g_0)."""
        alpha = np.exp(self.log_alpha)
        L_ = len(self.ws)
        dalpha = float(np.sum(q * u))
        grads = [np.zeros_like(w) for w in self.ws]
        # adjoint chain: dL/dg_l = conv_l(dL/dg_{l-1}); dW_l += wgrad(g_l, dL/dg_{l-1})
        a = q
        for l in range(1, L_ + 1):
            grads[l - 1] += wgrad(gs[l], a, self.dils[l - 1])
            a = conv3x3(a, self.ws[l - 1], self.dils[l - 1])
        # forward chain: dL/dh_L = a; dW_l += wgrad(dL/dh_l, h_{l-1})
        bgrad = a
        for l in range(L_, 0, -1):
            grads[l - 1] += wgrad(bgrad, hs[l - 1], self.dils[l - 1])
            if l > 1:
                bgrad = conv3x3(bgrad, adjoint_weights(self.ws[l - 1]), self.dils[l - 1])
        return grads, dalpha * alpha


def restrict(x, f):
    """Box restriction by f, stride f: y[b,c,I,J] = f^-2 sum of block (I, J),
    zero beyond the grid; output ceil(H/f) x ceil(W/f)."""
    x = np.asarray(x, dtype=np.float64)
    B, C, H, W = x.shape
    Hc, Wc = -(-H // f), -(-W // f)
    pad = np.zeros((B, C, Hc * f, Wc * f))
    pad[:, :, :H, :W] = x
    return pad.reshape(B, C, Hc, f, Wc, f).sum(axis=(3, 5)) / (f * f)


def prolong(y, f, H, W):
    """Adjoint of restrict: x[b,c,i,j] = f^-2 y[b,c,i//f,j//f]."""
    y = np.asarray(y, dtype=np.float64)
    x = np.repeat(np.repeat(y, f, axis=2), f, axis=3)[:, :, :H, :W]
    return x / (f * f)


class LevelsNet(Net):
    """K u = [C_0 u; C_1 R^1 u; ...]: one 3x3 conv (Cin = 1) per level, level l
    on the map restricted by factors[l] (successive box restrictions by
    factors[l] / factors[l-1]); K^T K = sum_l (R^l)^T C_l^T C_l R^l."""

    def __init__(self, ws, factors, log_alpha=0.0):
        super().__init__(ws, [1] * len(ws), log_alpha)
        self.factors = list(factors)
        self.steps = [1] + [self.factors[i] // self.factors[i - 1] for i in range(1, len(self.factors))]

    def _levels(self, u):
        us = [u]
        for st in self.steps[1:]:
            us.append(restrict(us[-1], st))
        return us

    def _ktk(self, us):
        gs = [conv3x3(conv3x3(ul, w, 1), adjoint_weights(w), 1) for ul, w in zip(us, self.ws)]
        acc = gs[-1]
        for l in range(len(gs) - 1, 0, -1):
            H, W = gs[l - 1].shape[-2:]
            acc = gs[l - 1] + prolong(acc, self.steps[l], H, W)
        return acc

    def apply(self, r, s):
        shape = np.shape(r)
        r = np.asarray(r, dtype=np.float64).reshape((-1, 1) + shape[-2:])
        s = np.asarray(s, dtype=np.float64).reshape((-1, 1) + shape[-2:])
        u = s * r
        z = s * (np.exp(self.log_alpha) * u + self._ktk(self._levels(u)))
        return z.reshape(shape)

    def _forward(self, r, s):
        u = s * r
        us = self._levels(u)
        return u, us, [self._ktk(us)]

    def _backward(self, q, u, us, gs):
        alpha = np.exp(self.log_alpha)
        dalpha = float(np.sum(q * u))
        qs = self._levels(q)
        grads = []
        for ul, ql, w in zip(us, qs, self.ws):
            grads.append(wgrad(conv3x3(ul, w, 1), ql, 1) + wgrad(conv3x3(ql, w, 1), ul, 1))
        return grads, dalpha * alpha


def tent_matrix(H, f):
    """(ceil(H/f), H) full-weighting weights: w[I, i] = max(0, 2f - |2i - (2fI + f - 1)|) / (2 f^2)."""
    Hc = -(-H // f)
    I = np.arange(Hc)[:, None]
    i = np.arange(H)[None, :]
    d = np.abs(2 * i - (2 * f * I + f - 1))
    return np.where(d < 2 * f, (2 * f - d) / (2.0 * f * f), 0.0)


def restrict_tent(x, f):
    x = np.asarray(x, dtype=np.float64)
    Wh = tent_matrix(x.shape[2], f)
    Ww = tent_matrix(x.shape[3], f)
    return np.einsum("Ii,bcij,Jj->bcIJ", Wh, x, Ww)


def prolong_tent(y, f, H, W):
    y = np.asarray(y, dtype=np.float64)
    return np.einsum("Ii,bcIJ,Jj->bcij", tent_matrix(H, f), y, tent_matrix(W, f))


class TentLevelsNet(LevelsNet):
    """LevelsNet with full-weighting restriction and bilinear prolongation."""

    def _levels(self, u):
        us = [u]
        for st in self.steps[1:]:
            us.append(restrict_tent(us[-1], st))
        return us

    def _ktk(self, us):
        gs = [conv3x3(conv3x3(ul, w, 1), adjoint_weights(w), 1) for ul, w in zip(us, self.ws)]
        acc = gs[-1]
        for l in range(len(gs) - 1, 0, -1):
            H, W = gs[l - 1].shape[-2:]
            acc = gs[l - 1] + prolong_tent(acc, self.steps[l], H, W)
        return acc
