"""Pins the learned-preconditioner oracle (CPU only)."""
import numpy as np
import pytest
from scipy.signal import correlate2d

from oracle import learned as L


def dilated_kernel(w9, dil):
    k = np.zeros((2 * dil + 1, 2 * dil + 1))
    for d in range(9):
        k[(d // 3) * dil, (d % 3) * dil] = w9[d]
    return k


@pytest.mark.parametrize("dil", [1, 2, 5])
def test_conv_matches_correlate2d(dil):
    rng = np.random.default_rng(dil)
    x = rng.standard_normal((2, 3, 13, 17))
    w = rng.standard_normal((4, 3, 9))
    y = L.conv3x3(x, w, dil)
    for b in range(2):
        for o in range(4):
            ref = sum(correlate2d(x[b, c], dilated_kernel(w[o, c], dil), mode="same") for c in range(3))
            assert np.allclose(y[b, o], ref, atol=1e-12)


@pytest.mark.parametrize("dil", [1, 3])
def test_adjoint_identity(dil):
    rng = np.random.default_rng(10 + dil)
    x = rng.standard_normal((2, 3, 11, 9))
    y = rng.standard_normal((2, 5, 11, 9))
    w = rng.standard_normal((5, 3, 9))
    lhs = np.sum(L.conv3x3(x, w, dil) * y)
    rhs = np.sum(x * L.conv3x3(y, L.adjoint_weights(w), dil))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


@pytest.mark.parametrize("dil", [1, 4])
def test_wgrad_brute_force(dil):
    rng = np.random.default_rng(20 + dil)
    x = rng.standard_normal((2, 3, 10, 12))
    a = rng.standard_normal((2, 2, 10, 12))
    g = L.wgrad(a, x, dil)
    for o in range(2):
        for c in range(3):
            for d in range(9):
                e = np.zeros((2, 3, 9))
                e[o, c, d] = 1.0
                assert abs(g[o, c, d] - np.sum(a * L.conv3x3(x, e, dil))) < 1e-10


def small_net(rng, dils=(1, 2, 1, 3), ch=4):
    ws = [rng.standard_normal((ch, 1, 9)) * 0.3]
    for _ in dils[1:]:
        ws.append(rng.standard_normal((ch, ch, 9)) * 0.3)
    return L.Net(ws, dils, log_alpha=0.2)


def test_loss_grad_finite_differences():
    rng = np.random.default_rng(3)
    net = small_net(rng)
    shp = (2, 1, 12, 10)
    r, e = rng.standard_normal(shp), rng.standard_normal(shp)
    s = rng.uniform(0.2, 1.5, shp)
    s[0, 0, 0, :3] = 0.0
    m = (rng.uniform(size=shp) > 0.2).astype(float)
    loss, grads, dla = net.loss_and_grad(r, e, s, m)
    h = 1e-6
    for l in range(len(net.ws)):
        for idx in [(0, 0, 0), (1, 0, 4), (net.ws[l].shape[0] - 1, net.ws[l].shape[1] - 1, 8)]:
            old = net.ws[l][idx]
            net.ws[l][idx] = old + h
            lp = net.loss_and_grad(r, e, s, m)[0]
            net.ws[l][idx] = old - h
            lm = net.loss_and_grad(r, e, s, m)[0]
            net.ws[l][idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - grads[l][idx]) < 1e-6 * max(1.0, abs(fd)), (l, idx, fd, grads[l][idx])
    old = net.log_alpha
    net.log_alpha = old + h
    lp = net.loss_and_grad(r, e, s, m)[0]
    net.log_alpha = old - h
    lm = net.loss_and_grad(r, e, s, m)[0]
    net.log_alpha = old
    assert abs((lp - lm) / (2 * h) - dla) < 1e-6 * max(1.0, abs(dla))


def test_loss_matches_apply():
    rng = np.random.default_rng(4)
    net = small_net(rng)
    shp = (1, 1, 9, 11)
    r, e = rng.standard_normal(shp), rng.standard_normal(shp)
    s = rng.uniform(0.2, 1.5, shp)
    m = rng.uniform(size=shp)
    loss = net.loss_and_grad(r, e, s, m)[0]
    z = net.apply(r, s)
    assert abs(loss - np.sum(m * (z - e) ** 2)) < 1e-10 * loss


def test_precon_symmetric_psd_and_jacobi_limit():
    rng = np.random.default_rng(5)
    net = small_net(rng)
    H, W = 8, 7
    s = rng.uniform(0.1, 2.0, (H, W))
    s[2, 3] = 0.0  # unhit pixel
    M = np.zeros((H * W, H * W))
    for k in range(H * W):
        e = np.zeros(H * W)
        e[k] = 1.0
        M[:, k] = net.apply(e.reshape(H, W), s).ravel()
    assert np.allclose(M, M.T, atol=1e-12 * np.abs(M).max())
    ev = np.linalg.eigvalsh(M)
    assert ev.min() > -1e-12 * ev.max()
    hit = s.ravel() > 0
    assert np.linalg.eigvalsh(M[np.ix_(hit, hit)]).min() > 0  # alpha > 0: SPD on hit pixels
    assert np.all(M[~hit] == 0) and np.all(M[:, ~hit] == 0)
    zero = L.Net([np.zeros_like(w) for w in net.ws], net.dils, log_alpha=0.0)
    r = rng.standard_normal((H, W))
    assert np.allclose(zero.apply(r, s), s * s * r, atol=1e-15)


def test_energy_loss_grad_finite_differences():
    rng = np.random.default_rng(6)
    net = small_net(rng, dils=(1, 2, 1, 1), ch=3)
    H, W = 7, 6
    n = H * W
    Q = rng.standard_normal((n, n))
    A = Q @ Q.T / n + 0.1 * np.eye(n)
    e = rng.standard_normal((H, W))
    r = (A @ e.ravel()).reshape(H, W)
    s = rng.uniform(0.3, 1.2, (H, W))
    loss, grads, dla = net.energy_loss_and_grad(r, e, s, A)
    z = net.apply(r, s).ravel()
    assert abs(loss - (z - e.ravel()) @ A @ (z - e.ravel())) < 1e-10 * loss
    assert abs(loss - (z - e.ravel()) @ (A @ z - r.ravel())) < 1e-9 * loss
    h = 1e-6
    for l in range(len(net.ws)):
        idx = (0, 0, 4)
        old = net.ws[l][idx]
        net.ws[l][idx] = old + h
        lp = net.energy_loss_and_grad(r, e, s, A)[0]
        net.ws[l][idx] = old - h
        lm = net.energy_loss_and_grad(r, e, s, A)[0]
        net.ws[l][idx] = old
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grads[l][idx]) < 1e-6 * max(1.0, abs(fd)), (l, fd, grads[l][idx])
    old = net.log_alpha
    net.log_alpha = old + h
    lp = net.energy_loss_and_grad(r, e, s, A)[0]
    net.log_alpha = old - h
    lm = net.energy_loss_and_grad(r, e, s, A)[0]
    net.log_alpha = old
    assert abs((lp - lm) / (2 * h) - dla) < 1e-6 * max(1.0, abs(dla))


def test_restrict_prolong_adjoint_and_means():
    rng = np.random.default_rng(11)
    x = rng.standard_normal((2, 1, 13, 10))
    y = rng.standard_normal((2, 1, 4, 3))
    rx = L.restrict(x, 4)
    assert rx.shape == (2, 1, 4, 3)
    assert abs(rx[1, 0, 1, 2] - x[1, 0, 4:8, 8:10].sum() / 16) < 1e-14  # partial block: zeros count
    assert abs(np.sum(rx * y) - np.sum(x * L.prolong(y, 4, 13, 10))) < 1e-12


def levels_net(rng, factors=(1, 2, 6), ch=3):
    ws = [rng.standard_normal((ch, 1, 9)) * 0.4 for _ in factors]
    return L.LevelsNet(ws, factors, log_alpha=-0.2)


def test_levels_grad_finite_differences_and_spd():
    rng = np.random.default_rng(12)
    net = levels_net(rng)
    shp = (2, 1, 14, 13)
    r, e = rng.standard_normal(shp), rng.standard_normal(shp)
    s = rng.uniform(0.2, 1.5, shp)
    m = rng.uniform(size=shp)
    loss, grads, dla = net.loss_and_grad(r, e, s, m)
    assert abs(loss - np.sum(m * (net.apply(r, s) - e) ** 2)) < 1e-10 * loss
    h = 1e-6
    for l in range(len(net.ws)):
        for idx in [(0, 0, 0), (2, 0, 7)]:
            old = net.ws[l][idx]
            net.ws[l][idx] = old + h
            lp = net.loss_and_grad(r, e, s, m)[0]
            net.ws[l][idx] = old - h
            lm = net.loss_and_grad(r, e, s, m)[0]
            net.ws[l][idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - grads[l][idx]) < 1e-6 * max(1.0, abs(fd)), (l, idx, fd, grads[l][idx])
    H, W = 9, 8
    s2 = rng.uniform(0.1, 2.0, (H, W))
    M = np.zeros((H * W, H * W))
    for k in range(H * W):
        v = np.zeros(H * W)
        v[k] = 1.0
        M[:, k] = net.apply(v.reshape(H, W), s2).ravel()
    assert np.allclose(M, M.T, atol=1e-12 * np.abs(M).max())
    assert np.linalg.eigvalsh(M).min() > 0


@pytest.mark.parametrize("f", [2, 3, 8])
def test_tent_restriction_adjoint_and_partition_of_unity(f):
    rng = np.random.default_rng(20 + f)
    H, W = 5 * f + 3, 4 * f
    x = rng.standard_normal((1, 1, H, W))
    y = rng.standard_normal((1, 1, -(-H // f), -(-W // f)))
    assert abs(np.sum(L.restrict_tent(x, f) * y) - np.sum(x * L.prolong_tent(y, f, H, W))) < 1e-12
    # interior coarse points average a constant to itself (weights sum to one)
    c = L.restrict_tent(np.ones((1, 1, 8 * f, 8 * f)), f)
    assert np.allclose(c[0, 0, 2:-2, 2:-2], 1.0)
    # bilinear prolongation of a constant is constant in the interior (up to the 1/f^2 scale)
    p = L.prolong_tent(np.ones((1, 1, 8, 8)), f, 8 * f, 8 * f)
    assert np.allclose(p[0, 0, 2 * f:-2 * f, 2 * f:-2 * f], 1.0 / (f * f))


def test_tent_levels_grad_finite_differences():
    rng = np.random.default_rng(21)
    ws = [rng.standard_normal((3, 1, 9)) * 0.4 for _ in range(3)]
    net = L.TentLevelsNet(ws, (1, 2, 6), log_alpha=-0.1)
    shp = (1, 1, 15, 13)
    r, e = rng.standard_normal(shp), rng.standard_normal(shp)
    s = rng.uniform(0.2, 1.5, shp)
    m = rng.uniform(size=shp)
    loss, grads, dla = net.loss_and_grad(r, e, s, m)
    h = 1e-6
    for l in range(3):
        idx = (1, 0, 4)
        old = net.ws[l][idx]
        net.ws[l][idx] = old + h
        lp = net.loss_and_grad(r, e, s, m)[0]
        net.ws[l][idx] = old - h
        lm = net.loss_and_grad(r, e, s, m)[0]
        net.ws[l][idx] = old
        assert abs((lp - lm) / (2 * h) - grads[l][idx]) < 1e-6 * max(1.0, abs(grads[l][idx]))


def test_level_scales_scale_each_level_term():
    """scale_levels (the per-level calibration of the learned preconditioner)
    multiplies level l's K^T K term by c_l: with the oracle's tent-levels net
    M(c) r = alpha S^2 r + sum_l c_l T_l r, each T_l measured as M(c = 2 at l,
    1 elsewhere) r - M(1) r."""
    from minkasi_amd.learned.nets import param_layout
    from minkasi_amd.learned.run import scale_levels
    rng = np.random.default_rng(21)
    factors = (1, 2, 4)
    lay, n = param_layout("tent", len(factors))
    p = rng.standard_normal(n).astype(np.float32) * 0.3
    c = [0.5, 3.0, 0.25]
    q = scale_levels(p, "tent", factors, c)
    assert q[-1] == p[-1]  # log alpha kept
    for (off, co, ci), cl in zip(lay, c):
        assert np.allclose(q[off:off + co * ci * 9], p[off:off + co * ci * 9] * np.sqrt(cl), rtol=1e-6)

    def apply(params, r, s):
        ws = [np.asarray(params[off:off + co * ci * 9], dtype=np.float64).reshape(co, ci, 9) for off, co, ci in lay]
        return L.TentLevelsNet(ws, factors, log_alpha=float(params[-1])).apply(r, s)

    r = rng.standard_normal((1, 1, 12, 11))
    s = rng.uniform(0.3, 1.2, r.shape)
    base = apply(p, r, s)
    terms = [apply(scale_levels(p, "tent", factors, [2.0 if k == l else 1.0 for k in range(3)]), r, s) - base
             for l in range(3)]
    alpha_part = base - sum(terms)
    assert np.allclose(alpha_part, np.exp(p[-1]) * s * s * r, rtol=0, atol=1e-5 * np.abs(base).max())
    assert min(np.abs(t).max() for t in terms) > 1e-3 * np.abs(base).max()
    want = alpha_part + sum(cl * t for cl, t in zip(c, terms))
    assert np.allclose(apply(q, r, s), want, rtol=0, atol=1e-5 * np.abs(want).max())
    with pytest.raises(ValueError):
        scale_levels(p, "chain", factors, c)
