"""The guide oracle pinned before it is trusted: finite differences, quadrature of the mixture, Adam's first step,
the encoding at exact points. CPU only."""
import math

import numpy as np

import guide

RNG = np.random.default_rng(7)


def random_theta(scale=0.3):
    return RNG.normal(0, scale, guide.nparams())


def random_batch(n):
    pos = RNG.random((n, 3))
    wo = RNG.normal(size=(n, 3)); wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    nn = RNG.normal(size=(n, 3)); nn /= np.linalg.norm(nn, axis=1, keepdims=True)
    rough = RNG.random(n)
    w = RNG.normal(size=(n, 3)); w /= np.linalg.norm(w, axis=1, keepdims=True)
    F = RNG.random(n) + 0.1
    b = np.maximum(np.einsum('ij,ij->i', w, nn), 0) / np.pi     # a Lambertian BSDF's density (0 below)
    return guide.encode(pos, wo, nn, rough), w, F, b


def test_parameter_count():
    # 80*64+64 + 2*(64*64+64) + 64*41+41
    assert guide.nparams() == 16169


def test_encoding_at_exact_points():
    # octave k of x = 2^-(k+1) is the angle pi/2: sin 1, cos 0 on that axis
    for k in range(12):
        x = 2.0 ** -(k + 1)
        X = guide.encode([[x, 0, 0]], [[0, 0, 1]], [[0, 1, 0]], [0.25])
        assert abs(X[0, 6 * k + 0] - 1) < 1e-12 and abs(X[0, 6 * k + 1]) < 1e-12
        assert abs(X[0, 6 * k + 2]) < 1e-12 and abs(X[0, 6 * k + 3] - 1) < 1e-12     # y = 0: sin 0, cos 1
    assert X[0, 72:75].tolist() == [0, 0, 1] and X[0, 75:78].tolist() == [0, 1, 0]
    assert X[0, 78] == 0.25 and X[0, 79] == 1


def test_backward_matches_finite_differences():
    # the surrogate L_q + L_c with q frozen at the point of differentiation
    theta = random_theta()
    X, w, F, b = random_batch(24)
    g = guide.backward(theta, X, w, F, b)
    q0 = guide.mixture_pdf(guide.forward(theta, X)[0], w)
    idx = np.concatenate([RNG.choice(len(theta), 300, replace=False), np.arange(len(theta) - 41, len(theta))])
    h = 1e-6
    L = lambda th: guide.loss_surrogate(guide.forward(th, X)[0], w, F, b, q0)
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h; tm[i] -= h
        fd = (L(tp) - L(tm)) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_head_gradient_matches_finite_differences_at_high_and_low_kappa():
    # the head alone, with concentrations near both ends of the clamp and lobes near and far from w
    o = RNG.normal(0, 1, (6, 41))
    o[:, 32:40] = np.linspace(-2.9, 6.9, 8)
    w = RNG.normal(size=(6, 3)); w /= np.linalg.norm(w, axis=1, keepdims=True)
    F = RNG.random(6) + 0.5
    b = np.array([0.0, 0.05, 0.2, 0.0, 1.0, 0.3])        # including directions the BSDF cannot sample
    g = guide.dloss_do(o, w, F, b)
    q0 = guide.mixture_pdf(o, w)
    h = 1e-6
    for n in range(6):
        for j in range(41):
            op, om = o.copy(), o.copy()
            op[n, j] += h; om[n, j] -= h
            fd = (guide.loss_surrogate(op, w, F, b, q0) - guide.loss_surrogate(om, w, F, b, q0)) / (2 * h)
            assert abs(fd - g[n, j]) < 1e-6 + 1e-5 * abs(fd), (n, j, fd, g[n, j])


def _sphere_quadrature(n):
    # midpoint rule in (cos theta, phi): equal-area cells
    ct = (np.arange(n) + 0.5) / n * 2 - 1
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * math.pi
    C, P = np.meshgrid(ct, ph, indexing='ij')
    s = np.sqrt(1 - C * C)
    w = np.stack([s * np.cos(P), s * np.sin(P), C], axis=-1).reshape(-1, 3)
    return w, 4 * math.pi / len(w)


def test_mixture_integrates_to_one():
    w, dA = _sphere_quadrature(600)
    for kap in (-2.5, 0.0, 3.0, 5.5):
        o = RNG.normal(0, 1, (1, 41))
        o[0, 32:40] = kap
        q = guide.mixture_pdf(np.repeat(o, len(w), axis=0), w)
        assert abs(q.sum() * dA - 1) < 2e-3, (kap, q.sum() * dA)


def test_vmf_sampler_follows_its_density():
    mu = np.array([0.3, -0.5, 0.8]); mu /= np.linalg.norm(mu)
    for kappa in (0.2, 4.0, 60.0):
        n = 200000
        d = guide.sample_vmf(np.tile(mu, (n, 1)), np.full(n, kappa), RNG.random(n), RNG.random(n))
        assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-9)
        c = d @ mu
        # the density of c = mu . w is k e^{k (c - 1)} / (1 - e^{-2k}) on [-1, 1]
        lo = max(-1.0, 1 - 8 / kappa)
        edges = np.linspace(lo, 1, 21)
        hist, _ = np.histogram(c, bins=edges)
        cdf = lambda x: (np.exp(kappa * (x - 1)) - math.exp(-2 * kappa)) / (1 - math.exp(-2 * kappa))
        expected = n * np.diff(cdf(edges))
        ok = expected > 50
        chi = ((hist[ok] - expected[ok]) ** 2 / expected[ok]).sum()
        assert chi < 60, (kappa, chi)


def test_adam_first_step_moves_each_weight_by_lr():
    # t = 1: m_hat = g and v_hat = g^2, so the step is lr * g / (|g| + eps), lr * sign(g) but for tiny g
    theta = random_theta()
    g = RNG.normal(0, 1e-2, theta.shape)
    step = 1e-2 * g / (np.abs(g) + guide.EPS)
    t2, m, v = guide.adam(theta, g, np.zeros_like(theta), np.zeros_like(theta), 1, 1e-2)
    assert np.allclose(theta - t2, step, rtol=1e-9, atol=1e-15)
    assert np.mean(np.abs(np.abs(theta - t2) - 1e-2) < 1e-4) > 0.99
    # a second step with the same gradient moves by the same amount (m_hat = g, v_hat = g^2 still)
    t3, _, _ = guide.adam(t2, g, m, v, 2, 1e-2)
    assert np.allclose(t2 - t3, step, rtol=1e-9, atol=1e-15)


def test_training_lowers_the_loss_on_a_fixed_target():
    # the loss of a lobe-shaped target falls under plain Adam on the oracle's own gradients, and the network learns
    # to trust itself (c -> 1) where the BSDF (here uniform over the sphere) is the poorer density
    theta = random_theta(0.1)
    n = 256
    pos = RNG.random((n, 3)); X = guide.encode(pos, np.tile([0, 0, 1.0], (n, 1)), np.tile([0, 0, 1.0], (n, 1)), np.full(n, 0.5))
    target = np.array([0.6, 0.0, 0.8])
    w = guide.sample_vmf(np.tile(target, (n, 1)), np.full(n, 20.0), RNG.random(n), RNG.random(n))
    F = np.ones(n)
    b = np.full(n, 1 / (4 * np.pi))
    l0 = guide.loss(guide.forward(theta, X)[0], w, F, b)
    m = np.zeros_like(theta); v = np.zeros_like(theta)
    for t in range(1, 201):
        theta, m, v = guide.adam(theta, guide.backward(theta, X, w, F, b), m, v, t, 1e-2)
    o1 = guide.forward(theta, X)[0]
    l1 = guide.loss(o1, w, F, b)
    # the target's own entropy bound: -E log vMF(20) = -(log C(20) + E[20 (c - 1)]) = -(log C(20) - 1)
    floor = -(guide.log_c(np.array(20.0)) - 1.0)
    assert l1 < l0 - 1.0 and l1 < floor + 0.5, (l0, l1, floor)
    assert guide.frac(o1).mean() > 0.9
