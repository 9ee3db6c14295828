"""THE GUIDE ORACLE -- test infrastructure only: an FP64 numpy restatement of the path-guiding network the engine
trains and samples on the GPU (BF16 weights and activations, FP32 accumulation there). tests/ use it as the checker;
the product never imports it.

The reference dropped its guiding (lib/pathtracer/README.md:853-861), so nothing here restates reference code; it is
the method the engine implements, written from the mathematics:

  * input: a vertex's position in the scene's box (x in [0, 1]^3) encoded at 12 octaves (sin, cos of 2^k pi x_i),
    its view direction, its normal, its roughness and a constant 1 -- 80 features;
  * network: 80 -> 64 -> 64 -> 64 -> 41, ReLU between (weights W[in, out], biases b[out]);
  * head: an 8-lobe von Mises-Fisher mixture on the sphere -- weights softmax(o[0:8]), mean directions
    normalize(o[8 + 3j : 11 + 3j]), concentrations exp(clamp(o[32 + j], KAPPA_LO, KAPPA_HI));
    vMF(w; mu, k) = k / (2 pi (1 - e^{-2k})) exp(k (mu . w - 1)) -- and the guided fraction c = sigmoid(o[40]);
  * the sampling density is the two strategies' mixture p(w) = c q(w) + (1 - c) b(w), b the BSDF's density;
  * losses over a batch of sampled directions w_i with weights F_i = luminance(f cos L_i) / p_used(w_i): the mixture
    learns the normalised product by its own KL divergence, L_q = -sum_i F_i log q(w_i) / sum_i F_i, and the fraction
    learns how far to trust it by the combined density's, L_c = -sum_i F_i log(c q(w_i) + (1 - c) b(w_i)) / sum_i F_i
    with q held fixed there (trained through L_c alone, a mixture the fraction has turned off would never learn
    again); the loss reported is L_c;
  * Adam (beta1 0.9, beta2 0.99, eps 1e-8) on the FP32 master weights.

Pinned by (tests/test_guide_oracle.py): central finite differences of the loss against `backward`; the mixture's
density integrates to 1 over the sphere by quadrature and its sampler's histogram follows it; one Adam step against
its closed form on the first step (each weight moves by lr * sign(g)); the encoding against its definition at exact
points (octave k of x = 1/2^(k+1) is sin = 1, cos = 0)."""
import math

import numpy as np

SIZES = [80, 64, 64, 64, 41]
LOBES = 8
OCTAVES = 12
KAPPA_LO, KAPPA_HI = -3.0, 7.0
BETA1, BETA2, EPS = 0.9, 0.99, 1e-8


def nparams():
    return sum(a * b + b for a, b in zip(SIZES[:-1], SIZES[1:]))


def unpack(theta):
    """Flat parameters -> [(W[in, out], b[out])] in the engine's order: W1, b1, W2, b2, W3, b3, W4, b4."""
    out, o = [], 0
    for a, b in zip(SIZES[:-1], SIZES[1:]):
        W = theta[o:o + a * b].reshape(a, b); o += a * b
        bb = theta[o:o + b]; o += b
        out.append((W, bb))
    return out


def encode(pos01, wo, n, rough):
    """[N, 80] features from positions already mapped into [0, 1]^3, view directions, normals and roughness."""
    pos01 = np.asarray(pos01, np.float64).reshape(-1, 3)
    N = len(pos01)
    X = np.zeros((N, 80))
    for k in range(OCTAVES):
        a = (2.0 ** k) * math.pi * pos01
        X[:, 6 * k + 0:6 * k + 6:2] = np.sin(a)
        X[:, 6 * k + 1:6 * k + 6:2] = np.cos(a)
    X[:, 72:75] = np.asarray(wo, np.float64).reshape(-1, 3)
    X[:, 75:78] = np.asarray(n, np.float64).reshape(-1, 3)
    X[:, 78] = np.asarray(rough, np.float64).reshape(-1)
    X[:, 79] = 1.0
    return X


def forward(theta, X):
    """Head outputs [N, 40] and the activations backward needs."""
    acts = [X]
    h = X
    layers = unpack(theta)
    for i, (W, b) in enumerate(layers):
        z = h @ W + b
        h = np.maximum(z, 0.0) if i < len(layers) - 1 else z
        acts.append(h)
    return h, acts


def frac(o):
    """The guided fraction c = sigmoid(o[40])."""
    return 1.0 / (1.0 + np.exp(-np.asarray(o, np.float64)[:, 40]))


def mixture(o):
    """(pi [N, 8], mu [N, 8, 3], kappa [N, 8], |m| [N, 8], clamp mask [N, 8]) from head outputs."""
    o = np.asarray(o, np.float64)
    a = o[:, :LOBES]
    a = a - a.max(axis=1, keepdims=True)
    e = np.exp(a)
    pi = e / e.sum(axis=1, keepdims=True)
    m = o[:, LOBES:4 * LOBES].reshape(-1, LOBES, 3)
    ml = np.maximum(np.linalg.norm(m, axis=2), 1e-6)
    mu = m / ml[:, :, None]
    s = o[:, 4 * LOBES:5 * LOBES]
    inside = (s > KAPPA_LO) & (s < KAPPA_HI)
    kappa = np.exp(np.clip(s, KAPPA_LO, KAPPA_HI))
    return pi, mu, kappa, ml, inside


def log_c(kappa):
    """log of the vMF normaliser k / (2 pi (1 - e^{-2k}))."""
    return np.log(kappa) - math.log(2 * math.pi) - np.log(-np.expm1(-2 * kappa))


def vmf_pdf(mu, kappa, w):
    return np.exp(log_c(kappa) + kappa * (np.einsum('...i,...i->...', mu, w) - 1.0))


def mixture_pdf(o, w):
    """q(w) for each row's mixture at its direction w [N, 3]."""
    pi, mu, kappa, _, _ = mixture(o)
    return (pi * vmf_pdf(mu, kappa, np.asarray(w, np.float64)[:, None, :])).sum(axis=1)


def combined_pdf(o, w, b):
    c = frac(o)
    return c * mixture_pdf(o, w) + (1 - c) * np.asarray(b, np.float64)


def loss(o, w, F, b):
    """L_c: the combined density's loss (the one reported)."""
    p = combined_pdf(o, w, b)
    F = np.asarray(F, np.float64)
    return -(F * np.log(p)).sum() / F.sum()


def loss_q(o, w, F):
    """L_q: the mixture's own loss."""
    F = np.asarray(F, np.float64)
    return -(F * np.log(mixture_pdf(o, w))).sum() / F.sum()


def loss_surrogate(o, w, F, b, q0):
    """L_q + L_c with q fixed at q0 inside L_c: what dloss_do differentiates (at q = q0)."""
    c = frac(o)
    F = np.asarray(F, np.float64)
    return loss_q(o, w, F) - (F * np.log(c * q0 + (1 - c) * np.asarray(b, np.float64))).sum() / F.sum()


def dloss_do(o, w, F, b):
    """d/do [N, 41] of L_q + L_c, q held fixed in L_c."""
    pi, mu, kappa, ml, inside = mixture(o)
    w = np.asarray(w, np.float64)
    F = np.asarray(F, np.float64)
    b = np.asarray(b, np.float64)
    comp = pi * vmf_pdf(mu, kappa, w[:, None, :])
    q = comp.sum(axis=1, keepdims=True)
    r = comp / q                                       # responsibilities
    c = frac(o)[:, None]
    p = c * q + (1 - c) * b[:, None]
    scale = (F / F.sum())[:, None]
    g = np.zeros_like(np.asarray(o, np.float64))
    g[:, :LOBES] = -(r - pi) * scale
    dot = np.einsum('nji,ni->nj', mu, w)
    tang = w[:, None, :] - dot[:, :, None] * mu        # (I - mu mu^T) w
    gm = -(r * kappa / ml)[:, :, None] * tang
    g[:, LOBES:4 * LOBES] = (gm * scale[:, :, None]).reshape(-1, 3 * LOBES)
    # d log C / dk = 1/k - 2 / (e^{2k} - 1) = 1/k + 2 e^{-2k} / expm1(-2k); dk/ds = k inside the clamp
    dlogc = 1.0 / kappa + 2.0 * np.exp(-2 * kappa) / np.expm1(-2 * kappa)
    g[:, 4 * LOBES:5 * LOBES] = -(r * (dlogc + (dot - 1.0)) * kappa * inside) * scale
    # dc/ds = c (1 - c): -F (q - b) c (1 - c) / p
    g[:, 40] = (-(F / F.sum())[:, None] * (q - b[:, None]) * c * (1 - c) / p)[:, 0]
    return g


def backward(theta, X, w, F, b):
    """dL/dtheta (flat, the engine's order)."""
    o, acts = forward(theta, X)
    d = dloss_do(o, w, F, b)
    layers = unpack(theta)
    grads = []
    for i in range(len(layers) - 1, -1, -1):
        W, b = layers[i]
        h_in = acts[i]
        gW = h_in.T @ d
        gb = d.sum(axis=0)
        grads.append((gW, gb))
        if i > 0:
            d = (d @ W.T) * (acts[i] > 0)
    grads.reverse()
    return np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])


def adam(theta, g, m, v, t, lr):
    """One Adam step (t from 1): new theta, m, v."""
    m = BETA1 * m + (1 - BETA1) * g
    v = BETA2 * v + (1 - BETA2) * g * g
    mh = m / (1 - BETA1 ** t)
    vh = v / (1 - BETA2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + EPS), m, v


def sample_vmf(mu, kappa, u1, u2):
    """Directions from vMF lobes (Jakob 2012's stable inversion): w = 1 + log(u1 + (1 - u1) e^{-2k}) / k."""
    mu = np.asarray(mu, np.float64)
    kappa = np.asarray(kappa, np.float64)
    wz = 1.0 + np.log(u1 + (1 - u1) * np.exp(-2 * kappa)) / kappa
    sr = np.sqrt(np.maximum(0.0, 1 - wz * wz))
    phi = 2 * math.pi * u2
    # a frame about mu
    t = np.where(np.abs(mu[..., 2:3]) < 0.999, np.cross(mu, [0.0, 0.0, 1.0]), np.cross(mu, [1.0, 0.0, 0.0]))
    t /= np.linalg.norm(t, axis=-1, keepdims=True)
    b = np.cross(mu, t)
    return t * (sr * np.cos(phi))[..., None] + b * (sr * np.sin(phi))[..., None] + mu * wz[..., None]
