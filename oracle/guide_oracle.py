"""FP64 CPU oracle of the path-guide network (test infrastructure only).

The guide has no counterpart in the reference: it is the learned proposal for scattering
directions.  This module fixes its mathematics so that the HIP kernels (BF16 MFMA with FP32
accumulation) can be checked against an FP64 statement of the same function and of its
gradient:

  features (9)   x/h, y/h, z/h, kx, ky, kz, (ln lambda - c)/s, |r|/h, |z|/h
  network        9 -> 64 -> 64 -> 64 -> 40, ReLU on the hidden layers
  output (40)    8 lobes x (mu_x, mu_y, mu_z, kappa_raw, weight_logit)
  mixture        mu = normalize(mu_raw); kappa = KMIN + softplus(kappa_raw), clamped at KMAX;
                 w = softmax(logits); q(k) = sum_j w_j vMF(k; mu_j, kappa_j)
  vMF density    kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.k - 1))
  loss           weighted negative log-likelihood  -sum_i c_i ln q(k_i | x_i) / sum_i c_i

The gradient is written out by hand (no autograd).  tests/test_guide_oracle.py pins it by
central finite differences, the density by sphere quadrature and the sampler by
Kolmogorov-Smirnov tests and the importance identity E_q[1/q] = 4 pi.
"""

from __future__ import annotations

import numpy as np

NIN, NH, NOUT, NLOBE = 9, 64, 40, 8
KMIN, KMAX = 0.2, 400.0


def init_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    p = {}
    dims = [(NIN, NH), (NH, NH), (NH, NH), (NH, NOUT)]
    for i, (a, b) in enumerate(dims):
        p[f"W{i}"] = rng.normal(0, scale * np.sqrt(2.0 / a), (b, a))
        p[f"b{i}"] = np.zeros(b)
    p["W3"] *= 0.1
    return p


def features(r, k, lam, half, lam_c=np.log(0.775e-6), lam_s=1.0):
    r = np.atleast_2d(r)
    k = np.atleast_2d(k)
    lam = np.atleast_1d(lam)
    f = np.zeros((len(r), NIN))
    f[:, 0:3] = r / half
    f[:, 3:6] = k
    f[:, 6] = (np.log(lam) - lam_c) / lam_s
    f[:, 7] = np.linalg.norm(r, axis=1) / half
    f[:, 8] = np.abs(r[:, 2]) / half
    return f


def softplus(x):
    return np.where(x > 20, x, np.log1p(np.exp(np.minimum(x, 20))))


def sigmoid(x):
    return 1 / (1 + np.exp(-x))


def forward(p, X):
    """Returns the raw outputs (n, 40) and the cache for the backward pass."""
    a0 = X
    z1 = a0 @ p["W0"].T + p["b0"]
    a1 = np.maximum(z1, 0)
    z2 = a1 @ p["W1"].T + p["b1"]
    a2 = np.maximum(z2, 0)
    z3 = a2 @ p["W2"].T + p["b2"]
    a3 = np.maximum(z3, 0)
    out = a3 @ p["W3"].T + p["b3"]
    return out, (a0, z1, a1, z2, a2, z3, a3)


def mixture(out):
    o = out.reshape(-1, NLOBE, 5)
    mu_raw = o[:, :, 0:3]
    nrm = np.linalg.norm(mu_raw, axis=2, keepdims=True)
    mu = mu_raw / np.maximum(nrm, 1e-12)
    kap = np.minimum(KMIN + softplus(o[:, :, 3]), KMAX)
    lg = o[:, :, 4]
    lg = lg - lg.max(axis=1, keepdims=True)
    w = np.exp(lg)
    w /= w.sum(axis=1, keepdims=True)
    return mu, kap, w


def vmf_pdf(k, mu, kap):
    """vMF density of unit vectors k (n,3) for lobes mu (n,L,3), kap (n,L)."""
    c = np.einsum("nj,nlj->nl", k, mu)
    return kap / (2 * np.pi * (-np.expm1(-2 * kap))) * np.exp(kap * (c - 1))


def mixture_pdf(out, k):
    mu, kap, w = mixture(out)
    return (w * vmf_pdf(k, mu, kap)).sum(axis=1)


def loss(p, X, K, C):
    out, _ = forward(p, X)
    q = mixture_pdf(out, K)
    return -np.sum(C * np.log(q)) / np.sum(C)


def grad(p, X, K, C):
    """Analytic gradient of loss(p, X, K, C) with respect to every parameter."""
    out, (a0, z1, a1, z2, a2, z3, a3) = forward(p, X)
    n = len(X)
    o = out.reshape(n, NLOBE, 5)
    mu_raw = o[:, :, 0:3]
    nrm = np.maximum(np.linalg.norm(mu_raw, axis=2), 1e-12)
    mu = mu_raw / nrm[:, :, None]
    sp = KMIN + softplus(o[:, :, 3])
    kap = np.minimum(sp, KMAX)
    lg = o[:, :, 4] - o[:, :, 4].max(axis=1, keepdims=True)
    w = np.exp(lg)
    w /= w.sum(axis=1, keepdims=True)
    c = np.einsum("nj,nlj->nl", K, mu)
    e2 = -np.expm1(-2 * kap)
    f = kap / (2 * np.pi * e2) * np.exp(kap * (c - 1))
    q = (w * f).sum(axis=1)
    g = -(C / np.sum(C))[:, None]                  # dL/d ln q per sample
    r = w * f / q[:, None]                        # lobe responsibilities
    # d ln q / d logit_j = r_j - w_j
    d_lg = g * (r - w)
    # d ln f / d kappa = 1/kappa - 2 e^{-2k}/(1-e^{-2k}) + (c - 1)
    dlnf_dk = 1 / kap - 2 * np.exp(-2 * kap) / e2 + (c - 1)
    d_kap = g * r * dlnf_dk * (sp < KMAX)
    d_kraw = d_kap * sigmoid(o[:, :, 3])
    # d ln f / d mu = kappa * K ; through the normalisation: (I - mu mu^T)/|raw|
    d_mu = (g * r * kap)[:, :, None] * K[:, None, :]
    d_raw = (d_mu - np.sum(d_mu * mu, axis=2, keepdims=True) * mu) / nrm[:, :, None]
    d_out = np.concatenate([d_raw, d_kraw[:, :, None], d_lg[:, :, None]], axis=2).reshape(n, NOUT)
    gr = {}
    gr["W3"] = d_out.T @ a3
    gr["b3"] = d_out.sum(0)
    d3 = (d_out @ p["W3"]) * (z3 > 0)
    gr["W2"] = d3.T @ a2
    gr["b2"] = d3.sum(0)
    d2 = (d3 @ p["W2"]) * (z2 > 0)
    gr["W1"] = d2.T @ a1
    gr["b1"] = d2.sum(0)
    d1 = (d2 @ p["W1"]) * (z1 > 0)
    gr["W0"] = d1.T @ a0
    gr["b0"] = d1.sum(0)
    return gr, d_out


def d_out_wrt_loss(p, X, K, C):
    """dL/d(raw outputs), the quantity the device backward pass starts from."""
    return grad(p, X, K, C)[1]


def sample_vmf(rng, mu, kap):
    """Direct 3-D vMF sampling: w = 1 + ln(u + (1-u) e^{-2 kappa}) / kappa."""
    u = rng.random(len(mu))
    v = rng.random(len(mu))
    wv = 1 + np.log(u + (1 - u) * np.exp(-2 * kap)) / kap
    phi = 2 * np.pi * v
    s = np.sqrt(np.maximum(0, 1 - wv * wv))
    # orthonormal frame around mu
    a = np.where(np.abs(mu[:, 2:3]) < 0.9, np.array([[0, 0, 1.0]]), np.array([[1.0, 0, 0]]))
    e1 = np.cross(mu, a)
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = np.cross(mu, e1)
    return wv[:, None] * mu + s[:, None] * (np.cos(phi)[:, None] * e1 + np.sin(phi)[:, None] * e2)


def hg_pdf(g, c):
    """Henyey-Greenstein density per steradian: Phi(cos) / (4 pi)."""
    t = 1 + g * g - 2 * g * c
    return (1 - g) * (1 + g) / np.sqrt(t * t * t) / (4 * np.pi)


def flatten(p):
    """Parameters in the device order W0, b0, W1, b1, W2, b2, W3, b3 (row-major)."""
    return np.concatenate([np.asarray(p[k], dtype=np.float64).reshape(-1) for k in ("W0", "b0", "W1", "b1", "W2", "b2", "W3", "b3")])


def unflatten(theta):
    shapes = [("W0", (NH, NIN)), ("b0", (NH,)), ("W1", (NH, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
              ("W3", (NOUT, NH)), ("b3", (NOUT,))]
    p, o = {}, 0
    for k, sh in shapes:
        n = int(np.prod(sh))
        p[k] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(sh)
        o += n
    return p


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def forward_bf16(p, X):
    """The device's arithmetic: BF16 weights, inputs and hidden activations, wide accumulation."""
    a = bf16(X)
    for i in range(3):
        a = bf16(np.maximum(a @ bf16(p[f"W{i}"]).T + np.float32(p[f"b{i}"]), 0))
    return a @ bf16(p["W3"]).T + p["b3"]
