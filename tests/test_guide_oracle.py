"""Pins the FP64 statement of the path guide (oracle/guide_oracle.py) on the CPU, so that the
device kernels can be checked against it:

* the hand-written gradient against central finite differences of the loss, for entries of
  every parameter block, with a lobe whose kappa sits at the KMAX clamp (zero derivative) and
  a lobe in softplus's linear branch (raw value > 20);
* the mixture density integrates to 1 over the sphere (Fibonacci quadrature), for random
  network outputs spanning kappa from KMIN to KMAX;
* the vMF sampler: mu.k follows the marginal CDF (e^{kappa (w - 1)} - e^{-2 kappa}) /
  (1 - e^{-2 kappa}) (Kolmogorov-Smirnov), the azimuth about mu is uniform, and the
  importance identity E_q[1/q] = 4 pi holds for the whole mixture sampled lobe by weight;
* planted mistakes in the gradient (the responsibility sign, the normalisation derivative,
  the clamp mask, softplus's derivative) and in the density (a missing 2 pi) fail these
  checks, so a plausible slip in the statement cannot pass.
"""

import inspect
import textwrap

import numpy as np
import pytest
from scipy import stats

import guide_oracle as G


def _case(seed=3, n=48):
    """Parameters, features, directions and weights with lobe 0 clamped at KMAX and lobe 1 in
    softplus's linear branch for every sample (set through the output biases)."""
    rng = np.random.default_rng(seed)
    p = G.init_params(seed, scale=1.0)
    p["W3"] = rng.normal(0, 0.3, p["W3"].shape)
    for k in ("b0", "b1", "b2"):
        p[k] = rng.normal(0, 0.1, p[k].shape)
    b3 = rng.normal(0, 0.5, G.NOUT)
    b3[0 * 5 + 3] = 600.0   # lobe 0: kappa = min(KMIN + 600, KMAX) = KMAX
    b3[1 * 5 + 3] = 30.0    # lobe 1: softplus(x) = x branch, kappa ~ 30
    b3[1 * 5 + 4] = 1.5     # and a large weight, so its gradient matters
    p["b3"] = b3
    r = rng.uniform(-1, 1, (n, 3))
    k = rng.normal(size=(n, 3))
    k /= np.linalg.norm(k, axis=1, keepdims=True)
    X = G.features(r, k, rng.uniform(0.3e-6, 2e-6, n), 1.0)
    K = rng.normal(size=(n, 3))
    K /= np.linalg.norm(K, axis=1, keepdims=True)
    # directions close to lobe 1's axis (a third) and to the narrow lobe 0's axis (a third),
    # so that their responsibilities, and the kappa and mu derivatives, are large
    out, _ = G.forward(p, X)
    for lobe, spread, part in ((1, 0.15, slice(0, n // 3)), (0, 0.03, slice(n // 3, 2 * n // 3))):
        mu = out.reshape(n, G.NLOBE, 5)[:, lobe, :3]
        mu = mu / np.linalg.norm(mu, axis=1, keepdims=True)
        near = mu + spread * rng.normal(size=(n, 3))
        near /= np.linalg.norm(near, axis=1, keepdims=True)
        K[part] = near[part]
    C = rng.uniform(0.1, 2.0, n)
    return p, X, K, C


def _fd_errors(grad_fn, loss_fn=G.loss, seed=3, per_block=10, h=1e-6):
    """Largest |analytic - central difference| / (|central difference| + 1e-3 * block scale)
    over sampled entries of every block, and the entries used."""
    p, X, K, C = _case(seed)
    gr, _ = grad_fn(p, X, K, C)
    rng = np.random.default_rng(seed + 100)
    worst = {}
    for name in ("W0", "b0", "W1", "b1", "W2", "b2", "W3", "b3"):
        a = p[name]
        flat = a.reshape(-1)
        idx = list(rng.choice(flat.size, size=min(per_block, flat.size), replace=False))
        if name == "b3":
            idx = list(range(flat.size))  # every output: the clamp, the linear branch, the logits
        if name == "W3":
            idx += [0 * 5 * 64 + 3 * 64 + j for j in range(0, 64, 8)]  # lobe 0's kappa row
            idx += [1 * 5 * 64 + 3 * 64 + j for j in range(0, 64, 8)]  # lobe 1's kappa row
        scale = np.abs(gr[name]).max()
        err = 0.0
        for i in idx:
            old = flat[i]
            flat[i] = old + h
            lp = loss_fn(p, X, K, C)
            flat[i] = old - h
            lm = loss_fn(p, X, K, C)
            flat[i] = old
            fd = (lp - lm) / (2 * h)
            an = gr[name].reshape(-1)[i]
            err = max(err, abs(an - fd) / (abs(fd) + 1e-3 * scale))
        worst[name] = err
    return worst


def test_case_reaches_clamp_and_linear_branch():
    p, X, K, C = _case()
    out, _ = G.forward(p, X)
    o = out.reshape(len(X), G.NLOBE, 5)
    assert np.all(G.KMIN + G.softplus(o[:, 0, 3]) > G.KMAX)   # lobe 0 clamped for every sample
    assert np.all(o[:, 1, 3] > 20)                             # lobe 1 in the linear branch
    assert np.all(G.KMIN + o[:, 1, 3] < G.KMAX)


def test_gradient_matches_central_differences():
    worst = _fd_errors(G.grad)
    print({k: f"{v:.2e}" for k, v in worst.items()})
    for name, e in worst.items():
        assert e < 1e-5, (name, e)


def test_clamped_lobe_has_zero_kappa_gradient():
    p, X, K, C = _case()
    _, d_out = G.grad(p, X, K, C)
    assert np.all(d_out[:, 0 * 5 + 3] == 0.0)
    assert np.any(np.abs(d_out[:, 1 * 5 + 3]) > 1e-6)


def _planted(old, new):
    """G.grad with one expression replaced: a mistake planted in the oracle, on the CPU."""
    src = textwrap.dedent(inspect.getsource(G.grad))
    assert old in src, old
    ns = dict(vars(G))
    exec(compile(src.replace(old, new), "<planted>", "exec"), ns)
    return ns["grad"]


@pytest.mark.parametrize("old,new", [
    ("(r - w)", "(w - r)"),                                    # responsibility sign
    ("1 / kap - 2 * np.exp(-2 * kap) / e2", "1 / kap + 2 * np.exp(-2 * kap) / e2"),  # normalisation derivative
    ("* (sp < KMAX)", ""),                                     # clamp mask dropped
    ("sigmoid(o[:, :, 3])", "(1 - sigmoid(o[:, :, 3]))"),      # softplus derivative
    ("- np.sum(d_mu * mu, axis=2, keepdims=True) * mu", ""),   # normalisation projection dropped
])
def test_planted_gradient_mistakes_fail(old, new):
    worst = _fd_errors(_planted(old, new))
    assert max(worst.values()) > 1e-3, (old, worst)


def _fibonacci(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    phi = np.pi * (1 + 5 ** 0.5) * i
    s = np.sqrt(1 - z * z)
    return np.column_stack([s * np.cos(phi), s * np.sin(phi), z])


def _random_outputs(rng, n):
    out = rng.normal(0, 1.0, (n, G.NOUT))
    o = out.reshape(n, G.NLOBE, 5)
    o[:, :, 3] = rng.uniform(-4, 8, (n, G.NLOBE))      # kappa from ~0.22 to ~8
    o[:, 0, 3] = 500.0                                   # a lobe at KMAX
    o[:, 1, 3] = rng.uniform(25, 120, n)                 # linear branch, kappa 25-120
    return out


def test_mixture_density_integrates_to_one():
    rng = np.random.default_rng(11)
    out = _random_outputs(rng, 6)
    pts = _fibonacci(1_000_000)
    for j in range(len(out)):
        q = G.mixture_pdf(np.repeat(out[j:j + 1], len(pts), axis=0), pts)
        integral = q.sum() * 4 * np.pi / len(pts)
        assert abs(integral - 1) < 2e-4, (j, integral)


def test_planted_density_mistake_fails():
    rng = np.random.default_rng(11)
    out = _random_outputs(rng, 1)
    pts = _fibonacci(200_000)
    mu, kap, w = G.mixture(np.repeat(out, len(pts), axis=0))
    c = np.einsum("nj,nlj->nl", pts, mu)
    wrong = (w * kap / (-np.expm1(-2 * kap)) * np.exp(kap * (c - 1))).sum(axis=1)  # 2 pi dropped
    assert abs(wrong.sum() * 4 * np.pi / len(pts) - 1) > 1


def vmf_marginal_cdf(x, kap):
    """CDF of mu.k for the 3-D vMF: (e^{kap (x - 1)} - e^{-2 kap}) / (1 - e^{-2 kap})."""
    return (np.exp(kap * (x - 1)) - np.exp(-2 * kap)) / (-np.expm1(-2 * kap))


@pytest.mark.parametrize("kap", [0.2, 1.0, 5.0, 50.0, 400.0])
def test_vmf_sampler_marginal_and_azimuth(kap):
    rng = np.random.default_rng(int(kap * 10) + 1)
    n = 40000
    mu = rng.normal(size=(n, 3))
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    mu[: n // 4] = [0.0, 0.0, 1.0]   # the frame's other branch (|mu_z| >= 0.9)
    k = G.sample_vmf(rng, mu, np.full(n, kap))
    np.testing.assert_allclose(np.linalg.norm(k, axis=1), 1.0, atol=1e-12)
    w = np.sum(k * mu, axis=1)
    ks = stats.kstest(w, lambda x: vmf_marginal_cdf(x, kap))
    assert ks.pvalue > 1e-3, ks
    # azimuth about mu: uniform on [0, 2 pi) in a fixed frame around the z-axis lobes
    kz = k[: n // 4]
    sel = kz[:, 2] < 1 - 1e-9
    phi = np.arctan2(kz[sel, 1], kz[sel, 0]) % (2 * np.pi)
    assert stats.kstest(phi / (2 * np.pi), "uniform").pvalue > 1e-3


def test_planted_sampler_mistake_fails():
    rng = np.random.default_rng(5)
    n, kap = 40000, 5.0
    mu = np.tile([0.0, 0.0, 1.0], (n, 1))
    k = G.sample_vmf(rng, mu, np.full(n, kap))
    # a sampler whose cosine is the wrong one (here: kappa doubled) fails the same test
    assert stats.kstest(k[:, 2], lambda x: vmf_marginal_cdf(x, 2 * kap)).pvalue < 1e-6


def sample_mixture(rng, out):
    """Lobe by weight, then the vMF of that lobe (the life cycle's guide sampler)."""
    mu, kap, w = G.mixture(out)
    n = len(out)
    cum = np.cumsum(w, axis=1)
    j = np.minimum((rng.random(n)[:, None] > cum).sum(axis=1), G.NLOBE - 1)
    return G.sample_vmf(rng, mu[np.arange(n), j], kap[np.arange(n), j])


def test_mixture_importance_identity():
    """E_q[1/q(k)] = 4 pi for directions sampled from the mixture itself: ties the sampler to
    the density (a sampler of another distribution, or a density off by a factor, fails)."""
    rng = np.random.default_rng(21)
    base = rng.normal(0, 1.0, (4, G.NOUT))
    b = base.reshape(4, G.NLOBE, 5)
    # a broad, heavy lobe 0 bounds q from below (1/q then has a finite variance); the other
    # lobes are narrower and lighter, so a sampler that ignores the weights is far off
    b[:, 0, 3] = -0.5
    b[:, 0, 4] = 2.5
    b[:, 1:, 3] = rng.uniform(2, 4, (4, G.NLOBE - 1))
    b[:, 1:, 4] = rng.uniform(-1, 1, (4, G.NLOBE - 1))
    n = 200_000
    for j in range(4):
        out = np.repeat(base[j:j + 1], n, axis=0)
        k = sample_mixture(rng, out)
        inv = 1.0 / G.mixture_pdf(out, k)
        m, se = inv.mean(), inv.std() / np.sqrt(n)
        assert abs(m - 4 * np.pi) < 5 * se + 1e-9, (j, m, se)
        # a sampler that ignores the weights (uniform lobe choice) fails
        mu, kap, w = G.mixture(out)
        jj = rng.integers(0, G.NLOBE, n)
        kb = G.sample_vmf(rng, mu[np.arange(n), jj], kap[np.arange(n), jj])
        inv_b = 1.0 / G.mixture_pdf(out, kb)
        assert abs(inv_b.mean() - 4 * np.pi) > 5 * inv_b.std() / np.sqrt(n), j
