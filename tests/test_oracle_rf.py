"""The oracle's radiation-field storage (MonteCarloSimulation::storeRadiationField) against
what the geometry fixes: for a point source in a homogeneous cube without scattering, the
stored field summed over cells is N L E_dir[(1 - exp(-kappa s_exit)) / kappa] -- the
per-segment logarithmic mean of the extinction factor integrates exp(-kappa s) exactly --
with the direction average taken by quadrature on a Fibonacci sphere.  Both branches of
lnmean are exercised: kappa = 1 per half-size (series branch unused) and kappa = 1e-4
(every segment on the series branch)."""

import numpy as np
import pytest

import oracle as O


def _fib_dirs(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    phi = np.pi * (1 + 5 ** 0.5) * i
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


@pytest.mark.parametrize("kappa", [1.0, 1e-4])
def test_rf_total_matches_path_integral(kappa):
    n = 8
    b = np.linspace(-1.0, 1.0, n + 1)
    g = O.Grid("cartesian", xb=b, yb=b, zb=b)
    sigma = 1e-3
    dens = np.full(n ** 3, kappa / sigma)
    npk = 200000
    frames = [dict(inclination=0.5, azimuth=0.0, roll=0.0, nx=4, ny=4, fovx=2.0, fovy=2.0)]
    _, cnt, rf = O.simulate(g, dens, [1.0], [sigma], [0.0], [0.5], [1.0 / npk], frames, 0, npk, threads=8,
                            seed=4, store_rf=True)
    assert cnt[0] == npk and cnt[3] == 0  # albedo 0: one forced path per packet
    k = _fib_dirs(2_000_000)
    s_exit = 1.0 / np.abs(k).max(axis=1)
    expect = np.mean(-np.expm1(-kappa * s_exit) / kappa)
    # direction-sampling error of the packets: std of the per-direction integral / sqrt(npk)
    sd = np.std(-np.expm1(-kappa * s_exit) / kappa) / np.sqrt(npk)
    assert abs(rf.sum() - expect) < 5 * sd + 1e-9 * expect, (rf.sum(), expect, sd)
    # the field is symmetric under the cube's reflections: octants hold the same total
    R = rf[:, 0].reshape(n, n, n)
    h = n // 2
    octs = [R[i * h:(i + 1) * h, j * h:(j + 1) * h, q * h:(q + 1) * h].sum() for i in (0, 1) for j in (0, 1) for q in (0, 1)]
    assert np.std(octs) / np.mean(octs) < 0.02


def test_lnmean_known_answers():
    """lnmean(x1, x1 (1 + x)) = x1 x / log1p(x): series branch (x < 1e-3, accurate to ~1e-15)
    and log branch (the difference of two logarithms loses ~|ln x1| / x ulps), both argument
    orders, and zero for a non-positive argument."""
    rng = np.random.default_rng(2)
    for x in np.concatenate([10 ** rng.uniform(-12, -3, 200), 10 ** rng.uniform(-3, 1, 200)]):
        for x1 in (1e-30, 0.37, 5.0):
            want = x1 * x / np.log1p(x)
            tol = 1e-13 if x < 1e-3 else 1e-13 + 4e-16 * (2 + abs(np.log(x1))) / x
            assert abs(O.lnmean(x1, x1 * (1 + x)) / want - 1) < tol
            assert abs(O.lnmean(x1 * (1 + x), x1) / want - 1) < tol
    assert O.lnmean(0.0, 1.0) == 0.0 and O.lnmean(2.0, 2.0) == 2.0
