"""The oracle's imported particle source (CPU): the launch preparation of
ImportedSource::prepareForLaunch against a hand-computed case, the luminosity bookkeeping in a
dust-free grid, and the positions drawn from the smoothing kernels against the kernels' column
densities (the projected surface brightness of one particle seen face-on), a closed form
independent of the radius-sampling table (CubicSplineSmoothingKernel::columnDensity,
CubicSplineSmoothingKernel.cpp:56-77; the uniform kernel's chord length)."""

import numpy as np

import oracle as O
from helpers import PC, cart_borders


def test_launch_ranges_known_answer():
    """Three particles with luminosities 1 : 2 : 1, source bias 0.5, 1000 packets: weights
    0.5 * (0.25, 0.5, 0.25) + 0.5 / 3, ranges [0, 292), [292, 708), [708, 1000), and per-packet
    luminosity L_m,l * nlambda / (N W_m)."""
    Iv, lp, W = O.particle_launch([1.0, 2.0, 1.0], [[1.0], [2.0], [1.0]], 1000, 0.5)
    np.testing.assert_allclose(W, [0.5 * 0.25 + 0.5 / 3, 0.5 * 0.5 + 0.5 / 3, 0.5 * 0.25 + 0.5 / 3], rtol=1e-15)
    np.testing.assert_array_equal(Iv, [0, 292, 708, 1000])
    np.testing.assert_allclose(lp[:, 0], np.array([1.0, 2.0, 1.0]) / (1000 * W), rtol=1e-15)
    # bias 0: ranges in proportion to luminosity; bias 1: equal ranges
    np.testing.assert_array_equal(O.particle_launch([1.0, 2.0, 1.0], [[1], [2], [1]], 1000, 0.0)[0], [0, 250, 750, 1000])
    np.testing.assert_array_equal(O.particle_launch([1.0, 2.0, 1.0], [[1], [2], [1]], 999, 1.0)[0], [0, 333, 666, 999])


def _dust_free(n=8):
    b = cart_borders(n, -10 * PC, 10 * PC)
    return O.Grid("cartesian", xb=b, yb=b, zb=b), np.zeros(n ** 3)


def test_luminosity_bookkeeping_dust_free():
    """One wavelength, no dust: every packet deposits its launch luminosity once (emission
    peel-off into an SED without aperture), so the SED equals sum_m count_m L_m / (N W_m)
    exactly, and sum_m L_m up to the rounding of the ranges."""
    grid, dens = _dust_free()
    rng = np.random.default_rng(4)
    n = 40
    xyzh = np.column_stack([rng.uniform(-6, 6, (n, 3)) * PC, rng.uniform(0.2, 2.0, n) * PC])
    L = rng.uniform(0.1, 3.0, n)
    N = 20000
    instr = [dict(sed=True, inclination=0.7, azimuth=0.2, radius=0.0)]
    fr, cnt = O.simulate(grid, dens, [0.55e-6], [1e-29], [0.0], [0.5], [1.0], instr, 0, N, threads=4, seed=5,
                         particles=dict(xyzh=xyzh, lbin=L[:, None], ltot=L, num_packets=N, source_bias=0.5))
    Iv, lp, W = O.particle_launch(L, L[:, None], N, 0.5)
    want = np.sum(np.diff(Iv) * lp[:, 0])
    assert abs(fr[0][0, 0, 0] / want - 1) < 1e-12
    assert abs(want / L.sum() - 1) < n / N


def _annulus_profile(kernel, h, nbins=8, N=300000):
    grid, dens = _dust_free()
    npix = 200
    fr, _ = O.simulate(grid, dens, [0.55e-6], [1e-29], [0.0], [0.5], [1.0],
                       [dict(inclination=0.0, azimuth=0.0, roll=0.0, nx=npix, ny=npix, fovx=20 * PC, fovy=20 * PC)],
                       0, N, threads=8, seed=11,
                       particles=dict(xyzh=[[0.0, 0.0, 0.0, h]], lbin=[[1.0]], ltot=[1.0], num_packets=N, kernel=kernel))
    img = fr[0][0]
    c = (np.arange(npix) + 0.5) / npix * 20 * PC - 10 * PC
    R = np.hypot(c[None, :], c[:, None]) / h
    edges = np.linspace(0, 1, nbins + 1)
    got = np.array([img[(R >= a) & (R < b)].sum() for a, b in zip(edges[:-1], edges[1:])]) / img.sum()
    return edges, got


def _column_density_spline(q):
    """CubicSplineSmoothingKernel::columnDensity (CubicSplineSmoothingKernel.cpp:56-77)."""
    q2 = q * q
    s = np.sqrt((1 - q) * (1 + q))
    if q < 0.5:
        t = np.sqrt((1 - 2 * q) * (1 + 2 * q))
        return 2 / np.pi * ((4 + 26 * q2) * s - (1 + 26 * q2) * t - 18 * q2 * q2 * np.log(2 * q / (1 + t))
                            - 6 * q2 * (4 + q2) * np.log(2 * (1 + s) / (1 + t)))
    return 4 / np.pi * ((2 + 13 * q2) * s + 3 * q2 * (4 + q2) * np.log(q / (1 + s)))


def _expected(edges, sigma):
    out = []
    for a, b in zip(edges[:-1], edges[1:]):
        q = np.linspace(a, b, 2001)[1:-1]
        out.append(np.trapezoid([2 * np.pi * x * sigma(x) for x in q], q))
    out = np.array(out)
    return out / out.sum()


def test_spline_kernel_positions_follow_column_density():
    """One particle (h = 8 pc) seen face-on: the fraction of its light in annuli of projected
    radius q = R / h follows the cubic spline's column density (within 1.5 % of each annulus's
    share); a wrong table (uniform kernel) misses it by far more."""
    h = 8 * PC
    edges, got = _annulus_profile(1, h)
    want = _expected(edges, _column_density_spline)
    np.testing.assert_allclose(got, want, rtol=0.015, atol=2e-4)
    wrong = _expected(edges, lambda q: np.sqrt(1 - q * q))
    assert np.max(np.abs(got - wrong) / want) > 0.1


def test_uniform_kernel_positions_follow_chord_length():
    """The uniform kernel (u = X^(1/3)): column density proportional to the chord sqrt(1 - q^2)."""
    h = 8 * PC
    edges, got = _annulus_profile(0, h)
    want = _expected(edges, lambda q: np.sqrt(1 - q * q))
    np.testing.assert_allclose(got, want, rtol=0.015, atol=2e-4)
