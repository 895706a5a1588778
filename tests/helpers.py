"""Shared helpers for the tests: package loading, golden fixtures and small workloads."""

from __future__ import annotations

import importlib.util
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
PC = 3.08567758e16  # parsec (m)


def load_pkg():
    name = "skirt9_gpu_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "skirt9-gpu_amd", "__init__.py"),
                                                  submodule_search_locations=[os.path.join(ROOT, "skirt9-gpu_amd")])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def golden(name):
    return np.load(os.path.join(GOLDEN, name))


def cart_borders(n, lo, hi):
    """LinMesh borders scaled as CartesianSpatialGrid::setupSelfAfter (i/N)*(max-min)+min."""
    t = np.array([0.0 + i * ((1.0 - 0.0) / n) for i in range(n + 1)])
    return t * (hi - lo) + lo


def child_first(child_begin, child_count, child_index):
    cb, cc, ci = map(np.asarray, (child_begin, child_count, child_index))
    out = np.full(len(cb), -1, np.int32)
    has = cc > 0
    out[has] = ci[cb[has]]
    return out


def disk_density_cartesian(n=64, half=10 * PC, tau_v=1.0, sext=1.5e-29, samples=100, seed=5):
    """Cell-averaged number density of the exponential disk (scale length 3 pc, height 0.3 pc)
    normalised to face-on tau_V (ExpDiskGeometry.cpp, OpticalDepthMaterialNormalization.cpp),
    sampled at `samples` random points per cell like MediumSystem's PropertySampler."""
    b = cart_borders(n, -half, half)
    hR, hz = 3 * PC, 0.3 * PC
    rho0 = 1 / (4 * np.pi * hR * hR * hz)
    N = tau_v / sext / (2 * rho0 * hz)
    rng = np.random.default_rng(seed)
    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    acc = np.zeros((n, n, n))
    for _ in range(samples):
        x = b[ii] + rng.random(ii.shape) * (b[ii + 1] - b[ii])
        y = b[jj] + rng.random(ii.shape) * (b[jj + 1] - b[jj])
        z = b[kk] + rng.random(ii.shape) * (b[kk + 1] - b[kk])
        acc += rho0 * np.exp(-np.sqrt(x * x + y * y) / hR) * np.exp(-np.abs(z) / hz)
    return b, (N * acc / samples).reshape(-1)


def frame60(npix=200, fov=20 * PC):
    return dict(inclination=np.radians(60.0), azimuth=0.0, roll=0.0, nx=npix, ny=npix, fovx=fov, fovy=fov)


def pixel_parity(imgs, per, ref, ref_sem, ref_runs, ref_packets, max_rel=0.1):
    """Per-pixel comparison at native resolution of nb independent images (each of `per`
    packets) with a reference mean image and its standard error over `ref_runs` runs of
    `ref_packets` packets.  Per pixel, the pooled per-packet variance of both sets gives a
    two-sample t statistic with nu = ref_runs + nb - 2 degrees of freedom, so that, when the
    two agree, E[t^2] = nu / (nu - 2) and P(|t| > 4) is the t distribution's tail.  Only
    pixels whose mean is known to max_rel relative standard error on both sides take part
    (enough deposits for the run sums to be near normal).

    Returns the number of pixels, the mean of t^2 with its expectation and a 5-sigma
    tolerance (from the t distribution's fourth moment), and the fraction with |t| > 4 with
    its expectation and a 5-sigma binomial tolerance."""
    from scipy import stats
    imgs = np.asarray(imgs, dtype=np.float64)
    nb = len(imgs)
    om = imgs.mean(0)
    var_o = imgs.var(0, ddof=1)
    var_r = ref_sem ** 2 * ref_runs
    nu = ref_runs + nb - 2
    vpp = ((ref_runs - 1) * var_r * ref_packets + (nb - 1) * var_o * per) / nu
    var = vpp / (nb * per) + vpp / (ref_runs * ref_packets)
    sel = (ref > 0) & (om > 0) & (ref_sem < max_rel * ref) & (np.sqrt(var_o / nb) < max_rel * om) & (var > 0)
    z = (om - ref)[sel] / np.sqrt(var[sel])
    m = int(sel.sum())
    e2 = nu / (nu - 2)
    sd2 = np.sqrt(3 * nu * nu / ((nu - 2) * (nu - 4)) - e2 * e2)
    p4 = 2 * stats.t.sf(4.0, nu)
    return dict(pixels=m, chi2_dof=float(np.mean(z ** 2)) if m else np.nan, expected=e2,
                tol=5 * sd2 / np.sqrt(max(m, 1)), frac4=float(np.mean(np.abs(z) > 4)) if m else np.nan,
                expected_frac4=p4, tol_frac4=5 * np.sqrt(p4 * (1 - p4) / max(m, 1)), z=z)
