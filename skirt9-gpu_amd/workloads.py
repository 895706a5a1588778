"""The benchmark and test workloads, set up the way the corresponding ski files configure the
reference (an isotropic point source at the origin, an exponential dust disk, one HG dust
mix, forced scattering with path-length bias 0.5, distant frame instruments).

  cartesian_disk   64^3 Cartesian grid over +-10 pc, disk normalised to face-on tau_V = 1,
                   0.55 um only, albedo 0.6, g 0.6, 200x200 frame at 60 deg
  octree_spiral    spiral-perturbed disk on a density-refined octree (levels 3..10, max dust
                   fraction 1e-6, ~2.9e6 leaves), power-law dust kappa ~ lambda^-1.5 with
                   albedo 0.5 -> 0.1 over 0.1..10 um, at a chosen face-on tau_V and wavelengths
  voronoi_clumpy   Voronoi grid of density-sampled sites in a lognormal clumpy medium: a
                   Gaussian random field with power spectrum P(k) ~ k^-3.7 on a lattice,
                   sigma_ln = 1, mean tau_V along z; 0.55 um, albedo 0.6, g 0.6; three
                   200x200 frames at 0, 45 and 90 deg
  tetra_batch      256 independent variants of cartesian_disk in one launch, on a Delaunay
                   tetrahedral mesh of 2e5 points distributed as the disk density
                   (~1.3e6 tetrahedra): per instance tau_V log-uniform in 0.1..50,
                   inclination uniform in 0..90 deg, 1e5..1e7 packets (log-uniform)
"""

from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

PC = 3.08567758e16
MU_DUST = 1.5e-29  # MeanListDustMix mass per hydrogen atom (kg)


def _pkg():
    import sys
    return sys.modules[__package__] if __package__ else __import__("skirt9_gpu_amd")


@dataclass
class Workload:
    name: str
    grid: str
    lam: np.ndarray
    sext: np.ndarray
    ssca: np.ndarray
    g: np.ndarray
    density: np.ndarray
    tables: dict = field(default_factory=dict)  # grid tables (borders or octree export)
    octree: object = None
    frames: list = field(default_factory=list)
    tau_v: float = 1.0
    setup_s: float = 0.0
    instances: dict = None  # batched variants: packets, seeds, tau_scale, frame per instance

    def apply(self, eng, npk_norm=1.0):
        if self.grid == "cartesian":
            eng.grid_cartesian(self.tables["xb"], self.tables["yb"], self.tables["zb"])
        elif self.grid in ("voronoi", "tetra"):
            self.tables["mesh"].upload(eng)
        else:
            self.octree.upload(eng)
        eng.medium(self.density)
        eng.dust(self.lam, self.sext, self.ssca, self.g)
        eng.frame_remove_all()
        for f in self.frames:
            eng.frame_add(f["inclination"], f["azimuth"], f["roll"], f["nx"], f["ny"], f["fovx"], f["fovy"])
        eng.source_point((0.0, 0.0, 0.0), np.full(len(self.lam), 1.0 / npk_norm))
        if self.instances is not None:
            I = self.instances
            n = len(I["packets"])
            lp = np.repeat((1.0 / np.maximum(I["packets"], 1))[:, None], len(self.lam), axis=1)
            eng.instances(I["packets"], I["seeds"], I["tau_scale"], lp, np.arange(n), np.ones(n))
        else:
            eng.instances([])

    def total_packets(self):
        return int(np.sum(self.instances["packets"])) if self.instances is not None else 0


def frame(inc_deg, npix=200, fov=20 * PC, az_deg=0.0):
    return dict(inclination=np.radians(inc_deg), azimuth=np.radians(az_deg), roll=0.0, nx=npix, ny=npix, fovx=fov,
                fovy=fov)


def log_wavelengths(lo_um, hi_um, n):
    """Log-spaced wavelengths (m), as the ski files list them (9 significant digits)."""
    v = [lo_um * (hi_um / lo_um) ** (i / (n - 1)) for i in range(n)] if n > 1 else [lo_um]
    return np.array([float(f"{x:.9g}") for x in v]) * 1e-6


def cartesian_disk(n=64, tau_v=1.0, samples=100, seed=1, threads=0):
    import time
    pkg = _pkg()
    t0 = time.time()
    h = 10 * PC
    t = np.array([0.0 + i * ((1.0 - 0.0) / n) for i in range(n + 1)])
    b = t * (h - (-h)) + (-h)
    sext = MU_DUST * 1.0
    model = pkg.disk_model(tau_v=tau_v, sext_v=sext)
    dens = pkg.sample_density_cartesian(b, b, b, model, samples, seed, threads)
    return Workload("cartesian_disk", "cartesian", np.array([0.55e-6]), np.array([sext]), np.array([0.6 * sext]),
                    np.array([0.6]), dens, dict(xb=b, yb=b, zb=b), frames=[frame(60.0)], tau_v=tau_v,
                    setup_s=time.time() - t0)


def octree_spiral(tau_v=1.0, nlambda=40, lam_range_um=(0.1, 10.0), min_level=3, max_level=10, max_frac=1e-6,
                  samples=100, seed=1, threads=0, frames=None, engine=None):
    """With an engine, the tree's per-level density estimates and the cell densities are
    computed on its GPU (the same random positions as on the host)."""
    import time
    pkg = _pkg()
    t0 = time.time()
    h = 10 * PC
    box = [-h, -h, -h, h, h, h]
    lam = log_wavelengths(lam_range_um[0], lam_range_um[1], nlambda)
    # power-law dust: kappa 1000 -> 1 m2/kg and albedo 0.5 -> 0.1 between 0.1 and 10 um, g = 0.6
    tab = np.array([[0.1e-6, 1000.0, 0.5, 0.6], [10e-6, 1.0, 0.1, 0.6]])
    sext, ssca, g = pkg.dust_meanlist(tab[:, 0], tab[:, 1], tab[:, 2], tab[:, 3], lam, MU_DUST)
    sv, _, _ = pkg.dust_meanlist(tab[:, 0], tab[:, 1], tab[:, 2], tab[:, 3], [0.55e-6], MU_DUST)
    model = pkg.disk_model(tau_v=tau_v, sext_v=float(sv[0]), spiral=True)
    if engine is not None:
        tree = pkg.Octree.build_gpu(engine, box, model, min_level, max_level, max_frac, samples, seed, threads)
        tree.upload(engine)
        dens = engine.sample_density(model, samples, seed + 1)
    else:
        tree = pkg.Octree.build(box, model, min_level, max_level, max_frac, samples, seed, threads)
        dens = tree.sample_density(model, samples, seed + 1, threads)
    return Workload("octree_spiral", "octree", lam, sext, ssca, g, dens, octree=tree,
                    frames=frames if frames is not None else [frame(60.0)], tau_v=tau_v, setup_s=time.time() - t0)


def lognormal_field(n=256, slope=-3.7, sigma_ln=1.0, seed=1):
    """exp(sigma_ln * G - sigma_ln^2 / 2) for a unit-variance Gaussian random field G with
    power spectrum P(k) ~ k^slope on a periodic n^3 lattice (mean 1)."""
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(n)[:, None, None]
    ky = np.fft.fftfreq(n)[None, :, None]
    kz = np.fft.rfftfreq(n)[None, None, :]
    k = np.sqrt(kx * kx + ky * ky + kz * kz)
    k[0, 0, 0] = 1.0
    amp = k ** (slope / 2.0)
    amp[0, 0, 0] = 0.0
    noise = np.fft.rfftn(rng.standard_normal((n, n, n)))
    g = np.fft.irfftn(noise * amp, s=(n, n, n))
    g = (g - g.mean()) / g.std()
    return np.exp(sigma_ln * g - 0.5 * sigma_ln * sigma_ln)


def voronoi_clumpy(num_sites=1_000_000, lattice=256, tau_v=5.0, sigma_ln=1.0, slope=-3.7, samples=20, seed=1,
                   threads=0, frames=None):
    import time
    pkg = _pkg()
    t0 = time.time()
    h = 10 * PC
    box = np.array([-h, -h, -h, h, h, h])
    rho = lognormal_field(lattice, slope, sigma_ln, seed)
    # sites distributed as the density: a lattice cell by weight, then uniform inside it
    rng = np.random.default_rng(seed + 1)
    w = rho.ravel() / rho.sum()
    cells = rng.choice(rho.size, size=num_sites, p=w)
    ijk = np.stack(np.unravel_index(cells, rho.shape), axis=1).astype(np.float64)
    sites = box[:3] + (ijk + rng.random((num_sites, 3))) * (2 * h / lattice)
    mesh = pkg.Voronoi.build(sites, box, threads)
    mean_rho = pkg.Voronoi.sample_density(mesh, rho, samples, seed + 2, threads)
    sext = MU_DUST * 1.0
    # mean optical depth along z through the box: tau = sext * n0 * <rho> * 2h with <rho> = 1
    n0 = tau_v / (sext * 2 * h)
    return Workload("voronoi_clumpy", "voronoi", np.array([0.55e-6]), np.array([sext]), np.array([0.6 * sext]),
                    np.array([0.6]), n0 * mean_rho, dict(mesh=mesh), tau_v=tau_v,
                    frames=frames if frames is not None else [frame(0.0), frame(45.0), frame(90.0)],
                    setup_s=time.time() - t0)


def disk_points(n, box, hR=3 * PC, hz=0.3 * PC, seed=1, spiral=None):
    """n positions distributed as the exponential disk: the distributions of
    SepAxGeometry::generatePosition (SepAxGeometry.cpp:11-19) with ExpDiskGeometry's
    randomCylRadius / randomZ (ExpDiskGeometry.cpp:46-68) -- cylindrical radius with density
    R exp(-R/hR), uniform azimuth, Laplace z -- drawn with numpy's samplers; positions outside
    the box are discarded and redrawn, as TetraMeshSpatialGrid's vertex sampling does
    (TetraMeshSpatialGrid.cpp:23-40).  With spiral parameters (arms, pitch, radius_zero,
    phase_zero, weight, index), the azimuth is redrawn until accepted with probability
    perturbation(R, phi) / max, as SpiralStructureGeometryDecorator::generatePosition does
    (SpiralStructureGeometryDecorator.cpp:33-46)."""
    rng = np.random.default_rng(seed)
    box = np.asarray(box, dtype=np.float64)
    out = np.empty((0, 3))
    while len(out) < n:
        m = 2 * (n - len(out)) + 1024
        R = rng.gamma(2.0, hR, size=m)
        phi = 2 * np.pi * rng.random(m)
        if spiral:
            import math
            arms, N, w = spiral["arms"], spiral["index"], spiral["weight"]
            cn = math.sqrt(math.pi) * math.gamma(N + 1.0) / math.gamma(N + 0.5)
            c = 1.0 + (cn - 1.0) * w
            gam = np.log(R / spiral["radius_zero"]) / math.tan(spiral["pitch"]) + spiral["phase_zero"] + 0.5 * math.pi / arms
            todo = np.ones(m, bool)
            while todo.any():
                k = np.flatnonzero(todo)
                ph = 2 * np.pi * rng.random(len(k))
                pert = (1 - w) + w * cn * np.sin(0.5 * arms * (gam[k] - ph)) ** (2 * N)
                ok = rng.random(len(k)) * c / pert <= 1
                phi[k[ok]] = ph[ok]
                todo[k[ok]] = False
        z = rng.exponential(hz, size=m) * np.where(rng.random(m) < 0.5, -1.0, 1.0)
        p = np.stack([R * np.cos(phi), R * np.sin(phi), z], axis=1)
        p = p[np.all((p > box[:3]) & (p < box[3:]), axis=1)]
        out = np.concatenate([out, p])
    return out[:n]


def batch_instances(n=256, packet_range=(1e5, 1e7), tau_range=(0.1, 50.0), inc_range_deg=(0.0, 90.0), seed=3,
                    packet_scale=1.0):
    """Per-instance parameters of the batched workload from a fixed seed: tau_V log-uniform,
    inclination uniform, packet count log-uniform (times packet_scale, at least 1)."""
    rng = np.random.default_rng(seed)
    tau = np.exp(rng.uniform(np.log(tau_range[0]), np.log(tau_range[1]), n))
    inc = rng.uniform(inc_range_deg[0], inc_range_deg[1], n)
    npk = np.exp(rng.uniform(np.log(packet_range[0]), np.log(packet_range[1]), n))
    npk = np.maximum(1, np.round(npk * packet_scale)).astype(np.uint64)
    seeds = (1000 + np.arange(n)).astype(np.uint64)
    return dict(packets=npk, seeds=seeds, tau_scale=tau, inclination_deg=inc)


def tetra_batch(num_points=200_000, num_instances=256, packet_scale=1.0, npix=200, samples=100, seed=1, threads=0):
    import time
    pkg = _pkg()
    t0 = time.time()
    h = 10 * PC
    box = np.array([-h, -h, -h, h, h, h])
    mesh = pkg.Tetra.build(disk_points(num_points, box, seed=seed), box)
    ex = mesh.export()
    sext = MU_DUST * 1.0
    # the disk normalised to face-on tau_V = 1; instance i scales the optical depths to its tau_V
    model = pkg.disk_model(tau_v=1.0, sext_v=sext)
    dens = pkg.sample_density_tetra(ex["verts"], ex["tet_verts"], model, samples, seed + 1, threads)
    inst = batch_instances(num_instances, seed=seed + 2, packet_scale=packet_scale)
    frames = [frame(i, npix=npix) for i in inst["inclination_deg"]]
    return Workload("tetra_batch", "tetra", np.array([0.55e-6]), np.array([sext]), np.array([0.6 * sext]),
                    np.array([0.6]), dens, dict(mesh=mesh, ntets=len(ex["tet_verts"])), frames=frames, tau_v=1.0,
                    instances=inst, setup_s=time.time() - t0)
