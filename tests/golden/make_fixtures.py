"""Write the golden vectors under tests/golden/ by running the reference SKIRT build.

Run here only (needs /root/reference and `make -C oracle/ref -j8`):

    python tests/golden/make_fixtures.py [--images]

Every fixture is data: seeded inputs plus the reference's outputs for them.

  cart_paths.npz   Cartesian grids (config-0 64^3 and an 8^3 grid): raw
                   (m, ds) sequences of CartesianSpatialGrid::MySegmentGenerator::next()
                   for seeded rays, and MediumSystem cumulative / observer optical depths
                   (the 64^3 grid omits tau because its densities are 2 MB).
  tree_paths.npz   a spiral-disk octree (DensityTreePolicy, levels 2..5): node bounds,
                   children, cell indices, per-wall sorted neighbour lists, per-cell
                   densities, and the same path / tau outputs for seeded rays.
  voronoi_paths.npz  a Voronoi grid of 1500 density-sampled sites: site coordinates,
                   neighbour lists (walls as -1..-6), block search lists, densities, the
                   path / tau outputs for seeded rays and SpatialGrid::cellIndex for points.
  tetra_paths.npz  a Delaunay tetrahedral grid of 500 density-sampled vertices plus the
                   domain corners: vertices, tetra vertex indices, face anchors, neighbours
                   and normals, centroids, densities, paths / tau and cellIndex for points.
  dust.npz         DustMix sections, asymmetry parameters and peel-off phase values
                   (HG and 4-degree mean HG) for a ten-wavelength MeanListDustMix, and the
                   40-wavelength power-law mix of the octree workload.
  image_cfg0.npz   (--images) reference frame images for the Cartesian workload:
                   per-pixel mean and standard error over independent runs, plus a
                   near-transparent calibration run giving the per-wavelength factor
                   from recorded luminosity to the FITS surface brightness.
  sed_cfg0.npz     (--seds) the same workload with two SEDInstruments at 60 deg (no
                   aperture, 3 pc aperture): F_nu mean and standard error over independent
                   runs; sed_cfg0_sed60.dat is one run's SED text output as written.
  image_mesh.npz   (--meshes) the config-0 disk on a Voronoi grid of 20000 sites and on a
                   tetrahedral grid of 20000 vertices (DustDensity policy), 100x100 frame at
                   60 deg: per-pixel mean and standard error over independent runs, each run
                   with its own random sites and densities.
  image_mesh_file.npz (--mesh-files) the same on a Voronoi grid and a tetrahedral grid whose
                   20000 sites / vertices are read from a file (policy File; the disk's
                   distribution, workloads.disk_points seed 77): every run on the same mesh.
  image_persp.npz  (--persp) the config-0 workload seen by a PerspectiveInstrument (eye
                   outside the domain, 100x100 pixels, viewport 20 pc wide at 30 pc focal
                   length, from (10, -30, 20) pc toward (1, 0, 0) pc, so the image has no mirror
                   symmetry): per-pixel mean and
                   standard error of the FITS surface brightness over independent runs.
  image_particles.npz  (--particles) the config-0 disk lit by 30 black-body stellar particles
                   (ParticleSource, skigen.disk_particles) instead of the point source, with the
                   cubic-spline and the uniform smoothing kernel, 100x100 frame at 60 deg:
                   per-pixel mean and standard error over independent runs.
  image_clumpy.npz (--clumpy) the config-2 medium at fixture size: the lognormal clumpy
                   medium of skirt9-gpu_amd/workloads.py voronoi_clumpy (20000 sites, seed 1),
                   written to a text file and imported by the reference as a VoronoiMeshMedium
                   on a VoronoiMeshSpatialGrid with policy ImportedMesh, three 100x100 frames at
                   0/45/90 deg: per-pixel mean and standard error over independent runs.
  rf_cfg0.npz      (--rf) the same workload on a 16^3 grid (1e4 density samples per cell)
                   storing the radiation field:
                   the RadiationFieldProbe's mean intensity J_nu per cell (PerCellForm),
                   mean and standard error over independent runs, and the file's header.
"""

from __future__ import annotations

import argparse
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import skigen  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "bin", "skirt", "ref", "skirt_ref")
PC = 3.08567758e16  # parsec in metres (Constants.hpp:50)


def run(*args):
    res = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"skirt_ref {args[0]} failed (status {res.returncode}):\n{res.stdout}\n{res.stderr}")
    return res.stdout


class Reader:
    def __init__(self, path):
        self.b = open(path, "rb").read()
        self.o = 0

    def take(self, dtype, n):
        dt = np.dtype(dtype)
        a = np.frombuffer(self.b, dtype=dt, count=n, offset=self.o).copy()
        self.o += dt.itemsize * n
        return a

    def i32(self):
        return int(self.take("<i4", 1)[0])

    def i64(self):
        return int(self.take("<i8", 1)[0])

    def magic(self, m):
        assert self.b[self.o:self.o + 4] == m.encode(), m
        self.o += 4
        assert self.i32() == 1


def write_ski(tmp, name, spec):
    path = os.path.join(tmp, name + ".ski")
    with open(path, "w") as f:
        f.write(skigen.render(spec))
    return path


def trace(tmp, ski, rays, lam):
    rp = os.path.join(tmp, "rays.bin")
    with open(rp, "wb") as f:
        f.write(struct.pack("<q", len(rays)))
        f.write(np.ascontiguousarray(rays, dtype="<f8").tobytes())
    op = os.path.join(tmp, "paths.bin")
    run("paths", ski, rp, op, repr(lam))
    r = Reader(op)
    r.magic("SKPT")
    n = r.i64()
    off = r.take("<i8", n + 1)
    m = r.take("<i4", off[-1])
    ds = r.take("<f8", off[-1])
    toff = r.take("<i8", n + 1)
    tau = r.take("<f8", toff[-1])
    tauobs = r.take("<f8", n)
    return dict(off=off, m=m, ds=ds, toff=toff, tau=tau, tauobs=tauobs)


def grid_dump(tmp, ski):
    op = os.path.join(tmp, "grid.bin")
    run("grid", ski, op)
    r = Reader(op)
    r.magic("SKGD")
    nc = r.i64()
    out = dict(density=r.take("<f8", nc))
    kind = r.i32()
    out["kind"] = kind
    if kind in (2, 3):
        out["box"] = r.take("<f8", 6)
        out["eps"] = r.take("<f8", 1)[0]
    if kind == 2:
        ns = r.i64()
        out["sites"] = r.take("<f8", 3 * ns).reshape(ns, 3)
        out["nb_begin"] = r.take("<i4", ns)
        out["nb_count"] = r.take("<i4", ns)
        out["nb_index"] = r.take("<i4", r.i64())
        nbk = r.i32()
        out["block_n"] = np.int32(nbk)
        out["block_begin"] = r.take("<i4", nbk ** 3)
        out["block_count"] = r.take("<i4", nbk ** 3)
        out["block_index"] = r.take("<i4", r.i64())
    if kind == 3:
        nv = r.i64()
        out["verts"] = r.take("<f8", 3 * nv).reshape(nv, 3)
        nt = r.i64()
        out["tet_verts"] = r.take("<i4", 4 * nt).reshape(nt, 4)
        out["face_anchor"] = r.take("<i4", 4 * nt).reshape(nt, 4)
        out["face_nb"] = r.take("<i4", 4 * nt).reshape(nt, 4)
        out["face_normal"] = r.take("<f8", 12 * nt).reshape(nt, 4, 3)
        out["centroid"] = r.take("<f8", 3 * nt).reshape(nt, 3)
    if kind == 1:
        nn = r.i64()
        out["bounds"] = r.take("<f8", 6 * nn).reshape(nn, 6)
        out["child_begin"] = r.take("<i4", nn)
        out["child_count"] = r.take("<i4", nn)
        nch = r.i64()
        out["child_index"] = r.take("<i4", nch)
        out["cell_index"] = r.take("<i4", nn)
        nbo = r.i64()
        out["nb_off"] = r.take("<i8", nbo)
        out["nb_id"] = r.take("<i4", out["nb_off"][-1])
        out["bounds_match"] = np.int32(r.i32())
    return out


def unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def make_rays(rng, lo, hi, borders=None, n_random=600):
    """Seeded rays: inside/outside starts, axis-aligned and grazing directions, border starts."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    c = 0.5 * (lo + hi)
    w = hi - lo
    rays = []
    # random inside, isotropic
    r = lo + w * rng.random((n_random, 3))
    k = unit(rng.normal(size=(n_random, 3)))
    rays.append(np.hstack([r, k]))
    # outside, aimed at a random inside point
    tgt = lo + w * rng.random((n_random // 3, 3))
    src = c + unit(rng.normal(size=(n_random // 3, 3))) * 1.5 * np.linalg.norm(w)
    rays.append(np.hstack([src, unit(tgt - src)]))
    # outside, pointing away (empty paths)
    src = c + unit(rng.normal(size=(20, 3))) * 1.5 * np.linalg.norm(w)
    rays.append(np.hstack([src, unit(src - c)]))
    # axis-aligned and one-zero-component directions from random inside points
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
            (np.sqrt(0.5), np.sqrt(0.5), 0), (0, -np.sqrt(0.5), np.sqrt(0.5)), (np.sqrt(0.5), 0, -np.sqrt(0.5)),
            (np.sin(np.pi / 3), 0, np.cos(np.pi / 3)), (1 / np.sqrt(3),) * 3, (-1 / np.sqrt(3), 1 / np.sqrt(3), -1 / np.sqrt(3)),
            (1 - 1e-17, 1e-16, 0.0), (1e-16, 1e-16, 1.0)]
    dirs = np.array(dirs, float)
    for d in dirs:
        r = lo + w * rng.random((6, 3))
        rays.append(np.hstack([r, np.tile(d, (6, 1))]))
        # from the centre and from the origin
        rays.append(np.hstack([c[None, :], d[None, :]]))
        rays.append(np.hstack([np.zeros((1, 3)), d[None, :]]))
    # starts exactly on grid borders
    if borders is not None:
        for ax in range(3):
            b = borders[ax]
            for _ in range(40):
                r = lo + w * rng.random(3)
                r[ax] = b[rng.integers(0, len(b))]
                k = unit(rng.normal(size=3))
                rays.append(np.hstack([r, k])[None, :])
    # starts exactly on the outer box faces, pointing in and out
    for ax in range(3):
        for side in (lo[ax], hi[ax]):
            r = lo + w * rng.random((5, 3))
            r[:, ax] = side
            k = unit(rng.normal(size=(5, 3)))
            rays.append(np.hstack([r, k]))
    return np.vstack(rays)


def cart_borders(n, lo, hi):
    # LinMesh (NR::buildLinearGrid) followed by CartesianSpatialGrid::setupSelfAfter scaling
    t = np.array([0.0 + i * ((1.0 - 0.0) / n) for i in range(n + 1)])
    return t * (hi - lo) + lo


def fixtures_cart(tmp):
    out = {}
    rng = np.random.default_rng(20261202)
    h = 10 * PC
    # config-0 grid: geometry only (64^3)
    spec = skigen.config0(num_packets=1)
    ski = write_ski(tmp, "cart64", spec)
    bx = [cart_borders(64, -h, h)] * 3
    rays = make_rays(rng, [-h] * 3, [h] * 3, borders=bx, n_random=240)
    p = trace(tmp, ski, rays, 0.55e-6)
    out.update({"c64_rays": rays, "c64_off": p["off"], "c64_m": p["m"], "c64_ds": p["ds"],
                "c64_n": np.array([64, 64, 64]), "c64_half": np.array(h)})
    # small asymmetric grid with tau (densities are small enough to commit)
    spec = skigen.SkiSpec(num_packets=1, grid=("cartesian", 8))
    ski = write_ski(tmp, "cart8", spec)
    g = grid_dump(tmp, ski)
    bx = [cart_borders(8, -h, h)] * 3
    rays = make_rays(rng, [-h] * 3, [h] * 3, borders=bx, n_random=400)
    p = trace(tmp, ski, rays, 0.55e-6)
    dm = dust_tables(tmp, ski)
    out.update({"c8_rays": rays, "c8_off": p["off"], "c8_m": p["m"], "c8_ds": p["ds"], "c8_toff": p["toff"],
                "c8_tau": p["tau"], "c8_tauobs": p["tauobs"], "c8_density": g["density"],
                "c8_n": np.array([8, 8, 8]), "c8_half": np.array(h), "c8_sext": dm["sext"]})
    np.savez_compressed(os.path.join(HERE, "cart_paths.npz"), **out)
    print("cart_paths.npz:", {k: v.shape for k, v in out.items()})


def fixtures_tree(tmp):
    rng = np.random.default_rng(7)
    h = 10 * PC
    spec = skigen.SkiSpec(num_packets=1, spiral=True, grid=("octree", 2, 5, 3e-4))
    ski = write_ski(tmp, "tree6", spec)
    g = grid_dump(tmp, ski)
    assert int(g["bounds_match"]) == 1, "replayed octree differs from the reference grid"
    rays = make_rays(rng, [-h] * 3, [h] * 3, n_random=600)
    # rays from leaf corners and face centres along diagonal / axis directions (neighbour tie-breaks)
    leaves = np.nonzero(g["child_count"] == 0)[0]
    pick = rng.choice(leaves, 80, replace=False)
    extra = []
    for l in pick:
        b = g["bounds"][l]
        corner = b[:3]
        face = np.array([b[0], 0.5 * (b[1] + b[4]), 0.5 * (b[2] + b[5])])
        for r0 in (corner, face):
            for d in ((1, 1, 1), (1, 0, 0), (0, 1, 1), (np.sin(np.pi / 3), 0, np.cos(np.pi / 3))):
                d = np.array(d, float)
                extra.append(np.hstack([r0, d / np.linalg.norm(d)]))
    rays = np.vstack([rays, np.array(extra)])
    p = trace(tmp, ski, rays, 0.55e-6)
    dm = dust_tables(tmp, ski)
    out = {"rays": rays, "off": p["off"], "m": p["m"], "ds": p["ds"], "toff": p["toff"], "tau": p["tau"],
           "tauobs": p["tauobs"], "density": g["density"], "bounds": g["bounds"], "child_begin": g["child_begin"],
           "child_count": g["child_count"], "child_index": g["child_index"], "cell_index": g["cell_index"],
           "nb_off": g["nb_off"], "nb_id": g["nb_id"], "half": np.array(h), "sext": dm["sext"],
           "tree_levels": np.array([2, 5]), "max_dust_fraction": np.array(3e-4)}
    np.savez_compressed(os.path.join(HERE, "tree_paths.npz"), **out)
    print("tree_paths.npz:", {k: v.shape for k, v in out.items()}, "leaves", len(leaves))


def dust_tables(tmp, ski, ncos=201):
    op = os.path.join(tmp, "dust.bin")
    run("dust", ski, op, ncos)
    r = Reader(op)
    r.magic("SKDM")
    nl = r.i64()
    nc = r.i64()
    return dict(lam=r.take("<f8", nl), sext=r.take("<f8", nl), ssca=r.take("<f8", nl), g=r.take("<f8", nl),
                cos=r.take("<f8", nc), phase=r.take("<f8", nl * nc).reshape(nl, nc))


def fixtures_dust(tmp):
    lam = [0.1, 0.2, 0.3, 0.5, 0.55, 0.7, 1.0, 2.0, 5.0, 10.0]
    g = [-0.97, -0.5, 0.0, 1e-7, 0.3, 0.6, 0.9, 0.96, 0.99, 0.9999999]
    kap = [500.0 * (l / 0.1) ** -1.5 for l in lam]
    alb = [0.7, 0.65, 0.6, 0.55, 0.5, 0.45, 0.4, 0.3, 0.2, 0.1]
    spec = skigen.SkiSpec(num_packets=1, wavelengths_um=lam, dust=tuple(zip(lam, kap, alb, g)),
                          grid=("cartesian", 4))
    d10 = dust_tables(tmp, write_ski(tmp, "dust10", spec))
    spec1 = skigen.config1(num_packets=1, max_level=3, min_level=2)
    d40 = dust_tables(tmp, write_ski(tmp, "dust40", spec1))
    out = {"in_lam_um": np.array(lam), "in_g": np.array(g), "in_kappa": np.array(kap), "in_albedo": np.array(alb)}
    out.update({"d10_" + k: v for k, v in d10.items()})
    out.update({"d40_" + k: v for k, v in d40.items()})
    out["d40_in"] = np.array([[0.1, 1000.0, 0.5, 0.6], [10.0, 1.0, 0.1, 0.6]])
    np.savez_compressed(os.path.join(HERE, "dust.npz"), **out)
    print("dust.npz:", {k: v.shape for k, v in out.items()})


def fits_image(path):
    """Minimal reader for the reference's float32 FITS cubes (BITPIX=-32, NAXIS 2 or 3)."""
    raw = open(path, "rb").read()
    hdr = {}
    pos = 0
    while True:
        card = raw[pos:pos + 80].decode("ascii")
        pos += 80
        key = card[:8].strip()
        if key == "END":
            break
        if card[8:10] == "= ":
            hdr[key] = card[10:].split("/")[0].strip()
    pos = (pos + 2879) // 2880 * 2880
    assert int(hdr["BITPIX"]) == -32
    dims = [int(hdr[f"NAXIS{i + 1}"]) for i in range(int(hdr["NAXIS"]))]
    n = int(np.prod(dims))
    data = np.frombuffer(raw, dtype=">f4", count=n, offset=pos).astype(np.float64)
    return data.reshape(dims[::-1])


def fixtures_images(tmp, runs=8, packets=1e6):
    outdir = os.path.join(tmp, "img")
    os.makedirs(outdir, exist_ok=True)
    imgs = []
    for s in range(runs):
        spec = skigen.config0(num_packets=packets, seed=1000 + s)
        ski = write_ski(tmp, f"img{s}", spec)
        run("run", ski, outdir, 8)
        imgs.append(fits_image(os.path.join(outdir, f"img{s}_i60_total.fits")))
    imgs = np.array(imgs)
    # near-transparent calibration run: recorded luminosity -> surface brightness factor
    spec = skigen.config0(num_packets=1e4, seed=1)
    spec.tau_v = 1e-9
    ski = write_ski(tmp, "calib", spec)
    run("run", ski, outdir, 8)
    cal = fits_image(os.path.join(outdir, "calib_i60_total.fits"))
    out = {"mean": imgs.mean(axis=0).astype(np.float32), "sem": (imgs.std(axis=0, ddof=1) / np.sqrt(runs)).astype(np.float32),
           "runs": np.array(runs), "packets": np.array(packets), "calib_total": np.array(cal.sum()),
           "calib_packets": np.array(1e4)}
    np.savez_compressed(os.path.join(HERE, "image_cfg0.npz"), **out)
    print("image_cfg0.npz", out["mean"].shape, "calib", out["calib_total"])


def cells(tmp, ski, pts):
    pp = os.path.join(tmp, "points.bin")
    with open(pp, "wb") as f:
        f.write(struct.pack("<q", len(pts)))
        f.write(np.ascontiguousarray(pts, dtype="<f8").tobytes())
    op = os.path.join(tmp, "cells.bin")
    run("cells", ski, pp, op)
    r = Reader(op)
    r.magic("SKCI")
    return r.take("<i4", r.i64())


def fixtures_voronoi(tmp):
    """Voronoi grid of 1500 density-sampled sites (VoronoiMeshSpatialGrid, DustDensity policy)."""
    rng = np.random.default_rng(11)
    h = 10 * PC
    spec = skigen.SkiSpec(num_packets=1, grid=("voronoi", 1500, "DustDensity"))
    ski = write_ski(tmp, "vor", spec)
    g = grid_dump(tmp, ski)
    rays = make_rays(rng, [-h] * 3, [h] * 3, n_random=600)
    # rays from sites, towards neighbouring sites (through a shared face) and along bisector planes
    sites = g["sites"]
    extra = []
    for m in rng.choice(len(sites), 60, replace=False):
        nb = g["nb_index"][g["nb_begin"][m]:g["nb_begin"][m] + g["nb_count"][m]]
        nb = nb[nb >= 0]
        j = nb[rng.integers(0, len(nb))]
        extra.append(np.hstack([sites[m], unit(sites[j] - sites[m])]))
        mid = 0.5 * (sites[m] + sites[j])
        t = unit(np.cross(sites[j] - sites[m], rng.normal(size=3)))
        extra.append(np.hstack([mid, t]))
    rays = np.vstack([rays, np.array(extra)])
    p = trace(tmp, ski, rays, 0.55e-6)
    pts = np.vstack([-h + 2 * h * rng.random((2000, 3)), sites[:200],
                     0.5 * (sites[:100] + sites[g["nb_index"][g["nb_begin"][:100]].clip(0)])])
    ci = cells(tmp, ski, pts)
    dm = dust_tables(tmp, ski)
    out = {"rays": rays, "off": p["off"], "m": p["m"], "ds": p["ds"], "toff": p["toff"], "tau": p["tau"],
           "tauobs": p["tauobs"], "density": g["density"], "box": g["box"], "eps": np.array(g["eps"]),
           "sites": sites, "nb_begin": g["nb_begin"], "nb_count": g["nb_count"], "nb_index": g["nb_index"],
           "block_n": g["block_n"], "block_begin": g["block_begin"], "block_count": g["block_count"],
           "block_index": g["block_index"], "points": pts, "point_cell": ci, "sext": dm["sext"]}
    np.savez_compressed(os.path.join(HERE, "voronoi_paths.npz"), **out)
    print("voronoi_paths.npz:", {k: v.shape for k, v in out.items()})


def traces_safely(tmp, ski, ray, mem_kib=4_000_000, seconds=20):
    rp = os.path.join(tmp, "one_ray.bin")
    with open(rp, "wb") as f:
        f.write(struct.pack("<q", 1))
        f.write(np.ascontiguousarray(ray, dtype="<f8").tobytes())
    cmd = f"ulimit -v {mem_kib}; exec timeout {seconds} {EXE} paths {ski} {rp} {os.path.join(tmp, 'one.bin')} 5.5e-7"
    return subprocess.run(["bash", "-c", cmd], capture_output=True).returncode == 0


def fixtures_tetra(tmp):
    """Delaunay tetrahedral grid of 500 density-sampled vertices plus the 8 domain corners
    (TetraMeshSpatialGrid, DustDensity policy, no refinement)."""
    rng = np.random.default_rng(12)
    h = 10 * PC
    spec = skigen.SkiSpec(num_packets=1, grid=("tetra", 500, "DustDensity"))
    ski = write_ski(tmp, "tet", spec)
    g = grid_dump(tmp, ski)
    rays = make_rays(rng, [-h] * 3, [h] * 3, n_random=600)
    # rays from vertices along edges and in random directions, from face centroids and from
    # tetra centroids (zero Pluecker products).  A ray running exactly along a mesh edge can make
    # the reference generator loop without end (its path grows until memory runs out); such rays
    # are detected with a memory and time limit and left out.
    V, T = g["verts"], g["tet_verts"]
    extra = []
    for t in rng.choice(len(T), 60, replace=False):
        a, b, c, d = V[T[t]]
        extra.append(np.hstack([a, unit(b - a)]))
        extra.append(np.hstack([a, unit(rng.normal(size=3))]))
        extra.append(np.hstack([(a + b + c) / 3, unit(d - (a + b + c) / 3)]))
        extra.append(np.hstack([g["centroid"][t], unit(rng.normal(size=3))]))
    keep = [ray for ray in extra if traces_safely(tmp, ski, ray)]
    print("tetra: left out", len(extra) - len(keep), "runaway rays of", len(extra))
    rays = np.vstack([rays, np.array(keep)])
    p = trace(tmp, ski, rays, 0.55e-6)
    pts = np.vstack([-h + 2 * h * rng.random((2000, 3)), g["centroid"][:200]])
    ci = cells(tmp, ski, pts)
    dm = dust_tables(tmp, ski)
    out = {"rays": rays, "off": p["off"], "m": p["m"], "ds": p["ds"], "toff": p["toff"], "tau": p["tau"],
           "tauobs": p["tauobs"], "density": g["density"], "box": g["box"], "eps": np.array(g["eps"]),
           "verts": V, "tet_verts": T, "face_anchor": g["face_anchor"], "face_nb": g["face_nb"],
           "face_normal": g["face_normal"], "centroid": g["centroid"], "points": pts, "point_cell": ci,
           "sext": dm["sext"]}
    np.savez_compressed(os.path.join(HERE, "tetra_paths.npz"), **out)
    print("tetra_paths.npz:", {k: v.shape for k, v in out.items()})


def sed_values(path):
    return np.array([float(ln.split()[1]) for ln in open(path) if ln.strip() and not ln.startswith("#")])


def fixtures_seds(tmp, runs=8, packets=1e6):
    outdir = os.path.join(tmp, "sed")
    os.makedirs(outdir, exist_ok=True)
    names = ("sed60", "ap60")
    vals = {n: [] for n in names}
    for s in range(runs):
        spec = skigen.config0(num_packets=packets, seed=2000 + s)
        spec.sed_instruments = (("sed60", 60.0, 0.0, 0.0), ("ap60", 60.0, 0.0, 3.0))
        ski = write_ski(tmp, f"sed{s}", spec)
        run("run", ski, outdir, 8)
        for n in names:
            vals[n].append(sed_values(os.path.join(outdir, f"sed{s}_{n}_sed.dat")))
    out = {"runs": np.array(runs), "packets": np.array(packets)}
    for n in names:
        v = np.array(vals[n])
        out[n + "_mean"] = v.mean(axis=0)
        out[n + "_sem"] = v.std(axis=0, ddof=1) / np.sqrt(runs)
    np.savez(os.path.join(HERE, "sed_cfg0.npz"), **out)
    with open(os.path.join(outdir, "sed0_sed60_sed.dat")) as f, open(os.path.join(HERE, "sed_cfg0_sed60.dat"), "w") as g:
        g.write(f.read())


def fixtures_rf(tmp, runs=8, packets=1e6):
    outdir = os.path.join(tmp, "rf")
    os.makedirs(outdir, exist_ok=True)
    Js, header = [], None
    for s in range(runs):
        spec = skigen.config0(num_packets=packets, seed=4000 + s)
        spec.grid = ("cartesian", 16)
        spec.store_rf = True
        # many density samples per cell: the sampled densities of the coarse cells around the
        # thin disk's midplane otherwise differ enough from run to run to shift the field of
        # a whole half of the grid together
        spec.num_density_samples = 10000
        ski = write_ski(tmp, f"rf{s}", spec)
        run("run", ski, outdir, 8)
        lines = open(os.path.join(outdir, f"rf{s}_rf_J.dat")).read().splitlines()
        header = [ln for ln in lines if ln.startswith("#")]
        Js.append(np.array([[float(x) for x in ln.split()[1:]] for ln in lines if not ln.startswith("#")]))
    Js = np.array(Js)
    np.savez(os.path.join(HERE, "rf_cfg0.npz"), mean=Js.mean(axis=0), sem=Js.std(axis=0, ddof=1) / np.sqrt(runs),
             runs=np.array(runs), packets=np.array(packets), header=np.array("\n".join(header)))


def fixtures_mesh_images(tmp, runs=8, packets=1e6):
    outdir = os.path.join(tmp, "mesh")
    os.makedirs(outdir, exist_ok=True)
    out = {"runs": np.array(runs), "packets": np.array(packets)}
    for kind in ("voronoi", "tetra"):
        imgs = []
        for s in range(runs):
            spec = skigen.config0(num_packets=packets, seed=6000 + s)
            spec.grid = (kind, 20000, "DustDensity")
            spec.num_pixels = 100
            ski = write_ski(tmp, f"{kind}{s}", spec)
            run("run", ski, outdir, 8)
            imgs.append(fits_image(os.path.join(outdir, f"{kind}{s}_i60_total.fits")))
        imgs = np.array(imgs)
        out[kind + "_mean"] = imgs.mean(axis=0).astype(np.float32)
        out[kind + "_sem"] = (imgs.std(axis=0, ddof=1) / np.sqrt(runs)).astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "image_mesh.npz"), **out)


MESH_FILE_POINTS = dict(n=20000, seed=77)


def mesh_file_points():
    """The 20000 sites / vertices of the File-policy mesh fixtures: the exponential disk's
    distribution (workloads.disk_points, seed 77) in the +-10 pc box."""
    import importlib.util
    name = "skirt9_gpu_amd"
    if name not in sys.modules:
        spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "skirt9-gpu_amd", "__init__.py"),
                                                      submodule_search_locations=[os.path.join(ROOT, "skirt9-gpu_amd")])
        mod = importlib.util.module_from_spec(spec)
        sys.modules[name] = mod
        spec.loader.exec_module(mod)
    from skirt9_gpu_amd import workloads
    h = 10 * PC
    return workloads.disk_points(MESH_FILE_POINTS["n"], np.array([-h, -h, -h, h, h, h]), seed=MESH_FILE_POINTS["seed"])


def fixtures_mesh_file_images(tmp, runs=8, packets=1e6):
    """The config-0 disk on a Voronoi grid and on a tetrahedral grid whose 20000 sites /
    vertices are read from a file (policy File; mesh_file_points), 100x100 frame at 60 deg:
    per-pixel mean and standard error over independent runs, all on the same mesh."""
    outdir = os.path.join(tmp, "meshfile")
    os.makedirs(outdir, exist_ok=True)
    skigen.write_points(os.path.join(tmp, "pts.txt"), mesh_file_points())
    out = {"runs": np.array(runs), "packets": np.array(packets)}
    for kind in ("voronoi", "tetra"):
        imgs = []
        for s in range(runs):
            spec = skigen.config0(num_packets=packets, seed=6100 + s)
            spec.grid = (kind + "_file", "pts.txt")
            spec.num_pixels = 100
            ski = write_ski(tmp, f"{kind}f{s}", spec)
            run("run", ski, outdir, 8)
            imgs.append(fits_image(os.path.join(outdir, f"{kind}f{s}_i60_total.fits")))
        imgs = np.array(imgs)
        out[kind + "_mean"] = imgs.mean(axis=0).astype(np.float32)
        out[kind + "_sem"] = (imgs.std(axis=0, ddof=1) / np.sqrt(runs)).astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "image_mesh_file.npz"), **out)


CAMERA = ("cam", (10.0, -30.0, 20.0), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), 30.0, 20.0, 100)


def fixtures_persp(tmp, runs=8, packets=1e6):
    outdir = os.path.join(tmp, "persp")
    os.makedirs(outdir, exist_ok=True)
    imgs = []
    for s in range(runs):
        spec = skigen.config0(num_packets=packets, seed=8000 + s)
        spec.persp_instruments = (CAMERA,)
        ski = write_ski(tmp, f"persp{s}", spec)
        run("run", ski, outdir, 8)
        imgs.append(fits_image(os.path.join(outdir, f"persp{s}_cam_total.fits")))
    imgs = np.array(imgs)
    np.savez_compressed(os.path.join(HERE, "image_persp.npz"), mean=imgs.mean(axis=0).astype(np.float32),
                        sem=(imgs.std(axis=0, ddof=1) / np.sqrt(runs)).astype(np.float32),
                        runs=np.array(runs), packets=np.array(packets))


def fixtures_particles(tmp, runs=8, packets=1e6):
    outdir = os.path.join(tmp, "particles")
    os.makedirs(outdir, exist_ok=True)
    skigen.write_particles(os.path.join(tmp, "parts.txt"), skigen.disk_particles())
    out = {"runs": np.array(runs), "packets": np.array(packets)}
    for kern in ("CubicSplineSmoothingKernel", "UniformSmoothingKernel"):
        imgs = []
        for s in range(runs):
            spec = skigen.config0(num_packets=packets, seed=10000 + s)
            spec.particle_file = "parts.txt"
            spec.particle_kernel = kern
            spec.num_pixels = 100
            ski = write_ski(tmp, f"pt{s}", spec)
            run("run", ski, outdir, 8)
            imgs.append(fits_image(os.path.join(outdir, f"pt{s}_i60_total.fits")))
        imgs = np.array(imgs)
        out[kern + "_mean"] = imgs.mean(axis=0).astype(np.float32)
        out[kern + "_sem"] = (imgs.std(axis=0, ddof=1) / np.sqrt(runs)).astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "image_particles.npz"), **out)


CLUMPY_FRAMES = (("i0", 0.0, 0.0), ("i45", 45.0, 0.0), ("i90", 90.0, 0.0))


def clumpy_ski(dirname, packets, seed):
    """The config-2 medium at fixture size as a ski file with its VoronoiMeshMedium sites file
    (shared by this generator and the GPU test)."""
    import importlib.util
    name = "skirt9_gpu_amd"
    if name not in sys.modules:
        spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "skirt9-gpu_amd", "__init__.py"),
                                                      submodule_search_locations=[os.path.join(ROOT, "skirt9-gpu_amd")])
        mod = importlib.util.module_from_spec(spec)
        sys.modules[name] = mod
        spec.loader.exec_module(mod)
    from skirt9_gpu_amd import workloads
    wl = workloads.voronoi_clumpy(num_sites=20000, threads=8)
    skigen.write_voronoi_medium(os.path.join(dirname, "clumpy_sites.txt"), wl.tables["mesh"].export()["sites"],
                                wl.density)
    spec = skigen.SkiSpec(num_packets=packets, seed=seed, grid=("voronoi_imported",), medium_file="clumpy_sites.txt",
                          num_pixels=100, instruments=CLUMPY_FRAMES)
    path = os.path.join(dirname, f"clumpy{seed}.ski")
    with open(path, "w") as f:
        f.write(skigen.render(spec))
    return path


def fixtures_clumpy(tmp, runs=8, packets=1e6):
    outdir = os.path.join(tmp, "clumpy")
    os.makedirs(outdir, exist_ok=True)
    imgs = {n: [] for n, _, _ in CLUMPY_FRAMES}
    for s in range(runs):
        ski = clumpy_ski(outdir, packets, 9000 + s)
        run("run", ski, outdir, 8)
        for n, _, _ in CLUMPY_FRAMES:
            imgs[n].append(fits_image(os.path.join(outdir, f"clumpy{9000 + s}_{n}_total.fits")))
    out = {"runs": np.array(runs), "packets": np.array(packets)}
    for n, v in imgs.items():
        v = np.array(v)
        out[n + "_mean"] = v.mean(axis=0).astype(np.float32)
        out[n + "_sem"] = (v.std(axis=0, ddof=1) / np.sqrt(runs)).astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "image_clumpy.npz"), **out)
    print("image_clumpy.npz", {k: v.shape for k, v in out.items()})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--images", action="store_true")
    ap.add_argument("--seds", action="store_true")
    ap.add_argument("--rf", action="store_true")
    ap.add_argument("--meshes", action="store_true")
    ap.add_argument("--mesh-files", action="store_true")
    ap.add_argument("--persp", action="store_true")
    ap.add_argument("--particles", action="store_true")
    ap.add_argument("--clumpy", action="store_true")
    ap.add_argument("--only", default="")
    a = ap.parse_args()
    with tempfile.TemporaryDirectory() as tmp:
        parts = a.only.split(",") if a.only else ["cart", "tree", "dust", "voronoi", "tetra"]
        if "cart" in parts:
            fixtures_cart(tmp)
        if "tree" in parts:
            fixtures_tree(tmp)
        if "dust" in parts:
            fixtures_dust(tmp)
        if "voronoi" in parts:
            fixtures_voronoi(tmp)
        if "tetra" in parts:
            fixtures_tetra(tmp)
        if a.images:
            fixtures_images(tmp)
        if a.seds:
            fixtures_seds(tmp)
        if a.rf:
            fixtures_rf(tmp)
        if a.meshes:
            fixtures_mesh_images(tmp)
        if a.mesh_files:
            fixtures_mesh_file_images(tmp)
        if a.persp:
            fixtures_persp(tmp)
        if a.particles:
            fixtures_particles(tmp)
        if a.clumpy:
            fixtures_clumpy(tmp)


if __name__ == "__main__":
    main()
