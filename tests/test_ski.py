"""The ski-file front end and FITS writer (CPU): parsing of the reference's configuration
items, the absolute calibration against the reference's own FITS output, and the FITS
round trip."""

import os
import sys

import numpy as np
import pytest

from helpers import GOLDEN, PC, ROOT, golden, load_pkg

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import skigen  # noqa: E402


@pytest.fixture(scope="module")
def ski():
    load_pkg()
    from skirt9_gpu_amd import ski as S
    return S


def _write(tmp_path, name, spec):
    p = os.path.join(tmp_path, name + ".ski")
    with open(p, "w") as f:
        f.write(skigen.render(spec))
    return p


def test_parse_config0(ski, tmp_path):
    sim = ski.load(_write(str(tmp_path), "c0", skigen.config0(num_packets=1e6, seed=3)))
    assert sim.num_packets == 1000000 and sim.seed == 3
    np.testing.assert_allclose(sim.wavelengths, [0.55e-6])
    assert sim.grid == dict(kind="cartesian", n=[64, 64, 64])
    assert sim.box == (-10 * PC, -10 * PC, -10 * PC, 10 * PC, 10 * PC, 10 * PC)
    assert sim.tau_v == 1.0 and not sim.spiral
    assert abs(sim.scale_length - 3 * PC) < 1e-6 * PC and abs(sim.scale_height - 0.3 * PC) < 1e-6 * PC
    (f,) = sim.frames
    assert f.name == "i60" and f.nx == f.ny == 200
    assert abs(f.inclination - np.radians(60)) < 1e-15 and abs(f.distance / (1e7 * PC) - 1) < 1e-15
    assert sim.min_weight_reduction == 1e4 and sim.path_length_bias == 0.5 and sim.force_scattering


def test_parse_config1(ski, tmp_path):
    sim = ski.load(_write(str(tmp_path), "c1", skigen.config1(num_packets=1e8, tau_v=50, nlambda=5, lam_range=(0.3, 2.0))))
    assert len(sim.wavelengths) == 5 and sim.spiral
    assert sim.grid == dict(kind="octree", min_level=3, max_level=10, max_fraction=1e-6)
    assert sim.spiral_params["arms"] == 2 and abs(sim.spiral_params["pitch"] - np.radians(15)) < 1e-15


def test_calibration_matches_reference_fits(ski, tmp_path):
    """A near-transparent reference run puts the whole source luminosity into the frame: its
    FITS total equals our bin luminosity times our calibration factor."""
    sim = ski.load(_write(str(tmp_path), "c0", skigen.config0()))
    total = ski.bin_luminosities(sim)[0] * ski.calibration(sim, sim.frames[0])[0]
    ref = float(golden("image_cfg0.npz")["calib_total"])
    assert abs(total / ref - 1) < 1e-6


def test_multiwavelength_bins_follow_black_body(ski, tmp_path):
    """Bin luminosities: constant bin width 2e-3 lambda_min and the 5800 K black body shape,
    equal to the normalisation at the normalisation wavelength."""
    sim = ski.load(_write(str(tmp_path), "c1", skigen.config1(nlambda=5, lam_range=(0.3, 2.0))))
    w = ski.bin_widths(sim.wavelengths)
    np.testing.assert_allclose(w, 2e-3 * sim.wavelengths[0], rtol=1e-9)
    L = ski.bin_luminosities(sim)
    lam = sim.wavelengths
    h, c, k, T = 6.62606957e-34, 2.99792458e8, 1.3806488e-23, 5800.0
    B = lambda l: 1 / l ** 5 / np.expm1(h * c / (l * k * T))  # noqa: E731
    np.testing.assert_allclose(L / w, 3.839e26 * 1e6 * B(lam) / B(0.55e-6), rtol=1e-12)


def test_fits_round_trip(ski, tmp_path):
    sim = ski.load(_write(str(tmp_path), "c0", skigen.config0()))
    cube = np.random.default_rng(1).random((3, 4, 5))
    p = os.path.join(str(tmp_path), "x.fits")
    ski.write_fits(p, cube, sim.frames[0], np.array([1e-6, 2e-6, 3e-6]))
    back = ski.read_fits(p)
    np.testing.assert_array_equal(back, cube.astype(np.float32))
    assert os.path.getsize(p) % 2880 == 0


def test_unsupported_items_raise(ski, tmp_path):
    spec = skigen.SkiSpec(num_packets=1, grid=("voronoi", 100, "Uniform"))
    with pytest.raises(ski.UnsupportedSki):
        ski.load(_write(str(tmp_path), "v", spec))
    p = _write(str(tmp_path), "f", skigen.SkiSpec(num_packets=1))
    text = open(p).read().replace("<FrameInstrument ", "<FullInstrument ")
    with open(p, "w") as f:
        f.write(text)
    with pytest.raises(ski.UnsupportedSki):
        ski.load(p)


def test_sed_instruments_parse_and_write(ski, tmp_path):
    """SEDInstrument items: parsed with their aperture; the SED file reproduces the reference's
    header, column lines and number format (a reference output kept as data in
    tests/golden/sed_cfg0_sed60.dat), and reads back; the F_nu calibration is
    L / (4 pi d^2) / (bin width) * lambda^2 / c * 1e26."""
    spec = skigen.config0(num_packets=1e5)
    spec.sed_instruments = (("sed60", 60.0, 0.0, 0.0), ("ap60", 60.0, 0.0, 3.0))
    p = os.path.join(str(tmp_path), "s.ski")
    with open(p, "w") as f:
        f.write(skigen.render(spec))
    sim = ski.load(p)
    names = [f.name for f in sim.frames]
    assert names == ["i60", "sed60", "ap60"]
    s0, s1 = sim.frames[1], sim.frames[2]
    assert isinstance(s0, ski.Sed) and s0.radius == 0.0 and abs(s1.radius / (3 * ski.PC) - 1) < 1e-15
    ref = open(os.path.join(GOLDEN, "sed_cfg0_sed60.dat")).read().splitlines()
    out = os.path.join(str(tmp_path), "x_sed.dat")
    v = float(ref[3].split()[1])
    ski.write_sed(out, [v], s0, sim.wavelengths)
    assert open(out).read().splitlines() == ref
    np.testing.assert_array_equal(ski.read_sed(out), [[0.55, v]])
    lam = sim.wavelengths[0]
    want = 1 / (4 * np.pi * s0.distance ** 2) / ski.bin_widths(sim.wavelengths)[0] * lam * lam / 2.99792458e8 * 1e26
    assert abs(ski.sed_calibration(sim, s0)[0] / want - 1) < 1e-15


def test_radiation_field_probe_parse_and_header(ski, tmp_path):
    """RadiationFieldOptions storeRadiationField with a RadiationFieldProbe (PerCellForm):
    parsed; the mean-intensity file's header equals the reference's (tests/golden/rf_cfg0.npz);
    J_nu = rf / (4 pi V dlambda) * lambda^2 / c; other probes are refused."""
    spec = skigen.config0(num_packets=1e5)
    spec.grid = ("cartesian", 16)
    spec.store_rf = True
    p = os.path.join(str(tmp_path), "rf.ski")
    with open(p, "w") as f:
        f.write(skigen.render(spec))
    sim = ski.load(p)
    assert sim.rf_probes == ["rf"]
    out = os.path.join(str(tmp_path), "rf_J.dat")
    ski.write_mean_intensity(out, np.array([[1.5e-26], [2.0e-27]]), sim.wavelengths)
    lines = open(out).read().splitlines()
    assert "\n".join(lines[:3]) == str(golden("rf_cfg0.npz")["header"])
    assert lines[3:] == ["0 1.500000000e-26", "1 2.000000000e-27"]
    sim.cell_volumes = np.array([2.0, 4.0])
    lam = sim.wavelengths[0]
    J = ski.mean_intensity(sim, np.array([[3.0], [3.0]]))
    want = 3.0 / (4 * np.pi * np.array([2.0, 4.0])) / ski.bin_widths(sim.wavelengths)[0] * lam * lam / 2.99792458e8
    np.testing.assert_allclose(J[:, 0], want, rtol=1e-15)
    with open(p, "w") as f:
        f.write(skigen.render(spec).replace("<PerCellForm/>", "<DefaultCutsForm/>"))
    with pytest.raises(ski.UnsupportedSki):
        ski.load(p)


def test_mesh_grids_parse(ski, tmp_path):
    """VoronoiMeshSpatialGrid / TetraMeshSpatialGrid with the DustDensity policy are read; other
    policies and relaxation / refinement are refused."""
    for kind in ("voronoi", "tetra"):
        spec = skigen.config0(num_packets=1e5)
        spec.grid = (kind, 5000, "DustDensity")
        p = os.path.join(str(tmp_path), kind + ".ski")
        text = skigen.render(spec)
        with open(p, "w") as f:
            f.write(text)
        assert ski.load(p).grid == dict(kind=kind, n=5000)
        with open(p, "w") as f:
            f.write(text.replace('policy="DustDensity"', 'policy="Uniform"'))
        with pytest.raises(ski.UnsupportedSki):
            ski.load(p)


def test_spiral_disk_points_follow_the_perturbation():
    """workloads.disk_points with spiral parameters: azimuths accepted with probability
    perturbation / max, so the azimuthal histogram at fixed radius follows the perturbation
    profile (checked against its mean over the arm phase)."""
    load_pkg()
    from skirt9_gpu_amd import workloads as W
    sp = dict(arms=2, pitch=np.radians(15.0), radius_zero=3 * PC, phase_zero=0.0, weight=0.5, index=1)
    h = 10 * PC
    pts = W.disk_points(200000, [-h, -h, -h, h, h, h], seed=5, spiral=sp)
    pts = pts[np.hypot(pts[:, 0], pts[:, 1]) < 9 * PC]  # the box clips larger radii by azimuth
    R = np.hypot(pts[:, 0], pts[:, 1])
    phi = np.arctan2(pts[:, 1], pts[:, 0])
    gam = np.log(R / sp["radius_zero"]) / np.tan(sp["pitch"]) + 0.5 * np.pi / 2
    ph = np.mod(0.5 * 2 * (gam - phi), np.pi)  # arm phase in [0, pi)
    hist, _ = np.histogram(ph, bins=8, range=(0, np.pi))
    a = np.arange(8) * np.pi / 8
    b = a + np.pi / 8
    mean_sin2 = 0.5 - (np.sin(2 * b) - np.sin(2 * a)) / (4 * (b - a))  # mean of sin^2 over each bin
    cn = np.sqrt(np.pi) * 1.0 / (np.sqrt(np.pi) / 2)  # Gamma(2) / Gamma(3/2)
    prof = 0.5 + 0.5 * cn * mean_sin2
    np.testing.assert_allclose(hist / hist.sum(), prof / prof.sum(), rtol=0.03)


CAMERA = ("cam", (10.0, -30.0, 20.0), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), 30.0, 20.0, 100)


def _project(M, p):
    v = np.array([p[0], p[1], p[2], 1.0]) @ M
    return v[0] / v[3], v[1] / v[3], v[2]


def test_perspective_parse_and_transform(ski, tmp_path):
    """PerspectiveInstrument items are read; the viewport transform puts the point source at
    the origin in the pixel where the reference's image of the config-0 workload peaks
    (tests/golden/image_persp.npz, camera off the scene's mirror plane), and for both
    rotation branches (viewing direction away from and near the x axis) maps the crosshair
    to the viewport centre, forward x up to +x, up to +y, with Fe / d / (pixel size) pixels
    per unit length at distance d from the eye; the eye lies one focal length behind the
    viewport origin."""
    spec = skigen.config0(num_packets=1e5)
    spec.persp_instruments = (CAMERA,)
    sim = ski.load(_write(str(tmp_path), "p", spec))
    cam = sim.frames[1]
    assert isinstance(cam, ski.Perspective) and cam.nx == cam.ny == 100
    assert cam.view == (10 * PC, -30 * PC, 20 * PC) and abs(cam.focal / (30 * PC) - 1) < 1e-15
    M, E, s = ski.perspective_setup(cam.view, cam.cross, cam.up, cam.focal, cam.width, cam.nx, cam.ny)
    ref = golden("image_persp.npz")["mean"][0]
    j, i = np.unravel_index(ref.argmax(), ref.shape)
    x, y, zp = _project(M, (0.0, 0.0, 0.0))
    assert (int(x), int(y)) == (i, j) and zp > 0
    for view, cross, up in (((10, -30, 20), (1, 0, 0), (0, 0, 1)), ((30, 1, 0), (0, 0, 0), (0, 0, 1)),
                            ((0.5, 0.2, 30), (0, 0, 0), (0, 1, 0)), ((-30, 0.5, 0.2), (1, 2, 3), (0, 0.3, 1))):
        view, cross, up = np.array(view, float) * PC, np.array(cross, float) * PC, np.array(up, float)
        M, E, s = ski.perspective_setup(view, cross, up, 25 * PC, 20 * PC, 100, 80)
        f = (cross - view) / np.linalg.norm(cross - view)
        np.testing.assert_allclose(E, view - 25 * PC * f, rtol=0, atol=1e-9 * PC)
        r = np.cross(f, up)
        r /= np.linalg.norm(r)
        u = np.cross(r, f)
        d = np.linalg.norm(cross - E)
        mag = 25 * PC / d / s * PC  # pixels per parsec at the crosshair
        x0, y0, zp = _project(M, cross)
        assert abs(x0 - 50) < 1e-9 and abs(y0 - 40) < 1e-9 and abs(zp - (d - 25 * PC)) < 1e-6 * PC
        x1, y1, _ = _project(M, cross + PC * r)
        x2, y2, _ = _project(M, cross + PC * u)
        assert abs(x1 - 50 - mag) < 1e-9 and abs(y1 - 40) < 1e-9
        assert abs(x2 - 50) < 1e-9 and abs(y2 - 40 - mag) < 1e-9


def test_particle_source_parse(ski, tmp_path):
    """ParticleSource with BlackBodySEDFamily: columns in the units of the '# column' header
    lines (positions and h in pc, radius in km, temperature in K); per-bin luminosity
    4 pi^2 R^2 B_lambda(T) times the bin width; unsupported options are refused."""
    rows = skigen.disk_particles()
    skigen.write_particles(os.path.join(str(tmp_path), "parts.txt"), rows)
    spec = skigen.config0(num_packets=1e5)
    spec.particle_file = "parts.txt"
    p = _write(str(tmp_path), "pt", spec)
    sim = ski.load(p)
    P = sim.particles
    assert P["kernel"] == 1 and len(P["xyzh"]) == len(rows)
    np.testing.assert_allclose(P["xyzh"], rows[:, :4] * PC, rtol=1e-9)
    np.testing.assert_allclose(P["radius"], rows[:, 4] * 1e3, rtol=1e-9)
    np.testing.assert_allclose(P["temperature"], rows[:, 5], rtol=1e-9)
    L = ski.particle_bin_luminosities(sim)
    lam, T, R = 0.55e-6, rows[0, 5], rows[0, 4] * 1e3
    h, c, k = 6.62607015e-34, 2.99792458e8, 1.380649e-23
    B = 2 * h * c * c / lam ** 5 / np.expm1(h * c / (lam * k * T))
    assert abs(L[0, 0] / (4 * np.pi ** 2 * R * R * B * 2e-3 * lam) - 1) < 1e-6
    text = open(p).read()
    spec.particle_kernel = "UniformSmoothingKernel"
    assert ski.load(_write(str(tmp_path), "pu", spec)).particles["kernel"] == 0
    for bad in (text.replace('importVelocity="false"', 'importVelocity="true"'),
                text.replace("CubicSplineSmoothingKernel", "ScaledGaussianSmoothingKernel"),
                text.replace("BlackBodySEDFamily", "BruzualCharlotSEDFamily")):
        with open(p, "w") as f:
            f.write(bad)
        with pytest.raises(ski.UnsupportedSki):
            ski.load(p)


@pytest.mark.parametrize("kind", ["voronoi", "tetra"])
def test_mesh_grid_file_policy_parse(ski, tmp_path, kind):
    """VoronoiMeshSpatialGrid / TetraMeshSpatialGrid with policy File: the sites or vertices
    are read from the column file (position x, y, z in the header's units, pc without a
    header line) and the grid keeps them in the file's order."""
    rng = np.random.default_rng(4)
    pts = (rng.random((300, 3)) * 2 - 1) * 9 * PC
    skigen.write_points(os.path.join(str(tmp_path), "pts.txt"), pts)
    spec = skigen.SkiSpec(num_packets=1e4, grid=("tetra_file", "pts.txt"))
    p = _write(str(tmp_path), "mf", spec)
    if kind == "voronoi":
        text = open(p).read().replace("<TetraMeshSpatialGrid", "<VoronoiMeshSpatialGrid").replace(
            'refine="false"', 'relaxSites="false"')
        with open(p, "w") as f:
            f.write(text)
    sim = ski.load(p)
    assert sim.grid["kind"] == kind and sim.grid["n"] == len(pts)
    np.testing.assert_allclose(sim.grid["points"], pts, rtol=1e-15, atol=64.0)
    # without header lines the columns are in pc; a header may give another length unit
    with open(os.path.join(str(tmp_path), "pts.txt"), "w") as f:
        np.savetxt(f, pts / PC, fmt="%.17g")
    np.testing.assert_allclose(ski.load(p).grid["points"], pts, rtol=1e-15, atol=64.0)
    with open(os.path.join(str(tmp_path), "pts.txt"), "w") as f:
        f.write("# column 1: position x (kpc)\n# column 2: position y (kpc)\n# column 3: position z (kpc)\n")
        np.savetxt(f, pts / (1e3 * PC), fmt="%.17g")
    np.testing.assert_allclose(ski.load(p).grid["points"], pts, rtol=1e-14, atol=64.0)


def test_voronoi_mesh_medium_parse(ski, tmp_path):
    """VoronoiMeshMedium on a VoronoiMeshSpatialGrid with policy ImportedMesh: the site and
    density columns in the units of the '# column' header lines (or the defaults pc and 1/cm3),
    a MassDensity column divided by the dust mix's mass per hydrogen atom, massFraction applied;
    the cells take their sites' densities in the mesh's order; unsupported options raise."""
    rng = np.random.default_rng(3)
    n = 400
    sites = (rng.random((n, 3)) * 2 - 1) * 9 * PC
    dens = rng.lognormal(0, 1, n) * 1e6
    skigen.write_voronoi_medium(os.path.join(str(tmp_path), "sites.txt"), sites, dens)
    spec = skigen.SkiSpec(num_packets=1e4, grid=("voronoi_imported",), medium_file="sites.txt")
    p = _write(str(tmp_path), "vm", spec)
    sim = ski.load(p)
    assert sim.grid == dict(kind="voronoi_imported")
    np.testing.assert_allclose(sim.imported["sites"], sites, rtol=1e-15, atol=64.0)
    np.testing.assert_allclose(sim.imported["density"], dens, rtol=1e-15)
    # defaults without header lines: pc and 1/cm3
    with open(os.path.join(str(tmp_path), "sites.txt"), "w") as f:
        np.savetxt(f, np.column_stack([sites / PC, dens]), fmt="%.17g")
    np.testing.assert_allclose(ski.load(p).imported["density"], dens * 1e6, rtol=1e-15)
    text = open(p).read()
    md = text.replace('massType="NumberDensity"', 'massType="MassDensity"').replace('massFraction="1"',
                                                                                  'massFraction="0.5"')
    with open(p, "w") as f:
        f.write(md)
    mu = 1.5e-29
    Msun = 1.9891e30
    np.testing.assert_allclose(ski.load(p).imported["density"], 0.5 * dens * Msun / PC ** 3 / mu, rtol=1e-12)
    # the cells' densities follow the builder's order of the sites
    with open(p, "w") as f:
        f.write(text)
    skigen.write_voronoi_medium(os.path.join(str(tmp_path), "sites.txt"), sites, dens)
    sim = ski.load(p)
    pkg = load_pkg()
    mesh = pkg.Voronoi.build(sim.imported["sites"], sim.box)
    ex = mesh.export()
    mesh.close()
    order = {tuple(s): d for s, d in zip(sim.imported["sites"], sim.imported["density"])}
    want = np.array([order[tuple(s)] for s in ex["sites"]])
    assert not np.array_equal(want, sim.imported["density"])  # the builder reorders (by x)
    np.testing.assert_array_equal(ski.imported_cell_densities(sim.imported["sites"], sim.imported["density"],
                                                               ex["sites"]), want)
    for bad in (text.replace('importVelocity="false"', 'importVelocity="true"'),
                text.replace('policy="ImportedMesh"', 'policy="DustDensity" numSites="100"'),
                text.replace('massType="NumberDensity"', 'massType="Mass"')):
        with open(p, "w") as f:
            f.write(bad)
        with pytest.raises(ski.UnsupportedSki):
            ski.load(p)
