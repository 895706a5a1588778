"""Host-side medium set-up of the engine (no GPU): the octree with its neighbour lists, the
dust tables and the density model, against the reference's golden vectors and closed forms."""

import numpy as np

from helpers import PC, cart_borders, child_first, golden, load_pkg


def test_octree_from_reference_topology_is_identical():
    """Replaying the reference grid's subdivision flags gives the same node ids, bounds, cells
    and (sorted) neighbour list of every leaf, bit for bit."""
    pkg = load_pkg()
    t = golden("tree_paths.npz")
    tr = pkg.Octree.from_flags(t["bounds"][0], (t["child_count"] > 0).astype(np.uint8))
    ex = tr.export()
    np.testing.assert_array_equal(ex["bounds"], t["bounds"])
    cf = child_first(t["child_begin"], t["child_count"], t["child_index"])
    np.testing.assert_array_equal(ex["child_first"], cf)
    np.testing.assert_array_equal(ex["cell_index"], t["cell_index"])
    eo, ei, ro, ri = ex["nb_off"], ex["nb_id"], t["nb_off"], t["nb_id"]
    for l in np.nonzero(cf < 0)[0]:
        for w in range(6):
            q = 6 * l + w
            assert np.array_equal(ei[eo[q]:eo[q + 1]], ri[ro[q]:ro[q + 1]]), (l, w)


def test_density_refined_octree_matches_reference_statistics():
    """Same policy as the reference fixture (spiral disk, levels 2..5, max dust fraction 3e-4):
    the leaf count agrees within 3 %, and the build is independent of the thread count."""
    pkg = load_pkg()
    t = golden("tree_paths.npz")
    m = pkg.disk_model(spiral=True)
    a = pkg.Octree.build(t["bounds"][0], m, 2, 5, 3e-4, 100, seed=3, threads=1)
    b = pkg.Octree.build(t["bounds"][0], m, 2, 5, 3e-4, 100, seed=3, threads=8)
    ref_leaves = int((t["child_count"] == 0).sum())
    assert abs(a.info()[1] / ref_leaves - 1) < 0.03
    ea, eb = a.export(), b.export()
    for k in ("bounds", "child_first", "cell_index", "nb_off", "nb_id"):
        np.testing.assert_array_equal(ea[k], eb[k])


def test_dust_tables_match_reference():
    pkg = load_pkg()
    d = golden("dust.npz")
    r = lambda a: np.array([float(f"{x:.9g}") for x in a])  # values as written in the ski file
    sext, ssca, g = pkg.dust_meanlist(r(d["in_lam_um"]) * 1e-6, r(d["in_kappa"]), r(d["in_albedo"]), r(d["in_g"]),
                                      d["d10_lam"])
    np.testing.assert_array_equal(sext, d["d10_sext"])
    np.testing.assert_array_equal(ssca, d["d10_ssca"])
    np.testing.assert_array_equal(g, d["d10_g"])
    i40 = d["d40_in"]
    sext, ssca, g = pkg.dust_meanlist(i40[:, 0] * 1e-6, i40[:, 1], i40[:, 2], i40[:, 3], d["d40_lam"])
    np.testing.assert_array_equal(sext, d["d40_sext"])
    np.testing.assert_array_equal(ssca, d["d40_ssca"])
    np.testing.assert_array_equal(g, d["d40_g"])


def test_disk_density_closed_form():
    pkg = load_pkg()
    m = pkg.disk_model(tau_v=2.0, sext_v=3e-29)
    hR, hz = 3 * PC, 0.3 * PC
    rho0 = 1 / (4 * np.pi * hR * hR * hz)
    nnorm = 2.0 / 3e-29 / (2 * rho0 * hz)
    for (x, y, z) in [(0.1 * PC, 0.2 * PC, 0.05 * PC), (4 * PC, -1 * PC, -0.7 * PC)]:
        ref = nnorm * rho0 * np.exp(-np.hypot(x, y) / hR) * np.exp(-abs(z) / hz)
        assert abs(pkg.disk_density(m, x, y, z) / ref - 1) < 1e-14
    # spiral perturbation averages to one over azimuth (cn normalisation)
    ms = pkg.disk_model(spiral=True)
    phis = np.linspace(0, 2 * np.pi, 4001)[:-1]
    R = 2 * PC
    v = np.array([pkg.disk_density(ms, R * np.cos(p), R * np.sin(p), 0.0) for p in phis])
    assert abs(v.mean() / pkg.disk_density(m, R, 0, 0) * (2.0 / 3e-29) / (1.0 / 1.5e-29) - 1) < 1e-9


def test_sampled_column_gives_face_on_tau():
    """Cell-averaged densities reproduce the face-on optical depth along Z through the centre."""
    pkg = load_pkg()
    n = 64
    b = cart_borders(n, -10 * PC, 10 * PC)
    m = pkg.disk_model(tau_v=1.0)
    d = pkg.sample_density_cartesian(b, b, b, m, num_samples=200, seed=5).reshape(n, n, n)
    # four cells around the axis, integrated along z
    dz = b[1] - b[0]
    col = 0.25 * (d[31, 31] + d[31, 32] + d[32, 31] + d[32, 32]).sum() * dz * 1.5e-29
    # the four cells average exp(-R/hR) over |x|,|y| < 0.3125 pc: ~0.917 of the on-axis value
    hR = 3 * PC
    xs = np.linspace(0, b[33], 401)
    X, Y = np.meshgrid(xs, xs)
    f = np.exp(-np.hypot(X, Y) / hR).mean()
    assert abs(col / f - 1) < 0.02
