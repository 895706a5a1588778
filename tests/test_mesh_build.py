"""The host-side Voronoi builder (skg_voronoi_build) against the reference's Voro++-built
tables and against what the geometry fixes (CPU only; the library's host code runs without
a GPU)."""

import numpy as np
import pytest

import oracle as O
from helpers import golden, load_pkg

KEYS = ("sites", "nb_begin", "nb_count", "nb_index", "block_n", "block_begin", "block_count", "block_index", "box")


@pytest.fixture(scope="module")
def ref():
    f = golden("voronoi_paths.npz")
    return {k: f[k] for k in f.files}


@pytest.fixture(scope="module")
def built(ref):
    pkg = load_pkg()
    v = pkg.Voronoi.build(ref["sites"], ref["box"], threads=4)
    out = v.export()
    v.close()
    return out


def _sets(t):
    return [set(t["nb_index"][b:b + c].tolist()) for b, c in zip(t["nb_begin"], t["nb_count"])]


def test_voronoi_builder_matches_reference_tables(ref, built):
    """Same cell order, the same neighbour set for every cell (walls included) and the same
    block search lists as VoronoiMeshSnapshot with Voro++; the cells tile the domain."""
    np.testing.assert_array_equal(built["sites"], ref["sites"])
    assert _sets(built) == _sets(ref)
    assert int(built["block_n"]) == int(ref["block_n"])
    np.testing.assert_array_equal(built["block_count"], ref["block_count"])
    np.testing.assert_array_equal(built["block_index"], ref["block_index"])
    vol = np.prod(ref["box"][3:] - ref["box"][:3])
    assert abs(built["volume"].sum() / vol - 1) < 1e-12


def test_voronoi_builder_paths_match_reference(ref, built):
    g = O.Grid("voronoi", **{k: built[k] for k in KEYS})
    off, m, ds = O.trace(g, ref["rays"])
    np.testing.assert_array_equal(off, ref["off"])
    np.testing.assert_array_equal(m, ref["m"])
    np.testing.assert_array_equal(ds, ref["ds"])


def test_voronoi_builder_invariants_random_sites():
    """2e4 uniform sites: neighbour relation mutual, volumes tile the box, the block search
    finds the brute-force nearest site, and the input order is sorted by x."""
    pkg = load_pkg()
    rng = np.random.default_rng(3)
    box = np.array([-2.0, -1.0, -1.5, 2.0, 1.0, 1.5])
    s = box[:3] + (box[3:] - box[:3]) * rng.random((20000, 3))
    v = pkg.Voronoi.build(s, box, threads=4)
    t = v.export()
    v.close()
    assert np.all(np.diff(t["sites"][:, 0]) >= 0)
    np.testing.assert_array_equal(np.sort(t["sites"][:, 0]), np.sort(s[:, 0]))
    sets = _sets(t)
    for i, nb in enumerate(sets):
        for j in nb:
            if j >= 0:
                assert i in sets[j]
    assert 13 < np.mean([len(x) for x in sets]) < 17  # Poisson-Voronoi mean is 15.54 faces
    vol = np.prod(box[3:] - box[:3])
    assert abs(t["volume"].sum() / vol - 1) < 1e-10
    g = O.Grid("voronoi", **{k: t[k] for k in KEYS})
    pts = box[:3] + (box[3:] - box[:3]) * rng.random((3000, 3))
    ci = O.cell_index(g, pts)
    d2 = ((pts[:, None, :] - t["sites"][None, :, :]) ** 2).sum(axis=2)
    np.testing.assert_array_equal(ci, d2.argmin(axis=1))


# ---------------------------------------------------------------- tetrahedra
def _tetra_ref():
    f = golden("tetra_paths.npz")
    return {k: f[k] for k in f.files}


def test_tetra_builder_matches_reference_delaunay():
    """Our Delaunay tetrahedralisation of the reference's vertices equals TetGen's except for
    the tetrahedra with three domain corners: the 4 corners of a box face are cocircular, so
    the face's diagonal is a free (degenerate) choice."""
    pkg = load_pkg()
    d = _tetra_ref()
    V, box = d["verts"], d["box"]
    corner = np.all((V == box[:3]) | (V == box[3:]), axis=1)
    t = pkg.Tetra.build(V[~corner], box)
    e = t.export()
    t.close()
    assert len(e["tet_verts"]) == len(d["tet_verts"])
    np.testing.assert_array_equal(np.sort(e["verts"], axis=0), np.sort(V, axis=0))

    def keyed(W, T):
        out = set()
        for tt in T:
            pts = [tuple(W[i]) for i in tt]
            if sum(all(c == box[0] or c == box[3] for c in p) for p in pts) >= 3:
                continue
            out.add(tuple(sorted(pts)))
        return out

    a, b = keyed(e["verts"], e["tet_verts"]), keyed(V, d["tet_verts"])
    assert a == b and len(a) > 0.99 * len(d["tet_verts"])


def _tetra_checks(e, box):
    W, T, nb = e["verts"], e["tet_verts"], e["face_nb"]
    a, b, c, d = (W[T[:, i]] for i in range(4))
    vol = np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a) / 6
    assert vol.min() > 0
    assert abs(vol.sum() / np.prod(box[3:] - box[:3]) - 1) < 1e-10
    for t in range(len(T)):
        for f in range(4):
            n = nb[t, f]
            if n >= 0:
                assert (nb[n] == t).sum() == 1
    assert (nb < 0).sum() == 12  # the two triangles of each box face
    # outward normals: the vertex opposite face f lies behind it
    for f in range(4):
        cv = [(f + 3) % 4, (f + 2) % 4, (f + 1) % 4]
        s = np.einsum("ij,ij->i", e["face_normal"][:, f], W[T[:, f]] - W[T[:, cv[0]]])
        assert np.all(s < 0)


def test_tetra_builder_invariants_random_points():
    """5e3 random points: positive volumes tiling the box, mutual neighbours, outward normals,
    and the empty-circumsphere property (brute force on a sample of tetrahedra)."""
    pkg = load_pkg()
    rng = np.random.default_rng(8)
    box = np.array([-1.0, -2.0, -0.5, 1.0, 2.0, 0.5])
    pts = box[:3] + (box[3:] - box[:3]) * rng.random((5000, 3))
    t = pkg.Tetra.build(pts, box)
    e = t.export()
    t.close()
    _tetra_checks(e, box)
    W, T = e["verts"], e["tet_verts"]
    for tt in rng.choice(len(T), 200, replace=False):
        P = W[T[tt]]
        A = 2 * (P[1:] - P[0])
        rhs = (P[1:] ** 2).sum(1) - (P[0] ** 2).sum()
        cen = np.linalg.solve(A, rhs)
        r2 = ((P[0] - cen) ** 2).sum()
        d2 = ((W - cen) ** 2).sum(1)
        d2[T[tt]] = np.inf
        assert d2.min() > r2 * (1 - 1e-5)  # exact on the predicate lattice (step 2^-23 of the box)


def test_tetra_builder_traversal_invariants():
    """Paths through our mesh (oracle generator): segment midpoints inside their tetrahedron,
    and cellIndex agrees with a brute-force containment search."""
    pkg = load_pkg()
    d = _tetra_ref()
    V, box = d["verts"], d["box"]
    corner = np.all((V == box[:3]) | (V == box[3:]), axis=1)
    t = pkg.Tetra.build(V[~corner], box)
    e = t.export()
    t.close()
    g = O.Grid("tetra", verts=e["verts"], tet_verts=e["tet_verts"], face_nb=e["face_nb"], face_normal=e["face_normal"],
               box=box)
    rays = d["rays"][:300]
    off, m, ds = O.trace(g, rays)
    W, T = e["verts"], e["tet_verts"]
    for i in range(len(rays)):
        r, k = rays[i, :3], rays[i, 3:]
        s = 0.0
        for q in range(off[i], off[i + 1]):
            mid = r + k * (s + 0.5 * ds[q])
            s += ds[q]
            if m[q] < 0:
                continue
            A = np.column_stack([W[T[m[q], j]] - W[T[m[q], 0]] for j in (1, 2, 3)])
            lam = np.linalg.solve(A, mid - W[T[m[q], 0]])
            assert lam.min() > -1e-9 and lam.sum() < 1 + 1e-9
    ci = O.cell_index(g, d["points"][:300])
    for p, c in zip(d["points"][:300], ci):
        A = np.column_stack([W[T[c, j]] - W[T[c, 0]] for j in (1, 2, 3)])
        lam = np.linalg.solve(A, p - W[T[c, 0]])
        assert lam.min() > -1e-9 and lam.sum() < 1 + 1e-9


def test_tetra_density_sampler_is_uniform_in_the_tetra():
    """Cell average of a steep exponential disk over single tetrahedra: the library's folded
    barycentric sampler against rejection sampling from the bounding box (an independent
    uniform sampler), to 1 % (both estimates have ~0.2 % MC error)."""
    pkg = load_pkg()
    pc = pkg.PC
    md = pkg.disk_model(tau_v=1.0, hR=3 * pc, hz=0.3 * pc)
    norm = pkg.disk_density(md, 0.0, 0.0, 0.0) * 1.0  # density at the origin = nnorm * rho0
    rng = np.random.default_rng(5)
    for V in (np.array([[0.1, 0.2, 0.05], [2.0, 0.3, 0.1], [0.5, 1.8, 0.2], [0.4, 0.6, 1.5]]) * pc,
              np.array([[-1.0, -1.0, -0.9], [1.5, -0.5, -0.2], [0.0, 2.0, -0.6], [0.3, 0.1, 0.8]]) * pc):
        got = pkg.sample_density_tetra(V, np.array([[0, 1, 2, 3]]), md, num_samples=200000, seed=3)[0] / norm
        lo, hi = V.min(0), V.max(0)
        T = np.column_stack([V[1] - V[0], V[2] - V[0], V[3] - V[0]])
        acc = []
        while sum(len(a) for a in acc) < 400000:
            p = lo + (hi - lo) * rng.random((1000000, 3))
            lam = np.linalg.solve(T, (p - V[0]).T).T
            inside = (lam.min(1) >= 0) & (lam.sum(1) <= 1)
            acc.append(p[inside])
        p = np.concatenate(acc)
        ref = np.mean(np.exp(-np.hypot(p[:, 0], p[:, 1]) / (3 * pc) - np.abs(p[:, 2]) / (0.3 * pc)))
        assert abs(got / ref - 1) < 0.01, (got, ref)
