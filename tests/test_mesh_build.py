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
