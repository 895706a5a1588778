"""Voronoi and tetrahedral grids on the GPU, through the C-ABI: parity with the reference
golden vectors (VoronoiMeshSpatialGrid / TetraMeshSpatialGrid tables and paths dumped from
the reference build) and with the FP64 CPU oracle for the full life cycle.

Bars: bit-exact cell indices and FP64 segment lengths; FP64 optical depths equal to the
reference's; FP32 Kahan optical depths within 1e-5 relative; frames of the same histories
agree with the oracle to 1e-6 in total and 1e-9 in per-pixel median.
"""

import numpy as np
import pytest

import oracle as O
from helpers import frame60, golden, load_pkg

pytestmark = pytest.mark.gpu

KINDS = ["voronoi", "tetra"]


@pytest.fixture(scope="module")
def eng():
    pkg = load_pkg()
    e = pkg.Engine(0)
    yield e
    e.close()


_F = {}


def _fixture(kind):
    if kind not in _F:
        f = golden(f"{kind}_paths.npz")
        _F[kind] = {k: f[k] for k in f.files}
    return _F[kind]


def _upload(eng, kind):
    d = _fixture(kind)
    if kind == "voronoi":
        eng.grid_voronoi(d["sites"], d["nb_begin"], d["nb_count"], d["nb_index"], d["block_n"], d["block_begin"],
                         d["block_count"], d["block_index"], d["box"])
    else:
        eng.grid_tetra(d["verts"], d["tet_verts"], d["face_nb"], d["face_normal"], d["box"])
    assert eng.num_cells() == len(d["density"])
    return d


@pytest.mark.parametrize("kind", KINDS)
def test_mesh_paths_bit_exact(eng, kind):
    d = _upload(eng, kind)
    print("case mesh paths", kind, flush=True)
    off, m, ds = eng.trace_paths(d["rays"])
    np.testing.assert_array_equal(off, d["off"])
    np.testing.assert_array_equal(m, d["m"])
    np.testing.assert_array_equal(ds, d["ds"])


@pytest.mark.parametrize("kind", KINDS)
def test_mesh_tau(eng, kind):
    d = _upload(eng, kind)
    eng.medium(d["density"])
    s = float(d["sext"][0])
    eng.dust([0.55e-6], [s], [0.6 * s], [0.6])
    print("case mesh tau", kind, flush=True)
    tp, to, cum = eng.trace_tau(d["rays"], toff=d["toff"])
    np.testing.assert_array_equal(cum, d["tau"])
    np.testing.assert_array_equal(to, d["tauobs"])
    tp32, to32 = eng.trace_tau(d["rays"], fp_mode=1)
    np.testing.assert_allclose(tp32, tp, rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(to32, to, rtol=1e-5, atol=1e-12)


@pytest.mark.parametrize("mode", [0, 1], ids=["pipeline", "persistent"])
@pytest.mark.parametrize("kind", KINDS)
def test_mesh_lifecycle_matches_oracle(eng, kind, mode):
    d = _upload(eng, kind)
    eng.medium(d["density"])
    s = float(d["sext"][0])
    eng.dust([0.55e-6], [s], [0.6 * s], [0.6])
    eng.frame_remove_all()
    fr = frame60(npix=100)
    eng.frame_add(fr["inclination"], fr["azimuth"], fr["roll"], fr["nx"], fr["ny"], fr["fovx"], fr["fovy"])
    npk = 20000
    eng.source_point((0.0, 0.0, 0.0), [1.0 / npk])
    eng.options(0.5, 1e4, 0, 0, 13)
    eng.tune_pipeline(mode, 1 << 18, 0, 8)
    eng.frame_clear()
    print("case mesh lifecycle", kind, mode, flush=True)
    eng.run_packets(0, npk)
    img = eng.frame_read(0)[0]
    cnt = eng.counters()
    g = O.Grid(kind, **d)
    ofr, ocnt = O.simulate(g, d["density"], [0.55e-6], [s], [0.6 * s], [0.6], [1.0 / npk], [fr], 0, npk, threads=8,
                           seed=13)
    oimg = ofr[0][0]
    assert cnt[0] == ocnt[0] == npk
    assert abs(int(cnt[2]) - int(ocnt[2])) <= 0.001 * ocnt[2]
    assert abs(img.sum() / oimg.sum() - 1) < 1e-6
    nz = oimg > 0
    rel = np.abs(img[nz] - oimg[nz]) / oimg[nz]
    assert np.median(rel) < 1e-9
    assert np.mean(rel < 1e-6) > 0.99


def test_built_meshes_trace_like_the_oracle(eng):
    """Meshes built by the library's host builders (skg_voronoi_build, skg_tetra_build) and
    uploaded from their handles: GPU paths bit-exact against the oracle on the exported tables."""
    pkg = load_pkg()
    rng = np.random.default_rng(21)
    box = np.array([-1.0, -2.0, -0.5, 1.0, 2.0, 0.5])
    pts = box[:3] + (box[3:] - box[:3]) * rng.random((3000, 3))
    k = rng.normal(size=(2000, 3))
    k /= np.linalg.norm(k, axis=1, keepdims=True)
    rays = np.hstack([box[:3] + (box[3:] - box[:3]) * rng.random((2000, 3)), k])
    for kind in KINDS:
        b = (pkg.Voronoi.build(pts, box, threads=4) if kind == "voronoi" else pkg.Tetra.build(pts, box))
        t = b.export()
        b.upload(eng)
        b.close()
        print("case built mesh paths", kind, flush=True)
        off, m, ds = eng.trace_paths(rays)
        if kind == "voronoi":
            g = O.Grid(kind, **{q: t[q] for q in ("sites", "nb_begin", "nb_count", "nb_index", "block_n", "block_begin",
                                                 "block_count", "block_index", "box")})
        else:
            g = O.Grid(kind, **t)
        ooff, om, ods = O.trace(g, rays)
        np.testing.assert_array_equal(off, ooff)
        np.testing.assert_array_equal(m, om)
        np.testing.assert_array_equal(ds, ods)
