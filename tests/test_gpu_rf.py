"""Radiation-field storage on the GPU (skg_rf_enable / skg_rf_read, persistent engine) against
the FP64 oracle on the same histories, with scattering: Cartesian disk, and the reference's
octree, Voronoi and tetrahedral fixture grids.  Bars: the same cells hold radiation field, each
within 1e-9 relative (FP64 atomics add in another order), the total within 1e-12; the
frames are unchanged by storing the field."""

import numpy as np
import pytest

import oracle as O
from helpers import child_first, disk_density_cartesian, frame60, golden, load_pkg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    pkg = load_pkg()
    e = pkg.Engine(0)
    yield e
    e.close()


def _setup(eng, kind):
    if kind == "cartesian":
        b, dens = disk_density_cartesian(n=24, samples=10, seed=5, tau_v=3.0)
        eng.grid_cartesian(b, b, b)
        g = O.Grid("cartesian", xb=b, yb=b, zb=b)
        s = 1.5e-29
    elif kind == "octree":
        d = golden("tree_paths.npz")
        cf = child_first(d["child_begin"], d["child_count"], d["child_index"])
        eng.grid_octree(d["bounds"], cf, d["cell_index"], d["nb_off"], d["nb_id"])
        g = O.Grid("octree", bounds=d["bounds"], child_first=cf, cell_index=d["cell_index"], nb_off=d["nb_off"],
                   nb_id=d["nb_id"])
        dens = d["density"]
        s = float(d["sext"][0])
    else:
        f = golden(f"{kind}_paths.npz")
        d = {k: f[k] for k in f.files}
        if kind == "voronoi":
            eng.grid_voronoi(d["sites"], d["nb_begin"], d["nb_count"], d["nb_index"], d["block_n"], d["block_begin"],
                             d["block_count"], d["block_index"], d["box"])
        else:
            eng.grid_tetra(d["verts"], d["tet_verts"], d["face_nb"], d["face_normal"], d["box"])
        g = O.Grid(kind, **d)
        dens = d["density"]
        s = float(d["sext"][0])
    return g, dens, s


@pytest.mark.parametrize("kind", ["cartesian", "octree", "voronoi", "tetra"])
def test_rf_matches_oracle(eng, kind):
    pkg = load_pkg()
    g, dens, s = _setup(eng, kind)
    eng.medium(dens)
    lam = [0.55e-6, 0.9e-6]
    sext, ssca, gpar = [s, 0.6 * s], [0.6 * s, 0.3 * s], [0.6, 0.4]
    eng.dust(lam, sext, ssca, gpar)
    eng.frame_remove_all()
    fr = frame60(npix=40)
    eng.frame_add(fr["inclination"], fr["azimuth"], fr["roll"], fr["nx"], fr["ny"], fr["fovx"], fr["fovy"])
    npk = 20000
    eng.source_point((0.0, 0.0, 0.0), [1.0 / npk, 2.0 / npk])
    eng.options(0.5, 1e4, 0, 0, 31)
    eng.tune_pipeline(pkg.SKG_MODE_PERSISTENT, 1 << 18, 0, 8)
    eng.frame_clear()
    eng.run_packets(0, npk)
    img_plain = eng.frame_read(0)
    eng.rf_enable(True)
    eng.frame_clear()
    print("case rf", kind, flush=True)
    eng.run_packets(0, npk)
    rf = eng.rf_read()
    img = eng.frame_read(0)
    eng.rf_enable(False)
    _, cnt, orf = O.simulate(g, dens, lam, sext, ssca, gpar, [1.0 / npk, 2.0 / npk], [fr], 0, npk, threads=8, seed=31,
                             store_rf=True)
    assert rf.shape == orf.shape
    np.testing.assert_array_equal(rf > 0, orf > 0)
    nz = orf > 0
    assert nz.sum() > 0.2 * nz.size or kind != "cartesian"
    assert np.max(np.abs(rf[nz] - orf[nz]) / orf[nz]) < 1e-9
    assert abs(rf.sum() / orf.sum() - 1) < 1e-12
    np.testing.assert_allclose(img, img_plain, rtol=1e-12, atol=0)


def test_rf_refused_where_not_supported(eng):
    pkg = load_pkg()
    g, dens, s = _setup(eng, "cartesian")
    eng.medium(dens)
    eng.dust([0.55e-6], [s], [0.6 * s], [0.6])
    eng.source_point((0.0, 0.0, 0.0), [1.0])
    eng.rf_enable(True)
    eng.tune_pipeline(pkg.SKG_MODE_PIPELINE, 1 << 18, 0, 8)
    with pytest.raises(pkg.EngineError):
        eng.run_packets(0, 100)
    eng.tune_pipeline(pkg.SKG_MODE_PERSISTENT, 1 << 18, 0, 8)
    eng.dust([0.55e-6, 0.9e-6], [s, s], [0.6 * s, 0.6 * s], [0.6, 0.6])  # the table no longer fits
    eng.source_point((0.0, 0.0, 0.0), [1.0, 1.0])
    with pytest.raises(pkg.EngineError):
        eng.run_packets(0, 100)
    eng.rf_enable(False)
    eng.run_packets(0, 100)


@pytest.mark.parametrize("kind", ["cartesian", "tetra"])
def test_rf_fp32_kahan_close_to_oracle(eng, kind):
    """The FP32 + Kahan variant: optical depths summed in FP32 with compensation, so the field
    differs from the FP64 oracle's by their rounding only: total within 1e-6, 99 % of the cells
    within 1e-4."""
    pkg = load_pkg()
    g, dens, s = _setup(eng, kind)
    eng.medium(dens)
    lam = [0.55e-6]
    eng.dust(lam, [s], [0.6 * s], [0.6])
    eng.frame_remove_all()
    fr = frame60(npix=20)
    eng.frame_add(fr["inclination"], fr["azimuth"], fr["roll"], fr["nx"], fr["ny"], fr["fovx"], fr["fovy"])
    npk = 20000
    eng.source_point((0.0, 0.0, 0.0), [1.0 / npk])
    eng.options(0.5, 1e4, 0, pkg.SKG_FP32_KAHAN, 41)
    eng.tune_pipeline(pkg.SKG_MODE_PERSISTENT, 1 << 18, 0, 8)
    eng.rf_enable(True)
    eng.frame_clear()
    print("case rf fp32", kind, flush=True)
    eng.run_packets(0, npk)
    rf = eng.rf_read()
    eng.rf_enable(False)
    eng.options(0.5, 1e4, 0, pkg.SKG_FP64, 41)
    _, _, orf = O.simulate(g, dens, lam, [s], [0.6 * s], [0.6], [1.0 / npk], [fr], 0, npk, threads=8, seed=41,
                           store_rf=True)
    assert abs(rf.sum() / orf.sum() - 1) < 1e-6
    nz = orf > 0
    rel = np.abs(rf[nz] - orf[nz]) / orf[nz]
    assert np.mean(rel < 1e-4) > 0.99
