"""Batched independent instances in one persistent launch (skg_instances) on a tetrahedral
grid: every instance's frames equal the FP64 oracle's single run with that instance's seed,
optical-depth scale, launch luminosities and instruments (bars as the other life-cycle parity
tests: total 1e-6, median pixel 1e-9, 99 % of pixels within 1e-6)."""

import numpy as np
import pytest

import oracle as O
from helpers import PC, golden, load_pkg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    pkg = load_pkg()
    e = pkg.Engine(0)
    yield e
    e.close()


def _frame(inc_deg, npix=60):
    return dict(inclination=np.radians(inc_deg), azimuth=0.3, roll=0.0, nx=npix, ny=npix, fovx=20 * PC, fovy=20 * PC)


def test_batched_instances_match_single_runs(eng):
    pkg = load_pkg()
    f = golden("tetra_paths.npz")
    d = {k: f[k] for k in f.files}
    eng.grid_tetra(d["verts"], d["tet_verts"], d["face_nb"], d["face_normal"], d["box"])
    eng.medium(d["density"])
    s = float(d["sext"][0])
    lam = np.array([0.55e-6, 1.0e-6])
    sext = np.array([s, 0.5 * s])
    ssca = np.array([0.6 * s, 0.25 * s])
    gpar = np.array([0.6, 0.4])
    eng.dust(lam, sext, ssca, gpar)
    eng.source_point((0.0, 0.0, 0.0), [1.0, 1.0])
    eng.options(0.5, 1e4, 0, 0, 99)
    eng.tune_pipeline(pkg.SKG_MODE_PERSISTENT, 1 << 18, 0, 8)
    # frames: instance 1 owns a same-observer pair; instance 2's only frame has the same
    # observer as the frame before it (the per-instance observer cache must start afresh)
    incs = [0.0, 30.0, 30.0, 30.0, 60.0, 90.0, 45.0]
    frames = [_frame(i) for i in incs]
    eng.frame_remove_all()
    for fr in frames:
        eng.frame_add(fr["inclination"], fr["azimuth"], fr["roll"], fr["nx"], fr["ny"], fr["fovx"], fr["fovy"])
    packets = np.array([4000, 9000, 2500, 0, 6000, 3000], np.uint64)
    seeds = np.array([11, 12, 13, 14, 15, 2**40 + 16], np.uint64)
    scale = np.array([0.3, 1.0, 3.0, 1.0, 0.1, 10.0])
    fbeg = np.array([0, 1, 3, 4, 5, 6], np.int32)
    fcnt = np.array([1, 2, 1, 1, 1, 1], np.int32)
    lp = np.array([[1.0, 0.5], [2.0, 1.0], [1.0, 3.0], [1.0, 1.0], [0.7, 0.2], [1.0, 1.0]]) / np.maximum(packets, 1)[:, None]
    total = eng.instances(packets, seeds, scale, lp, fbeg, fcnt)
    assert total == int(packets.sum())
    eng.frame_clear()
    print("case batched instances: tetra fixture, 6 instances,", total, "histories", flush=True)
    eng.run_packets(0, total)
    cnt = eng.counters()
    assert cnt[0] == total
    g = O.Grid("tetra", **d)
    for i in range(len(packets)):
        for q in range(fcnt[i]):
            img = eng.frame_read(int(fbeg[i] + q))
            if packets[i] == 0:
                assert not img.any()
                continue
        if packets[i] == 0:
            continue
        ofr, _ = O.simulate(g, d["density"], lam, sext * scale[i], ssca * scale[i], gpar, lp[i],
                            frames[fbeg[i]:fbeg[i] + fcnt[i]], 0, int(packets[i]), threads=8, seed=int(seeds[i]))
        for q in range(fcnt[i]):
            img = eng.frame_read(int(fbeg[i] + q))
            oimg = ofr[q]
            assert abs(img.sum() / oimg.sum() - 1) < 1e-6, (i, q)
            nz = oimg > 0
            rel = np.abs(img[nz] - oimg[nz]) / oimg[nz]
            assert np.median(rel) < 1e-9, (i, q)
            assert np.mean(rel < 1e-6) > 0.99, (i, q)
    # the pipeline engine refuses instances; removing them restores the plain run
    eng.tune_pipeline(pkg.SKG_MODE_PIPELINE, 1 << 18, 0, 8)
    with pytest.raises(pkg.EngineError):
        eng.run_packets(0, 100)
    eng.instances([])
    eng.run_packets(0, 100)
