"""SED instruments on the GPU: the engine against the FP64 oracle on the same histories (a
frame and three SED instruments, two sharing the frame's observer direction), and the ski
front end's SED output against the reference's own SED files for the config-0 workload
(mean and standard error over 8 runs of 1e6 packets; within 4 sigma)."""

import os
import sys

import numpy as np
import pytest

import oracle as O
from helpers import PC, ROOT, child_first, disk_density_cartesian, golden, load_pkg

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import skigen  # noqa: E402

pytestmark = pytest.mark.gpu


def test_sed_instruments_match_oracle():
    pkg = load_pkg()
    eng = pkg.Engine(0)
    b, dens = disk_density_cartesian(n=32, samples=10, seed=5, tau_v=2.0)
    eng.grid_cartesian(b, b, b)
    eng.medium(dens)
    s = 1.5e-29
    eng.dust([0.55e-6, 0.8e-6], [s, 0.7 * s], [0.6 * s, 0.35 * s], [0.6, 0.5])
    inc = np.radians(60.0)
    instr = [dict(inclination=inc, azimuth=0.0, roll=0.0, nx=40, ny=40, fovx=20 * PC, fovy=20 * PC),
             dict(sed=True, inclination=inc, azimuth=0.0, radius=0.0),
             dict(sed=True, inclination=inc, azimuth=0.0, radius=3 * PC),
             dict(sed=True, inclination=np.radians(20.0), azimuth=0.4, radius=1 * PC)]
    eng.frame_remove_all()
    for d in instr:
        if d.get("sed"):
            eng.sed_add(d["inclination"], d["azimuth"], d["radius"])
        else:
            eng.frame_add(d["inclination"], d["azimuth"], d["roll"], d["nx"], d["ny"], d["fovx"], d["fovy"])
    npk = 40000
    eng.source_point((0.0, 0.0, 0.0), [1.0 / npk, 1.0 / npk])
    eng.options(0.5, 1e4, 0, 0, 23)
    g = O.Grid("cartesian", xb=b, yb=b, zb=b)
    ofr, _ = O.simulate(g, dens, [0.55e-6, 0.8e-6], [s, 0.7 * s], [0.6 * s, 0.35 * s], [0.6, 0.5],
                        [1.0 / npk, 1.0 / npk], instr, 0, npk, threads=8, seed=23)
    for mode in (pkg.SKG_MODE_PIPELINE, pkg.SKG_MODE_PERSISTENT):
        eng.tune_pipeline(mode, 1 << 18, 0, 8)
        eng.frame_clear()
        print("case sed instruments", mode, flush=True)
        eng.run_packets(0, npk)
        for i in range(1, 4):
            v, ov = eng.frame_read(i)[:, 0, 0], ofr[i][:, 0, 0]
            assert np.all(ov > 0)
            np.testing.assert_allclose(v, ov, rtol=1e-9, atol=0)
    eng.close()
    # the SED without aperture holds every deposit toward its observer: the frame's total
    # (positions outside its field are not recorded) is a lower bound
    assert ofr[1][:, 0, 0].sum() >= ofr[0].sum() * (1 - 1e-12) and ofr[2][:, 0, 0].sum() < ofr[1][:, 0, 0].sum()


@pytest.mark.parametrize("kind", ["octree", "voronoi", "tetra"])
def test_sed_instruments_match_oracle_on_fixture_grids(kind):
    """The same instruments on the reference's octree, Voronoi and tetrahedral fixture grids
    (persistent engine): each SED within 1e-9 of the oracle's."""
    pkg = load_pkg()
    eng = pkg.Engine(0)
    d = golden(("tree" if kind == "octree" else kind) + "_paths.npz")
    d = {k: d[k] for k in d.files}
    if kind == "octree":
        cf = child_first(d["child_begin"], d["child_count"], d["child_index"])
        eng.grid_octree(d["bounds"], cf, d["cell_index"], d["nb_off"], d["nb_id"])
        grid = O.Grid("octree", bounds=d["bounds"], child_first=cf, cell_index=d["cell_index"], nb_off=d["nb_off"],
                      nb_id=d["nb_id"])
    elif kind == "voronoi":
        eng.grid_voronoi(d["sites"], d["nb_begin"], d["nb_count"], d["nb_index"], d["block_n"], d["block_begin"],
                         d["block_count"], d["block_index"], d["box"])
        grid = O.Grid(kind, **d)
    else:
        eng.grid_tetra(d["verts"], d["tet_verts"], d["face_nb"], d["face_normal"], d["box"])
        grid = O.Grid(kind, **d)
    dens, s = d["density"], float(d["sext"][0])
    eng.medium(dens)
    lam, sext, ssca, g = [0.55e-6, 0.9e-6], [s, 0.6 * s], [0.6 * s, 0.3 * s], [0.6, 0.4]
    eng.dust(lam, sext, ssca, g)
    inc = np.radians(60.0)
    instr = [dict(inclination=inc, azimuth=0.0, roll=0.0, nx=40, ny=40, fovx=20 * PC, fovy=20 * PC),
             dict(sed=True, inclination=inc, azimuth=0.0, radius=0.0),
             dict(sed=True, inclination=inc, azimuth=0.0, radius=3 * PC),
             dict(sed=True, inclination=np.radians(20.0), azimuth=0.4, radius=1 * PC)]
    eng.frame_remove_all()
    for q in instr:
        if q.get("sed"):
            eng.sed_add(q["inclination"], q["azimuth"], q["radius"])
        else:
            eng.frame_add(q["inclination"], q["azimuth"], q["roll"], q["nx"], q["ny"], q["fovx"], q["fovy"])
    npk = 20000
    lpack = [1.0 / npk, 2.0 / npk]
    eng.source_point((0.0, 0.0, 0.0), lpack)
    eng.options(0.5, 1e4, 0, 0, 37)
    ofr, _ = O.simulate(grid, dens, lam, sext, ssca, g, lpack, instr, 0, npk, threads=8, seed=37)
    eng.tune_pipeline(pkg.SKG_MODE_PERSISTENT, 1 << 18, 0, 8)
    eng.frame_clear()
    print("case sed instruments on", kind, flush=True)
    eng.run_packets(0, npk)
    for i in range(4):
        v, ov = eng.frame_read(i), ofr[i]
        assert ov.sum() > 0
        np.testing.assert_allclose(v, ov, rtol=1e-9, atol=1e-12 * ov.max(), err_msg=f"{kind} instrument {i}")
    eng.close()


def test_config0_sed_matches_reference(tmp_path):
    pkg = load_pkg()
    from skirt9_gpu_amd import ski as S
    spec = skigen.config0(num_packets=1e6)
    spec.sed_instruments = (("sed60", 60.0, 0.0, 0.0), ("ap60", 60.0, 0.0, 3.0))
    path = os.path.join(str(tmp_path), "cfg0.ski")
    with open(path, "w") as f:
        f.write(skigen.render(spec))
    sim = S.load(path)
    eng = pkg.Engine(0)
    nb = 6
    vals = {"sed60": [], "ap60": []}
    for bi in range(nb):
        sim.seed = 3000 + bi
        print("case ski sed batch", bi, flush=True)
        out = S.run(sim, str(tmp_path), f"b{bi}", engine=eng, log=lambda *a: None)
        for n in vals:
            vals[n].append(out[n])
    eng.close()
    back = S.read_sed(os.path.join(str(tmp_path), "b0_sed60_sed.dat"))
    assert abs(back[0, 0] - 0.55) < 1e-12 and abs(back[0, 1] / vals["sed60"][0][0] - 1) < 1e-9
    f = golden("sed_cfg0.npz")
    for n, v in vals.items():
        v = np.array(v)[:, 0]
        sem = np.hypot(v.std(ddof=1) / np.sqrt(nb), float(f[n + "_sem"][0]))
        assert abs(v.mean() - float(f[n + "_mean"][0])) < 4 * sem, (n, v.mean(), float(f[n + "_mean"][0]), sem)
