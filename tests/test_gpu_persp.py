"""Perspective instruments on the GPU: the engine against the FP64 oracle on the same histories
(a frame, a camera outside the domain, the same camera again, and a camera whose eye lies inside
the dust, so the optical depth toward it stops at the eye) on a Cartesian disk and on the
reference's octree, Voronoi and tetrahedral fixture grids, and the ski front end's FITS output
against the reference's images of the config-0 workload (mean and standard error over 8 runs of
1e6 packets, tests/golden/image_persp.npz): total within 4 sigma, 5x5-binned pooled chi-square
per degree of freedom < 1.5."""

import os
import sys

import numpy as np
import pytest

import oracle as O
from helpers import PC, ROOT, child_first, disk_density_cartesian, golden, load_pkg

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import skigen  # noqa: E402

pytestmark = pytest.mark.gpu

CAMERA = ("cam", (10.0, -30.0, 20.0), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), 30.0, 20.0, 100)


def _cameras(S):
    out = []
    for view, cross, up, fe, w, n in (((10, -30, 20), (1, 0, 0), (0, 0, 1), 30, 20, 60),
                                      ((10, -30, 20), (1, 0, 0), (0, 0, 1), 30, 20, 60),
                                      ((2, -5, 1), (0, 0, 0), (0, 0, 1), 3, 8, 50)):
        M, E, s = S.perspective_setup(np.array(view, float) * PC, np.array(cross, float) * PC, np.array(up, float),
                                      fe * PC, w * PC, n, n)
        out.append(dict(persp=True, M=M, eye=E, psize=s, nx=n, ny=n))
    return out


def test_perspective_instruments_match_oracle():
    pkg = load_pkg()
    from skirt9_gpu_amd import ski as S
    eng = pkg.Engine(0)
    b, dens = disk_density_cartesian(n=32, samples=10, seed=5, tau_v=2.0)
    eng.grid_cartesian(b, b, b)
    eng.medium(dens)
    s = 1.5e-29
    lam, sext, ssca, g = [0.55e-6, 0.8e-6], [s, 0.7 * s], [0.6 * s, 0.35 * s], [0.6, 0.5]
    eng.dust(lam, sext, ssca, g)
    instr = [dict(inclination=np.radians(60.0), azimuth=0.0, roll=0.0, nx=40, ny=40, fovx=20 * PC, fovy=20 * PC)]
    instr += _cameras(S)
    assert np.linalg.norm(instr[3]["eye"]) < 10 * PC  # the third camera's eye is inside the domain
    eng.frame_remove_all()
    for d in instr:
        if d.get("persp"):
            eng.perspective_add(d["M"], d["eye"], d["psize"], d["nx"], d["ny"])
        else:
            eng.frame_add(d["inclination"], d["azimuth"], d["roll"], d["nx"], d["ny"], d["fovx"], d["fovy"])
    npk = 40000
    lpack = [1.0 / npk, 1.0 / npk]
    eng.source_point((0.0, 0.0, 0.0), lpack)
    eng.options(0.5, 1e4, 0, 0, 23)
    grid = O.Grid("cartesian", xb=b, yb=b, zb=b)
    ofr, _ = O.simulate(grid, dens, lam, sext, ssca, g, lpack, instr, 0, npk, threads=8, seed=23)
    for mode in (pkg.SKG_MODE_PERSISTENT, pkg.SKG_MODE_PIPELINE):
        eng.tune_pipeline(mode, 1 << 18, 0, 8)
        eng.frame_clear()
        print("case perspective instruments, engine mode", mode, flush=True)
        eng.run_packets(0, npk)
        for i in range(4):
            v, ov = eng.frame_read(i), ofr[i]
            assert ov.sum() > 0
            np.testing.assert_allclose(v, ov, rtol=1e-9, atol=1e-12 * ov.max(), err_msg=f"mode {mode} instrument {i}")
    np.testing.assert_array_equal(ofr[1], ofr[2])
    # host-side argument checks: nothing is added
    d = instr[1]
    bad_eye = np.array(d["eye"], dtype=float)
    bad_eye[1] = np.nan
    bad_M = np.array(d["M"], dtype=float)
    bad_M[3, 3] = np.inf
    for args in ((d["M"], bad_eye, d["psize"], 10, 10), (bad_M, d["eye"], d["psize"], 10, 10),
                 (d["M"], d["eye"], 0.0, 10, 10), (d["M"], d["eye"], d["psize"], 0, 10),
                 (d["M"], d["eye"], d["psize"], 65536, 65536)):
        with pytest.raises(pkg.EngineError):
            eng.perspective_add(*args)
    assert len(eng.frames) == 4
    eng.close()


def _fixture_grid(eng, kind):
    f = golden(("tree" if kind == "octree" else kind) + "_paths.npz")
    d = {k: f[k] for k in f.files}
    if kind == "octree":
        cf = child_first(d["child_begin"], d["child_count"], d["child_index"])
        eng.grid_octree(d["bounds"], cf, d["cell_index"], d["nb_off"], d["nb_id"])
        g = O.Grid("octree", bounds=d["bounds"], child_first=cf, cell_index=d["cell_index"], nb_off=d["nb_off"],
                   nb_id=d["nb_id"])
    elif kind == "voronoi":
        eng.grid_voronoi(d["sites"], d["nb_begin"], d["nb_count"], d["nb_index"], d["block_n"], d["block_begin"],
                         d["block_count"], d["block_index"], d["box"])
        g = O.Grid(kind, **d)
    else:
        eng.grid_tetra(d["verts"], d["tet_verts"], d["face_nb"], d["face_normal"], d["box"])
        g = O.Grid(kind, **d)
    return g, d["density"], float(d["sext"][0])


@pytest.mark.parametrize("mode", [1, 0], ids=["persistent", "pipeline"])
@pytest.mark.parametrize("kind", ["octree", "voronoi", "tetra"])
def test_perspective_instruments_match_oracle_on_fixture_grids(kind, mode):
    pkg = load_pkg()
    from skirt9_gpu_amd import ski as S
    eng = pkg.Engine(0)
    grid, dens, s = _fixture_grid(eng, kind)
    eng.medium(dens)
    lam, sext, ssca, g = [0.55e-6, 0.9e-6], [s, 0.6 * s], [0.6 * s, 0.3 * s], [0.6, 0.4]
    eng.dust(lam, sext, ssca, g)
    instr = [dict(inclination=np.radians(60.0), azimuth=0.0, roll=0.0, nx=40, ny=40, fovx=20 * PC, fovy=20 * PC)]
    instr += _cameras(S)
    eng.frame_remove_all()
    for d in instr:
        if d.get("persp"):
            eng.perspective_add(d["M"], d["eye"], d["psize"], d["nx"], d["ny"])
        else:
            eng.frame_add(d["inclination"], d["azimuth"], d["roll"], d["nx"], d["ny"], d["fovx"], d["fovy"])
    npk = 20000
    lpack = [1.0 / npk, 2.0 / npk]
    eng.source_point((0.0, 0.0, 0.0), lpack)
    eng.options(0.5, 1e4, 0, 0, 29)
    ofr, _ = O.simulate(grid, dens, lam, sext, ssca, g, lpack, instr, 0, npk, threads=8, seed=29)
    eng.tune_pipeline(mode, 1 << 18, 0, 8)
    eng.frame_clear()
    print("case perspective instruments on", kind, "engine mode", mode, flush=True)
    eng.run_packets(0, npk)
    for i in range(4):
        v, ov = eng.frame_read(i), ofr[i]
        assert ov.sum() > 0
        np.testing.assert_allclose(v, ov, rtol=1e-9, atol=1e-12 * ov.max(), err_msg=f"{kind} mode {mode} instrument {i}")
    eng.close()


def test_config0_perspective_matches_reference(tmp_path):
    pkg = load_pkg()
    from skirt9_gpu_amd import ski as S
    spec = skigen.config0(num_packets=1e6)
    spec.persp_instruments = (CAMERA,)
    path = os.path.join(str(tmp_path), "cfg0.ski")
    with open(path, "w") as f:
        f.write(skigen.render(spec))
    sim = S.load(path)
    eng = pkg.Engine(0)
    nb = 6
    imgs = []
    for bi in range(nb):
        sim.seed = 9000 + bi
        print("case ski perspective batch", bi, flush=True)
        out = S.run(sim, str(tmp_path), f"b{bi}", engine=eng, log=lambda *a: None)
        imgs.append(out["cam"])
    eng.close()
    back = S.read_fits(os.path.join(str(tmp_path), "b0_cam_total.fits"))
    np.testing.assert_allclose(back, imgs[0].astype(np.float32), rtol=1e-6)
    imgs = np.array(imgs)[:, 0]
    f = golden("image_persp.npz")
    ref, rsem = f["mean"][0].astype(float), f["sem"][0].astype(float)
    nref = int(f["runs"])
    # the total within 4 sigma, the 5x5-binned image by a pooled chi-square per degree of freedom
    tot = imgs.sum(axis=(1, 2))
    ref_run_var_tot = (rsem ** 2).sum() * nref  # an underestimate of the run variance of the total (pixels correlate)
    pooled_tot = ((nref - 1) * ref_run_var_tot + (nb - 1) * tot.var(ddof=1)) / (nref + nb - 2)
    assert abs(tot.mean() - ref.sum()) < 4 * np.sqrt(pooled_tot * (1 / nb + 1 / nref)), (tot.mean(), ref.sum())

    def binned(a, B=5):
        n = a.shape[-1] // B
        return a.reshape(*a.shape[:-2], n, B, n, B).sum(axis=(-1, -3))

    oi = binned(imgs)
    var_ref_run = binned((rsem ** 2)[None])[0] * nref
    pooled = ((nref - 1) * var_ref_run + (nb - 1) * oi.var(0, ddof=1)) / (nref + nb - 2)
    var = pooled * (1 / nb + 1 / nref)
    m = var > 0
    z = (oi.mean(0) - binned(ref[None])[0])[m] / np.sqrt(var[m])
    print("bins", int(m.sum()), "chi2/dof", float(np.mean(z ** 2)), "max |z|", float(np.abs(z).max()), flush=True)
    assert m.sum() > 60  # the disk covers about a fifth of the viewport
    assert np.mean(z ** 2) < 1.5, np.mean(z ** 2)
    assert np.abs(z).max() < 4.5
