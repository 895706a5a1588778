"""Imported particle sources on the GPU (skg_source_particles, persistent engine) against the
FP64 oracle on the same histories: 60 particles with both smoothing kernels in the dusty
Cartesian disk, a frame and an SED; bars as for the point source (1e-9 relative per pixel).
The launch preparation (history ranges, per-packet luminosities, the spline table) is computed
independently on both sides.  The pipeline engine and unsupported combinations are refused,
and n = 0 restores the point source."""

import numpy as np
import pytest

import oracle as O
from helpers import PC, disk_density_cartesian, load_pkg

pytestmark = pytest.mark.gpu


def _particles(n=60, seed=6):
    rng = np.random.default_rng(seed)
    xyzh = np.column_stack([rng.uniform(-7, 7, (n, 2)) * PC, rng.uniform(-1, 1, n) * PC, rng.uniform(0.3, 2.0, n) * PC])
    lbin = rng.uniform(0.1, 3.0, (n, 2))
    lbin[5] = 0.0  # a particle without luminosity launches nothing
    return xyzh, lbin, lbin.sum(1)


@pytest.mark.parametrize("kernel", [1, 0])
def test_particle_source_matches_oracle(kernel):
    pkg = load_pkg()
    eng = pkg.Engine(0)
    b, dens = disk_density_cartesian(n=32, samples=10, seed=5, tau_v=2.0)
    eng.grid_cartesian(b, b, b)
    eng.medium(dens)
    s = 1.5e-29
    lam, sext, ssca, g = [0.55e-6, 0.8e-6], [s, 0.7 * s], [0.6 * s, 0.35 * s], [0.6, 0.5]
    eng.dust(lam, sext, ssca, g)
    instr = [dict(inclination=np.radians(60.0), azimuth=0.0, roll=0.0, nx=40, ny=40, fovx=20 * PC, fovy=20 * PC),
             dict(sed=True, inclination=np.radians(30.0), azimuth=0.5, radius=0.0)]
    eng.frame_remove_all()
    eng.frame_add(instr[0]["inclination"], 0.0, 0.0, 40, 40, 20 * PC, 20 * PC)
    eng.sed_add(instr[1]["inclination"], instr[1]["azimuth"], 0.0)
    npk = 30000
    xyzh, lbin, ltot = _particles()
    eng.source_point((0.0, 0.0, 0.0), [1.0 / npk, 1.0 / npk])
    eng.source_particles(xyzh, lbin, ltot, npk, source_bias=0.5, kernel=kernel)
    eng.options(0.5, 1e4, 0, 0, 41)
    eng.tune_pipeline(pkg.SKG_MODE_PERSISTENT, 1 << 18, 0, 8)
    eng.frame_clear()
    print("case particle source, kernel", kernel, flush=True)
    eng.run_packets(0, npk)
    grid = O.Grid("cartesian", xb=b, yb=b, zb=b)
    ofr, ocnt = O.simulate(grid, dens, lam, sext, ssca, g, [1.0 / npk, 1.0 / npk], instr, 0, npk, threads=8, seed=41,
                           particles=dict(xyzh=xyzh, lbin=lbin, ltot=ltot, num_packets=npk, source_bias=0.5,
                                          kernel=kernel))
    for i in range(2):
        v, ov = eng.frame_read(i), ofr[i]
        assert ov.sum() > 0
        np.testing.assert_allclose(v, ov, rtol=1e-9, atol=1e-12 * ov.max(), err_msg=f"instrument {i}")
    cnt = eng.counters()
    Iv, _, _ = O.particle_launch(ltot, lbin, npk, 0.5)
    # the oracle counts no history for the dark particle's range; the engine counts every launch
    assert cnt[0] == npk and ocnt[0] == npk - (Iv[6] - Iv[5])
    assert abs(int(cnt[3]) - int(ocnt[3])) <= 0.001 * ocnt[3]  # scattering events

    print("case particle source refused by the pipeline engine", flush=True)
    eng.tune_pipeline(pkg.SKG_MODE_PIPELINE, 1 << 18, 0, 8)
    with pytest.raises(pkg.EngineError):
        eng.run_packets(0, 1000)
    with pytest.raises(pkg.EngineError):
        eng.source_particles(xyzh[:2], lbin[:2] * np.nan, ltot[:2], npk)
    with pytest.raises(pkg.EngineError):
        eng.source_particles(xyzh[:2], lbin[:2], ltot[:2], npk, kernel=2)

    print("case particle luminosities made for another number of wavelengths", flush=True)
    eng.tune_pipeline(pkg.SKG_MODE_PERSISTENT, 1 << 18, 0, 8)
    eng.dust(lam + [1.2e-6], sext + [0.5 * s], ssca + [0.2 * s], g + [0.4])
    eng.source_point((0.0, 0.0, 0.0), [1.0 / npk] * 3)
    with pytest.raises(pkg.EngineError, match="wavelengths"):
        eng.run_packets(0, 1000)
    eng.dust(lam, sext, ssca, g)
    eng.source_point((0.0, 0.0, 0.0), [1.0 / npk, 1.0 / npk])

    print("case point source restored", flush=True)
    eng.source_particles(np.zeros((0, 4)), np.zeros((0, 2)), np.zeros(0), npk)
    eng.tune_pipeline(pkg.SKG_MODE_PERSISTENT, 1 << 18, 0, 8)
    eng.frame_clear()
    eng.run_packets(0, 5000)
    pfr, _ = O.simulate(grid, dens, lam, sext, ssca, g, [1.0 / npk, 1.0 / npk], instr, 0, 5000, threads=8, seed=41)
    np.testing.assert_allclose(eng.frame_read(0), pfr[0], rtol=1e-9, atol=1e-12 * pfr[0].max())
    eng.close()


@pytest.mark.parametrize("kernel", ["CubicSplineSmoothingKernel", "UniformSmoothingKernel"])
def test_config0_particle_source_matches_reference(tmp_path, kernel):
    """The config-0 disk lit by 30 black-body particles (ParticleSource read from a text file)
    through the ski front end, against 8 reference runs (tests/golden/image_particles.npz): total
    within 4 sigma, 5x5-binned pooled chi-square per degree of freedom < 1.5."""
    import os
    import sys
    from helpers import ROOT, golden
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import skigen
    pkg = load_pkg()
    from skirt9_gpu_amd import ski as S
    skigen.write_particles(os.path.join(str(tmp_path), "parts.txt"), skigen.disk_particles())
    spec = skigen.config0(num_packets=1e6)
    spec.particle_file = "parts.txt"
    spec.particle_kernel = kernel
    spec.num_pixels = 100
    path = os.path.join(str(tmp_path), "pt.ski")
    with open(path, "w") as f:
        f.write(skigen.render(spec))
    sim = S.load(path)
    assert sim.particles is not None and len(sim.particles["xyzh"]) == 30
    eng = pkg.Engine(0)
    nb = 6
    imgs = []
    for bi in range(nb):
        sim.seed = 11000 + bi
        print("case ski particle source", kernel, bi, flush=True)
        out = S.run(sim, str(tmp_path), f"b{bi}", engine=eng, log=lambda *a: None)
        imgs.append(out["i60"][0])
    eng.close()
    f = golden("image_particles.npz")
    ref, rsem = f[kernel + "_mean"][0].astype(float), f[kernel + "_sem"][0].astype(float)
    imgs = np.array(imgs)
    nref = int(f["runs"])
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
    assert m.sum() > 100
    assert np.mean(z ** 2) < 1.5, np.mean(z ** 2)
    assert np.mean(np.abs(z) > 4) < 0.01


@pytest.mark.parametrize("extra", ["rf", "persp", "rf+persp"])
def test_particle_source_with_rf_or_perspective_matches_oracle(extra):
    """Particles together with radiation-field storage, a perspective instrument, or both: the
    field per cell and every image within 1e-9 of the oracle's."""
    pkg = load_pkg()
    from skirt9_gpu_amd import ski as S
    eng = pkg.Engine(0)
    b, dens = disk_density_cartesian(n=24, samples=10, seed=5, tau_v=2.0)
    eng.grid_cartesian(b, b, b)
    eng.medium(dens)
    s = 1.5e-29
    lam, sext, ssca, g = [0.55e-6, 0.8e-6], [s, 0.7 * s], [0.6 * s, 0.35 * s], [0.6, 0.5]
    eng.dust(lam, sext, ssca, g)
    instr = [dict(inclination=np.radians(60.0), azimuth=0.0, roll=0.0, nx=40, ny=40, fovx=20 * PC, fovy=20 * PC)]
    if "persp" in extra:
        M, E, ps = S.perspective_setup(np.array([2.0, -5.0, 1.0]) * PC, np.zeros(3), np.array([0.0, 0.0, 1.0]),
                                       3 * PC, 8 * PC, 50, 50)
        instr.append(dict(persp=True, M=M, eye=E, psize=ps, nx=50, ny=50))
    eng.frame_remove_all()
    eng.frame_add(instr[0]["inclination"], 0.0, 0.0, 40, 40, 20 * PC, 20 * PC)
    if "persp" in extra:
        eng.perspective_add(instr[1]["M"], instr[1]["eye"], instr[1]["psize"], 50, 50)
    npk = 20000
    xyzh, lbin, ltot = _particles(n=25, seed=9)
    eng.source_point((0.0, 0.0, 0.0), [0.0, 0.0])
    eng.source_particles(xyzh, lbin, ltot, npk, source_bias=0.5, kernel=1)
    eng.options(0.5, 1e4, 0, 0, 43)
    eng.tune_pipeline(pkg.SKG_MODE_PERSISTENT, 1 << 18, 0, 8)
    if "rf" in extra:
        eng.rf_enable(True)
    eng.frame_clear()
    print("case particle source with", extra, flush=True)
    eng.run_packets(0, npk)
    grid = O.Grid("cartesian", xb=b, yb=b, zb=b)
    res = O.simulate(grid, dens, lam, sext, ssca, g, [0.0, 0.0], instr, 0, npk, threads=8, seed=43,
                     store_rf="rf" in extra,
                     particles=dict(xyzh=xyzh, lbin=lbin, ltot=ltot, num_packets=npk, source_bias=0.5, kernel=1))
    ofr = res[0]
    for i in range(len(instr)):
        v, ov = eng.frame_read(i), ofr[i]
        assert ov.sum() > 0
        np.testing.assert_allclose(v, ov, rtol=1e-9, atol=1e-12 * ov.max(), err_msg=f"instrument {i}")
    if "rf" in extra:
        rf, orf = eng.rf_read(), res[2]
        nz = orf > 0
        np.testing.assert_array_equal(rf > 0, nz)
        assert np.max(np.abs(rf[nz] - orf[nz]) / orf[nz]) < 1e-9
        eng.rf_enable(False)
    eng.close()


def test_autotune_with_particles_and_rf_keeps_state():
    """With a particle source and radiation-field storage the autotuner tries the persistent
    engine only (the pipeline engine runs neither) and restores the frame, the stored field and
    the counters after its trials."""
    pkg = load_pkg()
    eng = pkg.Engine(0)
    b, dens = disk_density_cartesian(n=16, samples=4, seed=5, tau_v=1.0)
    eng.grid_cartesian(b, b, b)
    eng.medium(dens)
    s = 1.5e-29
    eng.dust([0.55e-6], [s], [0.6 * s], [0.6])
    eng.frame_remove_all()
    eng.frame_add(np.radians(60.0), 0.0, 0.0, 30, 30, 20 * PC, 20 * PC)
    xyzh, lbin, ltot = _particles(n=10, seed=3)
    eng.source_point((0.0, 0.0, 0.0), [0.0])
    eng.source_particles(xyzh, lbin[:, :1], ltot, 4000, kernel=1)
    eng.options(0.5, 1e4, 0, 0, 47)
    eng.tune_pipeline(pkg.SKG_MODE_PERSISTENT, 1 << 18, 0, 8)
    eng.rf_enable(True)
    eng.frame_clear()
    eng.run_packets(0, 4000)
    img, rf, cnt = eng.frame_read(0), eng.rf_read(), eng.counters()
    print("case autotune with particles and rf", flush=True)
    cfg, rate = eng.autotune(10_000_000, trial_packets=20000)
    assert rate > 0 and cfg["mode"] == pkg.SKG_MODE_PERSISTENT
    np.testing.assert_array_equal(eng.frame_read(0), img)
    np.testing.assert_array_equal(eng.rf_read(), rf)
    np.testing.assert_array_equal(eng.counters(), cnt)
    eng.rf_enable(False)
    eng.close()
