"""GPU parity on the benchmark configurations' workloads at fixture size, through the C-ABI:

* config 1 -- the reference's spiral-disk octree (tests/golden/tree_paths.npz, 4320 leaves,
  face-on tau_V ~ 0.87) with the 41-wavelength power-law MeanListDustMix of the config
  (kappa ~ lambda^-1.5, albedo 0.5 -> 0.1, tests/golden/dust.npz) scaled to the fixture's
  V-band cross section: both engines in FP64 against the FP64 oracle on the same histories
  (total 1e-6, per-pixel median 1e-9), and in FP32 + Kahan, whose median per-pixel error
  against the FP64 oracle must be <= 1 % (printed; expected ~1e-7);
* config 2 -- the reference's 1500-site Voronoi mesh (tests/golden/voronoi_paths.npz) filled
  with a lognormal clumpy medium (P(k) ~ k^-3.7, sigma_ln = 1, seeded, sampled at the sites,
  mean tau_V = 5 along an axis), three frames at 0/45/90 degrees plus a second instrument of the
  45-degree observer (shared optical depth): both engines, FP64 against the oracle, FP32 bulk
  within 1 % per-pixel median;
* config 4 -- guided sampling on the fixture octree at face-on tau_V = 50 with the config's
  5 wavelengths 0.3-2 um: records from plain packets, Adam on the GPU, then the guided
  persistent engine equals the guided pipeline on the same histories, the guided total
  equals the plain total within Monte Carlo error, and the figure of merit of guided versus
  plain is printed with its 95 % interval beside the variance ceiling of any direction guide
  (two continuations of the same first scatterings).

References: TreeSpatialGrid.cpp:143-228, VoronoiMeshSnapshot.cpp:1089-1219,
MediumSystem.cpp:1785-1844 (optical depth), DustMix.cpp:47-150 (tables),
DistantInstrument.cpp:55-63 (same-observer sharing), MonteCarloSimulation.cpp:2034-2092."""

import os
import sys
import time

import numpy as np
import pytest

import guide_oracle as G
import oracle as O
from helpers import PC, ROOT, child_first, golden, load_pkg

pytestmark = pytest.mark.gpu

MU_DUST = 1.5e-29
TAU_FIXTURE_TREE = 0.8693823069720505  # face-on optical depth of the tree fixture at its sext


@pytest.fixture(scope="module")
def eng():
    pkg = load_pkg()
    e = pkg.Engine(0)
    yield e
    e.close()


def _tree(eng):
    d = golden("tree_paths.npz")
    cf = child_first(d["child_begin"], d["child_count"], d["child_index"])
    eng.grid_octree(d["bounds"], cf, d["cell_index"], d["nb_off"], d["nb_id"])
    g = O.Grid("octree", bounds=d["bounds"], child_first=cf, cell_index=d["cell_index"], nb_off=d["nb_off"],
               nb_id=d["nb_id"])
    return g, d["density"].copy(), float(d["sext"][0])


def _powerlaw_mix(lam, v_section):
    """The config's MeanListDustMix (kappa 1000 -> 1 m2/kg, albedo 0.5 -> 0.1 over 0.1-10 um,
    g = 0.6), resampled to lam by the engine's host code, scaled so sigma(0.55 um) = v_section."""
    pkg = load_pkg()
    tab = np.array([[0.1e-6, 1000.0, 0.5, 0.6], [10e-6, 1.0, 0.1, 0.6]])
    sext, ssca, g = pkg.dust_meanlist(tab[:, 0], tab[:, 1], tab[:, 2], tab[:, 3], lam, MU_DUST)
    sv, _, _ = pkg.dust_meanlist(tab[:, 0], tab[:, 1], tab[:, 2], tab[:, 3], [0.55e-6], MU_DUST)
    f = v_section / sv[0]
    return sext * f, ssca * f, g


def _frames(eng, frs):
    eng.frame_remove_all()
    for fr in frs:
        eng.frame_add(fr["inclination"], fr["azimuth"], fr.get("roll", 0.0), fr["nx"], fr["ny"], fr["fovx"], fr["fovy"])


def _frame(inc_deg, n=60, fov=20 * PC, az_deg=0.0):
    return dict(inclination=np.radians(inc_deg), azimuth=np.radians(az_deg), roll=0.0, nx=n, ny=n, fovx=fov, fovy=fov)


def _pixel_errors(img, ref):
    nz = ref > 0
    rel = np.abs(img[nz] - ref[nz]) / ref[nz]
    return np.median(rel), np.percentile(rel, 90), abs(img.sum() / ref.sum() - 1)


def _run(eng, mode, fp, npk, seed, nfr):
    eng.options(0.5, 1e4, 0, fp, seed)
    eng.tune_pipeline(mode, 1 << 18, 0, 8)
    eng.frame_clear()
    eng.run_packets(0, npk)
    return [eng.frame_read(i) for i in range(nfr)], eng.counters()


def test_config1_octree_40_wavelengths_fp64_and_fp32_kahan(eng):
    pkg = load_pkg()
    dust = golden("dust.npz")
    grid, dens, sv = _tree(eng)
    lam = dust["d40_lam"]
    sext, ssca, g = _powerlaw_mix(lam, sv)
    nl = len(lam)
    eng.medium(dens)
    eng.dust(lam, sext, ssca, g)
    frs = [_frame(60.0)]
    _frames(eng, frs)
    npk = 40000
    lpack = np.full(nl, 1.0 / npk)
    eng.source_point((0.0, 0.0, 0.0), lpack)
    ofr, ocnt = O.simulate(grid, dens, lam, sext, ssca, g, lpack, frs, 0, npk, threads=8, seed=43)
    ref = ofr[0]
    assert ref.shape == (nl, 60, 60) and np.all(ref.sum(axis=(1, 2)) > 0)
    for mode in (pkg.SKG_MODE_PIPELINE, pkg.SKG_MODE_PERSISTENT):
        print("case config 1 octree, 41 wavelengths, FP64, engine mode", mode, flush=True)
        (img,), cnt = _run(eng, mode, pkg.SKG_FP64, npk, 43, 1)
        assert cnt[0] == npk
        med, p90, tot = _pixel_errors(img, ref)
        print(f"  FP64 vs oracle: median {med:.2e}, p90 {p90:.2e}, total {tot:.2e}")
        assert tot < 1e-6 and med < 1e-9
        print("case config 1 octree, 41 wavelengths, FP32 + Kahan, engine mode", mode, flush=True)
        (img32,), _ = _run(eng, mode, pkg.SKG_FP32_KAHAN, npk, 43, 1)
        med, p90, tot = _pixel_errors(img32, ref)
        print(f"  FP32 + Kahan vs FP64 oracle: median {med:.2e}, p90 {p90:.2e}, total {tot:.2e}")
        assert med <= 0.01 and tot < 1e-3
        # per wavelength: every bin's image within the same bar
        for ell in (0, nl // 2, nl - 1):
            m_l, _, t_l = _pixel_errors(img32[ell], ref[ell])
            assert m_l <= 0.01 and t_l < 1e-2, (ell, m_l, t_l)


def _lognormal_at_sites(sites, box, n=32, seed=4):
    sys.path.insert(0, os.path.join(ROOT, "skirt9-gpu_amd"))
    from workloads import lognormal_field
    rho = lognormal_field(n, -3.7, 1.0, seed)
    lo, hi = box[:3], box[3:]
    ijk = np.clip(((sites - lo) / (hi - lo) * n).astype(int), 0, n - 1)
    return rho[ijk[:, 0], ijk[:, 1], ijk[:, 2]]


def test_config2_voronoi_lognormal_three_frames(eng):
    pkg = load_pkg()
    d = golden("voronoi_paths.npz")
    dd = {k: d[k] for k in d.files}
    eng.grid_voronoi(dd["sites"], dd["nb_begin"], dd["nb_count"], dd["nb_index"], dd["block_n"], dd["block_begin"],
                     dd["block_count"], dd["block_index"], dd["box"])
    grid = O.Grid("voronoi", **dd)
    box = dd["box"]
    rho = _lognormal_at_sites(dd["sites"], box)
    assert rho.std() / rho.mean() > 0.5  # a clumpy medium
    s = 1.5e-29
    n0 = 5.0 / (s * (box[5] - box[2]))  # mean tau_V = 5 along z for <rho> = 1
    dens = n0 * rho
    eng.medium(dens)
    lam, sext, ssca, g = [0.55e-6], [s], [0.6 * s], [0.6]
    eng.dust(lam, sext, ssca, g)
    frs = [_frame(0.0), _frame(45.0), _frame(90.0), _frame(45.0, n=30, fov=12 * PC)]
    _frames(eng, frs)
    npk = 30000
    eng.source_point((0.0, 0.0, 0.0), [1.0 / npk])
    ofr, _ = O.simulate(grid, dens, lam, sext, ssca, g, [1.0 / npk], frs, 0, npk, threads=8, seed=47)
    for mode in (pkg.SKG_MODE_PIPELINE, pkg.SKG_MODE_PERSISTENT):
        print("case config 2 Voronoi lognormal, FP64, engine mode", mode, flush=True)
        imgs, cnt = _run(eng, mode, pkg.SKG_FP64, npk, 47, len(frs))
        assert cnt[0] == npk
        for i, (img, ref) in enumerate(zip(imgs, ofr)):
            assert ref.sum() > 0
            med, p90, tot = _pixel_errors(img, ref)
            print(f"  frame {i} FP64 vs oracle: median {med:.2e}, total {tot:.2e}")
            assert tot < 1e-6 and med < 1e-9, i
        print("case config 2 Voronoi lognormal, FP32 bulk, engine mode", mode, flush=True)
        imgs32, _ = _run(eng, mode, pkg.SKG_FP32_KAHAN, npk, 47, len(frs))
        for i, (img, ref) in enumerate(zip(imgs32, ofr)):
            med, p90, tot = _pixel_errors(img, ref)
            print(f"  frame {i} FP32 bulk vs FP64 oracle: median {med:.2e}, p90 {p90:.2e}, total {tot:.2e}")
            assert med <= 0.01 and tot < 1e-3, i


def test_config4_guided_octree_tau50(eng):
    pkg = load_pkg()
    sys.path.insert(0, os.path.join(ROOT, "skirt9-gpu_amd"))
    from workloads import log_wavelengths
    grid, dens, sv = _tree(eng)
    lam = log_wavelengths(0.3, 2.0, 5)
    sext, ssca, g = _powerlaw_mix(lam, sv * 50.0 / TAU_FIXTURE_TREE)  # face-on tau_V = 50
    eng.medium(dens)
    eng.dust(lam, sext, ssca, g)
    _frames(eng, [_frame(60.0, n=40)])
    per = 50000
    eng.source_point((0.0, 0.0, 0.0), np.full(5, 1.0 / per))
    eng.options(0.5, 1e4, 0, pkg.SKG_FP64, 53)
    # records from plain packets, then Adam on the GPU
    eng.tune_pipeline(pkg.SKG_MODE_PIPELINE, 1 << 18, 0, 8)
    eng.guide_params(G.flatten(G.init_params(5)).astype(np.float32))
    eng.guide_collect(2_000_000)
    eng.guide_mode(2, half_size=float(golden("tree_paths.npz")["half"]))
    eng.frame_clear()
    print("case config 4 collect", flush=True)
    eng.run_packets(1 << 40, 200000)
    nrec = eng.guide_records()
    assert nrec > 200000
    print("case config 4 train", flush=True)
    l0, l1 = eng.guide_train(400, batch=4096, lr=2e-3)
    print(f"  {nrec} records, loss {l0:.4f} -> {l1:.4f}")
    assert l1 < l0
    # the guided engines on the same histories
    eng.guide_mode(1, mix=0.5, half_size=float(golden("tree_paths.npz")["half"]))
    imgs = []
    for mode in (pkg.SKG_MODE_PIPELINE, pkg.SKG_MODE_PERSISTENT):
        eng.tune_pipeline(mode, 1 << 18, 0, 8)
        eng.frame_clear()
        print("case config 4 guided, engine mode", mode, flush=True)
        eng.run_packets(0, per)
        imgs.append(eng.frame_read(0))
    a, c = imgs
    assert abs(c.sum() / a.sum() - 1) < 1e-9
    nz = a > 0
    assert np.max(np.abs(c[nz] - a[nz]) / a[nz]) < 1e-9
    # unbiased: the guided total equals the plain total within Monte Carlo error; the figure of
    # merit 1/(relative variance x time) of each sampler with a 95 % interval for their ratio
    # (the ratio of two batch variances follows an F distribution)
    from scipy import stats
    nb = 16
    tot, secs = {}, {}
    for gm in (0, 1):
        eng.guide_mode(gm, mix=0.5, half_size=float(golden("tree_paths.npz")["half"]))
        t = []
        t0 = time.perf_counter()
        for b in range(nb):
            eng.frame_clear()
            eng.run_packets(b * per, per)
            t.append(eng.frame_read(0).sum())
        secs[gm] = time.perf_counter() - t0
        tot[gm] = np.array(t)
    # the variance left once the first scattering is fixed: two continuations per batch
    cont = []
    eng.guide_mode(0)
    for key in (1, 2):
        eng.continuation_key(key)
        t = []
        for b in range(nb):
            eng.frame_clear()
            eng.run_packets(b * per, per)
            t.append(eng.frame_read(0).sum())
        cont.append(np.array(t))
    eng.continuation_key(0)
    m0, m1 = tot[0].mean(), tot[1].mean()
    v0, v1 = tot[0].var(ddof=1), tot[1].var(ddof=1)
    e0, e1 = np.sqrt(v0 / nb), np.sqrt(v1 / nb)
    print(f"  plain {m0:.6e} +- {e0:.1e}, guided {m1:.6e} +- {e1:.1e}")
    assert abs(m1 - m0) < 4 * np.hypot(e0, e1)
    fom = {gm: 1.0 / (tot[gm].var(ddof=1) / tot[gm].mean() ** 2 * secs[gm]) for gm in (0, 1)}
    gain = fom[1] / fom[0]
    lo, hi = gain / stats.f.ppf(0.975, nb - 1, nb - 1), gain / stats.f.ppf(0.025, nb - 1, nb - 1)
    ca, cb = cont
    ceiling = 0.5 * (ca.var(ddof=1) + cb.var(ddof=1)) / np.cov(ca, cb)[0, 1]
    print(f"  FOM plain {fom[0]:.3e}, guided {fom[1]:.3e}: gain {gain:.2f} (95 % {lo:.2f}-{hi:.2f}); "
          f"variance ceiling of any direction guide {ceiling:.2f}")
    assert np.isfinite(gain) and gain > 0
