"""Parity of the HIP engine (through the C-ABI) with the reference golden vectors and the
FP64 CPU oracle.

Bars: bit-exact cell indices and FP64 segment lengths for identical input rays (the engine
is compiled without FMA contraction); FP64 optical depths equal to the reference's; FP32
Kahan optical depths within 1e-5 relative; images of the same histories (same Philox
streams) agree with the oracle to 1e-9 relative in total and per-pixel median.
"""

import numpy as np
import pytest

import oracle as O
from helpers import PC, cart_borders, child_first, disk_density_cartesian, frame60, golden, load_pkg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    pkg = load_pkg()
    e = pkg.Engine(0)
    yield e
    e.close()


def _cart(eng, key):
    f = golden("cart_paths.npz")
    h = float(f[key + "_half"])
    n = f[key + "_n"]
    eng.grid_cartesian(cart_borders(n[0], -h, h), cart_borders(n[1], -h, h), cart_borders(n[2], -h, h))
    return f


@pytest.mark.parametrize("key", ["c64", "c8"])
def test_cartesian_paths_bit_exact(eng, key):
    f = _cart(eng, key)
    print("case cartesian paths", key, flush=True)
    off, m, ds = eng.trace_paths(f[key + "_rays"])
    np.testing.assert_array_equal(off, f[key + "_off"])
    np.testing.assert_array_equal(m, f[key + "_m"])
    np.testing.assert_array_equal(ds, f[key + "_ds"])


def test_cartesian_tau(eng):
    f = _cart(eng, "c8")
    eng.medium(f["c8_density"])
    s = float(f["c8_sext"][0])
    eng.dust([0.55e-6], [s], [0.6 * s], [0.6])
    print("case cartesian tau", flush=True)
    tp, to, cum = eng.trace_tau(f["c8_rays"], toff=f["c8_toff"])
    np.testing.assert_array_equal(cum, f["c8_tau"])
    np.testing.assert_array_equal(to, f["c8_tauobs"])
    last = np.array([f["c8_tau"][f["c8_toff"][i + 1] - 1] if f["c8_toff"][i + 1] > f["c8_toff"][i] else 0.0
                     for i in range(len(tp))])
    np.testing.assert_array_equal(tp, last)
    tp32, to32 = eng.trace_tau(f["c8_rays"], fp_mode=1)
    np.testing.assert_allclose(tp32, tp, rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(to32, to, rtol=1e-5, atol=1e-12)


def _tree(eng):
    t = golden("tree_paths.npz")
    eng.grid_octree(t["bounds"], child_first(t["child_begin"], t["child_count"], t["child_index"]), t["cell_index"],
                    t["nb_off"], t["nb_id"])
    return t


def test_octree_paths_bit_exact(eng):
    t = _tree(eng)
    print("case octree paths", flush=True)
    off, m, ds = eng.trace_paths(t["rays"])
    np.testing.assert_array_equal(off, t["off"])
    np.testing.assert_array_equal(m, t["m"])
    np.testing.assert_array_equal(ds, t["ds"])


def test_octree_tau(eng):
    t = _tree(eng)
    eng.medium(t["density"])
    s = float(t["sext"][0])
    eng.dust([0.55e-6], [s], [0.6 * s], [0.6])
    print("case octree tau", flush=True)
    tp, to, cum = eng.trace_tau(t["rays"], toff=t["toff"])
    np.testing.assert_array_equal(cum, t["tau"])
    np.testing.assert_array_equal(to, t["tauobs"])
    tp32, to32 = eng.trace_tau(t["rays"], fp_mode=1)
    np.testing.assert_allclose(tp32, tp, rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(to32, to, rtol=1e-5, atol=1e-12)


def _random_rays(rng, n, h):
    r = rng.uniform(-1.2 * h, 1.2 * h, (n, 3))
    k = rng.normal(size=(n, 3))
    # a share of axis-aligned and in-plane directions (exact ties between exit distances)
    m = rng.random(n) < 0.1
    k[m] = np.eye(3)[rng.integers(0, 3, m.sum())] * rng.choice([-1.0, 1.0], (m.sum(), 1))
    m2 = (rng.random(n) < 0.1) & ~m
    k[m2, rng.integers(0, 3, m2.sum())] = 0.0
    k /= np.linalg.norm(k, axis=1, keepdims=True)
    return np.hstack([r, k])


@pytest.mark.parametrize("grid", ["cartesian", "octree"])
def test_random_rays_match_oracle(eng, grid):
    """1e5 random rays (inside and outside starts, 20% with axis-aligned or in-plane
    directions): the device paths equal the oracle's bit for bit."""
    rng = np.random.default_rng(17)
    if grid == "octree":
        t = _tree(eng)
        og = O.Grid("octree", bounds=t["bounds"], child_first=child_first(t["child_begin"], t["child_count"], t["child_index"]),
                    cell_index=t["cell_index"], nb_off=t["nb_off"], nb_id=t["nb_id"])
        h = float(t["half"])
    else:
        f = _cart(eng, "c64")
        h = float(f["c64_half"])
        b = cart_borders(64, -h, h)
        og = O.Grid("cartesian", xb=b, yb=b, zb=b)
    rays = _random_rays(rng, 100000, h)
    print("case random rays", grid, flush=True)
    off, m, ds = eng.trace_paths(rays)
    ooff, om, ods = O.trace(og, rays)
    np.testing.assert_array_equal(off, ooff)
    np.testing.assert_array_equal(m, om)
    np.testing.assert_array_equal(ds, ods)


def test_empty_and_outside_rays(eng):
    _cart(eng, "c8")
    print("case empty rays", flush=True)
    off, m, ds = eng.trace_paths(np.zeros((0, 6)))
    assert len(off) == 1 and len(m) == 0
    far = np.array([[1e20, 1e20, 1e20, 1.0, 0.0, 0.0]])
    off, m, ds = eng.trace_paths(far)
    assert off[-1] == 0


def _lifecycle_case(eng, npk, seed, fp_mode=0, threshold=16, with_oracle=True, mode=0, pool=1 << 21, wgs=0, iters=8):
    b, dens = disk_density_cartesian(samples=10, seed=9)
    eng.grid_cartesian(b, b, b)
    eng.medium(dens)
    s = 1.5e-29
    eng.dust([0.55e-6], [s], [0.6 * s], [0.6])
    eng.frame_remove_all()
    fr = frame60()
    eng.frame_add(fr["inclination"], fr["azimuth"], fr["roll"], fr["nx"], fr["ny"], fr["fovx"], fr["fovy"])
    eng.source_point((0.0, 0.0, 0.0), [1.0 / npk])
    eng.options(0.5, 1e4, 0, fp_mode, seed)
    eng.tune(0, 256, threshold)
    eng.tune_pipeline(mode, pool, wgs, iters)
    eng.frame_clear()
    print("case lifecycle", npk, seed, fp_mode, threshold, mode, pool, wgs, iters, flush=True)
    eng.run_packets(0, npk)
    img = eng.frame_read(0)[0]
    cnt = eng.counters()
    if not with_oracle:
        return img, cnt, None, None
    g = O.Grid("cartesian", xb=b, yb=b, zb=b)
    ofr, ocnt = O.simulate(g, dens, [0.55e-6], [s], [0.6 * s], [0.6], [1.0 / npk], [fr], 0, npk, threads=8, seed=seed)
    return img, cnt, ofr[0][0], ocnt


@pytest.mark.parametrize("mode", [0, 1], ids=["pipeline", "persistent"])
def test_lifecycle_matches_oracle_fp64(eng, mode):
    """Same histories, same random streams: the frame equals the oracle's up to rare histories
    whose discrete decisions flip on last-bit differences of the device math library."""
    img, cnt, oimg, ocnt = _lifecycle_case(eng, 20000, 3, mode=mode)
    assert cnt[0] == ocnt[0] == 20000
    assert abs(int(cnt[2]) - int(ocnt[2])) <= 0.001 * ocnt[2]  # forced paths
    assert abs(img.sum() / oimg.sum() - 1) < 1e-6
    nz = oimg > 0
    rel = np.abs(img[nz] - oimg[nz]) / oimg[nz]
    assert np.median(rel) < 1e-9
    assert np.mean(rel < 1e-6) > 0.99


@pytest.mark.parametrize("mode", [0, 1], ids=["pipeline", "persistent"])
def test_lifecycle_fp32_kahan_close_to_fp64(eng, mode):
    img64, _, _, _ = _lifecycle_case(eng, 20000, 5, fp_mode=0, with_oracle=False, mode=mode)
    img32, _, _, _ = _lifecycle_case(eng, 20000, 5, fp_mode=1, with_oracle=False, mode=mode)
    assert abs(img32.sum() / img64.sum() - 1) < 1e-3
    nz = img64 > 0
    rel = np.abs(img32[nz] - img64[nz]) / img64[nz]
    assert np.median(rel) < 0.01


def test_lifecycle_independent_of_launch_shape(eng):
    """Results do not depend on scheduling: transition threshold, engine mode, pool size,
    trace workgroups and graph length change only the order of detector additions."""
    a, ca, _, _ = _lifecycle_case(eng, 5000, 7, threshold=1, with_oracle=False, mode=1)
    b, cb, _, _ = _lifecycle_case(eng, 5000, 7, threshold=64, with_oracle=False, mode=1)
    c, cc, _, _ = _lifecycle_case(eng, 5000, 7, with_oracle=False, mode=0, pool=512, wgs=3, iters=2)
    d, cd, _, _ = _lifecycle_case(eng, 5000, 7, with_oracle=False, mode=0, pool=1 << 16, wgs=0, iters=16)
    for x, cx in ((b, cb), (c, cc), (d, cd)):
        np.testing.assert_array_equal(ca[[0, 1, 2, 3]], cx[[0, 1, 2, 3]])
        np.testing.assert_allclose(a, x, rtol=1e-12, atol=0)


def test_pipeline_multiple_observers(eng):
    """Three instruments, two of them the same observer (shared optical depth): each frame
    matches the oracle's for the same histories."""
    b, dens = disk_density_cartesian(samples=10, seed=9)
    eng.grid_cartesian(b, b, b)
    eng.medium(dens)
    s = 1.5e-29
    eng.dust([0.55e-6], [s], [0.6 * s], [0.6])
    eng.frame_remove_all()
    frs = [frame60(), frame60(npix=100), dict(frame60(), inclination=np.radians(20.0), azimuth=np.radians(30.0))]
    for fr in frs:
        eng.frame_add(fr["inclination"], fr["azimuth"], fr["roll"], fr["nx"], fr["ny"], fr["fovx"], fr["fovy"])
    npk = 8000
    eng.source_point((0.0, 0.0, 0.0), [1.0 / npk])
    eng.options(0.5, 1e4, 0, 0, 21)
    eng.tune_pipeline(0, 1 << 21, 0, 8)
    eng.frame_clear()
    print("case multiple observers", flush=True)
    eng.run_packets(0, npk)
    ms, launches, iters = eng.last_run_stats()
    assert ms > 0 and launches >= 2 and iters >= launches
    assert eng.last_kernel_ms() >= ms
    g = O.Grid("cartesian", xb=b, yb=b, zb=b)
    ofr, _ = O.simulate(g, dens, [0.55e-6], [s], [0.6 * s], [0.6], [1.0 / npk], frs, 0, npk, threads=8, seed=21)
    for i in range(3):
        img = eng.frame_read(i)[0]
        assert abs(img.sum() / ofr[i][0].sum() - 1) < 1e-6, i


def test_autotune_keeps_detector_and_picks_valid_config(eng):
    """The autotuner's trial runs leave the frame and the counters as they were, the chosen
    configuration is one of the candidates, and a run after tuning still equals the oracle."""
    img0, cnt0, _, _ = _lifecycle_case(eng, 2000, 23, with_oracle=False)
    print("case autotune", flush=True)
    cfg, rate = eng.autotune(10_000_000, trial_packets=20000)
    assert rate > 0
    assert cfg["mode"] in (0, 1) and cfg["threshold"] in (12, 16, 20, 24, 32, 48, 64)
    np.testing.assert_array_equal(eng.frame_read(0)[0], img0)
    np.testing.assert_array_equal(eng.counters(), cnt0)
    eng.frame_clear()
    eng.run_packets(0, 2000)
    np.testing.assert_allclose(eng.frame_read(0)[0], img0, rtol=1e-12, atol=0)


def test_pipeline_many_observer_groups(eng):
    """40 distinct observers (more than a 32-bit mask could track) plus one instrument sharing
    the last observer: the pipeline's and the persistent engine's frames both equal the
    oracle's for the same histories, instrument by instrument."""
    b, dens = disk_density_cartesian(n=16, samples=4, seed=13)
    eng.grid_cartesian(b, b, b)
    eng.medium(dens)
    s = 1.5e-29
    eng.dust([0.55e-6], [s], [0.6 * s], [0.6])
    eng.frame_remove_all()
    frs = [dict(inclination=np.radians(2.0 + 2.2 * i), azimuth=np.radians(9.0 * i), roll=0.0, nx=12, ny=12,
                fovx=20 * PC, fovy=20 * PC) for i in range(40)]
    frs.append(dict(frs[-1], nx=6, ny=6))  # same observer as the previous instrument
    for fr in frs:
        eng.frame_add(fr["inclination"], fr["azimuth"], fr["roll"], fr["nx"], fr["ny"], fr["fovx"], fr["fovy"])
    npk = 3000
    eng.source_point((0.0, 0.0, 0.0), [1.0 / npk])
    eng.options(0.5, 1e4, 0, 0, 31)
    g = O.Grid("cartesian", xb=b, yb=b, zb=b)
    ofr, _ = O.simulate(g, dens, [0.55e-6], [s], [0.6 * s], [0.6], [1.0 / npk], frs, 0, npk, threads=8, seed=31)
    for mode in (0, 1):
        eng.tune_pipeline(mode, 1 << 16, 0, 8)
        eng.frame_clear()
        print("case 41 instruments, engine mode", mode, flush=True)
        eng.run_packets(0, npk)
        for i in range(len(frs)):
            img, oimg = eng.frame_read(i)[0], ofr[i][0]
            assert oimg.sum() > 0
            assert abs(img.sum() / oimg.sum() - 1) < 1e-6, (mode, i)
            nz = oimg > 0
            assert np.median(np.abs(img[nz] - oimg[nz]) / oimg[nz]) < 1e-9, (mode, i)


def test_pipeline_queue_overflow_fails_the_run(eng):
    """A request queue too small for one iteration's requests (forced from the host with
    skg_tune_queue) makes the pipeline run fail with an error instead of losing deposits;
    with the automatic size the same run equals the persistent engine."""
    pkg = load_pkg()
    b, dens = disk_density_cartesian(n=16, samples=4, seed=17)
    eng.grid_cartesian(b, b, b)
    eng.medium(dens)
    s = 1.5e-29
    eng.dust([0.55e-6], [s], [0.6 * s], [0.6])
    eng.frame_remove_all()
    fr = frame60(npix=20)
    eng.frame_add(fr["inclination"], fr["azimuth"], fr["roll"], fr["nx"], fr["ny"], fr["fovx"], fr["fovy"])
    npk = 4096
    eng.source_point((0.0, 0.0, 0.0), [1.0 / npk])
    eng.options(0.5, 1e4, 0, 0, 37)
    eng.tune_pipeline(0, 1024, 0, 8)
    eng.tune_queue(300)  # 1024 slots need up to 2048 requests per iteration
    eng.frame_clear()
    print("case queue overflow (expected error)", flush=True)
    with pytest.raises(pkg.EngineError, match="overflow"):
        eng.run_packets(0, npk)
    with pytest.raises(pkg.EngineError):
        eng.tune_queue(-1)
    eng.tune_queue(0)
    eng.frame_clear()
    eng.run_packets(0, npk)
    pipe = eng.frame_read(0)[0]
    eng.tune_pipeline(1, 1024, 0, 8)
    eng.frame_clear()
    eng.run_packets(0, npk)
    np.testing.assert_allclose(pipe, eng.frame_read(0)[0], rtol=1e-12, atol=0)


def test_continuation_streams(eng):
    """skg_continuation_key (persistent engine): key 0 is the default single stream; a key is
    deterministic; two keys share each history's emission and first-scattering peel-offs and
    continue independently, so their batch totals are correlated but not equal and agree in
    mean; the pipeline engine refuses a key."""
    pkg = load_pkg()
    b, dens = disk_density_cartesian(n=16, samples=4, seed=19)
    eng.grid_cartesian(b, b, b)
    eng.medium(dens)
    s = 1.5e-29
    eng.dust([0.55e-6], [s], [0.6 * s], [0.6])
    eng.frame_remove_all()
    fr = frame60(npix=30)
    eng.frame_add(fr["inclination"], fr["azimuth"], fr["roll"], fr["nx"], fr["ny"], fr["fovx"], fr["fovy"])
    per, nb = 20000, 24
    eng.source_point((0.0, 0.0, 0.0), [1.0 / per])
    eng.options(0.5, 1e4, 0, 0, 41)
    eng.tune_pipeline(1, 1 << 18, 0, 8)

    def run(key, first):
        eng.continuation_key(key)
        eng.frame_clear()
        eng.run_packets(first, per)
        return eng.frame_read(0)

    print("case continuation streams", flush=True)
    base = run(0, 0)
    eng.continuation_key(0)
    eng.frame_clear()
    eng.run_packets(0, per)
    np.testing.assert_allclose(eng.frame_read(0), base, rtol=1e-12, atol=0)
    np.testing.assert_allclose(run(1, 0), run(1, 0), rtol=1e-12, atol=0)
    t1 = np.array([run(1, k * per).sum() for k in range(nb)])
    t2 = np.array([run(2, k * per).sum() for k in range(nb)])
    eng.continuation_key(0)
    assert np.all(t1 != t2)
    corr = np.corrcoef(t1, t2)[0, 1]
    print(f"  correlation of batch totals between continuation streams: {corr:.3f}")
    assert corr > 0.2
    se = np.hypot(t1.std(ddof=1), t2.std(ddof=1)) / np.sqrt(nb)
    assert abs(t1.mean() - t2.mean()) < 4 * se
    eng.tune_pipeline(0, 1 << 18, 0, 8)
    eng.continuation_key(3)
    with pytest.raises(pkg.EngineError, match="persistent"):
        eng.run_packets(0, 1000)
    eng.continuation_key(0)
