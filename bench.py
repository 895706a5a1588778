#!/usr/bin/env python3
"""Photon-packet throughput of the MI355X engine on the octree disk at face-on tau_V = 50.

One step = one launch of the forced-scattering life cycle over a batch of photon packets
(launch, emission peel-off, forced paths, HG scattering, scattering peel-off into a 200x200
frame at 60 degrees) on the spiral-disk octree (levels 3..10, max dust fraction 1e-6,
~2.9e6 leaves), 5 wavelengths 0.3..2 um.  The grid, densities and dust tables are resident
in HBM before the timed region; `value` is completed histories per second.

Printed JSON line (rank 0, N = 1):
  roofline      dominant kernel: `lifecycle` (persistent engine, the default) or `pipe_trace`
                (the grid walk of the wavefront pipeline): algorithmic bytes per launch (64 B
                per octree segment crossing, plus 8 B per detector atomic for the persistent
                kernel or 96 B of request and packet state per traversal for the pipeline)
                divided by the kernel's average duration from HIP events on the engine's
                stream, against 8 TB/s HBM
  cpu_baseline  the reference SKIRT CPU path (oracle/_ref) on a bounded sample of the same
                workload on this host's cores; falls back to the FP64 oracle ("port")
  guide         plain HG versus the learned path guide: records from the first 1e7 plain
                packets, Adam training on the GPU, then 32 independent batches of each sampler;
                figure of merit 1/(relative variance of the mean x time) for the total frame
                flux and the median over pixels
"""

from __future__ import annotations

import argparse
import importlib.util
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
CPU_PACKETS_DEFAULT = 1.5e6  # packets of the reference CPU sample on the octree disk (about 12 s on 16 cores)
BYTES_PER_SEGMENT = 64.0  # octree: node bounds 48 B + leaf/cell 4 B + neighbour entry 4 B + density 8 B
BYTES_PER_ATOMIC = 8.0
BYTES_PER_REQUEST = 96.0  # per traversal: request 32 B + packet position/direction/bin 52 B + result 12 B


def load_pkg():
    name = "skirt9_gpu_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "skirt9-gpu_amd", "__init__.py"),
                                                  submodule_search_locations=[os.path.join(ROOT, "skirt9-gpu_amd")])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline_reference(tau_v, nlambda, lam_range, packets, threads, cartesian=False, voronoi=None, tetra=None):
    """Times the reference SKIRT build on the same workload (its own grid and densities; for
    a Voronoi workload our sites and cell densities, imported as a VoronoiMeshMedium whose mesh
    the grid employs; for the batched tetra workload a few of the instances, each on the
    Delaunay mesh of our vertices); packets/s over the primary emission phase, as logged by the
    reference."""
    exe = os.path.join(ROOT, "oracle", "_ref", "bin", "skirt", "ref", "skirt_ref")
    if not os.path.exists(exe):
        return None
    if tetra is not None:
        return cpu_baseline_reference_tetra(exe, tetra, threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import skigen
    if cartesian:
        spec = skigen.config0(num_packets=packets)
    elif voronoi is not None:
        spec = skigen.SkiSpec(num_packets=packets, grid=("voronoi_imported",), medium_file="sites.txt",
                              instruments=tuple((f"i{k}", np.degrees(f["inclination"]), np.degrees(f["azimuth"]))
                                                for k, f in enumerate(voronoi["frames"])))
    else:
        spec = skigen.config1(num_packets=packets, tau_v=tau_v, nlambda=nlambda, lam_range=lam_range)
    with tempfile.TemporaryDirectory() as tmp:
        if voronoi is not None:
            skigen.write_voronoi_medium(os.path.join(tmp, "sites.txt"), voronoi["sites"], voronoi["density"])
        ski = os.path.join(tmp, "bench.ski")
        with open(ski, "w") as f:
            f.write(skigen.render(spec))
        res = subprocess.run([exe, "run", ski, tmp, str(threads)], capture_output=True, text=True, timeout=900)
        if res.returncode != 0:
            return None
        log = open(os.path.join(tmp, "bench_log.txt")).read()
    m = re.search(r"Finished primary emission in ([0-9.]+) s", log)
    if not m:
        return None
    t = float(m.group(1))
    return dict(value=packets / t, unit="packets/s", cores=threads, kind="reference",
                sample=f"{packets:.0f} packets of the same workload ("
                       + ("our sites and cell densities imported, mesh built by the reference" if voronoi is not None
                          else "grid and densities built by the reference")
                       + f"), primary emission {t:.2f} s")


def cpu_baseline_reference_tetra(exe, tetra, threads):
    """The reference on a sample of the batched instances: each sampled instance (its tau_V and
    inclination, config 0's disk and dust otherwise) runs `packets` packets on the Delaunay
    mesh of our vertices (TetraMeshSpatialGrid policy File; the reference computes its own cell
    densities); packets/s = all packets / the summed primary emission times."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import skigen
    total_t, total_n, taus = 0.0, 0, []
    with tempfile.TemporaryDirectory() as tmp:
        skigen.write_points(os.path.join(tmp, "points.txt"), tetra["points"])
        for k, (tau, inc, n) in enumerate(tetra["runs"]):
            spec = skigen.SkiSpec(num_packets=n, tau_v=tau, grid=("tetra_file", "points.txt"),
                                  instruments=(("i0", inc, 0.0),))
            ski = os.path.join(tmp, f"inst{k}.ski")
            with open(ski, "w") as f:
                f.write(skigen.render(spec))
            res = subprocess.run([exe, "run", ski, tmp, str(threads)], capture_output=True, text=True, timeout=900)
            if res.returncode != 0:
                return None
            log = open(os.path.join(tmp, f"inst{k}_log.txt")).read()
            m = re.search(r"Finished primary emission in ([0-9.]+) s", log)
            if not m:
                return None
            total_t += float(m.group(1))
            total_n += n
            taus.append(f"{tau:.3g}")
    return dict(value=total_n / total_t, unit="packets/s", cores=threads, kind="reference",
                sample=f"{len(tetra['runs'])} instances (tau_V {', '.join(taus)}: the packet-weighted quartiles of "
                       f"the batch) x {tetra['runs'][0][2]} packets on the Delaunay mesh of our "
                       f"{len(tetra['points'])} vertices, built by the reference; primary emission {total_t:.2f} s")


def tetra_sample(wl, packets):
    """Four instances of the batch for the reference baseline: the instances at the packet-
    weighted quartiles of tau_V (the batch's cost is spread that way), `packets` // 4 packets
    each: (tau_V, inclination in degrees, packets)."""
    I = wl.instances
    order = np.argsort(I["tau_scale"])
    cum = np.cumsum(I["packets"][order].astype(np.float64))
    cum /= cum[-1]
    runs = []
    for q in (0.125, 0.375, 0.625, 0.875):
        i = order[min(int(np.searchsorted(cum, q)), len(order) - 1)]
        runs.append((float(I["tau_scale"][i] * wl.tau_v), float(I["inclination_deg"][i]), max(1, packets // 4)))
    return runs


def cpu_baseline_port(wl, packets, threads, seed):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    g = oracle_grid(wl)
    t0 = time.perf_counter()
    O.simulate(g, wl.density, wl.lam, wl.sext, wl.ssca, wl.g, np.ones(len(wl.lam)), wl.frames, 0, packets,
               threads=threads, seed=seed)
    t = time.perf_counter() - t0
    return dict(value=packets / t, unit="packets/s", cores=threads, kind="port",
                sample=f"{packets} packets of the same workload and grid through the FP64 oracle, {t:.2f} s")


def oracle_grid(wl):
    """The FP64 oracle's view of the workload's grid (the same tables the GPU traverses)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if wl.grid == "cartesian":
        return O.Grid("cartesian", xb=wl.tables["xb"], yb=wl.tables["yb"], zb=wl.tables["zb"])
    if wl.grid == "octree":
        ex = wl.octree.export()
        return O.Grid("octree", bounds=ex["bounds"], child_first=ex["child_first"], cell_index=ex["cell_index"],
                      nb_off=ex["nb_off"], nb_id=ex["nb_id"])
    ex = wl.tables["mesh"].export()
    if wl.grid == "voronoi":
        return O.Grid("voronoi", **{q: ex[q] for q in ("sites", "nb_begin", "nb_count", "nb_index", "block_n",
                                                        "block_begin", "block_count", "block_index", "box")})
    return O.Grid("tetra", **ex)


def error_vs_fp64(eng, pkg, wl, fp, sample, threads, seed):
    """Per-output error of the timed configuration against the FP64 oracle: the frames of
    `sample` histories run on the GPU in the bench's precision mode and by the oracle on the
    same histories (same Philox streams, MediumSystem.cpp:1785-1844 in FP64); per pixel with
    signal the relative difference (median, 90th percentile, largest) and of the total.  For
    batched instances, instance 0 alone."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    grid = oracle_grid(wl)
    lam, sext, ssca, g = wl.lam, wl.sext, wl.ssca, wl.g
    frames, first, oseed, nfr = wl.frames, 1 << 44, seed, len(wl.frames)
    lpack = np.ones(len(lam))
    if wl.instances is not None:
        I = wl.instances
        sample = int(min(sample, I["packets"][0]))
        sc = float(I["tau_scale"][0])
        sext, ssca = sext * sc, ssca * sc
        frames, first, oseed, nfr = wl.frames[:1], 0, int(I["seeds"][0]), 1
        lpack = np.full(len(lam), 1.0 / max(int(I["packets"][0]), 1))
    eng.frame_clear()
    eng.run_packets(first, sample)
    gpu = [eng.frame_read(i) for i in range(nfr)]
    eng.frame_clear()
    t0 = time.perf_counter()
    ofr, _ = O.simulate(grid, wl.density, lam, sext, ssca, g, lpack, frames, 0 if wl.instances is not None else first,
                        sample, threads=threads, seed=oseed)
    t = time.perf_counter() - t0
    a = np.concatenate([x.ravel() for x in gpu])
    b = np.concatenate([x.ravel() for x in ofr])
    nz = b > 0
    rel = np.abs(a[nz] - b[nz]) / b[nz]
    return dict(reference="FP64 CPU oracle (oracle/src/oracle.cpp), same histories", mode="fp64" if fp == pkg.SKG_FP64
                else "fp32 + Kahan", histories=int(sample), pixels=int(nz.sum()), median=float(np.median(rel)),
                p90=float(np.percentile(rel, 90)), max=float(rel.max()), total=float(abs(a.sum() / b.sum() - 1)),
                oracle_seconds=t)


def frame_batches(eng, first, batches, per):
    """Independent batches of `per` packets; returns (per-batch frames (B, nl, ny, nx), GPU
    wall time summed over the batches, next history index)."""
    frames, t = [], 0.0
    for _ in range(batches):
        eng.frame_clear()
        eng.sync()
        t0 = time.perf_counter()
        eng.run_packets(first, per)
        eng.sync()
        t += time.perf_counter() - t0
        frames.append(eng.frame_read(0))
        first += per
    return np.stack(frames), t, first


def figure_of_merit(X, t):
    """1/(relative variance of the mean x time) for the total frame flux and per pixel
    (median over pixels with signal)."""
    B = len(X)
    tot = X.reshape(B, -1).sum(axis=1)
    rv_tot = tot.var(ddof=1) / B / tot.mean() ** 2
    m = X.mean(axis=0)
    v = X.var(axis=0, ddof=1) / B
    ok = (m > 0) & (v > 0)
    rv_pix = v[ok] / m[ok] ** 2
    return dict(fom_total=1.0 / (rv_tot * t), fom_pixel_median=float(np.median(1.0 / (rv_pix * t))),
                rel_err_total=float(np.sqrt(rv_tot)), pixels=int(ok.sum()), seconds=t, mean_total=float(tot.mean()))


def fom_groups(X, t, groups):
    """Figures of merit of `groups` independent groups of the batches (equal time shares)."""
    B = len(X) // groups
    return [figure_of_merit(X[g * B:(g + 1) * B], t / groups) for g in range(groups)]


def gain_interval(plain, guided, key):
    """Mean over groups of guided / plain figure of merit and its 95 % interval (t with
    groups - 1 degrees of freedom on the log ratios)."""
    from scipy import stats
    r = np.log(np.array([gd[key] / pl[key] for pl, gd in zip(plain, guided)]))
    m, se = r.mean(), r.std(ddof=1) / np.sqrt(len(r))
    q = stats.t.ppf(0.975, len(r) - 1)
    return dict(gain=float(np.exp(m)), low=float(np.exp(m - q * se)), high=float(np.exp(m + q * se)), groups=len(r))


def guide_ceiling(eng, pkg, first, batches, per):
    """The variance no scattering-direction proposal can remove.  Each batch is run twice
    with histories that share their emission, first interaction and first-scattering
    peel-offs and continue on independent random streams (skg_continuation_key); the
    covariance of the two runs' batch means is Var(E[batch | first scattering states]), left
    by any guide, so Var / Cov bounds the variance reduction (and, at equal cost per packet,
    the figure-of-merit gain) of any direction guide applied from the first scattering on."""
    A, Bv = [], []
    for b in range(batches):
        runs = []
        for key in (1, 2):
            eng.continuation_key(key)
            eng.frame_clear()
            eng.run_packets(first + b * per, per)
            runs.append(eng.frame_read(0))
        A.append(runs[0])
        Bv.append(runs[1])
    eng.continuation_key(0)
    A, Bv = np.stack(A), np.stack(Bv)
    n = len(A)

    def ratio(x, y):
        xc, yc = x - x.mean(0), y - y.mean(0)
        var = 0.5 * ((xc * xc).sum(0) + (yc * yc).sum(0)) / (n - 1)
        cov = (xc * yc).sum(0) / (n - 1)
        return var, cov

    tA, tB = A.reshape(n, -1).sum(1), Bv.reshape(n, -1).sum(1)
    v, c = ratio(tA, tB)
    # bootstrap over batches for the total's interval
    rng = np.random.default_rng(3)
    bs = []
    for _ in range(400):
        i = rng.integers(0, n, n)
        vb, cb = ratio(tA[i], tB[i])
        bs.append(vb / cb if cb > 0 else np.inf)
    vp, cp = ratio(A.reshape(n, -1), Bv.reshape(n, -1))
    m = A.mean(0).reshape(-1) > 0
    rp = vp[m] / np.where(cp[m] > 0, cp[m], np.nan)
    return dict(batches=n, packets_per_batch=per, ceiling_total=float(v / c) if c > 0 else None,
                ceiling_total_95=[float(np.percentile(bs, 2.5)), float(np.percentile(bs, 97.5))],
                ceiling_pixel_median=float(np.nanmedian(rp)), correlation_total=float(c / v)), first + batches * per


def guided_comparison(eng, pkg, a, first):
    """Plain HG versus the learned guide (50/50 defensive mixture) on the same workload:
    records from the first `train_packets` plain packets, Adam on the GPU, then batches of
    both samplers in both engines; the gains compare each sampler in its faster engine."""
    out = {}
    eng.tune_pipeline(pkg.SKG_MODE_PIPELINE, a.pool, a.trace_wgs, a.graph_iters)
    theta = np.zeros(pkg.GUIDE_PARAMS, np.float32)
    rng = np.random.default_rng(7)
    # He-initialised weights, zero biases, small last layer (same layout as skg_guide_params)
    o = 0
    for (fan_in, fan_out) in ((9, 64), (64, 64), (64, 64), (64, 40)):
        w = rng.normal(0, np.sqrt(2.0 / fan_in), fan_in * fan_out) * (0.1 if fan_out == 40 else 1.0)
        theta[o:o + w.size] = w
        o += w.size + fan_out
    eng.guide_params(theta)
    t0 = time.perf_counter()
    if a.guide_target == "pixel":
        # plain packets for the image that weights the deposits (relative error per pixel)
        eng.guide_mode(0)
        eng.frame_clear()
        eng.run_packets(first, int(a.target_packets))
        first += int(a.target_packets)
        eng.guide_target(1, a.target_floor)
    else:
        eng.guide_target(0)
    eng.guide_mode(2, mix=a.mix)
    eng.guide_collect(int(a.record_capacity))
    eng.frame_clear()
    eng.run_packets(first, int(a.train_packets))
    first += int(a.train_packets)
    nrec = eng.guide_records()
    t1 = time.perf_counter()
    l0, l1 = eng.guide_train(a.train_steps, batch=a.train_batch, lr=a.lr)
    t2 = time.perf_counter()
    out.update(train_packets=int(a.train_packets), records=int(nrec), collect_s=t1 - t0, train_s=t2 - t1,
               train_steps=a.train_steps, loss_first=l0, loss_last=l1, mix=a.mix, target=a.guide_target)
    if a.guide_target == "pixel":
        out.update(target_packets=int(a.target_packets), target_floor=a.target_floor)
    print(f"guide: {nrec} records from {a.train_packets:.0f} packets in {t1 - t0:.2f} s, "
          f"{a.train_steps} Adam steps in {t2 - t1:.2f} s, loss {l0:.4f} -> {l1:.4f}", file=sys.stderr, flush=True)
    res, groups = {}, {}
    for name, mode, gm in (("plain_pipeline", pkg.SKG_MODE_PIPELINE, 0), ("guided_pipeline", pkg.SKG_MODE_PIPELINE, 1),
                           ("plain_persistent", pkg.SKG_MODE_PERSISTENT, 0),
                           ("guided_persistent", pkg.SKG_MODE_PERSISTENT, 1)):
        eng.tune_pipeline(mode, a.pool, a.trace_wgs, a.graph_iters)
        eng.guide_mode(gm, mix=a.mix)
        eng.run_packets(first, a.fom_packets)  # warm-up (graph capture, pool)
        first += a.fom_packets
        X, t, first = frame_batches(eng, first, a.fom_batches, a.fom_packets)
        res[name] = figure_of_merit(X, t)
        groups[name] = fom_groups(X, t, a.fom_groups)
        print(f"{name}: {a.fom_batches} x {a.fom_packets} packets in {t:.2f} s, rel. error of total "
              f"{res[name]['rel_err_total']:.3e}, FOM total {res[name]['fom_total']:.3e}, "
              f"pixel median {res[name]['fom_pixel_median']:.3e}", file=sys.stderr, flush=True)
    eng.guide_mode(0)
    # each sampler in its faster engine
    pn = max(("plain_pipeline", "plain_persistent"), key=lambda k: res[k]["fom_pixel_median"])
    gn = max(("guided_pipeline", "guided_persistent"), key=lambda k: res[k]["fom_pixel_median"])
    plain, g = res[pn], res[gn]
    out.update(batches=a.fom_batches, packets_per_batch=a.fom_packets, runs=res,
               gain_pixel_median=g["fom_pixel_median"] / plain["fom_pixel_median"],
               gain_total=g["fom_total"] / plain["fom_total"],
               gain_pixel_median_95=gain_interval(groups[pn], groups[gn], "fom_pixel_median"),
               gain_total_95=gain_interval(groups[pn], groups[gn], "fom_total"),
               guided_vs_plain_mean_total=g["mean_total"] / plain["mean_total"])
    print(f"guide gain (group means, 95 %): total {out['gain_total_95']}, pixel median "
          f"{out['gain_pixel_median_95']}", file=sys.stderr, flush=True)
    if a.ceiling_batches:
        eng.tune_pipeline(pkg.SKG_MODE_PERSISTENT, a.pool, a.trace_wgs, a.graph_iters)
        ceil, first = guide_ceiling(eng, pkg, first, a.ceiling_batches, a.fom_packets)
        out["ceiling"] = ceil
        print(f"guide ceiling: {ceil}", file=sys.stderr, flush=True)
    return out, first


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--packets", type=int, default=1 << 22, help="photon packets per step")
    ap.add_argument("--tau", type=float, default=50.0)
    ap.add_argument("--fp", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--cpu-packets", type=float, default=CPU_PACKETS_DEFAULT)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--threshold", type=int, default=16,
                    help="persistent engine: waiting lanes of a wavefront that start a transition round")
    ap.add_argument("--mode", choices=["pipeline", "persistent"], default="persistent")
    ap.add_argument("--pool", type=int, default=1 << 21, help="packet slots in flight (pipeline)")
    ap.add_argument("--trace-wgs", type=int, default=0)
    ap.add_argument("--graph-iters", type=int, default=8)
    ap.add_argument("--workload", choices=["octree", "cartesian", "voronoi", "tetra-batch"], default="octree",
                    help="octree: the spiral disk at face-on tau_V (the measured configuration); voronoi: "
                         "1e6-site Voronoi grid in a lognormal clumpy medium, mean tau_V 5, three frames; "
                         "cartesian: the 64^3 disk at face-on tau_V=1, 0.55 um, the reference's CPU-runnable case; "
                         "tetra-batch: 256 independent disk instances (tau_V 0.1..50, inclination 0..90 deg) on "
                         "a 2e5-point Delaunay mesh in one launch per step")
    ap.add_argument("--batch-scale", type=float, default=0.1,
                    help="tetra-batch: factor on the instances' packet counts (1e5..1e7 each at 1.0)")
    ap.add_argument("--sites", type=int, default=1_000_000)
    ap.add_argument("--nlambda", type=int, default=5, help="octree workload: number of wavelengths")
    ap.add_argument("--lam-range", type=float, nargs=2, default=(0.3, 2.0), metavar=("MIN_UM", "MAX_UM"))
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep --mode/--threshold/--pool/... instead of timing the candidate launch configurations")
    ap.add_argument("--no-guide", action="store_true", help="skip the guided-versus-plain comparison")
    ap.add_argument("--train-packets", type=float, default=1e7)
    ap.add_argument("--record-capacity", type=float, default=2.5e8)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--train-batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--mix", type=float, default=0.5)
    ap.add_argument("--guide-target", choices=["luminosity", "pixel"], default="luminosity",
                    help="training target of the records: the luminosity a packet deposits after the event, or "
                         "that luminosity divided per pixel by the pixel's flux (from --target-packets plain packets)")
    ap.add_argument("--target-packets", type=float, default=2e6)
    ap.add_argument("--target-floor", type=float, default=0.1,
                    help="pixel flux floor of the pixel target, relative to the median pixel of its plane")
    ap.add_argument("--fom-batches", type=int, default=256)
    ap.add_argument("--fom-packets", type=int, default=1 << 18)
    ap.add_argument("--fom-groups", type=int, default=8,
                    help="independent groups of the batches for the gain's 95 %% interval")
    ap.add_argument("--ceiling-batches", type=int, default=128,
                    help="batch pairs of the variance ceiling of any direction guide (0: skip)")
    ap.add_argument("--error-sample", type=int, default=1 << 15,
                    help="histories compared with the FP64 oracle for error_vs_fp64 (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy kernel of MIB mebibytes three times (the "
                         "reference reading for the FETCH_SIZE / WRITE_SIZE counters) and exit")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch
        n = a.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}))
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported "
              "(packets shard across GPUs with one detector reduction at the end).", file=sys.stderr)
        sys.exit(2)

    pkg = load_pkg()
    from skirt9_gpu_amd import workloads
    threads = host_threads()
    nlambda, lam_range = a.nlambda, tuple(a.lam_range)
    eng = pkg.Engine(0)
    if a.workload == "octree":
        # tree refinement estimates and cell densities on the GPU (the host builder's positions)
        wl = workloads.octree_spiral(tau_v=a.tau, nlambda=nlambda, lam_range_um=lam_range, threads=threads, engine=eng)
        ncells = int(wl.octree.info()[1])
        bytes_per_segment = BYTES_PER_SEGMENT
        wdesc = "octree spiral disk, face-on tauV=%g, %d wavelengths %g-%g um" % (a.tau, nlambda, *lam_range)
    elif a.workload == "cartesian":
        wl = workloads.cartesian_disk(threads=threads)
        ncells = 64 ** 3
        # Cartesian crossing: the cell's density (borders are cached)
        bytes_per_segment = 8.0
        nlambda, lam_range = 1, (0.55, 0.55)
        wdesc = "Cartesian 64^3 exponential disk, face-on tauV=1, 0.55 um, albedo 0.6, g 0.6"
        a.no_guide = True
        if a.cpu_packets == CPU_PACKETS_DEFAULT:
            a.cpu_packets = 1.5e7  # about 11 s of the reference on 16 cores
    elif a.workload == "tetra-batch":
        wl = workloads.tetra_batch(packet_scale=a.batch_scale, threads=threads)
        ncells = int(wl.tables["ntets"])
        # tetra crossing: the 256-B record (4 vertices, 4 outward normals, 4 neighbours, density)
        bytes_per_segment = 256.0
        a.packets = wl.total_packets()
        wdesc = ("%d batched instances of the tau_V=1 disk (tau_V 0.1-50, inclination 0-90 deg, %.0f packets in "
                 "all) on a %d-tetra Delaunay mesh of 2e5 density-sampled points" %
                 (len(wl.instances["packets"]), a.packets, ncells))
        a.no_guide = True
        if a.cpu_packets == CPU_PACKETS_DEFAULT:
            a.cpu_packets = 1e6  # four instances of 2.5e5 packets, about 11 s (their mesh builds excluded)
    else:
        wl = workloads.voronoi_clumpy(num_sites=a.sites, threads=threads)
        ncells, nnb = wl.tables["mesh"].info()[:2]
        # Voronoi crossing: the cell's run in the entry table, a 32 B header (density, entry
        # count) + one 32 B entry per neighbour (site, id, the neighbour's run)
        bytes_per_segment = 32.0 + 32.0 * nnb / ncells
        wdesc = "Voronoi %d sites, lognormal clumpy medium (P(k)~k^-3.7, sigma_ln 1), mean tauV=5, 3 frames" % ncells
        a.no_guide = True
        if a.cpu_packets == CPU_PACKETS_DEFAULT:
            a.cpu_packets = 2e5  # about 11 s of the reference on 16 cores (its mesh build excluded)
    print(f"setup: {ncells} cells, {wl.setup_s:.1f} s", file=sys.stderr, flush=True)

    wl.apply(eng, npk_norm=1.0)
    fp = pkg.SKG_FP64 if a.fp == "fp64" else pkg.SKG_FP32_KAHAN
    eng.options(0.5, 1e4, 0, fp, 12345)
    eng.tune(0, 256, a.threshold)
    mode = pkg.SKG_MODE_PIPELINE if a.mode == "pipeline" else pkg.SKG_MODE_PERSISTENT
    eng.tune_pipeline(mode, a.pool, a.trace_wgs, a.graph_iters)
    tuned = None
    if not a.no_autotune:
        # --mode pipeline keeps the autotuner to that engine; the default tries both
        cfg, rate = eng.autotune(1 << 40, trial_packets=1 << 20, engines=2 if a.mode == "pipeline" else 3)
        tuned = dict(cfg, trial_packets_per_s=rate)
        mode = cfg["mode"]
        print(f"autotune: {cfg}, {rate:.3e} packets/s in the trial", file=sys.stderr, flush=True)
    eng.frame_clear()
    first = 0
    # a batched step runs every instance's histories [0, total) once
    batched = wl.instances is not None
    for _ in range(a.warmup):
        eng.run_packets(0 if batched else first, a.packets)
        first += a.packets
    c0 = eng.counters()
    kernel_ms = []
    trace_ms, trace_launches, iterations = 0.0, 0, 0
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        eng.run_packets(0 if batched else first, a.packets)
        kernel_ms.append(eng.last_kernel_ms())
        if mode == pkg.SKG_MODE_PIPELINE:
            tm, tl, it = eng.last_run_stats()
            trace_ms += tm
            trace_launches += tl
            iterations += it
        first += a.packets
    eng.sync()
    t1 = time.perf_counter()
    c1 = eng.counters()
    dt = t1 - t0
    d = c1 - c0
    histories = int(d[0])
    value = histories / dt
    requests = int(d[1] + d[2] + d[6])
    if mode == pkg.SKG_MODE_PIPELINE:
        # dominant kernel: the trace kernel; its launches and summed duration from HIP events
        kname = "pipe_trace"
        launches = max(trace_launches, 1)
        avg_ms = trace_ms / launches
        total_bytes = bytes_per_segment * d[4] + BYTES_PER_REQUEST * requests
    else:
        kname = "lifecycle"
        launches = a.steps
        avg_ms = float(np.mean(kernel_ms))
        total_bytes = bytes_per_segment * d[4] + BYTES_PER_ATOMIC * d[5]
    bytes_per_launch = total_bytes / launches
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    print(f"timed: {histories} histories in {dt:.3f} s; {kname}: {launches} launches, avg {avg_ms:.3f} ms, "
          f"{d[4] / a.steps:.3e} segments and {requests / a.steps:.3e} traversals per step, "
          f"{iterations / a.steps:.0f} pipeline iterations per step", file=sys.stderr, flush=True)

    err = None
    if a.error_sample:
        err = error_vs_fp64(eng, pkg, wl, fp, a.error_sample, threads, 12345)
        print(f"error vs FP64 oracle: {err}", file=sys.stderr, flush=True)

    guide = None
    if not a.no_guide:
        guide, first = guided_comparison(eng, pkg, a, first)

    cpu = None
    if not a.no_cpu_baseline:
        try:
            vor = tet = None
            if a.workload == "voronoi":
                vor = dict(sites=wl.tables["mesh"].export()["sites"], density=wl.density, frames=wl.frames)
            if a.workload == "tetra-batch":
                tet = dict(points=wl.tables["mesh"].export()["verts"], runs=tetra_sample(wl, int(a.cpu_packets)))
            cpu = cpu_baseline_reference(a.tau, nlambda, lam_range, a.cpu_packets, threads,
                                         cartesian=a.workload == "cartesian", voronoi=vor, tetra=tet)
        except Exception as ex:  # the reference build may be absent or fail on this host
            print(f"reference CPU baseline unavailable: {ex}", file=sys.stderr)
        if cpu is None:
            cpu = cpu_baseline_port(wl, int(a.cpu_packets // 10), threads, 12345)

    # HBM bytes per launch of the dominant kernel from the committed rocprofv3 PMC passes
    # (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel's reading), when they cover this
    # kernel, workload and batch size
    traffic = None
    # named for the workload (and its tau_V for the octree disk, whose tau_V is an option)
    tag = f"{a.workload}_tau{a.tau:g}" if a.workload == "octree" else a.workload
    prof = os.path.join(ROOT, "profiles", f"{kname}_{tag}_traffic.json")
    if os.path.exists(prof):
        try:
            tj = json.load(open(prof))
            if (kname != "lifecycle" or tj.get("packets_per_launch") == a.packets) and tj.get("nlambda", 5) == len(wl.lam):
                traffic = tj.get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    line = {
        "metric": "photon packets/s + 1/(var·time) guided vs plain, 1 MI355X, octree τV=50",
        "value": value,
        "unit": "packets/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if fp == pkg.SKG_FP64 else "f32+kahan/f64",
        "data": "synthetic",
        "config": {"workload": wdesc + ", plain HG in the timed steps",
                   "cells": ncells, "packets_per_step": a.packets,
                   "frames": ("%d frames 200x200, one per instance" % len(wl.frames) if batched else
                              ", ".join("200x200 at %g deg" % np.degrees(f["inclination"]) for f in wl.frames)),
                   "engine": "pipeline" if mode == pkg.SKG_MODE_PIPELINE else "persistent",
                   "autotuned": tuned, "segments_per_packet": float(d[4] / max(histories, 1)),
                   "bytes_per_segment": bytes_per_segment},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kname, "kernel_ms": avg_ms, "launches_per_step": launches / a.steps},
        "cpu_baseline": cpu,
        "error_vs_fp64": err,
        "guide": guide,
    }
    print(json.dumps(line), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
