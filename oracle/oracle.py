"""ctypes binding of the FP64 CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker.  The product never imports this module.
"""

from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_trace.restype = C.c_int64
        _lib.orc_tau.restype = C.c_int64
        _lib.orc_phase_hg.restype = C.c_double
        _lib.orc_phase_hg.argtypes = [C.c_double, C.c_double]
        _lib.orc_frame_size.restype = C.c_int64
    return _lib


P = C.POINTER
f64p = P(C.c_double)
i32p = P(C.c_int32)
i64p = P(C.c_int64)


class OrcGrid(C.Structure):
    _fields_ = [("type", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("xb", f64p), ("yb", f64p), ("zb", f64p),
                ("nnodes", C.c_int64), ("bounds", f64p), ("child_first", i32p), ("cell_index", i32p),
                ("nb_off", i64p), ("nb_id", i32p), ("box", C.c_double * 6),
                ("sites", f64p), ("vnb_begin", i32p), ("vnb_count", i32p), ("vnb_index", i32p),
                ("block_n", C.c_int32), ("block_begin", i32p), ("block_count", i32p), ("block_index", i32p),
                ("verts", f64p), ("ntets", C.c_int64), ("tet_verts", i32p), ("face_nb", i32p),
                ("face_normal", f64p)]


class OrcInstrument(C.Structure):
    _fields_ = [("inclination", C.c_double), ("azimuth", C.c_double), ("roll", C.c_double),
                ("nx", C.c_int32), ("ny", C.c_int32),
                ("fovx", C.c_double), ("fovy", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
                ("sed", C.c_int32), ("radius", C.c_double), ("persp", C.c_int32), ("M", C.c_double * 16),
                ("eye", C.c_double * 3), ("psize", C.c_double)]


class OrcSim(C.Structure):
    _fields_ = [("grid", P(OrcGrid)), ("density", f64p), ("nlambda", C.c_int32),
                ("lam", f64p), ("sext", f64p), ("ssca", f64p), ("g", f64p), ("lpack", f64p),
                ("src", C.c_double * 3), ("xi", C.c_double), ("min_weight_reduction", C.c_double),
                ("min_scatt_events", C.c_int32), ("ninstr", C.c_int32), ("instr", P(OrcInstrument)),
                ("seed", C.c_uint64), ("ncells", C.c_int64), ("rf", f64p), ("nent", C.c_int64),
                ("ent_kernel", C.c_int32), ("ent", f64p), ("ent_first", i64p), ("ent_lpack", f64p), ("ent_X", f64p)]


def _p(a, t):
    return a.ctypes.data_as(t)


class Grid:
    """Keeps the numpy arrays alive behind an OrcGrid."""

    def __init__(self, kind, **kw):
        self.keep = []
        g = OrcGrid()
        if kind == "cartesian":
            g.type = 0
            xb, yb, zb = (np.ascontiguousarray(kw[k], dtype=np.float64) for k in ("xb", "yb", "zb"))
            self.keep += [xb, yb, zb]
            g.nx, g.ny, g.nz = len(xb) - 1, len(yb) - 1, len(zb) - 1
            g.xb, g.yb, g.zb = _p(xb, f64p), _p(yb, f64p), _p(zb, f64p)
            box = (xb[0], yb[0], zb[0], xb[-1], yb[-1], zb[-1])
        elif kind == "octree":
            g.type = 1
            bounds = np.ascontiguousarray(kw["bounds"], dtype=np.float64).reshape(-1, 6)
            cf = np.ascontiguousarray(kw["child_first"], dtype=np.int32)
            ci = np.ascontiguousarray(kw["cell_index"], dtype=np.int32)
            nbo = np.ascontiguousarray(kw["nb_off"], dtype=np.int64)
            nbi = np.ascontiguousarray(kw["nb_id"], dtype=np.int32)
            self.keep += [bounds, cf, ci, nbo, nbi]
            g.nnodes = len(bounds)
            g.bounds, g.child_first, g.cell_index = _p(bounds, f64p), _p(cf, i32p), _p(ci, i32p)
            g.nb_off, g.nb_id = _p(nbo, i64p), _p(nbi, i32p)
            box = tuple(bounds[0])
        elif kind == "voronoi":
            g.type = 2
            sites = np.ascontiguousarray(kw["sites"], dtype=np.float64)
            a = [np.ascontiguousarray(kw[k], dtype=np.int32) for k in
                 ("nb_begin", "nb_count", "nb_index", "block_begin", "block_count", "block_index")]
            self.keep += [sites] + a
            g.sites = _p(sites, f64p)
            g.vnb_begin, g.vnb_count, g.vnb_index, g.block_begin, g.block_count, g.block_index = (_p(x, i32p) for x in a)
            g.block_n = int(kw["block_n"])
            box = tuple(np.asarray(kw["box"], dtype=np.float64))
        elif kind == "tetra":
            g.type = 3
            verts = np.ascontiguousarray(kw["verts"], dtype=np.float64)
            tv = np.ascontiguousarray(kw["tet_verts"], dtype=np.int32)
            fn = np.ascontiguousarray(kw["face_nb"], dtype=np.int32)
            nrm = np.ascontiguousarray(kw["face_normal"], dtype=np.float64)
            self.keep += [verts, tv, fn, nrm]
            g.verts, g.tet_verts, g.face_nb, g.face_normal = _p(verts, f64p), _p(tv, i32p), _p(fn, i32p), _p(nrm, f64p)
            g.ntets = len(tv)
            box = tuple(np.asarray(kw["box"], dtype=np.float64))
        else:
            raise ValueError(kind)
        g.box = (C.c_double * 6)(*box)
        self.g = g


def trace(grid: Grid, rays, cap=None):
    rays = np.ascontiguousarray(rays, dtype=np.float64)
    n = len(rays)
    cap = cap or max(1024, n * 4096)
    off = np.zeros(n + 1, np.int64)
    m = np.zeros(cap, np.int32)
    ds = np.zeros(cap, np.float64)
    t = lib().orc_trace(C.byref(grid.g), _p(rays, f64p), C.c_int64(n), C.c_int64(cap), _p(off, i64p),
                        _p(m, i32p), _p(ds, f64p))
    if t < 0:
        raise RuntimeError("oracle trace capacity exceeded")
    return off, m[:t].copy(), ds[:t].copy()


def tau(grid: Grid, density, section, rays, L=1.0, cap=None):
    rays = np.ascontiguousarray(rays, dtype=np.float64)
    density = np.ascontiguousarray(density, dtype=np.float64)
    n = len(rays)
    cap = cap or max(1024, n * 4096)
    toff = np.zeros(n + 1, np.int64)
    tv = np.zeros(cap)
    sv = np.zeros(cap)
    mv = np.zeros(cap, np.int32)
    tobs = np.zeros(n)
    t = lib().orc_tau(C.byref(grid.g), _p(density, f64p), C.c_double(section), _p(rays, f64p), C.c_int64(n),
                      C.c_double(L), C.c_int64(cap), _p(toff, i64p), _p(tv, f64p), _p(sv, f64p), _p(mv, i32p),
                      _p(tobs, f64p))
    if t < 0:
        raise RuntimeError("oracle tau capacity exceeded")
    return toff, tv[:t].copy(), sv[:t].copy(), mv[:t].copy(), tobs


def cell_index(grid: Grid, pts):
    """SpatialGrid::cellIndex of each point (Voronoi and tetra grids)."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    out = np.zeros(len(pts), np.int32)
    lib().orc_cell_index(C.byref(grid.g), _p(pts, f64p), C.c_int64(len(pts)), _p(out, i32p))
    return out


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox(_p(c, P(C.c_uint32)), _p(k, P(C.c_uint32)), _p(out, P(C.c_uint32)))
    return out


def uniforms(seed, history, n, tag=0):
    out = np.zeros(n)
    lib().orc_uniforms(C.c_uint64(seed), C.c_uint64(history), C.c_uint32(tag), C.c_int64(n), _p(out, f64p))
    return out


def phase_hg(g, c):
    return lib().orc_phase_hg(float(g), float(c))


def scatter_hg(seed, history, g, k, n):
    k = np.ascontiguousarray(k, dtype=np.float64)
    out = np.zeros((n, 3))
    lib().orc_scatter_hg(C.c_uint64(seed), C.c_uint64(history), C.c_double(g), _p(k, f64p), C.c_int64(n),
                         _p(out, f64p))
    return out


def expon_cutoff(seed, history, xmax, n):
    out = np.zeros(n)
    lib().orc_expon_cutoff(C.c_uint64(seed), C.c_uint64(history), C.c_double(xmax), C.c_int64(n), _p(out, f64p))
    return out


def lnmean(x1, x2):
    """SpecialFunctions::lnmean as the oracle states it (radiation-field storage)."""
    f = lib().orc_lnmean
    f.restype = C.c_double
    return f(C.c_double(x1), C.c_double(x2))


def spline_table(n=400):
    """CubicSplineSmoothingKernel::setupSelfBefore (CubicSplineSmoothingKernel.cpp:22-40): the
    cumulative mass fraction inside radius u = k / n, k = 0..n."""
    du = 1.0 / n
    X = np.zeros(n + 1)
    for k in range(n + 1):
        u = k * du
        u2 = u * u
        u3 = u * u2
        if u < 0.5:
            X[k] = u3 * (32.0 / 3.0 - 192.0 / 5.0 * u2 + 32.0 * u3)
        else:
            X[k] = -1.0 / 15.0 + u3 * (64.0 / 3.0 - 48.0 * u + 192.0 / 5.0 * u2 - 32.0 / 3.0 * u3)
    return X


def _round_half_away(x):
    f = math.floor(x)
    return int(f) + (1 if x - f >= 0.5 else 0)


def particle_launch(ltot, lbin, num_packets, source_bias=0.5):
    """ImportedSource::prepareForLaunch (ImportedSource.cpp:290-322): launch weights and the
    first history index of each particle; and the launch luminosity per packet of particle m in
    bin l, L_m,l * nlambda / (N W_m) (the selection weight L_m / W_m of launchFromEntity,
    :660-661, times the oligochromatic bias weight s / b for a uniform bin choice, :670-705)."""
    ltot = np.asarray(ltot, dtype=np.float64)
    lbin = np.asarray(lbin, dtype=np.float64).reshape(len(ltot), -1)
    n, nl = lbin.shape
    total = 0.0
    for v in ltot:
        total += v
    Wv = (1 - source_bias) * (ltot / total) + source_bias / n
    Iv = np.zeros(n + 1, np.int64)
    W = 0.0
    for m in range(1, n):
        W += Wv[m - 1]
        Iv[m] = min(num_packets, _round_half_away(W * num_packets))
    Iv[n] = num_packets
    lp = np.where(ltot[:, None] > 0, (lbin * nl) / (float(num_packets) * Wv[:, None]), 0.0)
    return Iv, lp, Wv


def simulate(grid: Grid, density, lam, sext, ssca, g, lpack, instruments, first, count, threads=8, seed=1,
             xi=0.5, min_weight_reduction=1e4, min_scatt_events=0, src=(0.0, 0.0, 0.0), store_rf=False,
             particles=None):
    """Run histories [first, first+count); returns (frames list, counters), and the stored
    radiation field (ncells, nlambda) as a third item when store_rf.

    instruments: list of dicts with inclination, azimuth, roll (rad), nx, ny, fovx, fovy, cx, cy (m),
    or, for an SED instrument, sed=True, inclination, azimuth and an aperture radius (m, 0 = none),
    or, for a perspective instrument, persp=True, M (4x4 row-vector transform), eye, psize, nx, ny.
    Each frame has shape (nlambda, ny, nx) (an SED: (nlambda, 1, 1))."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (density, lam, sext, ssca, g, lpack)]
    ins = (OrcInstrument * len(instruments))()
    for i, d in enumerate(instruments):
        sed = bool(d.get("sed", False))
        persp = bool(d.get("persp", False))
        ins[i] = OrcInstrument(d.get("inclination", 0.0), d.get("azimuth", 0.0), d.get("roll", 0.0),
                               1 if sed else d["nx"], 1 if sed else d["ny"], d.get("fovx", 1.0), d.get("fovy", 1.0),
                               d.get("cx", 0.0), d.get("cy", 0.0), 1 if sed else 0, d.get("radius", 0.0),
                               1 if persp else 0,
                               (C.c_double * 16)(*np.asarray(d["M"], dtype=np.float64).ravel()) if persp
                               else (C.c_double * 16)(),
                               (C.c_double * 3)(*d["eye"]) if persp else (C.c_double * 3)(), d.get("psize", 0.0))
    s = OrcSim()
    s.grid = C.pointer(grid.g)
    s.density = _p(arrs[0], f64p)
    s.nlambda = len(arrs[1])
    s.lam, s.sext, s.ssca, s.g, s.lpack = (_p(a, f64p) for a in arrs[1:])
    s.src = (C.c_double * 3)(*src)
    s.xi = xi
    s.min_weight_reduction = min_weight_reduction
    s.min_scatt_events = min_scatt_events
    s.ninstr = len(instruments)
    s.instr = ins
    s.seed = seed
    s.ncells = len(arrs[0])
    rf = np.zeros((len(arrs[0]), len(arrs[1]))) if store_rf else None
    s.rf = _p(rf, f64p) if store_rf else None
    keep = []
    if particles is not None:
        # particles: dict(xyzh, lbin, ltot, num_packets, source_bias=0.5, kernel=1)
        xyzh = np.ascontiguousarray(particles["xyzh"], dtype=np.float64).reshape(-1, 4)
        Iv, lp, _ = particle_launch(particles["ltot"], particles["lbin"], int(particles["num_packets"]),
                                    particles.get("source_bias", 0.5))
        X = spline_table()
        keep = [xyzh, np.ascontiguousarray(Iv), np.ascontiguousarray(lp), X]
        s.nent = len(xyzh)
        s.ent_kernel = int(particles.get("kernel", 1))
        s.ent, s.ent_first, s.ent_lpack, s.ent_X = (_p(keep[0], f64p), _p(keep[1], i64p), _p(keep[2], f64p),
                                                    _p(keep[3], f64p))
    n = lib().orc_frame_size(C.byref(s))
    frame = np.zeros(n)
    cnt = np.zeros(4, np.int64)
    lib().orc_simulate(C.byref(s), C.c_uint64(first), C.c_uint64(count), C.c_int(threads), _p(frame, f64p),
                       _p(cnt, i64p))
    frames = []
    o = 0
    nl = len(arrs[1])
    for d in instruments:
        nx, ny = (1, 1) if d.get("sed", False) else (d["nx"], d["ny"])
        sz = nx * ny * nl
        frames.append(frame[o:o + sz].reshape(nl, ny, nx))
        o += sz
    if store_rf:
        return frames, cnt, rf
    return frames, cnt


def child_first_from(child_begin, child_count, child_index):
    """First child id per node (-1 for leaves) from the flattened reference children lists."""
    cb = np.asarray(child_begin)
    cc = np.asarray(child_count)
    ci = np.asarray(child_index)
    out = np.full(len(cb), -1, np.int32)
    has = cc > 0
    out[has] = ci[cb[has]]
    return out
