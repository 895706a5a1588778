"""Host-side binding of the MI355X photon-packet engine (libskirt_mi355x.so).

The engine is C++/HIP behind the C-ABI in include/skirt_mi355x.h; this module is the
thin ctypes layer the tests and the benchmark use.  Class and method names follow the
reference's seams: a SpatialGrid with path segments, a MediumSystem with cell densities
and dust, FrameInstrument detection and the MonteCarloSimulation life cycle.

The product path never falls back to a CPU implementation: if the shared library is
missing, or no GPU is present, the calls raise.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libskirt_mi355x.so")

SKG_OK = 0
SKG_GRID_CARTESIAN = 0
SKG_GRID_OCTREE = 1
SKG_FP64 = 0
SKG_FP32_KAHAN = 1

_lib = None

P = C.POINTER
f64p = P(C.c_double)
i32p = P(C.c_int32)
i64p = P(C.c_int64)


class EngineError(RuntimeError):
    pass


def build(arch: str = "gfx950") -> None:
    """Compile the HIP library in-tree (hipcc --offload-arch=gfx950)."""
    subprocess.run(["make", "-C", HERE, "-s", f"ARCH={arch}", "-B"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EngineError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.skg_last_error.restype = C.c_char_p
        L.skg_last_error.argtypes = [C.c_void_p]
        L.skg_num_cells.restype = C.c_int64
        L.skg_num_cells.argtypes = [C.c_void_p]
        L.skg_last_kernel_ms.restype = C.c_double
        L.skg_last_kernel_ms.argtypes = [C.c_void_p]
        for name in ("skg_create",):
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


EXPORTED = [
    "skg_create", "skg_destroy", "skg_last_error", "skg_grid_cartesian", "skg_grid_octree", "skg_grid_voronoi",
    "skg_grid_tetra", "skg_num_cells",
    "skg_medium", "skg_dust", "skg_trace_paths", "skg_trace_paths_noalloc", "skg_trace_tau", "skg_frame_add", "skg_frame_remove_all",
    "skg_frame_clear", "skg_frame_read", "skg_source_point", "skg_source_particles", "skg_options", "skg_run_packets",
    "skg_run_packets_async", "skg_sync", "skg_counters", "skg_counters_n", "skg_tune", "skg_last_kernel_ms",
    "skg_disk_density", "skg_octree_build", "skg_octree_from_flags", "skg_octree_info", "skg_octree_export",
    "skg_octree_upload", "skg_octree_free", "skg_sample_density_cartesian", "skg_sample_density_octree",
    "skg_dust_meanlist", "skg_tune_pipeline", "skg_tune_queue", "skg_continuation_key", "skg_last_run_stats",
    "skg_guide_params", "skg_guide_read_params", "skg_guide_infer", "skg_guide_gradient", "skg_guide_mode",
    "skg_guide_collect", "skg_guide_records", "skg_guide_train", "skg_guide_sample", "skg_guide_target",
    "skg_voronoi_build", "skg_voronoi_info", "skg_voronoi_export", "skg_voronoi_upload",
    "skg_sample_density_voronoi", "skg_voronoi_free", "skg_autotune",
    "skg_tetra_build", "skg_tetra_info", "skg_tetra_export", "skg_tetra_upload", "skg_tetra_free",
    "skg_sample_density_tetra", "skg_instances", "skg_rf_enable", "skg_rf_clear", "skg_rf_read",
    "skg_sed_add", "skg_sample_density", "skg_octree_build_gpu", "skg_sample_density_voronoi_model",
    "skg_perspective_add",
]

GUIDE_PARAMS = 11560

SKG_MODE_PIPELINE = 0
SKG_MODE_PERSISTENT = 1


def _p(a, t):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Engine:
    """One engine drives one GPU (GpuAcceleration's per-device runtime)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        st = lib().skg_create(C.c_int(device), C.byref(self._h))
        if st != SKG_OK:
            raise EngineError(lib().skg_last_error(None).decode())
        self.nlambda = 0
        self.frames = []

    def close(self):
        if self._h:
            lib().skg_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        if st < 0:
            raise EngineError(f"status {st}: {lib().skg_last_error(self._h).decode()}")
        return st

    # ---- grids
    def grid_cartesian(self, xb, yb, zb):
        xb, yb, zb = _f64(xb), _f64(yb), _f64(zb)
        self._chk(lib().skg_grid_cartesian(self._h, _p(xb, f64p), C.c_int32(len(xb) - 1), _p(yb, f64p),
                                           C.c_int32(len(yb) - 1), _p(zb, f64p), C.c_int32(len(zb) - 1)))

    def grid_octree(self, bounds, child_first, cell_index, nb_off, nb_id):
        bounds = _f64(bounds).reshape(-1, 6)
        cf = np.ascontiguousarray(child_first, dtype=np.int32)
        ci = np.ascontiguousarray(cell_index, dtype=np.int32)
        nbo = np.ascontiguousarray(nb_off, dtype=np.int64)
        nbi = np.ascontiguousarray(nb_id, dtype=np.int32)
        self._chk(lib().skg_grid_octree(self._h, C.c_int64(len(bounds)), _p(bounds, f64p), _p(cf, i32p),
                                        _p(ci, i32p), _p(nbo, i64p), _p(nbi, i32p)))

    def grid_voronoi(self, sites, nb_begin, nb_count, nb_index, block_n, block_begin, block_count, block_index, box):
        """Voronoi grid from VoronoiMeshSpatialGrid's traversal tables (walls in nb_index as -1..-6)."""
        sites = _f64(sites).reshape(-1, 3)
        a = [np.ascontiguousarray(x, dtype=np.int32) for x in
             (nb_begin, nb_count, nb_index, block_begin, block_count, block_index)]
        box = _f64(box)
        self._chk(lib().skg_grid_voronoi(self._h, C.c_int64(len(sites)), _p(sites, f64p), _p(a[0], i32p),
                                         _p(a[1], i32p), _p(a[2], i32p), C.c_int32(int(block_n)), _p(a[3], i32p),
                                         _p(a[4], i32p), _p(a[5], i32p), _p(box, f64p)))

    def grid_tetra(self, verts, tet_verts, face_nb, face_normal, box):
        """Tetrahedral grid from TetraMeshSpatialGrid's traversal tables."""
        verts = _f64(verts).reshape(-1, 3)
        tv = np.ascontiguousarray(tet_verts, dtype=np.int32).reshape(-1, 4)
        fn = np.ascontiguousarray(face_nb, dtype=np.int32).reshape(-1, 4)
        nrm = _f64(face_normal).reshape(-1, 12)
        box = _f64(box)
        self._chk(lib().skg_grid_tetra(self._h, C.c_int64(len(verts)), _p(verts, f64p), C.c_int64(len(tv)),
                                       _p(tv, i32p), _p(fn, i32p), _p(nrm, f64p), _p(box, f64p)))

    def num_cells(self):
        return int(lib().skg_num_cells(self._h))

    # ---- medium
    def medium(self, density):
        d = _f64(density)
        self._chk(lib().skg_medium(self._h, C.c_int64(len(d)), _p(d, f64p)))

    def dust(self, lam, sext, ssca, g):
        a = [_f64(x) for x in (lam, sext, ssca, g)]
        self.nlambda = len(a[0])
        self._chk(lib().skg_dust(self._h, C.c_int32(self.nlambda), *[_p(x, f64p) for x in a]))

    # ---- path service
    def trace_paths_noalloc(self, rays_ptr, n, cap, off_ptr, m_ptr, ds_ptr):
        """skg_trace_paths_noalloc on caller-owned DEVICE buffers (integer addresses): rays
        (n x 6 f64), off (n + 1 i64), m (cap i32), ds (cap f64); returns the segment total."""
        total = C.c_int64(0)
        self._chk(lib().skg_trace_paths_noalloc(self._h, C.c_void_p(rays_ptr), C.c_int64(n), C.c_int64(cap),
                                                C.c_void_p(off_ptr), C.c_void_p(m_ptr), C.c_void_p(ds_ptr),
                                                C.byref(total)))
        return total.value

    def trace_paths(self, rays, cap=None):
        rays = _f64(rays).reshape(-1, 6)
        n = len(rays)
        off = np.zeros(n + 1, np.int64)
        total = C.c_int64(0)
        cap = cap if cap is not None else n * 64
        m = np.zeros(max(cap, 1), np.int32)
        ds = np.zeros(max(cap, 1), np.float64)
        st = lib().skg_trace_paths(self._h, _p(rays, f64p), C.c_int64(n), C.c_int64(cap), _p(off, i64p),
                                   _p(m, i32p), _p(ds, f64p), C.byref(total))
        if st == -3:  # capacity: retry with the exact size
            cap = total.value
            m = np.zeros(max(cap, 1), np.int32)
            ds = np.zeros(max(cap, 1), np.float64)
            st = lib().skg_trace_paths(self._h, _p(rays, f64p), C.c_int64(n), C.c_int64(cap), _p(off, i64p),
                                       _p(m, i32p), _p(ds, f64p), C.byref(total))
        self._chk(st)
        t = total.value
        return off, m[:t].copy(), ds[:t].copy()

    def trace_tau(self, rays, ell=0, L=1.0, fp_mode=SKG_FP64, toff=None):
        rays = _f64(rays).reshape(-1, 6)
        n = len(rays)
        tp = np.zeros(n)
        to = np.zeros(n)
        if toff is not None:
            toff = np.ascontiguousarray(toff, dtype=np.int64)
            cum = np.zeros(max(int(toff[-1]), 1))
            st = lib().skg_trace_tau(self._h, _p(rays, f64p), C.c_int64(n), C.c_int32(ell), C.c_double(L),
                                     C.c_int32(fp_mode), _p(tp, f64p), _p(to, f64p), _p(toff, i64p), _p(cum, f64p))
            self._chk(st)
            return tp, to, cum[:int(toff[-1])]
        st = lib().skg_trace_tau(self._h, _p(rays, f64p), C.c_int64(n), C.c_int32(ell), C.c_double(L),
                                 C.c_int32(fp_mode), _p(tp, f64p), _p(to, f64p), None, None)
        self._chk(st)
        return tp, to

    # ---- instruments
    def frame_add(self, inclination, azimuth, roll, nx, ny, fovx, fovy, cx=0.0, cy=0.0):
        idx = self._chk(lib().skg_frame_add(self._h, C.c_double(inclination), C.c_double(azimuth), C.c_double(roll),
                                            C.c_int32(nx), C.c_int32(ny), C.c_double(fovx), C.c_double(fovy),
                                            C.c_double(cx), C.c_double(cy)))
        self.frames.append((nx, ny))
        return idx

    def sed_add(self, inclination, azimuth, radius=0.0):
        """SED instrument (skg_sed_add): frame_read returns its (nlambda, 1, 1) values."""
        idx = self._chk(lib().skg_sed_add(self._h, C.c_double(inclination), C.c_double(azimuth), C.c_double(radius)))
        self.frames.append((1, 1))
        return idx

    def perspective_add(self, transform, eye, pixel_size, nx, ny):
        """Perspective instrument (skg_perspective_add); transform: 4x4 row-vector matrix."""
        M = _f64(transform).reshape(16)
        E = _f64(eye).reshape(3)
        idx = self._chk(lib().skg_perspective_add(self._h, _p(M, f64p), _p(E, f64p), C.c_double(pixel_size),
                                                  C.c_int32(nx), C.c_int32(ny)))
        self.frames.append((nx, ny))
        return idx

    def frame_remove_all(self):
        self._chk(lib().skg_frame_remove_all(self._h))
        self.frames = []

    def frame_clear(self):
        self._chk(lib().skg_frame_clear(self._h))

    def frame_read(self, index):
        nx, ny = self.frames[index]
        out = np.zeros(self.nlambda * nx * ny)
        self._chk(lib().skg_frame_read(self._h, C.c_int32(index), _p(out, f64p), C.c_int64(len(out))))
        return out.reshape(self.nlambda, ny, nx)

    # ---- source, options, life cycle
    def source_point(self, pos, lpack):
        lp = _f64(lpack)
        self._chk(lib().skg_source_point(self._h, C.c_double(pos[0]), C.c_double(pos[1]), C.c_double(pos[2]),
                                         _p(lp, f64p)))

    def instances(self, packets, seeds=None, tau_scale=None, lpack=None, frame_begin=None, frame_count=None):
        """Batched independent instances in one persistent launch (skg_instances); an empty
        `packets` removes them.  lpack: (n, nlambda) launch luminosities per instance."""
        pk = np.ascontiguousarray(packets, dtype=np.uint64).ravel()
        n = len(pk)
        if n == 0:
            self._chk(lib().skg_instances(self._h, C.c_int32(0), None, None, None, None, None, None))
            return 0
        sd = np.ascontiguousarray(seeds, dtype=np.uint64).ravel()
        ts = _f64(tau_scale).ravel()
        lp = _f64(lpack).reshape(n, -1)
        fb = np.ascontiguousarray(frame_begin, dtype=np.int32).ravel()
        fc = np.ascontiguousarray(frame_count, dtype=np.int32).ravel()
        if not (len(sd) == len(ts) == len(fb) == len(fc) == n):
            raise EngineError("instances: all per-instance arrays need one entry per instance")
        u64p = P(C.c_uint64)
        self._chk(lib().skg_instances(self._h, C.c_int32(n), _p(pk, u64p), _p(sd, u64p), _p(ts, f64p), _p(lp, f64p),
                                      _p(fb, i32p), _p(fc, i32p)))
        return int(pk.sum())

    def sample_density(self, model, num_samples=100, seed=1):
        """Cell-averaged number densities of the resident grid's cells on the GPU
        (skg_sample_density; Cartesian, octree and tetrahedral grids)."""
        out = np.zeros(self.num_cells())
        self._chk(lib().skg_sample_density(self._h, C.byref(model), C.c_int32(num_samples), C.c_uint64(seed),
                                           _p(out, f64p), C.c_int64(len(out))))
        return out

    def rf_enable(self, on=True):
        """Radiation-field storage (skg_rf_enable): a cleared ncells x nlambda table."""
        self._chk(lib().skg_rf_enable(self._h, C.c_int32(1 if on else 0)))

    def rf_clear(self):
        self._chk(lib().skg_rf_clear(self._h))

    def rf_read(self):
        """The stored radiation field, shape (ncells, nlambda): sum of L * mean extinction * ds."""
        out = np.zeros(self.num_cells() * self.nlambda)
        self._chk(lib().skg_rf_read(self._h, _p(out, f64p), C.c_int64(len(out))))
        return out.reshape(-1, self.nlambda)

    def source_particles(self, xyzh, lbin, ltot, num_packets, source_bias=0.5, kernel=1):
        """Imported particle source (skg_source_particles); kernel 0 uniform, 1 cubic spline.
        An empty xyzh restores the point source."""
        X = _f64(xyzh).reshape(-1, 4)
        n = len(X)
        B = _f64(lbin).reshape(n, -1) if n else np.zeros(0)
        T = _f64(ltot).reshape(n)
        self._chk(lib().skg_source_particles(self._h, C.c_int64(n), _p(X, f64p), _p(B, f64p), _p(T, f64p),
                                             C.c_double(source_bias), C.c_int32(kernel), C.c_int64(num_packets)))

    def options(self, path_length_bias=0.5, min_weight_reduction=1e4, min_scatt_events=0, fp_mode=SKG_FP64, seed=1):
        self._chk(lib().skg_options(self._h, C.c_double(path_length_bias), C.c_double(min_weight_reduction),
                                    C.c_int32(min_scatt_events), C.c_int32(fp_mode), C.c_uint64(seed)))

    def tune(self, workgroups=0, threads=256, threshold=16):
        self._chk(lib().skg_tune(self._h, C.c_int32(workgroups), C.c_int32(threads), C.c_int32(threshold)))

    def tune_pipeline(self, mode=0, pool_slots=1 << 21, trace_workgroups=0, graph_iterations=8):
        self._chk(lib().skg_tune_pipeline(self._h, C.c_int32(mode), C.c_int64(pool_slots), C.c_int32(trace_workgroups),
                                          C.c_int32(graph_iterations)))

    def continuation_key(self, key=0):
        """Histories continue on Philox tag `key` from their first scattering on (0: off)."""
        self._chk(lib().skg_continuation_key(self._h, C.c_uint32(key)))

    def tune_queue(self, max_requests=0):
        """Cap the pipeline's request queue at max_requests entries (0: automatic, enough for any run)."""
        self._chk(lib().skg_tune_queue(self._h, C.c_int64(max_requests)))

    def last_run_stats(self):
        ms, launches, iters = C.c_double(), C.c_int64(), C.c_int64()
        self._chk(lib().skg_last_run_stats(self._h, C.byref(ms), C.byref(launches), C.byref(iters)))
        return ms.value, launches.value, iters.value

    # ---- path guide
    def guide_params(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        self._chk(lib().skg_guide_params(self._h, _p(t, P(C.c_float)), C.c_int32(len(t))))

    def guide_read_params(self):
        t = np.zeros(GUIDE_PARAMS, np.float32)
        self._chk(lib().skg_guide_read_params(self._h, _p(t, P(C.c_float)), C.c_int32(len(t))))
        return t

    def guide_infer(self, feats):
        f = np.ascontiguousarray(feats, dtype=np.float32).reshape(-1, 9)
        out = np.zeros((len(f), 40), np.float32)
        self._chk(lib().skg_guide_infer(self._h, _p(f, P(C.c_float)), C.c_int64(len(f)), _p(out, P(C.c_float))))
        return out

    def guide_sample(self, outputs, seed=1, at=None):
        """The life cycle's FP32 mixture sampler for rows of raw network outputs: returns the
        sampled directions (n x 3), the density there (n) and, with `at`, the density at those
        directions (n)."""
        o = np.ascontiguousarray(outputs, dtype=np.float32).reshape(-1, 40)
        n = len(o)
        dirs = np.zeros((n, 3))
        pdf = np.zeros(n)
        a = None if at is None else _f64(at).reshape(n, 3)
        pa = None if at is None else np.zeros(n)
        self._chk(lib().skg_guide_sample(self._h, _p(o, P(C.c_float)), C.c_int64(n), C.c_uint64(seed),
                                         None if a is None else _p(a, f64p), _p(dirs, f64p), _p(pdf, f64p),
                                         None if pa is None else _p(pa, f64p)))
        return (dirs, pdf) if at is None else (dirs, pdf, pa)

    def guide_gradient(self, records):
        r = np.ascontiguousarray(records, dtype=np.float32).reshape(-1, 16)
        g = np.zeros(GUIDE_PARAMS, np.float32)
        loss = C.c_double()
        self._chk(lib().skg_guide_gradient(self._h, _p(r, P(C.c_float)), C.c_int64(len(r)), _p(g, P(C.c_float)),
                                           C.byref(loss)))
        return g, loss.value

    def guide_mode(self, mode, mix=0.5, half_size=10 * 3.08567758e16, lambda_center=0.775e-6, lambda_scale=1.0):
        self._chk(lib().skg_guide_mode(self._h, C.c_int32(mode), C.c_double(mix), C.c_double(half_size),
                                       C.c_double(lambda_center), C.c_double(lambda_scale)))

    def guide_collect(self, capacity):
        self._chk(lib().skg_guide_collect(self._h, C.c_int64(capacity)))

    def guide_target(self, mode, floor_frac=0.1, weights=None):
        """Training target of collect runs (skg_guide_target): 0 the deposited luminosity, 1
        luminosity / max(F, floor) per frame entry from the image accumulated so far, 2 the
        given weights (one per frame entry, in frame-read order)."""
        w = None if weights is None else _f64(weights).ravel()
        self._chk(lib().skg_guide_target(self._h, C.c_int32(mode), C.c_double(floor_frac),
                                         None if w is None else _p(w, f64p), C.c_int64(0 if w is None else w.size)))

    def guide_records(self, copy=0):
        n = C.c_int64()
        if copy:
            out = np.zeros((copy, 16), np.float32)
            self._chk(lib().skg_guide_records(self._h, C.byref(n), _p(out, P(C.c_float)), C.c_int64(copy)))
            return n.value, out[:min(copy, n.value)]
        self._chk(lib().skg_guide_records(self._h, C.byref(n), None, C.c_int64(0)))
        return n.value

    def guide_train(self, steps, batch=4096, lr=1e-3):
        a, b = C.c_double(), C.c_double()
        self._chk(lib().skg_guide_train(self._h, C.c_int32(steps), C.c_int32(batch), C.c_float(lr), C.byref(a),
                                        C.byref(b)))
        return a.value, b.value

    def autotune(self, first, trial_packets=1 << 20, engines=3, cache=None, key=None):
        """Picks the fastest launch configuration (engine, transition threshold, workgroups,
        pool slots, trace workgroups, graph length) from timed trial runs of histories
        [first, ...) and keeps it; the detector contents are restored.  With `cache` (a JSON
        file) and `key`, a configuration stored under the key is applied without trials and
        new results are stored.  Returns (config dict, trial packets/s or None)."""
        import json
        names = ("mode", "threshold", "workgroups", "pool_slots", "trace_workgroups", "graph_iterations")
        if cache and key and os.path.exists(cache):
            try:
                entry = json.load(open(cache)).get(key)
            except (OSError, ValueError):
                entry = None
            if entry:
                self.tune(int(entry["workgroups"]), 256, int(entry["threshold"]))
                self.tune_pipeline(int(entry["mode"]), int(entry["pool_slots"]), int(entry["trace_workgroups"]),
                                   int(entry["graph_iterations"]))
                return {k: int(entry[k]) for k in names}, None
        best = np.zeros(6, np.int64)
        rate = C.c_double()
        self._chk(lib().skg_autotune(self._h, C.c_uint64(first), C.c_uint64(trial_packets), C.c_int32(engines),
                                     _p(best, i64p), C.byref(rate)))
        cfg = {k: int(v) for k, v in zip(names, best)}
        if cache and key:
            data = {}
            if os.path.exists(cache):
                try:
                    data = json.load(open(cache))
                except (OSError, ValueError):
                    data = {}
            data[key] = dict(cfg, rate=rate.value)
            with open(cache, "w") as f:
                json.dump(data, f, indent=1)
        return cfg, rate.value

    def run_packets(self, first, count):
        self._chk(lib().skg_run_packets(self._h, C.c_uint64(first), C.c_uint64(count)))

    def run_packets_async(self, first, count):
        self._chk(lib().skg_run_packets_async(self._h, C.c_uint64(first), C.c_uint64(count)))

    def sync(self):
        self._chk(lib().skg_sync(self._h))

    def counters(self, n=8):
        """Device event counters (skg_counters_n): [0] histories, [1] peel-off traversals,
        [2] first forced passes, [3] scatterings, [4] cell crossings, [5] detector atomics issued,
        [6] second forced passes, [7] crossings of second passes, [8] detector deposits."""
        out = np.zeros(n, np.int64)
        self._chk(lib().skg_counters_n(self._h, _p(out, i64p), C.c_int32(n)))
        return out

    def last_kernel_ms(self):
        return float(lib().skg_last_kernel_ms(self._h))


# ---------------------------------------------------------------------------- host set-up
class Voronoi:
    """Host-built Voronoi tessellation (neighbour lists, extents, block lists) in the engine's
    table layout; cell m is the m-th retained site in x order."""

    def __init__(self, handle, box):
        self._h = handle
        self.box = np.array(box, dtype=np.float64)

    @classmethod
    def build(cls, sites, box, threads=0):
        s = _f64(sites).reshape(-1, 3)
        b = _f64(box)
        h = C.c_void_p()
        st = lib().skg_voronoi_build(C.c_int64(len(s)), _p(s, f64p), _p(b, f64p), C.c_int32(threads or os.cpu_count() or 1),
                                     C.byref(h))
        if st != SKG_OK:
            raise EngineError(f"skg_voronoi_build failed with status {st}")
        return cls(h, b)

    def info(self):
        a, b, c, d = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int64()
        lib().skg_voronoi_info(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return a.value, b.value, c.value, d.value

    def export(self):
        n, nnb, bn, nbr = self.info()
        out = dict(sites=np.zeros((n, 3)), nb_begin=np.zeros(n, np.int32), nb_count=np.zeros(n, np.int32),
                   nb_index=np.zeros(max(nnb, 1), np.int32), block_n=np.int32(bn),
                   block_begin=np.zeros(bn ** 3, np.int32), block_count=np.zeros(bn ** 3, np.int32),
                   block_index=np.zeros(max(nbr, 1), np.int32), volume=np.zeros(n), box=self.box.copy())
        lib().skg_voronoi_export(self._h, _p(out["sites"], f64p), _p(out["nb_begin"], i32p), _p(out["nb_count"], i32p),
                                 _p(out["nb_index"], i32p), _p(out["block_begin"], i32p), _p(out["block_count"], i32p),
                                 _p(out["block_index"], i32p), _p(out["volume"], f64p))
        out["nb_index"] = out["nb_index"][:nnb]
        out["block_index"] = out["block_index"][:nbr]
        return out

    def upload(self, engine):
        engine._chk(lib().skg_voronoi_upload(engine._h, self._h))

    def sample_density_model(self, model, samples=100, seed=1, threads=0):
        """Cell-averaged number densities of a disk model (skg_sample_density_voronoi_model)."""
        n = self.info()[0]
        out = np.zeros(n)
        st = lib().skg_sample_density_voronoi_model(self._h, C.byref(model), C.c_int32(samples), C.c_uint64(seed),
                                                    C.c_int32(threads), _p(out, f64p))
        if st != SKG_OK:
            raise EngineError(f"skg_sample_density_voronoi_model failed with status {st}")
        return out

    def sample_density(self, lattice, samples=100, seed=1, threads=0):
        lat = _f64(lattice)
        nx, ny, nz = lat.shape
        out = np.zeros(self.info()[0])
        st = lib().skg_sample_density_voronoi(self._h, _p(lat, f64p), C.c_int32(nx), C.c_int32(ny), C.c_int32(nz),
                                              C.c_int32(samples), C.c_uint64(seed),
                                              C.c_int32(threads or os.cpu_count() or 1), _p(out, f64p))
        if st != SKG_OK:
            raise EngineError(f"skg_sample_density_voronoi failed with status {st}")
        return out

    def close(self):
        if self._h:
            lib().skg_voronoi_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DiskModel(C.Structure):
    """Exponential disk with optional spiral perturbation, normalised to face-on tau_v."""
    _fields_ = [("scale_length", C.c_double), ("scale_height", C.c_double), ("spiral_arms", C.c_int32),
                ("spiral_index", C.c_int32), ("pitch_angle", C.c_double), ("radius_zero", C.c_double),
                ("phase_zero", C.c_double), ("perturbation_weight", C.c_double), ("tau_v", C.c_double),
                ("sext_v", C.c_double)]


PC = 3.08567758e16  # parsec (m)


def disk_model(tau_v=1.0, sext_v=1.5e-29, spiral=False, hR=3 * PC, hz=0.3 * PC, scale_length=None, scale_height=None,
               arms=2, pitch=np.radians(15.0), radius_zero=3 * PC, phase_zero=0.0, weight=0.5, index=1):
    """Exponential disk normalised to face-on tau_V (the workloads' h_R = 3 pc, h_z = 0.3 pc), with
    an optional logarithmic spiral perturbation (default 2 arms, pitch 15 deg, R0 = 3 pc,
    weight 0.5, index 1)."""
    hR = scale_length if scale_length is not None else hR
    hz = scale_height if scale_height is not None else hz
    return DiskModel(hR, hz, arms if spiral else 0, index, pitch, radius_zero, phase_zero, weight, tau_v, sext_v)


def disk_density(model, x, y, z):
    lib().skg_disk_density.restype = C.c_double
    return lib().skg_disk_density(C.byref(model), C.c_double(x), C.c_double(y), C.c_double(z))


class Octree:
    """Host-built octree (DensityTreePolicy semantics) in the engine's table layout."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def build_gpu(cls, engine, box, model, min_level, max_level, max_dust_fraction, num_samples=100, seed=1, threads=0):
        """The same tree with each level's density estimates computed on the engine's GPU."""
        b = _f64(box)
        h = C.c_void_p()
        engine._chk(lib().skg_octree_build_gpu(engine._h, _p(b, f64p), C.byref(model), C.c_int32(min_level),
                                               C.c_int32(max_level), C.c_double(max_dust_fraction),
                                               C.c_int32(num_samples), C.c_uint64(seed), C.c_int32(threads),
                                               C.byref(h)))
        return cls(h)

    @classmethod
    def build(cls, box, model, min_level, max_level, max_dust_fraction, num_samples=100, seed=1, threads=0):
        b = _f64(box)
        h = C.c_void_p()
        st = lib().skg_octree_build(_p(b, f64p), C.byref(model), C.c_int32(min_level), C.c_int32(max_level),
                                    C.c_double(max_dust_fraction), C.c_int32(num_samples), C.c_uint64(seed),
                                    C.c_int32(threads), C.byref(h))
        if st != SKG_OK:
            raise EngineError(f"skg_octree_build failed with status {st}")
        return cls(h)

    @classmethod
    def from_flags(cls, box, divide):
        b = _f64(box)
        d = np.ascontiguousarray(divide, dtype=np.uint8)
        h = C.c_void_p()
        st = lib().skg_octree_from_flags(_p(b, f64p), C.c_int64(len(d)), _p(d, P(C.c_uint8)), C.byref(h))
        if st != SKG_OK:
            raise EngineError(f"skg_octree_from_flags failed with status {st}")
        return cls(h)

    def info(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        lib().skg_octree_info(self._h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def export(self):
        nn, nc, nnb = self.info()
        out = dict(bounds=np.zeros((nn, 6)), child_first=np.zeros(nn, np.int32), cell_index=np.zeros(nn, np.int32),
                   nb_off=np.zeros(6 * nn + 1, np.int64), nb_id=np.zeros(max(nnb, 1), np.int32))
        lib().skg_octree_export(self._h, _p(out["bounds"], f64p), _p(out["child_first"], i32p),
                                _p(out["cell_index"], i32p), _p(out["nb_off"], i64p), _p(out["nb_id"], i32p))
        out["nb_id"] = out["nb_id"][:nnb]
        out["ncells"] = nc
        return out

    def upload(self, engine):
        engine._chk(lib().skg_octree_upload(engine._h, self._h))

    def sample_density(self, model, num_samples=100, seed=1, threads=0):
        nn, nc, _ = self.info()
        out = np.zeros(nc)
        st = lib().skg_sample_density_octree(self._h, C.byref(model), C.c_int32(num_samples), C.c_uint64(seed),
                                             C.c_int32(threads), _p(out, f64p))
        if st != SKG_OK:
            raise EngineError(f"skg_sample_density_octree failed with status {st}")
        return out

    def close(self):
        if self._h:
            lib().skg_octree_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tetra:
    """Host-built Delaunay tetrahedral grid (the 8 domain corners added) in the engine's layout."""

    def __init__(self, handle, box):
        self._h = handle
        self.box = np.array(box, dtype=np.float64)

    @classmethod
    def build(cls, points, box):
        pts = _f64(points).reshape(-1, 3)
        b = _f64(box)
        h = C.c_void_p()
        st = lib().skg_tetra_build(C.c_int64(len(pts)), _p(pts, f64p), _p(b, f64p), C.byref(h))
        if st != SKG_OK:
            raise EngineError(f"skg_tetra_build failed with status {st}")
        return cls(h, b)

    def info(self):
        a, b = C.c_int64(), C.c_int64()
        lib().skg_tetra_info(self._h, C.byref(a), C.byref(b))
        return a.value, b.value

    def export(self):
        nv, nt = self.info()
        out = dict(verts=np.zeros((nv, 3)), tet_verts=np.zeros((nt, 4), np.int32), face_nb=np.zeros((nt, 4), np.int32),
                   face_normal=np.zeros((nt, 4, 3)), box=self.box.copy())
        lib().skg_tetra_export(self._h, _p(out["verts"], f64p), _p(out["tet_verts"], i32p), _p(out["face_nb"], i32p),
                               _p(out["face_normal"], f64p))
        return out

    def upload(self, engine):
        engine._chk(lib().skg_tetra_upload(engine._h, self._h))

    def close(self):
        if self._h:
            lib().skg_tetra_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sample_density_tetra(verts, tet_verts, model, num_samples=100, seed=1, threads=0):
    """Cell-averaged number density of a disk model in each tetrahedron (skg_sample_density_tetra)."""
    v = _f64(verts).reshape(-1, 3)
    tv = np.ascontiguousarray(tet_verts, dtype=np.int32).reshape(-1, 4)
    if len(tv) and (tv.min() < 0 or tv.max() >= len(v)):
        raise EngineError("sample_density_tetra: vertex index out of range")
    out = np.zeros(len(tv))
    st = lib().skg_sample_density_tetra(C.c_int64(len(tv)), _p(v, f64p), _p(tv, i32p), C.byref(model),
                                        C.c_int32(num_samples), C.c_uint64(seed), C.c_int32(threads), _p(out, f64p))
    if st != SKG_OK:
        raise EngineError(f"skg_sample_density_tetra failed with status {st}")
    return out


def sample_density_cartesian(xb, yb, zb, model, num_samples=100, seed=1, threads=0):
    xb, yb, zb = _f64(xb), _f64(yb), _f64(zb)
    out = np.zeros((len(xb) - 1) * (len(yb) - 1) * (len(zb) - 1))
    st = lib().skg_sample_density_cartesian(_p(xb, f64p), C.c_int32(len(xb) - 1), _p(yb, f64p), C.c_int32(len(yb) - 1),
                                            _p(zb, f64p), C.c_int32(len(zb) - 1), C.byref(model),
                                            C.c_int32(num_samples), C.c_uint64(seed), C.c_int32(threads),
                                            _p(out, f64p))
    if st != SKG_OK:
        raise EngineError(f"skg_sample_density_cartesian failed with status {st}")
    return out


def dust_meanlist(lam_in, kappa_in, albedo_in, g_in, lam_out, mu=1.5e-29):
    """Dust cross sections and g at lam_out from a MeanListDustMix table (mu: kg per H, as the reference)."""
    a = [_f64(x) for x in (lam_in, kappa_in, albedo_in, g_in)]
    lo = _f64(lam_out)
    sext, ssca, g = np.zeros(len(lo)), np.zeros(len(lo)), np.zeros(len(lo))
    st = lib().skg_dust_meanlist(C.c_int32(len(a[0])), *[_p(x, f64p) for x in a], C.c_double(mu), C.c_int32(len(lo)),
                                 _p(lo, f64p), _p(sext, f64p), _p(ssca, f64p), _p(g, f64p))
    if st != SKG_OK:
        raise EngineError(f"skg_dust_meanlist failed with status {st}")
    return sext, ssca, g
