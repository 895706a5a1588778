"""skg_trace_paths_noalloc (the path service without device allocations, for callers that
cannot allocate) against skg_trace_paths on the same rays: identical offsets, cells and
segment lengths, bit for bit, on the reference's octree and Cartesian fixture grids, for two
calls in a row with the second one smaller (the scratch array keeps the first call's
contents); too small a capacity and too many rays are refused."""

import numpy as np
import pytest

from helpers import cart_borders, child_first, golden, load_pkg

pytestmark = pytest.mark.gpu


class _Dev:
    """Device buffers for the test through the HIP runtime the library itself uses (another
    runtime in the same process, such as torch's, would not share its device context)."""

    def __init__(self):
        import ctypes as C
        load_pkg().lib()
        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def put(self, a):
        C = self.C
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))) == 0
        self.bufs.append(p)
        if a.nbytes:
            assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0
        return p.value

    def get(self, ptr, count, dtype):
        C = self.C
        out = np.zeros(count, dtype=dtype)
        if out.nbytes:
            assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _noalloc(eng, rays, cap):
    dev = _Dev()
    try:
        n = len(rays)
        r = dev.put(np.asarray(rays, dtype=np.float64))
        off = dev.put(np.full(n + 1, -7, np.int64))
        m = dev.put(np.full(max(cap, 1), -7, np.int32))
        ds = dev.put(np.full(max(cap, 1), -7.0))
        total = eng.trace_paths_noalloc(r, n, cap, off, m, ds)
        return dev.get(off, n + 1, np.int64), dev.get(m, total, np.int32), dev.get(ds, total, np.float64), total
    finally:
        dev.free()


@pytest.mark.parametrize("grid", ["octree", "cartesian"])
def test_noalloc_paths_equal_trace_paths(grid):
    pkg = load_pkg()
    eng = pkg.Engine(0)
    try:
        if grid == "octree":
            f = golden("tree_paths.npz")
            cf = child_first(f["child_begin"], f["child_count"], f["child_index"])
            eng.grid_octree(f["bounds"], cf, f["cell_index"], f["nb_off"], f["nb_id"])
            eng.medium(f["density"])
        else:
            f = golden("cart_paths.npz")
            h = float(f["c64_half"])
            n = f["c64_n"]
            eng.grid_cartesian(cart_borders(n[0], -h, h), cart_borders(n[1], -h, h), cart_borders(n[2], -h, h))
        rays = np.asarray(f["rays"] if grid == "octree" else f["c64_rays"], dtype=np.float64)
        for case, sub in (("all rays", rays), ("a smaller second call", rays[: len(rays) // 3])):
            print(f"case noalloc paths, {grid}, {case}: {len(sub)} rays", flush=True)
            off, m, ds = eng.trace_paths(sub)
            cap = int(off[-1]) + 10
            off2, m2, ds2, total = _noalloc(eng, sub, cap)
            assert total == off[-1]
            np.testing.assert_array_equal(off2, off)
            np.testing.assert_array_equal(m2, m)
            np.testing.assert_array_equal(ds2, ds)
        if grid == "cartesian":
            # more than 1024 scan blocks: the block-sum scan takes several sums per thread
            h8, n8 = float(f["c8_half"]), f["c8_n"]
            eng.grid_cartesian(cart_borders(n8[0], -h8, h8), cart_borders(n8[1], -h8, h8), cart_borders(n8[2], -h8, h8))
            rng = np.random.default_rng(4)
            nr = 1100 * 1024 + 17
            k = rng.normal(size=(nr, 3))
            k /= np.linalg.norm(k, axis=1, keepdims=True)
            big = np.hstack([rng.uniform(-1.2 * h8, 1.2 * h8, (nr, 3)), k])
            print(f"case noalloc paths, {nr} random rays on the 8^3 grid", flush=True)
            off, m, ds = eng.trace_paths(big, cap=nr * 40)
            off2, m2, ds2, total = _noalloc(eng, big, int(off[-1]))
            assert total == off[-1]
            np.testing.assert_array_equal(off2, off)
            np.testing.assert_array_equal(m2, m)
            np.testing.assert_array_equal(ds2, ds)
        print(f"case noalloc refusals, {grid}", flush=True)
        with pytest.raises(pkg.EngineError, match="capacity"):
            _noalloc(eng, rays[:50], 3)
        with pytest.raises(pkg.EngineError, match="rays per call"):
            eng.trace_paths_noalloc(0, 16384 * 1024 + 1, 0, 0, 0, 0)
    finally:
        eng.close()
