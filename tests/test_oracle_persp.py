"""The oracle's perspective instrument against closed forms (CPU): direct light from a point
source in a uniform grid without scattering.  Every packet deposits L e^-tau / zp^2 into the
pixel the source projects to and nothing else (PerspectiveInstrument::detect, local flux
recorder, PerspectiveInstrument.cpp:123-143, FluxRecorder.cpp:372-405).  The optical depth
toward the eye adds whole path segments until the travelled length exceeds zp
(MediumSystem::getExtinctionOpticalDepth with a distance, MediumSystem.cpp:2374-2382), so in a
uniform grid tau = kappa * (the first cell-wall crossing beyond zp, or the domain exit): a
segment clipped at zp, a limit at the eye distance, or no limit at all each fail."""

import numpy as np
import pytest

import oracle as O
from helpers import load_pkg


@pytest.fixture(scope="module")
def ski():
    load_pkg()
    from skirt9_gpu_amd import ski as S
    return S


def _project(M, p):
    v = np.array([p[0], p[1], p[2], 1.0]) @ M
    return v[0] / v[3], v[1] / v[3], v[2]


def _s_end(src, eye, walls, zp):
    """Cumulative path length at the end of the first segment that ends beyond zp."""
    k = (eye - src) / np.linalg.norm(eye - src)
    ts = []
    for a in range(3):
        if k[a] != 0:
            t = (walls - src[a]) / k[a]
            ts.extend(t[t > 0])
    t_exit = min(((walls[-1] if k[a] > 0 else walls[0]) - src[a]) / k[a] for a in range(3) if k[a] != 0)
    ts = np.sort([t for t in ts if t <= t_exit])
    assert np.abs(ts - zp).min() > 1e-6  # no wall crossing at zp itself (a rounding tie)
    beyond = ts[ts > zp]
    return beyond[0] if len(beyond) else t_exit


def test_perspective_direct_light_closed_form(ski):
    n = 8
    walls = np.linspace(-1.0, 1.0, n + 1)
    grid = O.Grid("cartesian", xb=walls, yb=walls, zb=walls)
    kappa = 0.7
    src = np.array([0.13, -0.07, 0.04])
    cams = []
    for view, cross, up, fe, w, npix in (((0.47, -0.61, 0.33), tuple(src), (0, 0, 1), 0.25, 0.5, 9),  # eye inside
                                         ((-0.2, 0.6, 0.5), (0, 0, 0), (0, 0.2, 1), 0.1, 0.4, 15),  # eye inside
                                         ((2.5, 0.4, -1.8), (0, 0, 0), (0, 0, 1), 0.5, 3.0, 31)):   # eye outside
        M, E, s = ski.perspective_setup(np.array(view, float), np.array(cross, float), np.array(up, float),
                                        fe, w, npix, npix)
        cams.append(dict(persp=True, M=M, eye=E, psize=s, nx=npix, ny=npix))
    assert np.all(np.abs(cams[0]["eye"]) < 1) and np.all(np.abs(cams[1]["eye"]) < 1)
    assert np.any(np.abs(cams[2]["eye"]) > 1)
    npk = 3000
    args = ([1e-6], [kappa], [0.0], [0.0], [1.0 / npk], cams, 0, npk)
    full, _ = O.simulate(grid, np.ones(n ** 3), *args, threads=4, seed=7, src=tuple(src))
    empty, _ = O.simulate(grid, np.zeros(n ** 3), *args, threads=4, seed=7, src=tuple(src))
    unit = None
    for c, f, e in zip(cams, full, empty):
        x, y, zp = _project(c["M"], src)
        i, j = int(x), int(y)
        assert 0 <= i < c["nx"] and 0 <= j < c["ny"] and zp > c["psize"] / 10
        assert np.count_nonzero(f) == 1 and f[0, j, i] > 0 and np.count_nonzero(e) == 1
        tau = kappa * _s_end(src, c["eye"], walls, zp)
        assert abs(f[0, j, i] / e[0, j, i] / np.exp(-tau) - 1) < 1e-11, (f[0, j, i] / e[0, j, i], np.exp(-tau))
        # clipping the last segment at zp would give exp(-kappa zp): far from the rule above here
        assert abs(np.exp(-kappa * zp) / np.exp(-tau) - 1) > 1e-3
        u = e[0, j, i] * zp * zp  # the same luminosity for every camera: 1 / zp^2
        unit = u if unit is None else unit
        assert abs(u / unit - 1) < 1e-11
