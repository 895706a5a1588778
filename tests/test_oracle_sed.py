"""The oracle's SED instrument (SEDInstrument with ApertureInstrument's aperture) against a
frame on the same histories: without an aperture, or with one wider than the domain, the
SED holds exactly what a frame covering every position receives; a 3 pc aperture holds less;
an SED sharing the frame's observer direction reuses its optical depths without changing
the result (the same histories give the same SED with the frame absent)."""

import numpy as np

import oracle as O
from helpers import PC, disk_density_cartesian


def test_sed_equals_frame_covering_all_positions():
    b, dens = disk_density_cartesian(n=32, samples=10, seed=5, tau_v=2.0)
    s = 1.5e-29
    inc = np.radians(60.0)
    instr = [dict(inclination=inc, azimuth=0.0, roll=0.0, nx=40, ny=40, fovx=60 * PC, fovy=60 * PC),
             dict(sed=True, inclination=inc, azimuth=0.0, radius=0.0),
             dict(sed=True, inclination=inc, azimuth=0.0, radius=3 * PC),
             dict(sed=True, inclination=inc, azimuth=0.0, radius=1e30)]
    g = O.Grid("cartesian", xb=b, yb=b, zb=b)
    args = (g, dens, [0.55e-6], [s], [0.6 * s], [0.6], [1e-4])
    fr, _ = O.simulate(*args, instr, 0, 10000, threads=8, seed=23)
    full = fr[1][0, 0, 0]
    assert abs(fr[0].sum() / full - 1) < 1e-12
    assert fr[3][0, 0, 0] == full
    assert 0.9 * full < fr[2][0, 0, 0] < full
    alone, _ = O.simulate(*args, instr[1:2], 0, 10000, threads=8, seed=23)
    assert abs(alone[0][0, 0, 0] / full - 1) < 1e-12
