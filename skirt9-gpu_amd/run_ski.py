#!/usr/bin/env python3
"""Runs a ski file on the GPU engine and writes the instruments' FITS cubes, as the reference
command line does for the supported configurations (skirt -o <outdir> <file.ski>).

    python skirt9-gpu_amd/run_ski.py <file.ski> [-o OUTDIR] [--packets N]

Exits with status 3 and the reason when the ski file uses something the engine does not run
(the caller then runs the reference instead)."""

import argparse
import importlib.util
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def load_pkg():
    name = "skirt9_gpu_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("ski")
    ap.add_argument("-o", "--outdir", default=".")
    ap.add_argument("--packets", type=float, default=None, help="override numPackets")
    a = ap.parse_args()
    load_pkg()
    from skirt9_gpu_amd import ski as S
    try:
        sim = S.load(a.ski)
    except S.UnsupportedSki as e:
        print(f"unsupported ski file: {e}", file=sys.stderr)
        sys.exit(3)
    if a.packets:
        sim.num_packets = int(a.packets)
    prefix = os.path.splitext(os.path.basename(a.ski))[0]
    S.run(sim, a.outdir, prefix, log=lambda m: print(m, flush=True))
    for f in sim.frames:
        suffix = "sed.dat" if isinstance(f, S.Sed) else "total.fits"
        print(os.path.join(a.outdir, f"{prefix}_{f.name}_{suffix}"))


if __name__ == "__main__":
    main()
