"""Drop-in for the reference CLI `sim <params.json> <facets.fct> <targets.txt> <outdir>`.

Usage: python -m europaraypaths_amd.sim params.json facets.fct targets.txt outdir
           [--npy] [--strict-count] [--device N]

Behaviour kept from sim.cu:92-1008 / file_io.cu: the output directory must
exist; existing s*.txt files are removed just before the run (other files
survive); one s%06d.txt per shot with nr lines of "%e %e"; a shot with no
illuminated facet writes no file.  --npy additionally writes the (nr, ns)
radargram as rdrgrm.npy.  Errors are reported and return exit status 1 (no
partial traces are written for a failed run).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="europaraypaths_amd.sim")
    ap.add_argument("params")
    ap.add_argument("facets")
    ap.add_argument("targets")
    ap.add_argument("outdir")
    ap.add_argument("--npy", action="store_true", help="also write rdrgrm.npy (nr, ns)")
    ap.add_argument("--strict-count", action="store_true",
                    help="count records as newlines + 1 like the reference")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)

    from . import formats
    from ._native import Context
    from .simulate import Simulation

    if not os.path.isdir(a.outdir):
        print(f"Directory not found: {a.outdir}", file=sys.stderr)
        return 1
    t0 = time.time()
    try:
        sim = Simulation.from_files(a.params, a.facets, a.targets, a.strict_count)
        print(f"Found {sim.mesh.shape[1]} facets, {sim.targets[0].shape[1]} targets, "
              f"{sim.shots.shape[0]} sources; {sim.par.nr} range bins")
        t1 = time.time()
        rdr, nvalid = sim.run(ctx=Context(a.device))
        t2 = time.time()
    except Exception as e:  # noqa: BLE001 - CLI boundary
        print(f"ERROR: {e}", file=sys.stderr)
        return 1
    formats.remove_s_txt_files(a.outdir)
    for s in range(rdr.shape[1]):
        if nvalid[s] == 0:
            continue
        formats.save_trace_txt(os.path.join(a.outdir, f"s{s:06d}.txt"), rdr[:, s])
    if a.npy:
        np.save(os.path.join(a.outdir, "rdrgrm.npy"), rdr)
    t3 = time.time()
    ns = rdr.shape[1]
    print("=== Run summary ===")
    print(f"load {t1 - t0:.3f} s, simulate {t2 - t1:.3f} s, write {t3 - t2:.3f} s")
    if ns:
        print(f"Time per source (trace): {1e3 * (t2 - t1) / ns:.3f} ms")
        print(f"Throughput: {ns / max(t2 - t1, 1e-9):.1f} traces/sec")
    return 0


if __name__ == "__main__":
    sys.exit(main())
