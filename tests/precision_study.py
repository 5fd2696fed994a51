"""Precision study of the subsurface-target path at orbital range.

A 2048 x 2048 fBm scene seen from 400 km puts every two-way range near
4e5 m, where FP32 resolves 0.03 m: a tenth of a cycle at 9 MHz and most of
one at 60 MHz.  This compares, against the FP64 oracle, the traces the GPU
produces with the arithmetic of Itd, Ttd, rngt, the bin coordinate and the
phase in FP64 (the product), double-float and FP32 (the reference's own
arithmetic, radar_kernels.cu:681-686, 811-813), each with the bin windows
summed in FP64 (the product), plain FP32 and compensated FP32.

The scene is the fBm geometry scaled down so the FP64 oracle finishes in
seconds: a 192 x 192 mesh at 50 m (Hurst 0.8, 10 m RMS), two target layers
at 1.5 km and 6 km depth on a 500 m grid over the mesh with eps 3.15 above
the 1.5 km interface and 3.4 below it, the surface echo off, 4096 bins at
48 MHz from 399 km, 10 deg aperture (every facet lit).
Errors are taken over bins within 60 dB of each trace's peak.

Run as a script on a GPU host:  python tests/precision_study.py [--json out]
(the GPU test test_gpu_precision.py runs the same study on one frequency).
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from europaraypaths_amd import workloads  # noqa: E402
from europaraypaths_amd._abi import (ERP_RUN_ACC_COMP, ERP_RUN_ACC_F32,  # noqa: E402
                                     ERP_RUN_PHASE_DF, ERP_RUN_PHASE_F32)
from europaraypaths_amd.params import make_shots, parse_parameters  # noqa: E402
from europaraypaths_amd.surfaces import fbm_heights, target_layer, terrain_soa  # noqa: E402

PHASES = {"fp64": 0, "double-float": ERP_RUN_PHASE_DF, "fp32": ERP_RUN_PHASE_F32}
ACCUMS = {"fp64": 0, "fp32": ERP_RUN_ACC_F32, "fp32-compensated": ERP_RUN_ACC_COMP}


def orbital_scene(platform, grid=192, ns=2, nr=4096, h=400e3, surface=False):
    """The orbital scene; surface=True keeps the surface echo on (the
    study's surface leg runs it without targets)."""
    fs = 50.0
    ox = oy = -fs * grid / 2
    z = fbm_heights(grid, grid, fs, 10.0, 0.8, seed=3)
    soa = terrain_soa(grid, grid, fs, ox, oy, z)
    p = workloads.make_params(platform, "planetary_ice", nr=nr, smpl=48e6, rst=h - 1000.0,
                              aperture=10.0, fs=fs, ns=ns, sx0=-25.0 * (ns - 1), sdx=50.0,
                              sy=0.0, sz=h, disable_surface=not surface, eps_3=3.4,
                              z_interface=-1500.0)
    par = parse_parameters(p)
    m = max(2, int(grid * fs / 500.0))
    lo = -500.0 * (m - 1) / 2
    a = target_layer(m, m, 500.0, lo, lo, -1500.0)
    b = target_layer(m, m, 500.0, lo, lo, -6000.0)
    tg = (np.concatenate([a[0], b[0]], axis=1), np.concatenate([a[1], b[1]]))
    return par, soa, tg, make_shots(par)


def errors(g, o, floor_db=60.0):
    """Worst and RMS phase (rad) and relative amplitude error on bins within
    floor_db of each oracle trace's peak, and the worst |g - o| / peak."""
    ph, am, ab = [], [], 0.0
    for x, y in zip(np.asarray(g, np.complex128), np.asarray(o, np.complex128)):
        pk = np.abs(y).max()
        if pk == 0:
            continue
        sel = np.abs(y) >= pk * 10 ** (-floor_db / 20)
        ph.append(np.abs(np.angle(x[sel] / y[sel])))
        am.append(np.abs(np.abs(x[sel]) / np.abs(y[sel]) - 1))
        ab = max(ab, float(np.abs(x - y).max() / pk))
    ph, am = np.concatenate(ph), np.concatenate(am)
    return {"phase_max": float(ph.max()), "phase_rms": float(np.sqrt((ph ** 2).mean())),
            "amp_max": float(am.max()), "amp_rms": float(np.sqrt((am ** 2).mean())),
            "abs_max": ab, "bins": int(ph.size)}


def run_study(ctx, platforms=("REASON_HF", "REASON_VHF"), grid=192, ns=2, combos=None,
              legs=("targets", "surface")):
    """legs: "targets" = the two target layers with the surface echo off,
    "surface" = the surface echo alone (no targets)."""
    from oracle import oracle
    rows = []
    combos = combos or [(p, a) for p in PHASES for a in ACCUMS]
    cases = [(leg, plat) for leg in legs for plat in platforms]
    for leg, plat in cases:
        par, soa, tg, shots = orbital_scene(plat, grid=grid, ns=ns, surface=leg == "surface")
        if leg == "surface":
            tg = None
        e = par.to_erp()
        ref, rv = oracle.run(e, soa, tg, shots)
        m = ctx.upload_mesh(soa)
        t = ctx.upload_targets(*tg) if tg is not None else None
        try:
            for pn, an in combos:
                g, gv = ctx.run(e, m, t, shots, flags=PHASES[pn] | ACCUMS[an])
                assert np.array_equal(gv, rv)
                r = {"leg": leg, "chirp_MHz": float(par.f0) / 1e6, "phase": pn,
                     "accumulation": an}
                r.update(errors(g, ref))
                rows.append(r)
                print(f"{leg:8s} {r['chirp_MHz']:5.0f} MHz  phase {pn:13s} acc {an:17s} "
                      f"phase max {r['phase_max']:.2e} rms {r['phase_rms']:.2e}  "
                      f"amp max {r['amp_max']:.2e} rms {r['amp_rms']:.2e}", flush=True)
        finally:
            m.release()
            if t is not None:
                t.release()
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--json")
    ap.add_argument("--grid", type=int, default=192)
    ap.add_argument("--shots", type=int, default=2)
    a = ap.parse_args()
    from europaraypaths_amd._native import Context
    ctx = Context(0)
    rows = run_study(ctx, grid=a.grid, ns=a.shots)
    print("\n| leg | chirp | phase arithmetic | accumulation | phase max (rad) | phase rms | "
          "amp max (rel) | amp rms |")
    print("|---|---|---|---|---|---|---|---|")
    for r in rows:
        print(f"| {r['leg']} | {r['chirp_MHz']:.0f} MHz | {r['phase']} | {r['accumulation']} | "
              f"{r['phase_max']:.2e} | {r['phase_rms']:.2e} | {r['amp_max']:.2e} | "
              f"{r['amp_rms']:.2e} |")
    if a.json:
        with open(a.json, "w") as f:
            json.dump(rows, f, indent=1)
    ctx.close()


if __name__ == "__main__":
    main()
