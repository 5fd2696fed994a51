"""Benchmark of the per-shot path on one MI355X.

A step is one pass of the path (surface + targets, range compression) over a
batch of shots of the workload; inputs (mesh, targets) are resident in HBM
before timing starts and the output stays in device memory.  The default
workload is the 2048 x 2048 fBm surface with two target layers, 16384 range
bins at 48 MHz, 60 MHz / 10 MHz, 10 deg aperture (workloads.fbm_two_layer);
consecutive steps walk along its 2000-shot track.

Prints ONE JSON line: throughput in facet-shot pairs/s (pairs = illuminated
facet x shot, counted by the kernels), the roofline of the dominant kernel
from live hipEvent timers, and the FP32 CPU restatement of the reference path
timed on a bounded sample of the same workload on the host cores.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("facet·shot pairs/s/GPU at ≤1e-3 rad phase err (served vs plain); "
          "p50/p99 latency")
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# Algorithmic operations per facet-target-shot triple of the target kernel
# (FP32 + FP64 adds, multiplies, divides, square roots and one per
# transcendental call of the formula), census in DESIGN.md.
TRIPLE_FLOPS = 121.0
# Algorithmic HBM bytes per facet-shot pair of the surface path: the 12
# float facet attributes streamed once per shot.
PAIR_BYTES = 48.0


def hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return ctypes.CDLL(name)
        except OSError:
            continue
    raise RuntimeError("libamdhip64.so not found")


def dev_alloc(h, nbytes):
    p = ctypes.c_void_p()
    rc = h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes))
    if rc != 0:
        raise RuntimeError(f"hipMalloc failed: {rc}")
    return p.value


def build_workload(name, ns):
    from europaraypaths_amd import workloads
    if name == "fbm_two_layer":
        return workloads.fbm_two_layer(ns=ns)
    if name == "fbm_surface_only":
        return workloads.fbm_two_layer(ns=ns, aperture=60.0, with_targets=False)
    if name == "gaussian_layer":
        return workloads.gaussian_layer(ns=ns)
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(p, soa, tg, shots, nshots):
    """FP32 CPU restatement of the reference path on `nshots` mid-track shots."""
    from europaraypaths_amd.params import parse_parameters
    from oracle import oracle
    par = parse_parameters(p)
    mid = len(shots) // 2
    sample = shots[mid:mid + nshots]
    cores = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", cores))
    t0 = time.perf_counter()
    _, nv = oracle.run(par.to_erp(), soa, tg, sample, precision=32, nthreads=threads)
    dt = time.perf_counter() - t0
    pairs = float(nv.sum())
    return {"value": pairs / dt, "unit": "facet·shot pairs/s", "cores": threads,
            "kind": "port",
            "sample": (f"{nshots} consecutive shots from track index {mid} with all their "
                       f"in-aperture targets, FP32 restatement of the reference, "
                       f"{dt:.1f} s on {threads} threads ({int(pairs)} pairs)")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shots-per-step", type=int, default=50)
    ap.add_argument("--workload", default="fbm_two_layer")
    ap.add_argument("--track-shots", type=int, default=2000)
    ap.add_argument("--cpu-shots", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--flags", type=int, default=0)
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the path shards by shot range with no "
              "exchange between GPUs, and this build runs on a single MI355X; "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from europaraypaths_amd._native import Context
    from europaraypaths_amd.params import parse_parameters

    t_build = time.perf_counter()
    p, soa, tg, shots = build_workload(a.workload, a.track_shots)
    par = parse_parameters(p)
    e = par.to_erp()
    nr = par.nr
    print(f"# workload {a.workload}: F={soa.shape[1]} T={0 if tg is None else tg[0].shape[1]} "
          f"ns={len(shots)} nr={nr} built in {time.perf_counter() - t_build:.1f} s",
          file=sys.stderr, flush=True)

    ctx = Context(0)
    mesh = ctx.upload_mesh(soa)
    targets = ctx.upload_targets(*tg) if tg is not None else None
    h = hip()
    B = a.shots_per_step
    out_dev = dev_alloc(h, B * nr * 8)
    nv_dev = dev_alloc(h, B * 4)

    def batch(i):
        s0 = (i * B) % len(shots)
        idx = (np.arange(B) + s0) % len(shots)
        return shots[idx]

    for i in range(a.warmup):
        ctx.run_dev(e, mesh, targets, batch(i), out_dev, nv_dev, flags=a.flags)
    h.hipDeviceSynchronize()
    ctx.profile(True)
    lat = []
    t0 = time.perf_counter()
    for i in range(a.warmup, a.warmup + a.steps):
        ts = time.perf_counter()
        ctx.run_dev(e, mesh, targets, batch(i), out_dev, nv_dev, flags=a.flags)
        lat.append(time.perf_counter() - ts)
        print(f"# step {i - a.warmup + 1}/{a.steps}: {1e3 * lat[-1]:.1f} ms", file=sys.stderr,
              flush=True)
    h.hipDeviceSynchronize()
    t1 = time.perf_counter()
    prof = ctx.profile_read()
    ctx.profile(False)
    dt = t1 - t0
    pairs = prof["surface"][2]
    triples = prof["target_accum"][2]
    value = pairs / dt

    stage_ms = {k: v[0] for k, v in prof.items()}
    dom = max(stage_ms, key=stage_ms.get)
    ms, launches, units = prof[dom]
    if dom == "target_accum":
        achieved = (units * TRIPLE_FLOPS) / (ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None}
    else:
        achieved = (units * PAIR_BYTES) / (ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None}
    roof.update({"kernel": dom, "avg_launch_ms": ms / max(launches, 1), "launches": launches})

    res = {
        "metric": METRIC, "value": value, "unit": "facet·shot pairs/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * dt / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (ranges and phase f64)", "data": "synthetic",
        "config": {"workload": a.workload, "facets": int(soa.shape[1]),
                   "targets": 0 if tg is None else int(tg[0].shape[1]),
                   "track_shots": len(shots), "shots_per_step": B, "nr": nr,
                   "flags": a.flags},
        "targets_per_pair": (triples / pairs) if pairs else 0.0,
        "triples_per_s": triples / dt,
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)),
                       "p99": 1e3 * float(np.percentile(lat, 99))},
        "stage_ms": stage_ms,
        "roofline": roof,
    }
    if not a.no_cpu:
        try:
            res["cpu_baseline"] = cpu_baseline(p, soa, tg, shots, a.cpu_shots)
        except Exception as ex:  # noqa: BLE001
            res["cpu_baseline"] = {"error": str(ex)}
    print(json.dumps(res))
    h.hipFree(ctypes.c_void_p(out_dev))
    h.hipFree(ctypes.c_void_p(nv_dev))


if __name__ == "__main__":
    main()
