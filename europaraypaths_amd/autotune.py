"""Shape autotuner of the multi-track batches (erp_run_tracks).

The shapes of a multi-track batch are the (track, shot) rows per batch
(ERP_TUNE_TRACK_ROWS) and the (row, target) pairs per target launch
(ERP_TUNE_PAIRS_PER_LAUNCH).  `tune` times erp_run_tracks on a sample of
jobs for every candidate pair of shapes on the device at hand, keeps the
fastest, sets it on the context and returns the record; `save` / `apply`
persist it in a JSON file keyed by device name, so that later contexts on
the same kind of device start from the tuned shapes without timing again.
Results do not depend on the shapes beyond rounding (in the deterministic
mode not at all), so the tuner only trades time.
"""
from __future__ import annotations

import ctypes
import json
import os
import time

from ._abi import ERP_TUNE_PAIRS_PER_LAUNCH, ERP_TUNE_TRACK_ROWS

DEFAULT_PATH = os.path.join(os.path.expanduser("~"), ".cache", "europaraypaths_amd",
                            "tuning.json")
ROWS = (16, 32, 64, 128)
PAIRS = (0, 4096)  # 0: the default (every pair row that fits 6 GB, at most 16384)


def device_name(device: int = 0) -> str:
    """The HIP device name (hipDeviceGetName), the key of the persisted shapes."""
    h = ctypes.CDLL("libamdhip64.so")
    buf = ctypes.create_string_buffer(256)
    if h.hipDeviceGetName(buf, ctypes.c_int(256), ctypes.c_int(device)) != 0:
        return f"device{device}"
    return buf.value.decode(errors="replace") or f"device{device}"


def sample_jobs(jobs, tracks: int, shots: int):
    """The first `tracks` jobs with their first `shots` shots: a short run
    with the jobs' own meshes, targets and range lengths."""
    out = []
    for params, mesh, targets, sh, dev in jobs[:tracks]:
        out.append((params, mesh, targets, sh[:shots], dev))
    return out


def tune(ctx, jobs, rows=ROWS, pairs=PAIRS, repeats: int = 2) -> dict:
    """Time erp_run_tracks over jobs = [(params, mesh, targets, shots,
    DeviceTraces)] for each (rows, pairs) candidate; the fastest is set on
    ctx and returned with every candidate's time (seconds, best of
    `repeats` after one untimed call that captures the graphs)."""
    from ._native import DeviceTraces
    # outputs sized for the sampled shot counts
    runs, bufs = [], []
    try:
        for params, mesh, targets, sh, _ in jobs:
            d = DeviceTraces(len(sh), params.nr)
            bufs.append(d)
            runs.append((params, mesh, targets, sh, d))
        cands = []
        for r in rows:
            for c in pairs:
                ctx.set_tuning(ERP_TUNE_TRACK_ROWS, r)
                ctx.set_tuning(ERP_TUNE_PAIRS_PER_LAUNCH, c)
                ctx.run_tracks_dev(runs)
                best = float("inf")
                for _ in range(repeats):
                    t0 = time.perf_counter()
                    ctx.run_tracks_dev(runs)
                    best = min(best, time.perf_counter() - t0)
                cands.append({"rows": r, "pairs": c, "seconds": best})
    finally:
        for d in bufs:
            d.free()
    win = min(cands, key=lambda x: x["seconds"])
    ctx.set_tuning(ERP_TUNE_TRACK_ROWS, win["rows"])
    ctx.set_tuning(ERP_TUNE_PAIRS_PER_LAUNCH, win["pairs"])
    return {"rows": win["rows"], "pairs": win["pairs"], "seconds": win["seconds"],
            "candidates": cands,
            "sample": {"tracks": len(jobs), "shots": int(sum(len(j[3]) for j in jobs))}}


def save(rec: dict, device: int = 0, path: str = DEFAULT_PATH) -> str:
    """Persist the tuned shapes for this device's name; returns the path."""
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[device_name(device)] = {"rows": rec["rows"], "pairs": rec["pairs"],
                                 "seconds": rec["seconds"], "sample": rec.get("sample")}
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1)
    os.replace(tmp, path)
    return path


def apply(ctx, device: int = 0, path: str = DEFAULT_PATH) -> dict | None:
    """Set the persisted shapes of this device's name on ctx; None if there
    are none."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get(device_name(device))
    if not rec:
        return None
    ctx.set_tuning(ERP_TUNE_TRACK_ROWS, int(rec["rows"]))
    ctx.set_tuning(ERP_TUNE_PAIRS_PER_LAUNCH, int(rec["pairs"]))
    return rec
