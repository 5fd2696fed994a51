"""Radargram assembly (src/PYTHON/output_handling.py:4-72, 97-107).

The radargram layout is (nr range bins, ns shots), complex, columns in
sorted s%06d.txt order — the same array the in-process API returns.
"""
from __future__ import annotations

import glob
import os
import pickle

import numpy as np


def compile_rdrgrm(path, par, rx_win_file=None):
    """Stack s*.txt traces into an (nr, ns) complex array (output_handling.py:4-34).

    A shot whose file is missing is skipped, so later columns shift left —
    the reference behaviour for shots with no illuminated facet.
    """
    filenames = sorted(glob.glob(f"{path}/s*.txt"))
    if len(filenames) == 0:
        raise ValueError(f"No files found in {path} with pattern s*.txt")
    if rx_win_file:
        rx_win_adj = np.load(rx_win_file)
        rx_win_adj_rb = ((rx_win_adj - np.min(rx_win_adj)) //
                         ((1 / par["rx_sample_rate"]) * 299792458)).astype(int)
    cols = []
    for i, f in enumerate(filenames):
        if i < par["ns"]:
            index = int(os.path.basename(f)[1:-4])
            arr = np.loadtxt(f).T
            col = arr[0] + 1j * arr[1]
            if rx_win_file:
                col = np.roll(col, rx_win_adj_rb[index])
            cols.append(col)
    return np.array(cols).T


def get_simulation_range(path):
    """(min, max) shot index present in a directory (output_handling.py:37-49)."""
    idx = [int(os.path.basename(f)[1:-4]) for f in sorted(glob.glob(f"{path}/s*.txt"))]
    return np.min(idx), np.max(idx)


def load_params(param_path, target_path, source_path=None):
    """Parameter pickle + first target position (output_handling.py:52-72)."""
    with open(param_path, "rb") as hdl:
        params = pickle.load(hdl)
    targets = np.genfromtxt(target_path, delimiter=",")
    if len(targets.shape) > 1:
        targets = targets[0]
    params["tx"], params["ty"], params["tz"] = targets[0], targets[1], targets[2]
    if source_path:
        params["ns"] = np.genfromtxt(source_path, delimiter=",").shape[0]
    return params


def load_attenuation_geom(filepath):
    """Prism file columns (output_handling.py:74-95)."""
    cols = [[] for _ in range(7)]
    with open(filepath) as f:
        for line in f:
            if not line.strip():
                continue
            v = [float(x) for x in line.strip().split(",")]
            for k in range(7):
                cols[k].append(v[k])
    cond, xmin, ymin, zmin, xmax, ymax, zmax = cols
    return (cond, xmin, xmax, ymin, ymax, zmin, zmax)


def roll_rdrgrm(rdrgrm, shifts):
    """Undo per-shot rx window shifts (output_handling.py:97-107)."""
    out = np.zeros_like(rdrgrm)
    for i, s in enumerate(np.asarray(shifts).astype(int)):
        out[:, i] = np.roll(rdrgrm[:, i], -1 * s)
    return out
