"""Synthetic workloads as concrete inputs (params dict, mesh SoA, targets, shots).

plain_small      : tiny scenes for parity tests (flat / rough, optional targets).
gaussian_layer   : 200 x 200 facets at 100 m, Gaussian heights (RMS 1 m,
                   correlation 500 m, seed 0), one target layer at 2 km depth
                   (same grid), h = 25 km, 100 shots every 50 m, REASON_HF
                   (9 MHz / 1 MHz), 4096 bins at 48 MHz, aperture 10 deg.
fbm_two_layer    : 2048 x 2048 facets at 50 m, fBm (Hurst 0.8, RMS 10 m,
                   seed 2), target layers at 1.5 km and 6 km depth on a 500 m
                   grid (205 x 205 each), h = 50 km, 2000 shots, REASON_VHF
                   (60 MHz / 10 MHz), 16384 bins at 48 MHz, aperture 10 deg.
The reference models one subsurface medium (a single eps_2, file_io.cu:81).
fbm_two_layer adds the second one of this build's extension (erp_params
eps_3 / z_int): ice of eps 3.15 down to the 1.5 km interface and eps 3.4
below it, so the 6 km layer's rays travel partly through each; with
eps2_only=True both layers sit in eps 3.15, the reference's model.
"""
from __future__ import annotations

import numpy as np

from .params import NR_LIGHT, gen_params, make_shots, parse_parameters
from .surfaces import fbm_heights, gaussian_heights, target_layer, terrain_soa


def rx_window_for(nr: int, smpl: float) -> float:
    """rx_window_m giving exactly nr bins under file_io.cu:212's float maths."""
    f32 = np.float32
    w = (nr + 0.5) * NR_LIGHT / smpl
    for _ in range(8):
        got = int(f32(f32(f32(w) / f32(NR_LIGHT)) * f32(smpl)))
        if got == nr:
            return float(f32(w))
        w *= (nr + 0.5) / max(got, 1)
    raise ValueError(f"cannot realise nr={nr} at {smpl} Hz")


def make_params(platform="REASON_HF", media="planetary_ice", *, nr, smpl=48e6, rst,
                aperture=10.0, fs=100.0, ns=10, sx0=0.0, sdx=50.0, sy=0.0, sz=25e3,
                **over):
    domain = {"fs": fs}
    rec = {"rx_window_m": rx_window_for(nr, smpl), "rx_sample_rate": smpl,
           "rx_window_offset_m": rst}
    src = {"ns": ns, "aperture": aperture, "sx0": sx0, "sy": sy, "sz": sz, "sdx": sdx}
    p = gen_params(platform, media, domain, rec, src, par=over or None)
    return p


def flat_soa(nx, ny, fs, ox, oy, z=0.0):
    return terrain_soa(nx, ny, fs, ox, oy, np.full((ny, nx), float(z)))


def gaussian_layer(ns=100, scale=1.0, aperture=10.0, with_targets=True):
    """The 200 x 200 Gaussian-surface workload (scale < 1 shrinks the grid)."""
    n = max(8, int(round(200 * scale)))
    fs = 100.0
    ox = oy = -fs * n / 2
    h = gaussian_heights(n, n, fs, 1.0, 500.0, seed=0)
    soa = terrain_soa(n, n, fs, ox, oy, h)
    p = make_params("REASON_HF", "planetary_ice", nr=4096, smpl=48e6, rst=24000.0,
                    aperture=aperture, fs=fs, ns=ns, sx0=-2475.0, sdx=50.0, sy=0.0, sz=25e3)
    tg = target_layer(n, n, fs, ox, oy, -2000.0) if with_targets else None
    par = parse_parameters(p)
    return p, soa, tg, make_shots(par)


def fbm_two_layer(ns=2000, grid=2048, layer_grid=205, aperture=10.0, with_targets=True,
                  seed=2, eps2_only=False):
    """The 2048 x 2048 fBm workload with two target layers (eps 3.15 above
    the 1.5 km interface, 3.4 below it)."""
    fs = 50.0
    ox = oy = -fs * grid / 2
    h = fbm_heights(grid, grid, fs, 10.0, 0.8, seed=seed)
    soa = terrain_soa(grid, grid, fs, ox, oy, h)
    extra = {} if eps2_only else {"eps_3": 3.4, "z_interface": -1500.0}
    p = make_params("REASON_VHF", "planetary_ice", nr=16384, smpl=48e6, rst=49000.0,
                    aperture=aperture, fs=fs, ns=ns, sx0=-49975.0, sdx=50.0, sy=0.0,
                    sz=50e3, **extra)
    tg = None
    if with_targets:
        sp = 500.0
        lo = -sp * (layer_grid - 1) / 2
        a = target_layer(layer_grid, layer_grid, sp, lo, lo, -1500.0)
        b = target_layer(layer_grid, layer_grid, sp, lo, lo, -6000.0)
        tg = (np.concatenate([a[0], b[0]], axis=1), np.concatenate([a[1], b[1]]))
    par = parse_parameters(p)
    return p, soa, tg, make_shots(par)
