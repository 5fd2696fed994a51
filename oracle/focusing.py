"""CPU restatement of the reference's SAR focusing — TEST INFRASTRUCTURE ONLY.

Restates src/PYTHON/simple_focusing.py:47-113,310-404 (est_slant_range with
clutter=True, get_target_phase, focus_rdrgrm) and unit_convs.py
slantrange_to_rangebin, step for step: per range bin a straight-line slant
range over the synthetic aperture, its range-bin shifts and a matched filter
from the refracted two-layer phase; per pixel the aperture window with every
column rolled by its shift, the row of interest zero-padded to twice the
aperture, an FFT product with the padded filter, an inverse FFT, and the one
output sample that lands on the pixel.  The root of the refraction equation
uses scipy's brentq on the reference's bracket, as the reference does.

Only tests/ import this module; the product's focusing
(europaraypaths_amd/focusing.py) never does.
"""
from __future__ import annotations

import numpy as np
from scipy.optimize import root_scalar

C_PY = 299792458  # the Python side's speed of light (unit_convs.py:4)


def slantrange_to_rangebin(sltrng, params):
    """unit_convs.py slantrange_to_rangebin."""
    dm = C_PY / params["rx_sample_rate"]
    return ((sltrng - params["rx_window_offset_m"]) // dm).astype(int)


def _refracted_angle(th_r, xoff, h, d, eps_2):
    """simple_focusing.py:79-83."""
    y = np.arcsin(np.sqrt(eps_2) * np.sin(th_r))
    return h * np.tan(y) + d * np.tan(th_r) - xoff


def get_target_phase(params, h, d, xoff):
    """simple_focusing.py:86-112 with cmplx_out=True."""
    lam = params["lam"] if "lam" in params else C_PY / params["frequency"]
    th_r = np.zeros_like(xoff)
    for i, dx in enumerate(xoff):
        sol = root_scalar(_refracted_angle, args=(np.abs(dx), h, d, params["eps_2"]),
                          bracket=[0, np.radians(20)], method="brentq")
        th_r[i] = sol.root
    th_i = np.arcsin(np.sqrt(params["eps_2"]) * np.sin(th_r))
    r1 = h / np.cos(th_i)
    r2 = d / np.cos(th_r)
    corrected_rng = r1 + r2 * np.sqrt(params["eps_2"])
    return np.exp(((-4j * np.pi) / lam) * corrected_rng)


def row_tables(params, i, Nr):
    """Aperture, shifts and matched filter of range bin i
    (simple_focusing.py:333-353)."""
    rng = params["rx_window_m"] * (i / Nr) + params["rx_window_offset_m"]
    apt_m = rng * np.sin(np.radians(params["aperture"]))
    apt_smpl = int(2 * apt_m / params["spacing"])
    apt_hsmpl = int(apt_m / params["spacing"])
    sx = np.linspace(-1 * apt_m, 1 * apt_m, apt_smpl)
    sz = params["altitude"]
    tx = 0
    tz = params["altitude"] - rng
    sltrng = np.sqrt((sx - tx) ** 2 + (sz - tz) ** 2)   # est_slant_range, clutter=True
    sltrng_rb = slantrange_to_rangebin(sltrng, params)
    mth_filt = np.conj(get_target_phase(params, sz, tz, sx - tx))
    return apt_smpl, apt_hsmpl, sltrng_rb, mth_filt


def focus_rdrgrm(rdrgrm, params, st=None, en=None):
    """simple_focusing.py:310-404 (without the progress print)."""
    Nr, Na = rdrgrm.shape
    rdr_f = np.zeros_like(rdrgrm)
    for i in range(Nr):
        if st and en:
            if i < st or i > en:
                continue
        apt_smpl, apt_hsmpl, sltrng_rb, mth_filt = row_tables(params, i, Nr)
        for j in range(Na):
            az_st = max(0, j - apt_hsmpl)
            az_en = min(Na, j + apt_hsmpl)
            window = rdrgrm[:, az_st:az_en]
            shift_amounts = sltrng_rb - i
            rdr_rcmc = np.array([np.roll(window[:, k], -int(shift_amounts[k]))
                                 for k in range(window.shape[1])]).T
            fft_len = int(2 * apt_smpl)
            pad = fft_len - (az_en - az_st)
            row_pad = np.pad(rdr_rcmc[i, :], (0, pad), mode="constant")
            filt_pad = np.pad(mth_filt, (0, pad), mode="constant")
            az_fft = np.fft.fft(row_pad, n=fft_len)
            ft_fft = np.fft.fft(filt_pad, n=fft_len)
            focused_row = np.fft.ifft(az_fft * ft_fft)
            rdr_f[i, j] = focused_row[(j - az_st) + pad // 2]
    return rdr_f
