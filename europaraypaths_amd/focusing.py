"""SAR focusing of a radargram on the GPU (simple_focusing.py:310-404).

`focus_rdrgrm(rdrgrm, par, st, en)` has the reference's signature and
result.  Per range bin i the reference builds, on the host, a synthetic
aperture of traces around each pixel, the range-bin shift of each aperture
sample (a straight-line slant range, est_slant_range with clutter=True) and
a matched filter from the refracted two-layer phase (get_target_phase); per
pixel it then rolls the aperture columns, zero-pads row i, and keeps one
sample of an FFT product with the padded filter.  Here the per-row tables
are built on the host with numpy (the refraction angle by bisection on the
reference's bracket) and every pixel is evaluated on the device as the dot
product that sample equals (erp_focus).
"""
from __future__ import annotations

import numpy as np

C_PY = 299792458  # the Python side's speed of light (unit_convs.py:4)


def slantrange_to_rangebin(sltrng, params):
    """unit_convs.py slantrange_to_rangebin."""
    dm = C_PY / params["rx_sample_rate"]
    return ((sltrng - params["rx_window_offset_m"]) // dm).astype(int)


def refraction_angle(xoff, h, d, eps_2, bisect=12, newton=8):
    """Root th_r in [0, 20 deg] of h tan(asin(sqrt(eps_2) sin th)) +
    d tan(th) - |xoff| (refracted_angle, simple_focusing.py:79-83), for every
    element of xoff (h and d scalars or arrays of the same shape) at once:
    bisection steps to get close, then Newton steps kept inside the
    bracket.  Raises ValueError where the bracket holds no sign change, as
    the reference's brentq does."""
    x = np.abs(np.asarray(xoff, dtype=np.float64))
    h = np.broadcast_to(np.asarray(h, dtype=np.float64), x.shape)
    d = np.broadcast_to(np.asarray(d, dtype=np.float64), x.shape)
    n2 = np.sqrt(eps_2)

    def f(th):
        return h * np.tan(np.arcsin(n2 * np.sin(th))) + d * np.tan(th) - x

    def df(th):
        s = n2 * np.sin(th)
        c = np.sqrt(1.0 - s * s)
        return h * (n2 * np.cos(th) / c) / (c * c) + d / np.cos(th) ** 2

    lo = np.zeros_like(x)
    hi = np.full_like(x, np.radians(20))
    flo, fhi = f(lo), f(hi)
    if np.any(np.sign(flo) * np.sign(fhi) > 0):
        raise ValueError("f(a) and f(b) must have different signs")
    root0 = flo == 0.0
    for _ in range(bisect):
        mid = 0.5 * (lo + hi)
        fm = f(mid)
        left = np.sign(fm) == np.sign(flo)
        lo = np.where(left, mid, lo)
        flo = np.where(left, fm, flo)
        hi = np.where(left, hi, mid)
    th = 0.5 * (lo + hi)
    for _ in range(newton):
        step = f(th) / df(th)
        nxt = th - step
        th = np.where((nxt > lo) & (nxt < hi), nxt, 0.5 * (lo + hi))
    return np.where(root0, 0.0, th)


def get_target_phase(params, h, d, xoff, cmplx_out=False):
    """simple_focusing.py:86-112."""
    lam = params["lam"] if "lam" in params else C_PY / params["frequency"]
    th_r = refraction_angle(xoff, h, d, params["eps_2"])
    th_i = np.arcsin(np.sqrt(params["eps_2"]) * np.sin(th_r))
    r1 = h / np.cos(th_i)
    r2 = d / np.cos(th_r)
    corrected_rng = r1 + r2 * np.sqrt(params["eps_2"])
    cmplx = np.exp(((-4j * np.pi) / lam) * corrected_rng)
    if cmplx_out:
        return cmplx
    return np.degrees(np.angle(cmplx))


def row_tables(params, i, Nr):
    """(n_sx, half aperture, range-bin shifts, matched filter) of range bin i
    (simple_focusing.py:333-353)."""
    rng = params["rx_window_m"] * (i / Nr) + params["rx_window_offset_m"]
    apt_m = rng * np.sin(np.radians(params["aperture"]))
    apt_smpl = int(2 * apt_m / params["spacing"])
    apt_hsmpl = int(apt_m / params["spacing"])
    sx = np.linspace(-1 * apt_m, 1 * apt_m, apt_smpl)
    sz = params["altitude"]
    tz = params["altitude"] - rng
    sltrng = np.sqrt((sx - 0) ** 2 + (sz - tz) ** 2)
    sltrng_rb = slantrange_to_rangebin(sltrng, params)
    mth_filt = np.conj(get_target_phase(params, sz, tz, sx - 0, cmplx_out=True))
    return apt_smpl, apt_hsmpl, sltrng_rb, mth_filt


def tables(params, lo, hi, Nr):
    """row_tables of bins lo .. hi packed for erp_focus: half apertures,
    aperture lengths, offsets, radargram rows (shifted bins modulo Nr) and
    matched filters, with one refraction solve over every sample."""
    i = np.arange(lo, hi + 1)
    rng = params["rx_window_m"] * (i / Nr) + params["rx_window_offset_m"]
    apt_m = rng * np.sin(np.radians(params["aperture"]))
    nsx = (2 * apt_m / params["spacing"]).astype(np.int64)
    hs = (apt_m / params["spacing"]).astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(nsx)[:-1]]).astype(np.int64)
    sx = np.concatenate([np.linspace(-1 * a, 1 * a, n) for a, n in zip(apt_m, nsx)])
    row_of = np.repeat(np.arange(len(i)), nsx)
    sz = params["altitude"]
    tz = params["altitude"] - rng[row_of]
    sltrng = np.sqrt((sx - 0) ** 2 + (sz - tz) ** 2)
    rows = np.mod(slantrange_to_rangebin(sltrng, params), Nr).astype(np.int32)
    lam = params["lam"] if "lam" in params else C_PY / params["frequency"]
    th_r = refraction_angle(sx - 0, sz, tz, params["eps_2"])
    th_i = np.arcsin(np.sqrt(params["eps_2"]) * np.sin(th_r))
    corrected_rng = sz / np.cos(th_i) + tz / np.cos(th_r) * np.sqrt(params["eps_2"])
    filt = np.conj(np.exp(((-4j * np.pi) / lam) * corrected_rng)).astype(np.complex64)
    return hs.astype(np.int32), nsx.astype(np.int32), offs, rows, filt


def _rows(Nr, st, en):
    # the reference skips rows outside [st, en] only when both are truthy
    if st and en:
        lo, hi = max(0, int(st)), min(Nr - 1, int(en))
        return lo, hi
    return 0, Nr - 1


def focus_rdrgrm(rdrgrm, par, st=None, en=None, return_shifts=False,
                 return_match_filter=False, ctx=None):
    """Focus the (Nr, Na) radargram like simple_focusing.focus_rdrgrm: each
    range bin i in [st, en] (every bin unless both are given) gets the
    aperture-synthesis sum over the traces within the aperture of each trace.
    `par` needs rx_window_m, rx_window_offset_m, rx_sample_rate, aperture,
    spacing (trace spacing, m), altitude (m), eps_2 and lam or frequency.
    return_shifts / return_match_filter return the first processed row's
    range-bin shifts or matched filter, as the reference does."""
    rdrgrm = np.asarray(rdrgrm)
    Nr, Na = rdrgrm.shape
    lo, hi = _rows(Nr, st, en)
    if return_shifts or return_match_filter:
        _, _, rb, mf = row_tables(par, lo, Nr)
        return rb if return_shifts else mf
    rdr_f = np.zeros_like(rdrgrm)
    if hi < lo:
        return rdr_f
    hs, nsx, offs, rows, filt = tables(par, lo, hi, Nr)
    if ctx is None:
        from .simulate import default_context
        ctx = default_context()
    out = ctx.focus(np.ascontiguousarray(rdrgrm, dtype=np.complex64), lo, hi - lo + 1,
                    np.asarray(hs, np.int32), np.asarray(nsx, np.int32),
                    np.asarray(offs, np.int64), rows, filt)
    rdr_f[lo:hi + 1] = out
    return rdr_f
