"""Parameter objects of the simulator and the derivations the path consumes.

Mirrors the reference's Python parameter layer (src/PYTHON/param_gen.py:4-152)
and the C++ parser that turns the exported JSON into the simulation struct
(src/CUDA/file_io.cu:40-249).  The derived values are computed with the
reference's FP32 operation order so that the `erp_params` handed to the HIP
kernels holds the same float values the reference's SimulationParameters
would hold.
"""
from __future__ import annotations

import json
import math
import os
import pickle
from dataclasses import dataclass

import numpy as np

from ._abi import (ERP_MODE_CONV_LINEAR, ERP_MODE_CONVOLUTION,
                   ERP_MODE_DISABLE_SURFACE, ERP_MODE_INCOHERENT,
                   ERP_MODE_LOSSLESS, ERP_MODE_SPECULAR, ErpParams, SHOT_DTYPE)

f32 = np.float32

C_LIGHT = 299792458  # param_gen.py:123 (Python side, exact)
# file_io.cu:212 derives the bin count with this constant, not c.
NR_LIGHT = 299792480.0


# --------------------------------------------------------------------------
# Python-side parameter generation (param_gen.py)
# --------------------------------------------------------------------------

_PLATFORMS = {
    # param_gen.py:45-65: power [W], centre frequency [Hz], bandwidth [Hz],
    # surface gain [dB]; the subsurface gain equals the surface gain.
    "REASON_VHF": (100, 60e6, 10e6, 7.3),
    "REASON_HF": (100, 9e6, 1e6, 4.0),
    "1MHz": (100, 1e6, 1e5, 4.0),
}

_MEDIA = {
    # param_gen.py:70-79: eps_2, sig_2 [S/m], rms_height [m]
    "planetary_ice": (3.15, 1e-6, 0.4),
    "vacuum": (1.0, 0, 0),
}


def export_params(params: dict, filename: str, directory: str = "inputs") -> None:
    """Write `<directory>/<filename>.json` and `.pkl` (param_gen.py:4-15).

    Raises ValueError for a None value, as the reference does.
    """
    for k, v in params.items():
        if v is None:
            raise ValueError(f"Found null for key [{k}] when exporting parameter dict")
    with open(os.path.join(directory, f"{filename}.json"), "w") as f:
        json.dump(params, f, indent=4)
    with open(os.path.join(directory, f"{filename}.pkl"), "wb") as f:
        pickle.dump(params, f, protocol=pickle.HIGHEST_PROTOCOL)


def gen_params(platform, media, domainpar, recpar, sourcepar, par=None) -> dict:
    """Build the parameter dictionary the simulator reads (param_gen.py:17-129).

    Same keys, defaults, presets and error behaviour as the reference: the
    plain method is convolution=True, convolution_linear=True,
    specular=False, lossless=False, incoherent=False.
    """
    pars = {
        "eps_1": 1.0, "sig_1": 0.0, "mu_1": 1.0, "mu_2": 1.0,
        "convolution": True, "convolution_linear": True, "specular": False,
        "lossless": False, "incoherent": False, "debug_surface": False,
        "rerad_funct": 2, "buff": 1.1,
    }
    if platform not in _PLATFORMS:
        raise ValueError(f"platform {platform} not recognized.")
    P, f0, B, g = _PLATFORMS[platform]
    pars.update({"power": P, "frequency": f0, "bandwidth": B,
                 "surface_gain": g, "subsurface_gain": g, "polarization": "HH"})
    if media not in _MEDIA:
        raise ValueError(f"media {media} not recognized.")
    e2, s2, rms = _MEDIA[media]
    pars.update({"eps_2": e2, "sig_2": s2, "rms_height": rms})

    if "fs" not in domainpar:
        raise ValueError("Facet size (fs) missing from input dictionary domainpar")
    pars["fs"] = domainpar["fs"]
    if "ox" in domainpar:
        for k, v in domainpar.items():
            if k not in pars:
                pars[k] = v

    for k in ("rx_window_m", "rx_sample_rate"):
        if k not in recpar:
            raise ValueError(f"key {k} neccessary in recpar which has keys {recpar.keys()}")
        pars[k] = recpar[k]
    if "rx_window_position_file" in recpar:
        pars["rx_window_position_file"] = recpar["rx_window_position_file"]
    if "rx_window_offset_m" in recpar:
        pars["rx_window_offset_m"] = recpar["rx_window_offset_m"]
    if "rx_window_position_file" not in recpar and "rx_window_offset_m" not in recpar:
        raise ValueError("input dictionary recpar needs either a rx window position "
                         "file or a static rx window offset")

    for k in ("ns", "aperture"):
        if k not in sourcepar:
            raise ValueError(f"key {k} necessary in sourcepar which has keys {sourcepar.keys()}")
        pars[k] = sourcepar[k]
    if "source_path_file" in sourcepar:
        pars["source_path_file"] = sourcepar["source_path_file"]
    else:
        for k in ("sx0", "sy", "sz", "sdx"):
            pars[k] = sourcepar[k]

    if par is not None:
        pars.update(par)

    pars["lam"] = C_LIGHT / pars["frequency"]
    pars["k"] = (2 * np.pi) / pars["lam"]
    if "sz" in pars:
        pars["altitude"] = pars["sz"]
    return pars


def vert_source_path(params, minZ, maxZ, filename, direct="inputs", xoff=0, yoff=0):
    """Vertical source track file, no trailing newline (param_gen.py:131-144)."""
    sz = np.linspace(minZ, maxZ, params["ns"])
    sy = np.zeros_like(sz) + yoff
    sx = np.zeros_like(sz) + xoff
    lines = [f"{sx[i]},{sy[i]},{sz[i]}" for i in range(params["ns"])]
    with open(os.path.join(direct, f"{filename}.txt"), "w") as f:
        f.write("\n".join(lines))
    return sz


def track_Z_rxwin(sz, ahead, filename, direct="inputs"):
    """Rx window start per shot = sz - ahead, one per line (param_gen.py:146-152)."""
    with open(os.path.join(direct, f"{filename}.txt"), "w") as f:
        for pos in np.asarray(sz) - ahead:
            f.write(f"{pos}\n")


# --------------------------------------------------------------------------
# JSON -> simulation struct (file_io.cu:122-249)
# --------------------------------------------------------------------------

_REQUIRED = ("frequency", "bandwidth", "power", "aperture", "rms_height", "buff",
             "eps_1", "eps_2", "sig_1", "sig_2", "rx_sample_rate", "rx_window_m",
             "fs", "rerad_funct", "convolution", "convolution_linear",
             "specular", "lossless")


def _need(j: dict, key: str):
    if key not in j:
        # nlohmann::json's j["key"] on a missing key throws (file_io.cu:138-241)
        raise KeyError(f"required parameter '{key}' missing from parameter file")
    return j[key]


def _alpha(sig, eps, f0, lam, eps_0):
    """Attenuation constant, file_io.cu:196-205 (and :502-505 for prisms)."""
    eps_pp = f32(float(f32(sig)) / (2 * 3.14159 * float(f0) * float(eps_0)))
    a = f32(np.sqrt(f32(f32(1) + (eps_pp / eps) * (eps_pp / eps))) - f32(1))
    a = f32(math.sqrt(0.5 * float(eps) * float(a)))
    return f32((float(a) * 2 * 3.14159) / float(lam))


@dataclass
class SimulationParameters:
    """The parsed + derived parameter set (file_io.cu:40-120).

    Float fields hold float32 values computed in the reference's order.
    """
    raw: dict
    c: np.float32
    lam: np.float32
    k: np.float32
    rng_res: np.float32
    P: np.float32
    f0: np.float32
    B: np.float32
    aperture: np.float32
    Grefl: np.float32
    Grefr: np.float32
    Grefl_lin: np.float32
    Grefr_lin: np.float32
    rms_h: np.float32
    ks: np.float32
    buff: np.float32
    eps_1: np.float32
    eps_2: np.float32
    sig_1: np.float32
    sig_2: np.float32
    c_1: np.float32
    c_2: np.float32
    nu_1: np.float32
    nu_2: np.float32
    alpha1: np.float32
    alpha2: np.float32
    eps_0: np.float32
    smpl: np.float32
    nr: int
    dr: np.float32
    rst: np.float32
    fs: np.float32
    rerad_funct: int
    convolution: bool
    convolution_linear: bool
    specular: bool
    lossless: bool
    incoherent: bool
    disable_surface: bool
    debug_surface: bool
    source_path_file: str
    gain_file: str
    rx_window_position_file: str
    atten_geom_path: str
    source_normal_multiplier: np.float32
    altitude: np.float32
    sy: np.float32 = f32(0)
    sz: np.float32 = f32(0)
    sx0: np.float32 = f32(0)
    sdx: np.float32 = f32(0)
    ns: int = 0
    seed: int = 0
    # second subsurface medium below z = z_int (an extension of the
    # reference's single eps_2; eps_3 = 0 disables it)
    eps_3: np.float32 = f32(0)
    z_int: np.float32 = f32(0)

    @property
    def mode(self) -> int:
        m = 0
        if self.convolution:
            m |= ERP_MODE_CONVOLUTION
        if self.convolution_linear:
            m |= ERP_MODE_CONV_LINEAR
        if self.specular:
            m |= ERP_MODE_SPECULAR
        if self.lossless:
            m |= ERP_MODE_LOSSLESS
        if self.incoherent:
            m |= ERP_MODE_INCOHERENT
        if self.disable_surface:
            m |= ERP_MODE_DISABLE_SURFACE
        return m

    @property
    def aperture_threshold(self) -> np.float32:
        """(pi/180)*aperture in float (geometry_kernels.cu:183, sim.cu:824)."""
        return f32(f32(f32(3.14159265) / f32(180)) * self.aperture)

    def to_erp(self) -> ErpParams:
        """The POD struct handed across the C ABI."""
        e = ErpParams()
        for name in ("lam", "k", "rng_res", "P", "fs", "ks", "eps_1", "eps_2",
                     "c", "c_1", "c_2", "nu_1", "nu_2", "alpha1", "alpha2",
                     "aperture", "smpl", "B", "f0", "dr"):
            setattr(e, name, float(getattr(self, name)))
        e.nr = int(self.nr)
        e.rerad_funct = int(self.rerad_funct)
        e.pol = 0
        e.mode = self.mode
        e.cos_ap = math.cos(float(self.aperture_threshold))
        e.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        e.eps_3 = float(self.eps_3)
        e.z_int = float(self.z_int)
        return e


def parse_parameters(j: dict | str) -> SimulationParameters:
    """parseSimulationParameters (file_io.cu:122-249) on a dict or JSON path.

    Raises KeyError for a missing required key (the reference throws), and
    ValueError where the reference would compute a non-positive bin count.
    Two optional keys extend the reference's single subsurface medium:
    "eps_3", the permittivity below the horizontal interface at
    z = "z_interface" (m); absent or 0, there is one medium as in the
    reference.
    """
    if isinstance(j, (str, os.PathLike)):
        with open(j) as f:
            j = json.load(f)
    c = f32(299792458.0)
    nu0 = f32(376.7)
    eps_0 = f32(8.85e-12)
    pi = f32(3.14159265)

    source_path_file = j.get("source_path_file", "NONE")
    sy = sz = sx0 = sdx = f32(0)
    ns = 0
    if source_path_file == "NONE":
        sy, sz = f32(_need(j, "sy")), f32(_need(j, "sz"))
        sx0, sdx = f32(_need(j, "sx0")), f32(_need(j, "sdx"))
        ns = int(_need(j, "ns"))
    mult = f32(j.get("source_normal_multiplier", 1.0))
    altitude = f32(j.get("altitude", 0.0))

    f0 = f32(_need(j, "frequency"))
    B = f32(_need(j, "bandwidth"))
    P = f32(_need(j, "power"))
    aperture = f32(_need(j, "aperture"))
    gain_file = j.get("gain_pattern_file", "NONE")
    if gain_file == "NONE":
        Grefr = f32(_need(j, "subsurface_gain"))
        Grefl = f32(_need(j, "surface_gain"))
    else:
        Grefr = Grefl = f32(0)

    rng_res = f32(c / B)
    lam = f32(c / f0)
    k = f32(f32(2 * pi) / lam)
    Grefr_lin = f32(10.0 ** float(f32(Grefr / f32(10))))
    Grefl_lin = f32(10.0 ** float(f32(Grefl / f32(10))))

    rms_h = f32(_need(j, "rms_height"))
    ks = f32(k * rms_h)
    buff = f32(_need(j, "buff"))
    eps_1 = f32(_need(j, "eps_1"))
    eps_2 = f32(_need(j, "eps_2"))
    sig_1 = f32(_need(j, "sig_1"))
    sig_2 = f32(_need(j, "sig_2"))
    c_1 = f32(c / np.sqrt(eps_1))
    c_2 = f32(c / np.sqrt(eps_2))
    nu_1 = f32(nu0 / np.sqrt(eps_1))
    nu_2 = f32(nu0 / np.sqrt(eps_2))
    alpha1 = _alpha(sig_1, eps_1, f0, lam, eps_0)
    alpha2 = _alpha(sig_2, eps_2, f0, lam, eps_0)

    smpl = f32(_need(j, "rx_sample_rate"))
    rx_file = j.get("rx_window_position_file", "NONE")
    win = f32(_need(j, "rx_window_m"))
    nr_f = f32(f32(win / f32(NR_LIGHT)) * smpl)
    if not np.isfinite(nr_f) or nr_f < 1:
        raise ValueError(f"rx_window_m/rx_sample_rate give {nr_f} range bins")
    nr = int(nr_f)  # float -> int truncation, file_io.cu:212
    dr = f32(win / f32(nr))
    rst = f32(0) if rx_file != "NONE" else f32(_need(j, "rx_window_offset_m"))

    return SimulationParameters(
        raw=dict(j), c=c, lam=lam, k=k, rng_res=rng_res, P=P, f0=f0, B=B,
        aperture=aperture, Grefl=Grefl, Grefr=Grefr, Grefl_lin=Grefl_lin,
        Grefr_lin=Grefr_lin, rms_h=rms_h, ks=ks, buff=buff, eps_1=eps_1,
        eps_2=eps_2, sig_1=sig_1, sig_2=sig_2, c_1=c_1, c_2=c_2, nu_1=nu_1,
        nu_2=nu_2, alpha1=alpha1, alpha2=alpha2, eps_0=eps_0, smpl=smpl, nr=nr,
        dr=dr, rst=rst, fs=f32(_need(j, "fs")),
        rerad_funct=int(_need(j, "rerad_funct")),
        convolution=bool(_need(j, "convolution")),
        convolution_linear=bool(_need(j, "convolution_linear")),
        specular=bool(_need(j, "specular")), lossless=bool(_need(j, "lossless")),
        incoherent=bool(j.get("incoherent", False)),
        disable_surface=bool(j.get("disable_surface", False)),
        debug_surface=bool(j.get("debug_surface", False)),
        source_path_file=source_path_file, gain_file=gain_file,
        rx_window_position_file=rx_file,
        atten_geom_path=j.get("attenuation_geometry_file", "NONE"),
        source_normal_multiplier=mult, altitude=altitude, sy=sy, sz=sz, sx0=sx0,
        sdx=sdx, ns=ns, seed=int(j.get("seed", 0)),
        eps_3=f32(j.get("eps_3", 0.0)), z_int=f32(j.get("z_interface", 0.0)),
    )


def prism_alpha(par: SimulationParameters, conductivity) -> np.float32:
    """Prism conductivity -> alpha with eps_2 (file_io.cu:502-505)."""
    return _alpha(conductivity, par.eps_2, par.f0, par.lam, par.eps_0)


def make_shots(par: SimulationParameters, sources=None, gains_db=None,
               rx_windows=None) -> np.ndarray:
    """Per-shot tables of sim.cu:542-651 as an array of erp_shot records.

    sources: optional (ns, 3|6) float32 array from a source path file; without
    it the reference's linear track sx = sx0 + is*sdx (float arithmetic) is
    used.  gains_db: optional (ns, 2) dB pairs (refl, refr); rx_windows:
    optional (ns,) rx window starts.
    """
    if sources is None:
        ns = int(par.ns)
        isx = np.arange(ns, dtype=np.float32)
        sx = (par.sx0 + isx * par.sdx).astype(np.float32)
        sy = np.full(ns, par.sy, np.float32)
        sz = np.full(ns, par.sz, np.float32)
        sn = np.zeros((ns, 3), np.float32)
        sn[:, 2] = f32(1) * par.source_normal_multiplier
    else:
        src = np.asarray(sources, dtype=np.float32)
        ns = src.shape[0]
        sx, sy, sz = src[:, 0], src[:, 1], src[:, 2]
        if src.shape[1] >= 6:
            sn = (src[:, 3:6] * par.source_normal_multiplier).astype(np.float32)
        else:
            sn = np.zeros((ns, 3), np.float32)
            sn[:, 2] = f32(1) * par.source_normal_multiplier
    shots = np.zeros(ns, dtype=SHOT_DTYPE)
    shots["sx"], shots["sy"], shots["sz"] = sx, sy, sz
    shots["snx"], shots["sny"], shots["snz"] = sn[:, 0], sn[:, 1], sn[:, 2]
    if rx_windows is not None:
        rw = np.asarray(rx_windows, dtype=np.float32)
        if rw.shape[0] != ns:
            raise ValueError(f"rx window file has {rw.shape[0]} entries, expected {ns}")
        shots["rst"] = rw
    else:
        shots["rst"] = par.rst
    if gains_db is not None:
        g = np.asarray(gains_db, dtype=np.float32)
        # sim.cu:630-633: powf(10, g/10) per shot
        shots["grefl_lin"] = (10.0 ** (g[:, 0] / np.float32(10)).astype(np.float64)).astype(np.float32)
        shots["grefr_lin"] = (10.0 ** (g[:, 1] / np.float32(10)).astype(np.float64)).astype(np.float32)
    else:
        shots["grefl_lin"] = par.Grefl_lin
        shots["grefr_lin"] = par.Grefr_lin
    return shots
