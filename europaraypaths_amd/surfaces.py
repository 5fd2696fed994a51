"""Facet meshes and target layers.

Mirrors src/PYTHON/simple_surfaces.py:3-254 (grid surfaces, central-
difference normals, cross-product tangents, target arrays) and adds the
seeded synthetic terrains the benchmark workloads need (Gaussian random
field, fBm), returned directly as the (12, F) float32 SoA the C ABI takes.
Facet index is row-major over the (ny, nx) grid with x fastest
(simple_surfaces.py:150-181).
"""
from __future__ import annotations

import numpy as np

from .formats import export_facets, export_targets


def gen_normals(params, fzzs):
    """Normals and tangents of a height grid (simple_surfaces.py:3-39)."""
    fs = params["fs"]
    dZdx = np.zeros_like(fzzs)
    dZdy = np.zeros_like(fzzs)
    if fzzs.shape[1] > 2:
        dZdx[:, 1:-1] = (fzzs[:, 2:] - fzzs[:, :-2]) / (2 * fs)
    if fzzs.shape[0] > 2:
        dZdy[1:-1, :] = (fzzs[2:, :] - fzzs[:-2, :]) / (2 * fs)
    if fzzs.shape[1] > 1:
        dZdx[:, 0] = dZdx[:, 1]
        dZdx[:, -1] = dZdx[:, -2]
    if fzzs.shape[0] > 1:
        dZdy[0, :] = dZdy[1, :]
        dZdy[-1, :] = dZdy[-2, :]
    norms = np.sqrt(dZdx ** 2 + dZdy ** 2 + 1)
    n_hat = np.zeros(fzzs.shape + (3,))
    n_hat[..., 0] = -dZdx / norms
    n_hat[..., 1] = -dZdy / norms
    n_hat[..., 2] = 1 / norms
    ref = np.zeros_like(n_hat)
    ref[..., 0] = 1
    ref[np.abs(n_hat[..., 0]) > 0.9, :] = [0, 1, 0]
    u_hat = np.cross(ref, n_hat)
    u_hat /= np.linalg.norm(u_hat, axis=-1, keepdims=True)
    v_hat = np.cross(n_hat, u_hat)
    return n_hat, u_hat, v_hat


def gen_flat(params, fzzs):
    return np.zeros_like(fzzs)


def gen_concave(params, fxxs, fyys, fzzs):
    R = params["sz"]
    return R - np.sqrt(np.maximum(0, R ** 2 - (fxxs ** 2 + fyys ** 2)))


def gen_convex(params, fxxs, fyys, fzzs):
    R = params["sz"]
    return np.sqrt(np.maximum(0, R ** 2 - (fxxs ** 2 + fyys ** 2))) - R


def calc_fresnel(params, debug=False):
    """First Fresnel zone radius over a flat surface (simple_surfaces.py:124-127)."""
    return np.sqrt((params["sz"] + params["lam"] / 4) ** 2 - params["sz"] ** 2)


def calc_fresnel_spherical(params, debug=False):
    """Fresnel radius over a sphere of radius = altitude (simple_surfaces.py:129-134)."""
    h = r = params["sz"]
    return np.sqrt((params["lam"] * h * r) / (2 * (h + r)))


def gen_fresnel(params, fxxs, fyys, fzzs, stype="flat"):
    fxys = np.sqrt(fxxs ** 2 + fyys ** 2)
    if stype == "flat":
        fzzs, Fr = gen_flat(params, fzzs), calc_fresnel(params)
    else:
        fzzs, Fr = gen_convex(params, fxxs, fyys, fzzs), calc_fresnel_spherical(params)
    return fzzs, fxys <= Fr


def _grid(params):
    fxs = np.arange(params["nx"]) * params["fs"] + params["ox"]
    fys = np.arange(params["ny"]) * params["fs"] + params["oy"]
    return np.meshgrid(fxs, fys)


def mesh_soa(fxxs, fyys, fzzs, params, mask=None) -> np.ndarray:
    """(12, F) float32 SoA of a height grid, optional boolean mask."""
    n_hat, u_hat, v_hat = gen_normals(params, fzzs)
    cols = [fxxs, fyys, fzzs, n_hat[..., 0], n_hat[..., 1], n_hat[..., 2],
            u_hat[..., 0], u_hat[..., 1], u_hat[..., 2], v_hat[..., 0],
            v_hat[..., 1], v_hat[..., 2]]
    soa = np.stack([np.asarray(c, dtype=np.float64).reshape(-1) for c in cols])
    if mask is not None:
        soa = soa[:, np.asarray(mask).reshape(-1)]
    return soa.astype(np.float32)


def surface_grid(params, stype):
    """Heights and mask of a reference surface type (simple_surfaces.py:136-170)."""
    fxxs, fyys = _grid(params)
    fzzs = np.zeros_like(fxxs)
    mask = None
    if stype == "flat":
        fzzs = gen_flat(params, fzzs)
    elif stype == "concave":
        fzzs = gen_concave(params, fxxs, fyys, fzzs)
    elif stype == "fresnel":
        fzzs, mask = gen_fresnel(params, fxxs, fyys, fzzs, "flat")
    elif stype == "fresnel-convex":
        fzzs, mask = gen_fresnel(params, fxxs, fyys, fzzs, "convex")
    else:
        raise ValueError(f"Unknown surface type: {stype}")
    return fxxs, fyys, fzzs, mask


def make_surface(params, stype, filename=None, obj=None):
    """Build a surface, write its .fct file (simple_surfaces.py:147-201).

    Returns the (12, F) float32 SoA (the reference returns nothing).  With
    filename=None nothing is written.
    """
    fxxs, fyys, fzzs, mask = surface_grid(params, stype)
    soa = mesh_soa(fxxs, fyys, fzzs, params, mask)
    if filename is not None:
        # the writer formats the float64 grid values like the reference does
        n_hat, u_hat, v_hat = gen_normals(params, fzzs)
        cols = [fxxs, fyys, fzzs, n_hat[..., 0], n_hat[..., 1], n_hat[..., 2],
                u_hat[..., 0], u_hat[..., 1], u_hat[..., 2], v_hat[..., 0],
                v_hat[..., 1], v_hat[..., 2]]
        cols = [np.asarray(c).reshape(-1) for c in cols]
        if mask is not None:
            m = np.asarray(mask).reshape(-1)
            cols = [c[m] for c in cols]
        export_facets(*cols, filename, obj=obj)
    return soa


def make_target_array(params, ttype, filename=None, obj=None, zoffset=0):
    """Layer of point targets on the surface grid (simple_surfaces.py:204-253).

    Returns (xyzn (6, T) float32, type (T,) uint8 == 1).
    """
    fxxs, fyys = _grid(params)
    fzzs = np.zeros_like(fxxs)
    if ttype == "flat":
        fzzs = gen_flat(params, fzzs)
    fzzs = fzzs + zoffset
    n_hat, _, _ = gen_normals(params, fzzs)
    cols = [fxxs, fyys, fzzs, n_hat[..., 0], n_hat[..., 1], n_hat[..., 2]]
    cols = [np.asarray(c).reshape(-1) for c in cols]
    print(f"Target count: ntargets = {len(cols[0])}")
    if filename is not None:
        export_targets(*cols, filename, obj=obj)
    xyzn = np.stack(cols).astype(np.float32)
    return xyzn, np.ones(xyzn.shape[1], dtype=np.uint8)


# ---------------------------------------------------------------------------
# Seeded synthetic terrains for the benchmark workloads
# ---------------------------------------------------------------------------

def gaussian_heights(nx, ny, fs, rms, corr_len, seed=0) -> np.ndarray:
    """Gaussian random field with a Gaussian autocorrelation of length
    corr_len, by periodic spectral filtering of seeded white noise, rescaled
    to RMS exactly `rms`."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((ny, nx))
    kx = np.fft.fftfreq(nx, d=fs) * 2 * np.pi
    ky = np.fft.fftfreq(ny, d=fs) * 2 * np.pi
    K2 = kx[None, :] ** 2 + ky[:, None] ** 2
    filt = np.exp(-K2 * corr_len ** 2 / 8.0)
    z = np.real(np.fft.ifft2(np.fft.fft2(w) * filt))
    z -= z.mean()
    s = z.std()
    return z * (rms / s) if s > 0 else z


def fbm_heights(nx, ny, fs, rms, hurst, seed=0) -> np.ndarray:
    """Fractional Brownian surface by spectral synthesis, PSD ~ |k|^-(2H+2),
    rescaled to RMS `rms`."""
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(nx, d=fs)
    ky = np.fft.fftfreq(ny, d=fs)
    K = np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2)
    K[0, 0] = np.inf
    amp = K ** (-(hurst + 1.0))
    phase = rng.uniform(0, 2 * np.pi, size=(ny, nx))
    spec = amp * (rng.standard_normal((ny, nx)) * np.exp(1j * phase))
    z = np.real(np.fft.ifft2(spec))
    z -= z.mean()
    s = z.std()
    return z * (rms / s) if s > 0 else z


def terrain_noise(kind: str, nx: int, ny: int, seed: int = 0) -> np.ndarray:
    """The draws of gaussian_heights (kind "gaussian": ny x nx normals) or
    fbm_heights (kind "fbm": ny x nx phases, then ny x nx normals) for a
    seed, flattened in that order: handed to the device generator
    (Context.mesh_from_terrain) they give the host generator's field."""
    rng = np.random.default_rng(seed)
    if kind == "gaussian":
        return rng.standard_normal((ny, nx)).reshape(-1)
    if kind == "fbm":
        phase = rng.uniform(0, 2 * np.pi, size=(ny, nx))
        return np.concatenate([phase.reshape(-1), rng.standard_normal((ny, nx)).reshape(-1)])
    raise ValueError(f"unknown terrain kind {kind}")


def device_terrain(ctx, kind: str, nx: int, ny: int, fs: float, ox: float, oy: float,
                   rms: float, *, corr_len: float = 0.0, hurst: float = 0.8, seed: int = 0,
                   host_draws: bool = True):
    """A terrain mesh generated on the device (erp_mesh_from_terrain): the
    gaussian_heights / fbm_heights field and its terrain_soa mesh, built
    without the host FFTs and normals.  host_draws: the host's seeded numpy
    draws (the same field as the host generators, to FP32 transform
    rounding); otherwise the device's own seeded draws."""
    from ._abi import ERP_TERRAIN_FBM, ERP_TERRAIN_GAUSSIAN
    code = {"gaussian": ERP_TERRAIN_GAUSSIAN, "fbm": ERP_TERRAIN_FBM}[kind]
    noise = terrain_noise(kind, nx, ny, seed) if host_draws else None
    return ctx.mesh_from_terrain(code, nx, ny, fs, ox, oy, rms, corr_len=corr_len, hurst=hurst,
                                 seed=seed, noise=noise)


def terrain_soa(nx, ny, fs, ox, oy, heights) -> np.ndarray:
    """(12, F) SoA of a height grid placed at (ox, oy) with spacing fs."""
    params = {"fs": fs, "nx": nx, "ny": ny, "ox": ox, "oy": oy}
    fxxs, fyys = _grid(params)
    return mesh_soa(fxxs, fyys, heights, params)


def target_layer(nx, ny, spacing, ox, oy, z, ttype=1):
    """Flat layer of point targets (the make_target_array form) at depth z."""
    xs = np.arange(nx) * spacing + ox
    ys = np.arange(ny) * spacing + oy
    X, Y = np.meshgrid(xs, ys)
    T = X.size
    xyzn = np.zeros((6, T), np.float32)
    xyzn[0] = X.reshape(-1)
    xyzn[1] = Y.reshape(-1)
    xyzn[2] = z
    xyzn[5] = 1.0
    return xyzn, np.full(T, ttype, np.uint8)
