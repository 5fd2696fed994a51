"""Readers and writers for the simulator's file formats.

Mirrors src/CUDA/file_io.cu:284-624 (facets, targets, sources, gains, rx
windows, attenuation prisms, per-shot trace text) and the Python writers
src/PYTHON/simple_surfaces.py:65-108.

Record counting: the reference's count_lines returns newlines + 1
(file_io.cu:314-335), so a file that ends with a newline gains one
uninitialised record.  `strict_count=False` (the default here) ignores empty
trailing lines instead; `strict_count=True` reproduces the reference count
and fills the phantom record with zeros (type 0).
"""
from __future__ import annotations

import glob
import os
import re

import numpy as np

from ._abi import PRISM_DTYPE


def count_lines(path: str) -> int:
    """count_lines of file_io.cu:314-335: number of '\\n' bytes + 1."""
    with open(path, "rb") as f:
        return f.read().count(b"\n") + 1


def _records(path: str, strict_count: bool) -> tuple[list[str], int]:
    with open(path) as f:
        text = f.read()
    lines = text.split("\n")
    n_ref = len(lines)  # == count_lines
    if not strict_count:
        while lines and lines[-1].strip() == "":
            lines.pop()
        return lines, len(lines)
    return lines, n_ref


def _parse_floats(text: str, n: int) -> np.ndarray:
    # %f parsing: decimal -> double -> float32 (see formats docstring).
    vals = np.array(text.split(","), dtype=np.float64)
    if vals.size != n:
        raise ValueError("malformed record")
    return vals.astype(np.float32)


def load_facets(path: str, strict_count: bool = False) -> np.ndarray:
    """loadFacetFile (file_io.cu:337-362): `x,y,z:nx,ny,nz:ux,uy,uz:vx,vy,vz`.

    Returns a (12, F) float32 SoA array in the order of include/erp.h.
    """
    lines, n = _records(path, strict_count)
    body = "\n".join(l for l in lines if l.strip() != "")
    if body:
        flat = np.fromstring(re.sub(r"[:\n]", ",", body), dtype=np.float64, sep=",")
    else:
        flat = np.zeros(0)
    nrec = sum(1 for l in lines if l.strip() != "")
    if flat.size != 12 * nrec:
        raise ValueError(f"{path}: expected 12 fields per facet line")
    soa = np.zeros((12, n), dtype=np.float32)
    soa[:, :nrec] = flat.reshape(nrec, 12).T.astype(np.float32)
    return soa


def load_targets(path: str, strict_count: bool = False):
    """loadTargetFile (file_io.cu:365-426): 3, 4, 6 or 7 comma fields.

    Returns (xyzn (6, T) float32, type (T,) uint8).  Missing normals default to
    (0, 0, 1) and a missing type to 0 (sim.cu:205-206).  A line with another
    field count stops the read with an error, as the reference does.
    """
    lines, n = _records(path, strict_count)
    xyzn = np.zeros((6, n), dtype=np.float32)
    xyzn[5, :] = 1.0
    ttype = np.zeros(n, dtype=np.uint8)
    for i, line in enumerate(lines[:n]):
        if line.strip() == "":
            continue
        f = line.strip().split(",")
        if len(f) == 3:
            xyzn[0:3, i] = _parse_floats(",".join(f), 3)
        elif len(f) == 4:
            xyzn[0:3, i] = _parse_floats(",".join(f[:3]), 3)
            ttype[i] = int(float(f[3])) & 0xFF
        elif len(f) == 6:
            xyzn[:, i] = _parse_floats(",".join(f), 6)
        elif len(f) == 7:
            xyzn[:, i] = _parse_floats(",".join(f[:6]), 6)
            ttype[i] = int(float(f[6])) & 0xFF
        else:
            raise ValueError(f"Target file line number={i + 1} has incorrect number of entries.")
    return xyzn, ttype


def load_sources(path: str, strict_count: bool = False) -> np.ndarray:
    """loadSourceFile (file_io.cu:429-474): 3 or 6 fields -> (ns, 6) float32."""
    lines, n = _records(path, strict_count)
    out = np.zeros((n, 6), dtype=np.float32)
    out[:, 5] = 1.0
    for i, line in enumerate(lines[:n]):
        if line.strip() == "":
            continue
        f = line.strip().split(",")
        if len(f) == 3:
            out[i, 0:3] = _parse_floats(line.strip(), 3)
        elif len(f) == 6:
            out[i, :] = _parse_floats(line.strip(), 6)
        else:
            raise ValueError(f"Source file line number={i + 1} has incorrect number of entries.")
    return out


def load_gains(path: str, ns: int) -> np.ndarray:
    """loadGainFile (file_io.cu:597-624): `g_refl_dB,g_refr_dB`, blank lines skipped."""
    rows = []
    with open(path) as f:
        for line in f:
            if line[:1] in ("\n", "\r") or line.strip() == "":
                continue
            rows.append(_parse_floats(line.strip(), 2))
    g = np.array(rows, dtype=np.float32).reshape(-1, 2)
    if g.shape[0] != ns:
        raise ValueError(f"Gain pattern file has {g.shape[0]} lines, but number of sources is {ns}")
    return g


def load_rx_windows(path: str, ns: int) -> np.ndarray:
    """loadRxWindowPositions (file_io.cu:520-551): exactly ns floats."""
    vals = []
    with open(path) as f:
        for i, line in enumerate(f):
            if len(vals) >= ns:
                raise ValueError(f"rx_window_positions file has MORE entries than sources (ns = {ns})")
            try:
                vals.append(np.float32(float(line.split()[0])))
            except (ValueError, IndexError):
                raise ValueError(f"Invalid float on line {i + 1}") from None
    if len(vals) != ns:
        raise ValueError(f"rx_window_positions file has {len(vals)} entries, but expected exactly {ns}")
    return np.array(vals, dtype=np.float32)


def load_prisms(path: str, par) -> np.ndarray:
    """loadAttenPrismFile (file_io.cu:477-517): `sigma,xmin,ymin,zmin,xmax,ymax,zmax`.

    Returns PRISM_DTYPE records with alpha derived from sigma and eps_2.
    """
    from .params import prism_alpha
    rows = []
    with open(path) as f:
        for line in f:
            if line.strip() == "":
                continue
            v = _parse_floats(line.strip(), 7)
            rows.append(v)
    out = np.zeros(len(rows), dtype=PRISM_DTYPE)
    for i, v in enumerate(rows):
        out[i]["alpha"] = prism_alpha(par, v[0])
        (out[i]["xmin"], out[i]["ymin"], out[i]["zmin"],
         out[i]["xmax"], out[i]["ymax"], out[i]["zmax"]) = v[1:7]
    return out


def export_facets(fxs, fys, fzs, fnxs, fnys, fnzs, fuxs, fuys, fuzs, fvxs, fvys,
                  fvzs, filename, obj=None):
    """Facet writer, `%.6f` with no trailing newline (simple_surfaces.py:65-86)."""
    cols = np.stack([fxs, fys, fzs, fnxs, fnys, fnzs, fuxs, fuys, fuzs, fvxs,
                     fvys, fvzs], axis=1).astype(np.float64)
    fmt = "%.6f,%.6f,%.6f:%.6f,%.6f,%.6f:%.6f,%.6f,%.6f:%.6f,%.6f,%.6f"
    with open(filename, "w") as f:
        f.write("\n".join(fmt % tuple(r) for r in cols))
    print(f"Exported facet data to: {filename}")


def export_targets(fxs, fys, fzs, fnxs, fnys, fnzs, filename, obj=None, ttype=1):
    """Target writer, type 1, no trailing newline (simple_surfaces.py:88-108)."""
    cols = np.stack([fxs, fys, fzs, fnxs, fnys, fnzs], axis=1).astype(np.float64)
    fmt = "%.6f,%.6f,%.6f,%.6f,%.6f,%.6f," + str(int(ttype))
    with open(filename, "w") as f:
        f.write("\n".join(fmt % tuple(r) for r in cols))


def save_trace_txt(filename: str, trace: np.ndarray) -> None:
    """saveSignalToFile (file_io.cu:284-296): nr lines of `%e %e`."""
    tr = np.asarray(trace)
    re_ = tr.real.astype(np.float32).astype(np.float64)
    im_ = tr.imag.astype(np.float32).astype(np.float64)
    with open(filename, "w") as f:
        f.write("".join("%e %e\n" % (a, b) for a, b in zip(re_, im_)))


def remove_s_txt_files(directory: str) -> None:
    """remove_s_txt_files (file_io.cu:554-593): delete s*.txt, keep the rest."""
    for p in glob.glob(os.path.join(directory, "s*.txt")):
        name = os.path.basename(p)
        if name.startswith("s") and name.endswith(".txt") and len(name) >= 5:
            os.unlink(p)
