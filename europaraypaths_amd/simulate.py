"""In-process entry points of the path.

`simulate(params, mesh, targets, shots)` returns the radargram (nr, ns) the
reference assembles from its s%06d.txt files (output_handling.py:4-34), with
the whole shot loop of sim.cu:615-1008 running in the HIP kernels of
liberp.so.  `Simulation.from_files` mirrors the reference CLI's input
handling (sim.cu:92-613): the same parameter file, facet file, target file and
optional source / gain / rx-window / attenuation files.
"""
from __future__ import annotations

import os

import numpy as np

from . import formats
from ._abi import PRISM_DTYPE
from ._native import Context
from .params import SimulationParameters, make_shots, parse_parameters

_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _as_params(params) -> SimulationParameters:
    if isinstance(params, SimulationParameters):
        return params
    return parse_parameters(params)


def simulate(params, mesh, targets=None, shots=None, prisms=None, ctx: Context | None = None,
             flags: int = 0, return_nvalid: bool = False):
    """Run the path for every shot and return the radargram (nr, ns) complex64.

    params  : gen_params dict, JSON path or SimulationParameters
    mesh    : (12, F) float32 SoA facet array (or an uploaded Mesh)
    targets : None, (xyzn (6, T), type (T,)) or an uploaded Targets
    shots   : erp_shot records (make_shots); default: the linear track of params
    prisms  : PRISM_DTYPE records (alpha already derived), optional
    """
    par = _as_params(params)
    ctx = ctx or default_context()
    m = mesh if hasattr(mesh, "handle") else ctx.upload_mesh(mesh)
    t = None
    if targets is not None:
        t = targets if hasattr(targets, "handle") else ctx.upload_targets(*targets)
    if shots is None:
        shots = make_shots(par)
    out, nvalid = ctx.run(par.to_erp(), m, t, shots, prisms=prisms, flags=flags)
    rdr = np.ascontiguousarray(out.T)
    return (rdr, nvalid) if return_nvalid else rdr


class Simulation:
    """Inputs of one `sim` invocation, loaded with the reference's semantics."""

    def __init__(self, par: SimulationParameters, mesh: np.ndarray, targets, shots,
                 prisms=None):
        self.par, self.mesh, self.targets, self.shots = par, mesh, targets, shots
        self.prisms = prisms if prisms is not None else np.zeros(0, PRISM_DTYPE)

    @classmethod
    def from_files(cls, param_file, facet_file, target_file, strict_count=False):
        for p in (param_file, facet_file, target_file):
            if not os.path.isfile(p):
                raise FileNotFoundError(f"File not found: {p}")
        par = parse_parameters(param_file)
        base = os.path.dirname(os.path.abspath(param_file))

        def resolve(p):
            # the reference opens paths relative to its working directory
            return p if os.path.isfile(p) else os.path.join(base, p)

        sources = None
        if par.source_path_file != "NONE":
            sources = formats.load_sources(resolve(par.source_path_file), strict_count)
            par.ns = sources.shape[0]
        gains = None
        if par.gain_file != "NONE":
            gains = formats.load_gains(resolve(par.gain_file), par.ns)
        rxw = None
        if par.rx_window_position_file != "NONE":
            rxw = formats.load_rx_windows(resolve(par.rx_window_position_file), par.ns)
        prisms = None
        if par.atten_geom_path != "NONE":
            prisms = formats.load_prisms(resolve(par.atten_geom_path), par)
        mesh = formats.load_facets(facet_file, strict_count)
        targets = formats.load_targets(target_file, strict_count)
        shots = make_shots(par, sources=sources, gains_db=gains, rx_windows=rxw)
        return cls(par, mesh, targets, shots, prisms)

    def run(self, ctx: Context | None = None, flags: int = 0):
        """-> (radargram (nr, ns) complex64, illuminated facet count per shot)."""
        return simulate(self.par, self.mesh, self.targets if self.targets[0].shape[1] else None,
                        self.shots, prisms=self.prisms, ctx=ctx, flags=flags,
                        return_nvalid=True)
