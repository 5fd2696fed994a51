"""MI355X-native per-shot raypath / phasor / range-compression path of the
EuropaRaypaths radar-sounder simulator.

The compute runs in hand-written HIP kernels for gfx950 behind the C ABI of
include/erp.h (liberp.so, loaded with ctypes).  This package is the host
side: the reference's parameter objects, file formats and radargram layout,
and the in-process `simulate` entry point.
"""
from . import params, formats, surfaces, output_handling  # noqa: F401
from .params import (gen_params, export_params, parse_parameters,  # noqa: F401
                     make_shots, vert_source_path, track_Z_rxwin)
from .surfaces import make_surface, make_target_array, gen_normals  # noqa: F401
from .output_handling import compile_rdrgrm, load_params  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    # the native layer is imported lazily so that CPU-only tooling works
    # without the shared library; the submodule `simulate` is replaced by
    # its function of the same name at package level
    if name in ("simulate", "Context", "Simulation"):
        import importlib
        mod = importlib.import_module(".simulate", __name__)
        for n in ("simulate", "Context", "Simulation"):
            globals()[n] = getattr(mod, n)
        return globals()[name]
    raise AttributeError(name)
