"""CPU tests of the host side: parameter derivations, file formats, and the
C-ABI library (loads, exports every declared symbol, struct layouts agree
with include/erp.h).  No GPU calls."""
import ctypes
import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from europaraypaths_amd import _abi, _native, formats, surfaces
from europaraypaths_amd.params import (NR_LIGHT, gen_params, make_shots, parse_parameters,
                                       prism_alpha)
from europaraypaths_amd.workloads import make_params, rx_window_for

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "erp.h")
GOLDEN = os.path.join(ROOT, "tests", "golden")
f32 = np.float32


# ---------------------------------------------------------------- parameters

def _vhf(**over):
    dom = {"fs": 2.5}
    rec = {"rx_window_m": 1250.0, "rx_sample_rate": 60e6, "rx_window_offset_m": 19000.0}
    src = {"ns": 4, "aperture": 70, "sx0": 0.0, "sy": 0.0, "sz": 20e3, "sdx": 10.0}
    return gen_params("REASON_VHF", "planetary_ice", dom, rec, src, par=over or None)


def test_parameter_derivations_against_hand_values():
    """file_io.cu:40-120,166-213 in float32: lam = c/f0, rng_res = c/B (one
    way), k = 2 pi/lam with pi = 3.14159265f, nr = int(win/299792480 * smpl)
    (not c), dr = win/nr, nu = 376.7/sqrt(eps), c_i = c/sqrt(eps), linear
    gains 10^(dB/10)."""
    par = parse_parameters(_vhf())
    assert par.c == f32(299792458.0)
    assert par.lam == f32(f32(299792458.0) / f32(60e6))
    assert float(par.lam) == pytest.approx(4.99654097, rel=1e-7)
    assert float(par.rng_res) == pytest.approx(29.9792458, rel=1e-7)
    assert par.k == f32(f32(2 * f32(3.14159265)) / par.lam)
    assert par.nr == 250  # 1250 / 299792480 * 60e6 = 250.17
    assert par.dr == f32(5.0)
    assert float(par.nu_1) == pytest.approx(376.7, rel=1e-7)
    assert float(par.nu_2) == pytest.approx(376.7 / math.sqrt(3.15), rel=1e-6)
    assert float(par.c_2) == pytest.approx(299792458.0 / math.sqrt(3.15), rel=1e-6)
    assert float(par.Grefr_lin) == pytest.approx(10 ** 0.73, rel=1e-6)
    assert float(par.ks) == pytest.approx(float(par.k) * 0.4, rel=1e-6)
    assert float(par.rst) == 19000.0
    e = par.to_erp()
    assert e.nr == 250 and e.pol == 0
    assert e.cos_ap == pytest.approx(math.cos(70 * float(f32(3.14159265)) / 180), rel=1e-6)


def test_nr_uses_299792480_not_c():
    """file_io.cu:212 divides by 299792480 (not c): the window
    rx_window_for() derives under that constant realises exactly nr bins,
    and the same window with the true c would give a different count."""
    w = rx_window_for(4096, 48e6)
    par = parse_parameters(make_params("REASON_HF", "planetary_ice", nr=4096, smpl=48e6,
                                       rst=24000.0))
    assert par.nr == 4096
    assert int(f32(f32(f32(w) / f32(NR_LIGHT)) * f32(48e6))) == 4096
    assert f32(f32(f32(w) / f32(299792458.0)) * f32(48e6)) != f32(f32(f32(w) / f32(NR_LIGHT)) * f32(48e6))


def test_attenuation_from_conductivity_float32_quirk():
    """alpha = (2 pi/lam) sqrt(eps/2 (sqrt(1 + (eps''/eps)^2) - 1)),
    eps'' = sigma/(2 pi f0 eps0) (file_io.cu:196-205).  In float32 the inner
    sqrt(1 + 9e-9) rounds to 1, so ice at 60 MHz with sigma = 1e-6 S/m gets
    alpha = 0 exactly, as the reference computes it; sigma = 1e-2 gives the
    closed-form value."""
    par = parse_parameters(_vhf())
    assert par.alpha2 == 0.0
    par = parse_parameters(_vhf(sig_2=1e-2))
    f0, eps, lam = 60e6, 3.15, float(par.lam)
    epp = 1e-2 / (2 * 3.14159 * f0 * 8.85e-12)
    want = (2 * 3.14159 / lam) * math.sqrt(eps / 2 * (math.sqrt(1 + (epp / eps) ** 2) - 1))
    assert float(par.alpha2) == pytest.approx(want, rel=1e-4)
    assert float(prism_alpha(par, 1e-2)) == pytest.approx(want, rel=1e-4)


def test_missing_required_key_raises():
    p = _vhf()
    del p["rx_sample_rate"]
    with pytest.raises(KeyError):
        parse_parameters(p)


def test_linear_track_shots():
    """sim.cu:542-651: sx = sx0 + is * sdx in float, source normal (0,0,1),
    per-shot rst and linear gains."""
    par = parse_parameters(_vhf())
    sh = make_shots(par)
    np.testing.assert_array_equal(sh["sx"], [0.0, 10.0, 20.0, 30.0])
    assert (sh["sz"] == 20e3).all() and (sh["snz"] == 1.0).all()
    assert (sh["rst"] == f32(19000.0)).all()
    assert (sh["grefr_lin"] == par.Grefr_lin).all()


# ---------------------------------------------------------------- formats

def test_reference_targets_fixture_newline_quirk():
    """The reference's only committed input (examples/UNIT_TESTS/basic/inputs/
    targets.txt, 18 bytes ending in a newline): count_lines reports 2
    (file_io.cu:314-335).  strict_count reproduces the phantom second
    target as zeros of type 0; the default ignores the empty line."""
    path = os.path.join(GOLDEN, "basic_targets.txt")
    with open(path, "rb") as f:
        assert f.read() == b"0,0,-1500,0,0,1,1\n"
    assert formats.count_lines(path) == 2
    xyzn, tt = formats.load_targets(path)
    assert xyzn.shape == (6, 1)
    np.testing.assert_array_equal(xyzn[:, 0], [0, 0, -1500, 0, 0, 1])
    assert tt[0] == 1
    xyzn2, tt2 = formats.load_targets(path, strict_count=True)
    assert xyzn2.shape == (6, 2) and tt2[1] == 0
    np.testing.assert_array_equal(xyzn2[:, 0], xyzn[:, 0])


def test_facet_file_round_trip(tmp_path):
    z = surfaces.gaussian_heights(7, 5, 10.0, 1.0, 20.0, seed=2)
    soa = surfaces.terrain_soa(7, 5, 10.0, -35.0, -25.0, z)
    path = str(tmp_path / "f.fct")
    formats.export_facets(*soa, path)
    with open(path, "rb") as f:
        data = f.read()
    assert not data.endswith(b"\n") and data.count(b"\n") == soa.shape[1] - 1
    back = formats.load_facets(path)
    # %.6f keeps six decimals: |error| <= 5e-7 plus float32 rounding
    assert np.all(np.abs(back - soa) <= 5e-7 + 1e-6 * np.abs(soa))


def test_target_file_field_counts(tmp_path):
    path = str(tmp_path / "t.txt")
    with open(path, "w") as f:
        f.write("1,2,3\n4,5,6,2\n7,8,9,0,1,0\n10,11,12,0,0,-1,3")
    xyzn, tt = formats.load_targets(path)
    np.testing.assert_array_equal(xyzn[:3].T, [[1, 2, 3], [4, 5, 6], [7, 8, 9], [10, 11, 12]])
    np.testing.assert_array_equal(xyzn[3:].T, [[0, 0, 1], [0, 0, 1], [0, 1, 0], [0, 0, -1]])
    np.testing.assert_array_equal(tt, [0, 2, 0, 3])
    with open(path, "w") as f:
        f.write("1,2\n")
    with pytest.raises(ValueError):
        formats.load_targets(path)


def test_rx_window_and_gain_files(tmp_path):
    p = str(tmp_path / "rx.txt")
    with open(p, "w") as f:
        f.write("100.5\n200.25\n")
    np.testing.assert_array_equal(formats.load_rx_windows(p, 2), [100.5, 200.25])
    with pytest.raises(ValueError):
        formats.load_rx_windows(p, 3)
    g = str(tmp_path / "g.txt")
    with open(g, "w") as f:
        f.write("3.0,4.0\n\n5.0,6.0\n")
    np.testing.assert_array_equal(formats.load_gains(g, 2), [[3, 4], [5, 6]])


def test_trace_text_format(tmp_path):
    p = str(tmp_path / "s000000.txt")
    formats.save_trace_txt(p, np.array([1 + 2j, -0.5 + 0.25j], np.complex64))
    with open(p) as f:
        assert f.read() == "1.000000e+00 2.000000e+00\n-5.000000e-01 2.500000e-01\n"


# ---------------------------------------------------------------- C ABI

def _header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(erp_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert _header_functions() == sorted(_native.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _native.lib()
    for name in _header_functions():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (erp_[a-z_]+)\b", out))
    assert set(_header_functions()) <= exported


def test_struct_layouts_match_header(tmp_path):
    """Compile a probe against include/erp.h with gcc and compare sizes and
    field offsets with the ctypes mirror."""
    src = tmp_path / "probe.c"
    src.write_text(r"""
#include <stdio.h>
#include <stddef.h>
#include "erp.h"
int main(void) {
  printf("%zu %zu %zu\n", sizeof(erp_params), sizeof(erp_shot), sizeof(erp_prism));
  printf("%zu %zu %zu %zu\n", offsetof(erp_params, nr), offsetof(erp_params, mode),
         offsetof(erp_params, cos_ap), offsetof(erp_params, seed));
  printf("%zu %zu\n", offsetof(erp_shot, snx), offsetof(erp_shot, grefr_lin));
  printf("%d %d %d %d %d\n", ERP_MODE_CONVOLUTION, ERP_MODE_SPECULAR, ERP_MODE_DISABLE_SURFACE,
         ERP_RUN_RAW_SURFACE, ERP_E_UNSUPPORTED);
  printf("%zu %zu %zu %d %d\n", sizeof(erp_track), offsetof(erp_track, ns),
         offsetof(erp_track, nvalid_dev), ERP_TUNE_TRACK_ROWS, ERP_TUNE_DIRECT_LAUNCH);
  printf("%zu %zu %zu %zu %d %d\n", sizeof(erp_terrain), offsetof(erp_terrain, nx),
         offsetof(erp_terrain, seed), offsetof(erp_terrain, noise), ERP_TERRAIN_GAUSSIAN,
         ERP_TERRAIN_FBM);
  return 0;
}
""")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                   check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True,
                           check=True).stdout.split("\n")
    sizes = [int(v) for v in lines[0].split()]
    assert sizes == [ctypes.sizeof(_abi.ErpParams), ctypes.sizeof(_abi.ErpShot),
                     ctypes.sizeof(_abi.ErpPrism)]
    offs = [int(v) for v in lines[1].split()]
    P = _abi.ErpParams
    assert offs == [P.nr.offset, P.mode.offset, P.cos_ap.offset, P.seed.offset]
    so = [int(v) for v in lines[2].split()]
    assert so == [_abi.ErpShot.snx.offset, _abi.ErpShot.grefr_lin.offset]
    consts = [int(v) for v in lines[3].split()]
    assert consts == [_abi.ERP_MODE_CONVOLUTION, _abi.ERP_MODE_SPECULAR,
                      _abi.ERP_MODE_DISABLE_SURFACE, _abi.ERP_RUN_RAW_SURFACE,
                      _abi.ERP_E_UNSUPPORTED]
    tr = [int(v) for v in lines[4].split()]
    T = _abi.ErpTrack
    assert tr == [ctypes.sizeof(T), T.ns.offset, T.nvalid_dev.offset,
                  _abi.ERP_TUNE_TRACK_ROWS, _abi.ERP_TUNE_DIRECT_LAUNCH]
    te = [int(v) for v in lines[5].split()]
    R = _abi.ErpTerrain
    assert te == [ctypes.sizeof(R), R.nx.offset, R.seed.offset, R.noise.offset,
                  _abi.ERP_TERRAIN_GAUSSIAN, _abi.ERP_TERRAIN_FBM]


def test_footprint_is_host_only_and_monotone():
    par = parse_parameters(make_params("REASON_VHF", "planetary_ice", nr=16384, smpl=48e6,
                                       rst=49000.0))
    e = par.to_erp()
    a = _native.Context.footprint(e, 4_194_304, 84_050, 50)
    b = _native.Context.footprint(e, 4_194_304, 84_050, 100)
    assert 0 < a <= b


def test_product_path_fails_loudly_without_the_library(tmp_path):
    """The package has no CPU fallback: with liberp.so absent, loading the
    binding raises."""
    code = ("import europaraypaths_amd._native as n; n.LIB_PATH = '/nonexistent/liberp.so'; "
            "n._lib = None\ntry:\n    n.lib()\nexcept RuntimeError as e:\n    print('raised')")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         cwd=ROOT).stdout
    assert "raised" in out


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, cwd=ROOT)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_package_level_entry_points():
    import europaraypaths_amd as e
    assert callable(e.simulate) and e.simulate.__name__ == "simulate"
    assert e.Context is _native.Context
    assert callable(e.gen_params) and callable(e.compile_rdrgrm)


def test_cli_rejects_missing_output_directory(tmp_path):
    """sim.cu:97 / file_io.cu:266-281: the output directory must exist; the
    drop-in CLI reports it and exits with status 1 before touching a GPU."""
    r = subprocess.run([sys.executable, "-m", "europaraypaths_amd.sim", "p.json", "f.fct",
                        "t.txt", str(tmp_path / "missing")], capture_output=True, text=True,
                       cwd=ROOT)
    assert r.returncode == 1 and "Directory not found" in r.stderr
