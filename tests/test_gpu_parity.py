"""GPU parity of the HIP path against the FP64 oracle (called through the C ABI).

Tolerance: per-bin phase within 1e-3 rad and amplitude within
1e-4 relative on bins within 40 dB of each trace's peak, plus an absolute
bound of 1e-4 x peak on every bin; illuminated-facet counts and, through the
raw surface trace, range-bin indices are compared exactly.
"""
import numpy as np
import pytest

from europaraypaths_amd import surfaces, workloads
from europaraypaths_amd._abi import (ERP_RUN_DETERMINISTIC, ERP_RUN_RAW_SURFACE,
                                    PRISM_DTYPE)
from europaraypaths_amd.params import make_shots, parse_parameters, prism_alpha
from oracle import oracle
from tests.helpers import compare_traces

pytestmark = pytest.mark.gpu


def scene(nx=48, ny=40, fs=20.0, rough=0.0, nr=512, h=5000.0, rst=None, aperture=20.0,
          ns=3, platform="REASON_VHF", seed=0, **over):
    ox, oy = -fs * nx / 2, -fs * ny / 2
    z = surfaces.gaussian_heights(nx, ny, fs, rough, 60.0, seed) if rough else np.zeros((ny, nx))
    soa = surfaces.terrain_soa(nx, ny, fs, ox, oy, z)
    if rst is None:
        rst = h - 400.0
    p = workloads.make_params(platform, "planetary_ice", nr=nr, smpl=48e6, rst=rst,
                              aperture=aperture, fs=fs, ns=ns, sx0=-30.0, sdx=25.0, sy=7.0,
                              sz=h, **over)
    par = parse_parameters(p)
    return par, soa, make_shots(par)


def run_both(ctx, par, soa, shots, targets=None, prisms=None, flags=0):
    e = par.to_erp()
    m = ctx.upload_mesh(soa)
    t = ctx.upload_targets(*targets) if targets is not None else None
    g, gv = ctx.run(e, m, t, shots, prisms=prisms, flags=flags)
    if flags & ERP_RUN_RAW_SURFACE:
        _, ov, parts = oracle.run(e, soa, targets, shots, prisms=prisms, parts=True)
        o = parts[:, 0, :]
    else:
        o, ov = oracle.run(e, soa, targets, shots, prisms=prisms)
    return g, gv, o, ov


def test_raw_surface_trace_bins_and_phasors(gpu_ctx):
    par, soa, shots = scene(rough=0.8)
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots, flags=ERP_RUN_RAW_SURFACE)
    np.testing.assert_array_equal(gv, ov)
    assert (gv > 0).all()
    # same occupied bins (bit-exact bin indices), then values
    np.testing.assert_array_equal(np.abs(g) > 0, np.abs(o) > 0)
    compare_traces(g, o, floor_db=60.0)


def test_surface_plain_pow2(gpu_ctx):
    par, soa, shots = scene(rough=0.5)
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots)
    np.testing.assert_array_equal(gv, ov)
    compare_traces(g, o)


def test_surface_plain_non_pow2_nr(gpu_ctx):
    par, soa, shots = scene(rough=0.5, nr=600)
    assert par.nr == 600
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots)
    compare_traces(g, o)


def test_surface_modes(gpu_ctx):
    for over in ({"lossless": True}, {"incoherent": True, "seed": 7}, {"disable_surface": True}):
        par, soa, shots = scene(rough=0.5, **over)
        g, gv, o, ov = run_both(gpu_ctx, par, soa, shots)
        compare_traces(g, o)


def _targets(par, n=5, depth=300.0, ttype=1):
    xs = np.linspace(-120, 120, n)
    xyzn = np.zeros((6, n), np.float32)
    xyzn[0] = xs
    xyzn[1] = 13.0
    xyzn[2] = -depth
    xyzn[5] = 1.0
    return xyzn, np.full(n, ttype, np.uint8)


def test_targets_plain(gpu_ctx):
    par, soa, shots = scene(rough=0.3, nr=1024, h=3000.0, rst=2600.0)
    tg = _targets(par)
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots, targets=tg)
    compare_traces(g, o)


def test_targets_plain_non_pow2(gpu_ctx):
    par, soa, shots = scene(rough=0.3, nr=700, h=3000.0, rst=2600.0)
    tg = _targets(par, n=3)
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots, targets=tg)
    compare_traces(g, o)


@pytest.mark.parametrize("nr", [2048, 4096, 8192, 16384, 12288, 5000, 20480, 20000])
def test_targets_plain_range_sizes(gpu_ctx, nr):
    """Every spectral-chain variant of plain-mode targets: the per-pair LDS
    chain (P = 2 nr, nr 2048..16384), its cropped variant for nr below the
    power of two (12288, 5000), and past 16384 bins the four-kernel two-pass
    chain (nr 20480: P = 65536, nr a multiple of the 256-point rows) and the
    unfused path (nr 20000)."""
    print(f"targets plain nr={nr}", flush=True)
    par, soa, shots = scene(rough=0.3, nr=nr, h=3000.0, rst=2600.0, ns=2)
    tg = _targets(par, n=3)
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots, targets=tg)
    np.testing.assert_array_equal(gv, ov)
    compare_traces(g, o)


@pytest.mark.parametrize("nr", [1024, 4096, 5000, 20000])
def test_deterministic_mode_matches_oracle_and_repeats_bitwise(gpu_ctx, nr):
    """ERP_RUN_DETERMINISTIC: fixed-point bins summed with integer atomics.
    Same tolerance against the FP64 oracle as the default mode, and two runs
    give the same bits.  nr 1024 and 4096: the per-pair LDS chain; 5000: its
    cropped variant; 20000: the unfused chain that reads whole rows."""
    print(f"deterministic nr={nr}", flush=True)
    par, soa, shots = scene(rough=0.3, nr=nr, h=3000.0, rst=2600.0, ns=2)
    tg = _targets(par, n=3)
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots, targets=tg, flags=ERP_RUN_DETERMINISTIC)
    np.testing.assert_array_equal(gv, ov)
    compare_traces(g, o)
    m = gpu_ctx.upload_mesh(soa)
    t = gpu_ctx.upload_targets(*tg)
    g2, _ = gpu_ctx.run(par.to_erp(), m, t, shots, flags=ERP_RUN_DETERMINISTIC)
    assert np.array_equal(g.view(np.uint32), g2.view(np.uint32))


def test_wide_range_spread_both_modes(gpu_ctx):
    """A 30 km wide mesh seen at 85 deg aperture from 3 km: one target's
    ranges span about 6200 bins, more than one LDS window of the target
    kernel, and surface blocks spill past their LDS window.  Default mode
    (window plus atomics in the trace row) and deterministic mode (two
    fixed-point windows) against the FP64 oracle."""
    par, soa, shots = scene(nx=300, ny=40, fs=100.0, rough=0.3, nr=16384, h=3000.0,
                            rst=2600.0, aperture=85.0, ns=2)
    tg = _targets(par, n=3)
    o, ov = oracle.run(par.to_erp(), soa, tg, shots)
    m = gpu_ctx.upload_mesh(soa)
    t = gpu_ctx.upload_targets(*tg)
    for flags in (0, ERP_RUN_DETERMINISTIC):
        print(f"wide spread flags={flags}", flush=True)
        g, gv = gpu_ctx.run(par.to_erp(), m, t, shots, flags=flags)
        np.testing.assert_array_equal(gv, ov)
        compare_traces(g, o)


def test_targets_plain_general_facet_frames(gpu_ctx):
    """Facet normals that are not +-(u x v) (a general mesh file): the target
    kernel must use the stored normal, not the one derived from u and v."""
    par, soa, shots = scene(rough=0.3, nr=1024, h=3000.0, rst=2600.0)
    soa = soa.copy()
    n = soa[3:6] + 0.05 * soa[6:9]
    soa[3:6] = (n / np.linalg.norm(n, axis=0)).astype(np.float32)
    tg = _targets(par, n=3)
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots, targets=tg)
    np.testing.assert_array_equal(gv, ov)
    compare_traces(g, o)


def test_targets_lossless_and_prisms(gpu_ctx):
    par, soa, shots = scene(rough=0.3, nr=1024, h=3000.0, rst=2600.0, lossless=True)
    tg = _targets(par, n=3)
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots, targets=tg)
    compare_traces(g, o)
    par, soa, shots = scene(rough=0.3, nr=1024, h=3000.0, rst=2600.0, sig_2=1e-3)
    pr = np.zeros(2, PRISM_DTYPE)
    pr[0] = (prism_alpha(par, 1e-2), -200, -200, -250, 0, 200, 0)
    pr[1] = (prism_alpha(par, 5e-3), 0, -200, -400, 200, 200, -100)
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots, targets=tg, prisms=pr)
    compare_traces(g, o)


def test_specular_linear(gpu_ctx):
    par, soa, shots = scene(rough=0.3, nr=1024, h=3000.0, rst=2600.0, specular=True)
    tg1 = _targets(par, n=3, ttype=1)
    tg0 = _targets(par, n=3, ttype=0)
    tg = (np.concatenate([tg1[0], tg0[0]], 1), np.concatenate([tg1[1], tg0[1]]))
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots, targets=tg)
    compare_traces(g, o)


def test_specular_circular_type0(gpu_ctx):
    par, soa, shots = scene(rough=0.3, nr=512, h=3000.0, rst=2600.0, specular=True,
                            convolution_linear=False)
    tg = _targets(par, n=3, ttype=0)
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots, targets=tg)
    compare_traces(g, o)


def test_specular_noconv(gpu_ctx):
    par, soa, shots = scene(nx=24, ny=20, rough=0.3, nr=256, h=3000.0, rst=2900.0,
                            specular=True, convolution=False)
    tg = _targets(par, n=2, ttype=0)
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots, targets=tg)
    compare_traces(g, o)


@pytest.mark.parametrize("case", ["specular_linear", "specular_circular", "specular_noconv",
                                  "plain_type0"])
def test_deterministic_specular_and_type0(gpu_ctx, case):
    """ERP_RUN_DETERMINISTIC beyond the plain method's typed targets: the
    specular method (typed targets: per-pair fixed-point windows; type-0
    targets: per-shot fixed-point rows scaled by a bound pass; the direct
    sums of the no-convolution method) and type-0 targets under the plain
    method.  Same tolerance against the FP64 oracle as the default mode, and
    two runs give the same bits."""
    print(f"deterministic {case}", flush=True)
    if case == "specular_linear":
        par, soa, shots = scene(rough=0.3, nr=1024, h=3000.0, rst=2600.0, specular=True)
        kinds = (1, 0)
    elif case == "specular_circular":
        par, soa, shots = scene(rough=0.3, nr=512, h=3000.0, rst=2600.0, specular=True,
                                convolution_linear=False)
        kinds = (0,)
    elif case == "specular_noconv":
        par, soa, shots = scene(nx=24, ny=20, rough=0.3, nr=256, h=3000.0, rst=2900.0,
                                specular=True, convolution=False)
        kinds = (0,)
    else:
        par, soa, shots = scene(rough=0.3, nr=1024, h=3000.0, rst=2600.0)
        kinds = (1, 0)
    parts = [_targets(par, n=3, ttype=k) for k in kinds]
    tg = (np.concatenate([p[0] for p in parts], 1), np.concatenate([p[1] for p in parts]))
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots, targets=tg, flags=ERP_RUN_DETERMINISTIC)
    np.testing.assert_array_equal(gv, ov)
    print(f"  worst errors {compare_traces(g, o)}", flush=True)
    m = gpu_ctx.upload_mesh(soa)
    t = gpu_ctx.upload_targets(*tg)
    try:
        g2, _ = gpu_ctx.run(par.to_erp(), m, t, shots, flags=ERP_RUN_DETERMINISTIC)
    finally:
        m.release()
        t.release()
    assert np.array_equal(g.view(np.uint32), g2.view(np.uint32))


def test_dark_shots_are_zero(gpu_ctx):
    par, soa, shots = scene(rough=0.3)
    shots["snz"][1] = -1.0  # antenna pointing up: nothing illuminated
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots)
    assert gv[1] == 0 and ov[1] == 0
    assert np.all(g[1] == 0)
    compare_traces(g, o)


def test_gaussian_layer_subset(gpu_ctx):
    """The 200 x 200 Gaussian-surface workload with its 2 km target layer,
    first shot, against the FP64 oracle."""
    p, soa, tg, shots = workloads.gaussian_layer(ns=1)
    par = parse_parameters(p)
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots[:1], targets=tg)
    np.testing.assert_array_equal(gv, ov)
    compare_traces(g, o)


def test_fbm_two_layer_full_scale_shot(gpu_ctx):
    """One shot of the benchmark workload at full size (2048 x 2048 fBm
    facets, 84 050 targets in two layers, 16 384 bins, 60 MHz; about 9.8e4
    lit facets and 2.2e3 targets in the aperture) against the FP64 oracle
    with the same tolerance as the small cases."""
    p, soa, tg, shots = workloads.fbm_two_layer(ns=2000)
    par = parse_parameters(p)
    s1 = shots[1000:1001]
    print("fbm_two_layer shot 1000", flush=True)
    g, gv, o, ov = run_both(gpu_ctx, par, soa, s1, targets=tg)
    np.testing.assert_array_equal(gv, ov)
    assert gv[0] > 50000
    w = compare_traces(g, o)
    print(f"worst errors {w}", flush=True)


@pytest.mark.parametrize("shape", [(40, 48), (1, 7), (2, 2), (3, 300)])
def test_mesh_from_heights_matches_host_soa(gpu_ctx, shape):
    """erp_mesh_from_heights builds on the device the SoA that
    surfaces.terrain_soa builds on the host (gen_normals, simple_surfaces.py:
    3-39): identical float32 values, including steep slopes where the
    reference vector switches to y."""
    ny, nx = shape
    fs, ox, oy = 12.5, -300.0, 125.0
    z = surfaces.gaussian_heights(nx, ny, fs, 4.0, 20.0, seed=5) if nx > 1 and ny > 1 \
        else np.linspace(0.0, 30.0, nx * ny).reshape(ny, nx)
    z[ny // 2, :] += np.linspace(-60.0, 60.0, nx) * 20.0  # slopes with |n_x| > 0.9
    host = surfaces.terrain_soa(nx, ny, fs, ox, oy, z)
    m = gpu_ctx.mesh_from_heights(z, fs, ox, oy)
    try:
        dev = m.download()
    finally:
        m.release()
    np.testing.assert_array_equal(dev, host)


def test_targets_nearly_aligned_with_facet_rows(gpu_ctx):
    """Targets 0.2 mm off a facet row and column: the azimuth of the facets
    in that row and column is within 1e-5 of an axis, so two of the four
    sinc arguments are about 1e-5, where sin(pi x) has to be relative-
    accurate (an absolute error of 1e-7 would be 1 % of it)."""
    par, soa, shots = scene(nx=48, ny=40, fs=20.0, rough=0.0, nr=1024, h=3000.0, rst=2600.0)
    xyzn = np.zeros((6, 3), np.float32)
    xyzn[0] = np.float32(20.0) + np.float32(2e-4) * np.array([1, -1, 1], np.float32)
    xyzn[1] = np.float32(-40.0) + np.float32(2e-4) * np.array([1, 1, -1], np.float32)
    xyzn[2] = -300.0
    xyzn[5] = 1.0
    tg = (xyzn, np.ones(3, np.uint8))
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots, targets=tg)
    np.testing.assert_array_equal(gv, ov)
    print(f"worst errors {compare_traces(g, o)}", flush=True)


def test_tiled_surface_order(gpu_ctx):
    """A 128 x 64 grid mesh at a 40 deg aperture is visited in 64 x 32 tiles;
    the same facets shuffled are visited in order.  The deterministic mode
    must give the same bits for both (the visit order is a bijection that
    no sum depends on), and both match the oracle."""
    par, soa, shots = scene(nx=128, ny=64, fs=20.0, rough=0.5, nr=1024, h=3000.0,
                            rst=2600.0, aperture=40.0)
    tg = _targets(par)
    perm = np.random.default_rng(11).permutation(soa.shape[1])
    e = par.to_erp()
    t = gpu_ctx.upload_targets(*tg)
    ga, gva = gpu_ctx.run(e, gpu_ctx.upload_mesh(soa), t, shots, flags=ERP_RUN_DETERMINISTIC)
    gb, gvb = gpu_ctx.run(e, gpu_ctx.upload_mesh(np.ascontiguousarray(soa[:, perm])), t, shots,
                          flags=ERP_RUN_DETERMINISTIC)
    np.testing.assert_array_equal(gva, gvb)
    np.testing.assert_array_equal(ga, gb)
    g, gv, o, ov = run_both(gpu_ctx, par, soa, shots, targets=tg)
    np.testing.assert_array_equal(gv, ov)
    compare_traces(g, o)
    compare_traces(ga, o)


@pytest.mark.parametrize("nr", [1024, 4096, 5000, 20000])
def test_second_subsurface_medium(gpu_ctx, nr):
    """Targets above and below an interface at z = -200 m with eps 3.4 below
    it (the extension of erp_params eps_3 / z_int): against the FP64 oracle
    in the default and the deterministic mode (per-pair chain at 1024 and
    4096 bins, its cropped variant at 5000, unfused chain at 20000)."""
    print(f"second medium nr={nr}", flush=True)
    par, soa, shots = scene(rough=0.3, nr=nr, h=3000.0, rst=2600.0, ns=2, eps_3=3.4,
                            z_interface=-200.0)
    assert float(par.eps_3) == pytest.approx(3.4) and float(par.z_int) == -200.0
    xyzn = np.zeros((6, 4), np.float32)
    xyzn[0] = [-100.0, -20.0, 60.0, 110.0]
    xyzn[1] = 13.0
    xyzn[2] = [-150.0, -300.0, -450.0, -200.0]
    xyzn[5] = 1.0
    tg = (xyzn, np.ones(4, np.uint8))
    for flags in (0, ERP_RUN_DETERMINISTIC):
        g, gv, o, ov = run_both(gpu_ctx, par, soa, shots, targets=tg, flags=flags)
        np.testing.assert_array_equal(gv, ov)
        print(f"  flags={flags} worst errors {compare_traces(g, o)}", flush=True)


def test_second_medium_equal_permittivity_is_bitwise_single_medium(gpu_ctx):
    """eps_3 = eps_2 runs the second-medium kernel variant, which must give
    the single-medium result bit for bit in the deterministic mode."""
    par, soa, shots = scene(rough=0.3, nr=1024, h=3000.0, rst=2600.0, ns=2)
    tg = _targets(par, n=3, depth=350.0)
    m = gpu_ctx.upload_mesh(soa)
    t = gpu_ctx.upload_targets(*tg)
    e = par.to_erp()
    a, _ = gpu_ctx.run(e, m, t, shots, flags=ERP_RUN_DETERMINISTIC)
    e.eps_3, e.z_int = e.eps_2, -200.0
    b, _ = gpu_ctx.run(e, m, t, shots, flags=ERP_RUN_DETERMINISTIC)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    e.eps_3 = 3.4
    c, _ = gpu_ctx.run(e, m, t, shots, flags=ERP_RUN_DETERMINISTIC)
    assert not np.array_equal(a, c)
