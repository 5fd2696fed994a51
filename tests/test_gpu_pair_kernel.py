"""The shot-pair target kernel (k_target_accum_pair), through the C ABI.

Plain runs of the default arithmetic take each target with two consecutive
shots and evaluate the facet -> target leg once for the facets either shot
lights.  Each case runs with it (the default) and with ERP_TUNE_SINGLE_SHOT
(each (shot, target) pair on its own) and compares both with the FP64
oracle (tests/helpers.py::compare_traces: 1e-3 rad, 1e-4 relative
amplitude within 40 dB of the peak, 1e-4 of the peak everywhere) and the
illuminated counts exactly.  Cases: odd shot counts (a block with one
shot), consecutive shots whose apertures light different facets and take
different targets, a dark shot inside a block, launches that cut a
block's items and batches of odd size, prisms, general facet frames,
tilted target normals, the second medium, pair windows wider than the
kernel's 5056 bins, and deterministic runs under several launch shapes.
"""
import numpy as np
import pytest

from europaraypaths_amd import surfaces, workloads
from europaraypaths_amd._abi import (ERP_RUN_DETERMINISTIC, ERP_TUNE_PAIRS_PER_LAUNCH,
                                    ERP_TUNE_SHOTS_PER_BATCH, ERP_TUNE_SINGLE_SHOT, PRISM_DTYPE)
from europaraypaths_amd.params import make_shots, parse_parameters, prism_alpha
from oracle import oracle
from tests.helpers import compare_traces

pytestmark = pytest.mark.gpu


def _scene(nx=64, ny=40, fs=20.0, rough=0.4, nr=1024, h=3000.0, rst=2600.0, aperture=20.0,
           ns=5, sdx=25.0, seed=0, **over):
    ox, oy = -fs * nx / 2, -fs * ny / 2
    z = surfaces.gaussian_heights(nx, ny, fs, rough, 60.0, seed)
    soa = surfaces.terrain_soa(nx, ny, fs, ox, oy, z)
    p = workloads.make_params("REASON_VHF", "planetary_ice", nr=nr, smpl=48e6, rst=rst,
                              aperture=aperture, fs=fs, ns=ns, sx0=-4.0 * sdx, sdx=sdx, sy=7.0,
                              sz=h, **over)
    par = parse_parameters(p)
    return par, soa, make_shots(par)


def _grid(nx, ny, span_x, span_y, depth):
    xs, ys = np.meshgrid(np.linspace(-span_x, span_x, nx), 13.0 + np.linspace(-span_y, span_y, ny))
    n = xs.size
    xyzn = np.zeros((6, n), np.float32)
    xyzn[0] = xs.ravel()
    xyzn[1] = ys.ravel()
    xyzn[2] = -depth
    xyzn[5] = 1.0
    return xyzn, np.ones(n, np.uint8)


def _both_kernels(ctx, par, soa, shots, tg, prisms=None, tune=()):
    """The run with the shot-pair kernel and with the single-shot kernel,
    each against the oracle; returns both radargrams."""
    e = par.to_erp()
    o, ov = oracle.run(e, soa, tg, shots, prisms=prisms)
    m = ctx.upload_mesh(soa)
    t = ctx.upload_targets(*tg)
    out = []
    try:
        for single in (0, 1):
            ctx.set_tuning(ERP_TUNE_SINGLE_SHOT, single)
            for k, v in tune:
                ctx.set_tuning(k, v)
            g, gv = ctx.run(e, m, t, shots, prisms=prisms)
            np.testing.assert_array_equal(gv, ov)
            w = compare_traces(g, o)
            print(f"  single={single} worst {w}", flush=True)
            out.append(g)
    finally:
        for k in (ERP_TUNE_SINGLE_SHOT, ERP_TUNE_PAIRS_PER_LAUNCH, ERP_TUNE_SHOTS_PER_BATCH):
            ctx.set_tuning(k, 0)
        m.release()
        t.release()
    return out


def test_narrow_aperture_shots_differ_in_facets_and_targets(gpu_ctx):
    """5 deg aperture, 40 m between shots: each shot's footprint (about 260 m
    across) loses and gains facets from one shot to the next, and the 7 x 3
    target grid enters and leaves the aperture; 7 shots (the last block
    holds one)."""
    print("narrow aperture, 7 shots", flush=True)
    par, soa, shots = _scene(aperture=5.0, ns=7, sdx=40.0)
    tg = _grid(7, 3, 300.0, 60.0, 300.0)
    _both_kernels(gpu_ctx, par, soa, shots, tg)


def test_dark_shot_inside_a_block(gpu_ctx):
    print("dark shot 2 of 5", flush=True)
    par, soa, shots = _scene(ns=5)
    shots["snz"][2] = -1.0  # pointing up: nothing lit, no target taken
    tg = _grid(3, 2, 120.0, 40.0, 300.0)
    g, _ = _both_kernels(gpu_ctx, par, soa, shots, tg)
    assert np.all(g[2] == 0)


@pytest.mark.parametrize("pairs,shots_per_batch", [(3, 3), (5, 32), (2, 1)])
def test_launches_that_cut_blocks(gpu_ctx, pairs, shots_per_batch):
    """Few rows per launch (a block's items over several launches, a launch
    ending inside a block) and odd batch sizes (blocks of one shot)."""
    print(f"pairs per launch {pairs}, shots per batch {shots_per_batch}", flush=True)
    par, soa, shots = _scene(ns=6, aperture=8.0, sdx=30.0)
    tg = _grid(4, 2, 150.0, 40.0, 280.0)
    _both_kernels(gpu_ctx, par, soa, shots, tg,
                  tune=((ERP_TUNE_PAIRS_PER_LAUNCH, pairs),
                        (ERP_TUNE_SHOTS_PER_BATCH, shots_per_batch)))


def test_prisms_frames_tilted_targets_and_second_medium(gpu_ctx):
    """Attenuating prisms, facet normals off u x v, target normals off +z,
    and targets on both sides of a second interface."""
    print("prisms, frames, tilted targets, second medium", flush=True)
    par, soa, shots = _scene(ns=4, sig_2=1e-3, eps_3=3.4, z_interface=-250.0)
    soa = soa.copy()
    n = soa[3:6] + 0.05 * soa[6:9]
    soa[3:6] = (n / np.linalg.norm(n, axis=0)).astype(np.float32)
    xyzn, ty = _grid(3, 2, 120.0, 40.0, 200.0)
    xyzn[2, 3:] = -380.0
    tn = np.array([0.2, -0.1, 1.0])
    xyzn[3:6, ::2] = (tn / np.linalg.norm(tn)).astype(np.float32)[:, None]
    pr = np.zeros(2, PRISM_DTYPE)
    pr[0] = (prism_alpha(par, 1e-2), -200, -200, -250, 0, 200, 0)
    pr[1] = (prism_alpha(par, 5e-3), 0, -200, -400, 200, 200, -100)
    _both_kernels(gpu_ctx, par, soa, shots, (xyzn, ty), prisms=pr)
    # the deterministic variant of the shot-pair kernel on the same scene:
    # its per-shot bound must hold with the prisms' and the second medium's
    # factors (a broken bound would overflow the fixed point or trip the
    # window check); against the oracle, and two runs give the same bits
    print("  deterministic", flush=True)
    e = par.to_erp()
    o, ov = oracle.run(e, soa, (xyzn, ty), shots, prisms=pr)
    m = gpu_ctx.upload_mesh(soa)
    t = gpu_ctx.upload_targets(xyzn, ty)
    try:
        a, av = gpu_ctx.run(e, m, t, shots, prisms=pr, flags=ERP_RUN_DETERMINISTIC)
        b, _ = gpu_ctx.run(e, m, t, shots, prisms=pr, flags=ERP_RUN_DETERMINISTIC)
    finally:
        m.release()
        t.release()
    np.testing.assert_array_equal(av, ov)
    print(f"  worst {compare_traces(a, o)}", flush=True)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_windows_wider_than_the_kernel(gpu_ctx):
    """85 deg aperture over a 30 km mesh from 3 km: a pair's ranges span
    about 6200 bins, past the kernel's 5056-bin windows (the row takes the
    rest through atomics)."""
    print("wide windows", flush=True)
    par, soa, shots = _scene(nx=300, ny=40, fs=100.0, rough=0.3, nr=16384, aperture=85.0,
                             ns=3)
    tg = _grid(3, 1, 120.0, 0.0, 300.0)
    _both_kernels(gpu_ctx, par, soa, shots, tg)


@pytest.mark.parametrize("pairs,shots_per_batch", [(7, 0), (40, 0), (70, 0), (150, 0),
                                                   (333, 0), (70, 3), (100, 1)])
def test_deterministic_launch_shapes_give_the_same_bits(gpu_ctx, pairs, shots_per_batch):
    """ERP_RUN_DETERMINISTIC with the shot-pair kernel: 7 shots 40 m apart
    at 5 deg aperture over a 40 x 5 target grid (about 92 targets per shot,
    the sets shifting from one shot to the next, so the two shots' row
    counts of a block differ).  Launches of fewer rows than a block holds
    cut its items; each shot's rows are still sliced at multiples of 32 of
    its own target index, so every launch shape and batch size gives the
    bits of the default run (7 rows: raised to one slice of 32 in this
    mode; 40 rows: below two slices, one shot at a time; 70 to 333: merged
    items and per-shot pads; 3 and 1 shots per batch: other pairings),
    which matches the FP64 oracle."""
    print(f"deterministic, pairs per launch {pairs}, shots per batch {shots_per_batch}",
          flush=True)
    par, soa, shots = _scene(aperture=5.0, ns=7, sdx=40.0)
    tg = _grid(40, 5, 600.0, 40.0, 280.0)
    e = par.to_erp()
    o, ov = oracle.run(e, soa, tg, shots)
    m = gpu_ctx.upload_mesh(soa)
    t = gpu_ctx.upload_targets(*tg)
    try:
        a, av = gpu_ctx.run(e, m, t, shots, flags=ERP_RUN_DETERMINISTIC)
        gpu_ctx.set_tuning(ERP_TUNE_PAIRS_PER_LAUNCH, pairs)
        gpu_ctx.set_tuning(ERP_TUNE_SHOTS_PER_BATCH, shots_per_batch)
        b, bv = gpu_ctx.run(e, m, t, shots, flags=ERP_RUN_DETERMINISTIC)
    finally:
        gpu_ctx.set_tuning(ERP_TUNE_PAIRS_PER_LAUNCH, 0)
        gpu_ctx.set_tuning(ERP_TUNE_SHOTS_PER_BATCH, 0)
        m.release()
        t.release()
    np.testing.assert_array_equal(av, ov)
    np.testing.assert_array_equal(bv, ov)
    print(f"  worst {compare_traces(a, o)}", flush=True)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_full_size_consecutive_shots(gpu_ctx):
    """Shots 1000-1002 of the 2048 x 2048 fBm workload with its two target
    layers (about 9.8e4 lit facets and 2.3e3 targets per shot): the shot-pair kernel
    against the single-shot kernel, each to the parity tolerance of the
    other."""
    p, soa, tg, shots = workloads.fbm_two_layer(ns=2000)
    par = parse_parameters(p)
    s3 = shots[1000:1003]
    e = par.to_erp()
    m = gpu_ctx.upload_mesh(soa)
    t = gpu_ctx.upload_targets(*tg)
    try:
        print("full size, shot-pair kernel", flush=True)
        a, av = gpu_ctx.run(e, m, t, s3)
        gpu_ctx.set_tuning(ERP_TUNE_SINGLE_SHOT, 1)
        print("full size, single-shot kernel", flush=True)
        b, bv = gpu_ctx.run(e, m, t, s3)
    finally:
        gpu_ctx.set_tuning(ERP_TUNE_SINGLE_SHOT, 0)
        m.release()
        t.release()
    np.testing.assert_array_equal(av, bv)
    assert (av > 50000).all()
    print(f"worst {compare_traces(a, b)}", flush=True)
