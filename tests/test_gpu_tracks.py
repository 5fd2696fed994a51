"""erp_run_tracks: several tracks in multi-track batches, through the C ABI.

Every track's radargram from one erp_run_tracks call must match the FP64
oracle within the parity tolerance (tests/helpers.py::compare_traces), and
its illuminated counts exactly.  In the deterministic mode it must equal
erp_run (the track alone) value for value, whichever way the batches and
launches fall: rows per batch and pairs per launch are set small here so
that batches mix tracks, end in padding rows and cut one shot's pairs over
several launches.  The captured graphs are replayed across batches and
across calls, with other eager work in between, and must give the bits of
the directly launched sequence.
"""
import numpy as np
import pytest

from europaraypaths_amd._abi import (ERP_RUN_DETERMINISTIC, ERP_TUNE_DIRECT_LAUNCH,
                                    ERP_TUNE_PAIRS_PER_LAUNCH, ERP_TUNE_SINGLE_SHOT,
                                    ERP_TUNE_TRACK_ROWS)
from oracle import oracle
from tests.helpers import compare_traces
from tests.test_gpu_parity import scene

pytestmark = pytest.mark.gpu


def _grid_targets(nx, ny, span, depth, ttype=1, y0=13.0):
    xs, ys = np.meshgrid(np.linspace(-span, span, nx), y0 + np.linspace(-60.0, 60.0, ny))
    n = xs.size
    xyzn = np.zeros((6, n), np.float32)
    xyzn[0] = xs.ravel()
    xyzn[1] = ys.ravel()
    xyzn[2] = -depth
    xyzn[5] = 1.0
    return xyzn, np.full(n, ttype, np.uint8)


def _tracks():
    """(name, par, soa, targets or None, shots): grouped tracks of three range
    lengths (one not a power of two), one with a second medium, and tracks
    that must run alone."""
    out = []
    par, soa, shots = scene(rough=0.3, nr=1024, h=3000.0, rst=2600.0, ns=5)
    out.append(("nr1024_a", par, soa, _grid_targets(3, 1, 120.0, 300.0), shots))
    # another mesh, the 9 MHz chirp (other range kernels and Bluestein band),
    # 45 targets: a shot's pairs span several launches of 40
    par, soa, shots = scene(nx=32, ny=56, fs=25.0, rough=0.6, seed=3, nr=1024, h=3000.0,
                            rst=2600.0, ns=23, platform="REASON_HF")
    out.append(("nr1024_hf_many", par, soa, _grid_targets(15, 3, 200.0, 250.0), shots))
    par, soa, shots = scene(rough=0.5, seed=5, nr=1024, h=3000.0, rst=2600.0, ns=9)
    out.append(("nr1024_surface_only", par, soa, None, shots))
    par, soa, shots = scene(rough=0.3, seed=6, nr=1024, h=3000.0, rst=2600.0, ns=4,
                            disable_surface=True)
    out.append(("nr1024_no_surface", par, soa, _grid_targets(2, 2, 80.0, 350.0), shots))
    par, soa, shots = scene(rough=0.3, seed=7, nr=2048, h=3000.0, rst=2600.0, ns=7)
    out.append(("nr2048", par, soa, _grid_targets(4, 1, 150.0, 300.0), shots))
    # a range length that is not a power of two: the cropped pair chain
    par, soa, shots = scene(rough=0.3, seed=11, nr=1500, h=3000.0, rst=2600.0, ns=5)
    out.append(("nr1500", par, soa, _grid_targets(3, 1, 150.0, 300.0), shots))
    par, soa, shots = scene(rough=0.3, seed=8, nr=1024, h=3000.0, rst=2600.0, ns=6, eps_3=3.4,
                            z_interface=-200.0)
    out.append(("nr1024_second_medium", par, soa, _grid_targets(2, 2, 100.0, 320.0), shots))
    # 8192 and 16384 bins: the multi-track chain with the wave-local
    # transforms (8 and 16 waves)
    par, soa, shots = scene(rough=0.3, seed=13, nr=8192, h=3000.0, rst=2600.0, ns=3)
    out.append(("nr8192", par, soa, _grid_targets(2, 1, 100.0, 300.0), shots))
    par, soa, shots = scene(rough=0.3, seed=12, nr=16384, h=3000.0, rst=2600.0, ns=3)
    out.append(("nr16384", par, soa, _grid_targets(2, 1, 100.0, 300.0), shots))
    # alone: type-0 targets, and a range length without the per-pair chain
    par, soa, shots = scene(rough=0.3, seed=9, nr=1024, h=3000.0, rst=2600.0, ns=3)
    out.append(("type0_alone", par, soa, _grid_targets(3, 1, 120.0, 300.0, ttype=0), shots))
    par, soa, shots = scene(rough=0.3, seed=10, nr=500, h=3000.0, rst=2600.0, ns=2)
    out.append(("nr500_alone", par, soa, _grid_targets(2, 1, 100.0, 300.0), shots))
    return out


class _Uploaded:
    def __init__(self, ctx, tracks):
        self.jobs, self.handles = [], []
        for name, par, soa, tg, shots in tracks:
            m = ctx.upload_mesh(soa)
            t = ctx.upload_targets(*tg) if tg is not None else None
            self.handles += [m] + ([t] if t is not None else [])
            self.jobs.append((par.to_erp(), m, t, shots))

    def release(self):
        for h in self.handles:
            h.release()


@pytest.fixture
def small_batches(gpu_ctx):
    """16 rows per batch and 40 pairs per launch on the shared context."""
    gpu_ctx.set_tuning(ERP_TUNE_TRACK_ROWS, 16)
    gpu_ctx.set_tuning(ERP_TUNE_PAIRS_PER_LAUNCH, 40)
    try:
        yield gpu_ctx
    finally:
        for key in (ERP_TUNE_TRACK_ROWS, ERP_TUNE_PAIRS_PER_LAUNCH, ERP_TUNE_DIRECT_LAUNCH):
            gpu_ctx.set_tuning(key, 0)


def test_tracks_match_the_oracle(small_batches):
    ctx = small_batches
    tracks = _tracks()
    up = _Uploaded(ctx, tracks)
    try:
        print("multi-track run, default mode", flush=True)
        res = ctx.run_tracks(up.jobs)
        for (name, par, soa, tg, shots), (e, _, _, _), (g, gv) in zip(tracks, up.jobs, res):
            o, ov = oracle.run(e, soa, tg, shots)
            np.testing.assert_array_equal(gv, ov, err_msg=name)
            print(f"  {name}: {compare_traces(g, o)}", flush=True)
    finally:
        up.release()


def test_tracks_shot_pair_and_single_shot_kernels(gpu_ctx):
    """The default arithmetic's multi-track batches pair each track's
    consecutive rows for the shot-pair kernel (a track with an odd shot count
    ends on a padding row; 15 rows per batch become 16) and replay the
    captured graphs; with ERP_TUNE_SINGLE_SHOT every (row, target) pair runs
    on its own.  Both against the FP64 oracle, replays included."""
    tracks = _tracks()
    up = _Uploaded(gpu_ctx, tracks)
    try:
        gpu_ctx.set_tuning(ERP_TUNE_TRACK_ROWS, 15)
        gpu_ctx.set_tuning(ERP_TUNE_PAIRS_PER_LAUNCH, 60)
        runs = []
        for single in (0, 0, 1):
            gpu_ctx.set_tuning(ERP_TUNE_SINGLE_SHOT, single)
            print(f"multi-track run, single-shot kernel {single}", flush=True)
            runs.append(gpu_ctx.run_tracks(up.jobs))
        for (name, par, soa, tg, shots), (e, _, _, _), *res in zip(tracks, up.jobs, *runs):
            o, ov = oracle.run(e, soa, tg, shots)
            for g, gv in res:
                np.testing.assert_array_equal(gv, ov, err_msg=name)
                compare_traces(g, o)
    finally:
        for key in (ERP_TUNE_TRACK_ROWS, ERP_TUNE_PAIRS_PER_LAUNCH, ERP_TUNE_SINGLE_SHOT):
            gpu_ctx.set_tuning(key, 0)
        up.release()


def test_tracks_record_buffer_grows_for_each_layout():
    """One context's multi-track record buffer serves both record layouts:
    the single-shot kernel's packed records (60 B each) and the shot-pair
    kernel's block records (84 B each).  A deterministic single-shot call
    that sizes it, then a shot-pair call with fewer records than that
    capacity, then the first call again: the buffer must grow for the larger
    layout (a capacity counted in records alone would not), every call must
    match the oracle, and the third the first bit for bit."""
    from europaraypaths_amd._native import Context
    ctx = Context(0)
    tracks = _tracks()[:4]
    up = _Uploaded(ctx, tracks)
    try:
        print("single-shot, shot-pair, single-shot on a fresh context", flush=True)
        ctx.set_tuning(ERP_TUNE_SINGLE_SHOT, 1)
        a = ctx.run_tracks(up.jobs, flags=ERP_RUN_DETERMINISTIC)
        ctx.set_tuning(ERP_TUNE_SINGLE_SHOT, 0)
        b = ctx.run_tracks(up.jobs)
        ctx.set_tuning(ERP_TUNE_SINGLE_SHOT, 1)
        c = ctx.run_tracks(up.jobs, flags=ERP_RUN_DETERMINISTIC)
        for (name, par, soa, tg, shots), (e, *_), (ga, _), (gb, vb), (gc, _) in zip(
                tracks, up.jobs, a, b, c):
            o, ov = oracle.run(e, soa, tg, shots)
            np.testing.assert_array_equal(vb, ov, err_msg=name)
            compare_traces(ga, o)
            compare_traces(gb, o)
            assert np.array_equal(ga.view(np.uint32), gc.view(np.uint32)), name
    finally:
        up.release()
        ctx.close()


def test_tracks_deterministic_equal_each_track_alone(small_batches):
    ctx = small_batches
    tracks = _tracks()
    up = _Uploaded(ctx, tracks)
    det = ERP_RUN_DETERMINISTIC
    try:
        print("multi-track run, deterministic, graphs", flush=True)
        first = ctx.run_tracks(up.jobs, flags=det)
        for (name, *_), (e, m, t, shots), (g, gv) in zip(tracks, up.jobs, first):
            a, av = ctx.run(e, m, t, shots, flags=det)
            np.testing.assert_array_equal(gv, av, err_msg=name)
            # the same values (a zero bin may differ in its sign)
            assert np.array_equal(g, a), name
        # replayed graphs, after eager work on the same stream
        print("multi-track run, deterministic, replay", flush=True)
        again = ctx.run_tracks(up.jobs, flags=det)
        ctx.set_tuning(ERP_TUNE_DIRECT_LAUNCH, 1)
        print("multi-track run, deterministic, direct launches", flush=True)
        direct = ctx.run_tracks(up.jobs, flags=det)
        for (name, *_), (g, _), (g2, _), (g3, _) in zip(tracks, first, again, direct):
            assert np.array_equal(g.view(np.uint32), g2.view(np.uint32)), name
            assert np.array_equal(g.view(np.uint32), g3.view(np.uint32)), name
    finally:
        up.release()


def test_tracks_batch_shapes_do_not_change_deterministic_results(gpu_ctx):
    """Default shapes (64 rows, all pairs of a batch in one launch) against
    8 rows and 33 pairs per launch: the same bits."""
    tracks = _tracks()[:3]
    up = _Uploaded(gpu_ctx, tracks)
    try:
        a = gpu_ctx.run_tracks(up.jobs, flags=ERP_RUN_DETERMINISTIC)
        gpu_ctx.set_tuning(ERP_TUNE_TRACK_ROWS, 8)
        gpu_ctx.set_tuning(ERP_TUNE_PAIRS_PER_LAUNCH, 33)
        b = gpu_ctx.run_tracks(up.jobs, flags=ERP_RUN_DETERMINISTIC)
        for (name, *_), (ga, va), (gb, vb) in zip(tracks, a, b):
            np.testing.assert_array_equal(va, vb, err_msg=name)
            assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32)), name
    finally:
        gpu_ctx.set_tuning(ERP_TUNE_TRACK_ROWS, 0)
        gpu_ctx.set_tuning(ERP_TUNE_PAIRS_PER_LAUNCH, 0)
        up.release()


def test_tracks_rejects_bad_arguments(gpu_ctx):
    from europaraypaths_amd._native import ErpError
    par, soa, shots = scene(nr=1024, h=3000.0, rst=2600.0, ns=2)
    m = gpu_ctx.upload_mesh(soa)
    try:
        with pytest.raises(ErpError):
            gpu_ctx.run_tracks([(par.to_erp(), m, None, shots)], flags=0x40)
        assert gpu_ctx.run_tracks([]) == []
    finally:
        m.release()


def test_tracks_at_full_size_deterministic(gpu_ctx):
    """The batching workload at its stated sizes: the first three tracks of
    the job stream (SeedSequence(4), full shot counts, F 7.3e5-3.6e6, two
    range lengths, one and two media) in one call, deterministic: each
    radargram equals the track run alone bit for bit."""
    from europaraypaths_amd import serve
    jobs = serve.mixed_job_stream(3, None, seed=4, reuse=0.0)
    handles, runs = [], []
    try:
        for j in jobs:
            m = gpu_ctx.upload_mesh(j.scene.soa)
            t = gpu_ctx.upload_targets(*j.scene.targets)
            handles += [m, t]
            runs.append((j.par.to_erp(), m, t, j.shots))
        got = gpu_ctx.run_tracks(runs, flags=ERP_RUN_DETERMINISTIC)
        for j, (e, m, t, sh), (g, gv) in zip(jobs, runs, got):
            a, av = gpu_ctx.run(e, m, t, sh, flags=ERP_RUN_DETERMINISTIC)
            print(f"track F={j.scene.F} ns={j.ns} nr={j.nr} eps3={j.scene.eps_3:.3f} "
                  f"pairs={int(av.sum())}", flush=True)
            np.testing.assert_array_equal(gv, av)
            assert np.array_equal(g, a)
    finally:
        for h in handles:
            h.release()
