"""erp_run_dev_noalloc against erp_run_dev, through the C ABI.

The allocation-free entry point runs the same kernels with every work buffer
carved from the library's per-device scratch array; batch shapes shrink to
fit it.  Its traces must match erp_run_dev's within the parity tolerance
(tests/helpers.py::compare_traces) and, in the deterministic mode, bit for
bit, also where its batches and launches shrink (each shot's fixed-point
scale and 32-row slices do not depend on them).  Calls follow each other with
smaller inputs (the array keeps the previous call's contents), a job whose
per-shot buffers cannot shrink below ERP_SCRATCH_BYTES is rejected with
ERP_E_NOMEM, and the entry point still works after that rejection.
"""
import ctypes

import numpy as np
import pytest

from europaraypaths_amd import surfaces, workloads
from europaraypaths_amd._abi import ERP_E_NOMEM, ERP_RUN_DETERMINISTIC
from europaraypaths_amd._native import ErpError
from europaraypaths_amd.params import make_shots, parse_parameters
from tests.helpers import compare_traces

pytestmark = pytest.mark.gpu


def _hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return h


class DevOut:
    """ns x nr complex64 traces and ns int32 counts in device memory."""

    def __init__(self, ns, nr):
        self.h = _hip()
        self.ns, self.nr = ns, nr
        self.out = ctypes.c_void_p()
        self.nv = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(self.out), max(1, ns * nr * 8)) == 0
        assert self.h.hipMalloc(ctypes.byref(self.nv), max(1, ns * 4)) == 0
        # garbage in the outputs: every value must be written by the run
        assert self.h.hipMemset(self.out, 0x7F, max(1, ns * nr * 8)) == 0
        assert self.h.hipMemset(self.nv, 0x7F, max(1, ns * 4)) == 0

    def read(self):
        o = np.zeros((self.ns, self.nr), np.complex64)
        v = np.zeros(self.ns, np.int32)
        assert self.h.hipMemcpy(o.ctypes.data, self.out, o.nbytes, 2) == 0  # device -> host
        assert self.h.hipMemcpy(v.ctypes.data, self.nv, v.nbytes, 2) == 0
        return o, v

    def free(self):
        self.h.hipFree(self.out)
        self.h.hipFree(self.nv)


def _both(ctx, par, soa, tg, shots, flags=0):
    """(noalloc traces, counts), (erp_run_dev traces, counts)."""
    e = par.to_erp()
    m = ctx.upload_mesh(soa)
    t = ctx.upload_targets(*tg) if tg is not None else None
    res = []
    try:
        for fn in (ctx.run_dev_noalloc, ctx.run_dev):
            d = DevOut(shots.shape[0], e.nr)
            try:
                fn(e, m, t, shots, d.out.value, d.nv.value, flags=flags)
                res.append(d.read())
            finally:
                d.free()
    finally:
        m.release()
        if t is not None:
            t.release()
    return res


def _scene(n, fs, nr, ns, h=3000.0, rst=2600.0, aperture=20.0, ntg=5, **over):
    ox = oy = -fs * n / 2
    z = surfaces.gaussian_heights(n, n, fs, 0.4, 6.0 * fs, seed=4)
    soa = surfaces.terrain_soa(n, n, fs, ox, oy, z)
    p = workloads.make_params("REASON_VHF", "planetary_ice", nr=nr, smpl=48e6, rst=rst,
                              aperture=aperture, fs=fs, ns=ns, sx0=-0.5 * fs * ns / 4,
                              sdx=fs / 4, sy=3.0, sz=h, **over)
    par = parse_parameters(p)
    xyzn = np.zeros((6, ntg), np.float32)
    xyzn[0] = np.linspace(-0.3, 0.3, ntg) * fs * n
    xyzn[1] = 0.1 * fs * n
    xyzn[2] = -300.0
    xyzn[5] = 1.0
    return par, soa, (xyzn, np.ones(ntg, np.uint8)), make_shots(par)


def test_noalloc_matches_run_dev_two_calls(gpu_ctx):
    """A 40-shot run (two batches, per-pair chain at nr = 1024), then a
    smaller one (fewer facets, shots and targets): same traces as
    erp_run_dev, and in the deterministic mode the same bits."""
    cases = [_scene(96, 10.0, 1024, 40), _scene(48, 10.0, 1024, 7, ntg=2)]
    for k, (par, soa, tg, shots) in enumerate(cases):
        print(f"noalloc call {k}: F={soa.shape[1]} ns={shots.shape[0]}", flush=True)
        (g, gv), (r, rv) = _both(gpu_ctx, par, soa, tg, shots)
        np.testing.assert_array_equal(gv, rv)
        assert (gv > 0).all()
        compare_traces(g, r)
        (g, gv), (r, rv) = _both(gpu_ctx, par, soa, tg, shots, flags=ERP_RUN_DETERMINISTIC)
        np.testing.assert_array_equal(gv, rv)
        assert np.array_equal(g.view(np.uint32), r.view(np.uint32))


def test_noalloc_cropped_chain_and_deterministic_specular(gpu_ctx):
    """The cropped per-pair chain (nr = 3000, whose range kernel spectrum the
    run builds in the scratch array) and a deterministic specular run with
    typed and type-0 targets (the fixed-point pair windows and the per-shot
    type-0 rows): the bits of erp_run_dev."""
    par, soa, tg, shots = _scene(64, 10.0, 3000, 6, ntg=4)
    assert par.nr == 3000
    print("noalloc cropped chain", flush=True)
    (g, gv), (r, rv) = _both(gpu_ctx, par, soa, tg, shots, flags=ERP_RUN_DETERMINISTIC)
    np.testing.assert_array_equal(gv, rv)
    assert np.array_equal(g.view(np.uint32), r.view(np.uint32))
    par, soa, (xyzn, ttype), shots = _scene(64, 10.0, 1024, 6, ntg=6, specular=True)
    ttype[::2] = 0
    print("noalloc deterministic specular", flush=True)
    (g, gv), (r, rv) = _both(gpu_ctx, par, soa, (xyzn, ttype), shots,
                             flags=ERP_RUN_DETERMINISTIC)
    np.testing.assert_array_equal(gv, rv)
    assert np.array_equal(g.view(np.uint32), r.view(np.uint32))


def test_noalloc_wave_local_chain(gpu_ctx):
    """nr = 6000 (n = 8192, the cropped wave-local chain, whose permuted
    range kernel spectrum the run builds in the scratch array): erp_run_dev's
    traces within tolerance, and its bits in the deterministic mode."""
    par, soa, tg, shots = _scene(64, 10.0, 6000, 6, ntg=4)
    assert par.nr == 6000
    print("noalloc wave-local chain, nr 6000", flush=True)
    (g, gv), (r, rv) = _both(gpu_ctx, par, soa, tg, shots)
    np.testing.assert_array_equal(gv, rv)
    compare_traces(g, r)
    (g, gv), (r, rv) = _both(gpu_ctx, par, soa, tg, shots, flags=ERP_RUN_DETERMINISTIC)
    np.testing.assert_array_equal(gv, rv)
    assert np.array_equal(g.view(np.uint32), r.view(np.uint32))


def test_noalloc_shrinks_batches_to_fit(gpu_ctx):
    """nr = 16384 (256 KB pair rows: the launch takes fewer pairs) and a
    mesh of 331 776 facets lit by every shot (their records, 20 MB per shot,
    cut each batch to the shots that fit, which pairs other shots), then a
    small call again; each also in the deterministic mode, where the
    smaller batches and launches must give erp_run_dev's bits."""
    cases = [_scene(128, 10.0, 16384, 5, ntg=40),
             _scene(576, 4.0, 2048, 12, aperture=80.0, ntg=3),
             _scene(32, 10.0, 2048, 3, ntg=2)]
    for k, (par, soa, tg, shots) in enumerate(cases):
        print(f"noalloc shrink call {k}: F={soa.shape[1]} nr={par.nr} ns={shots.shape[0]}",
              flush=True)
        (g, gv), (r, rv) = _both(gpu_ctx, par, soa, tg, shots)
        np.testing.assert_array_equal(gv, rv)
        assert (gv > 0).all()
        compare_traces(g, r)
        print(f"noalloc shrink call {k}, deterministic", flush=True)
        (g, gv), (r, rv) = _both(gpu_ctx, par, soa, tg, shots, flags=ERP_RUN_DETERMINISTIC)
        np.testing.assert_array_equal(gv, rv)
        assert np.array_equal(g.view(np.uint32), r.view(np.uint32))


def test_noalloc_rejects_a_shot_that_cannot_fit(gpu_ctx):
    """5 242 880 facets lit by one shot need 315 MB of records, more than
    ERP_SCRATCH_BYTES: ERP_E_NOMEM with the sizes; the next call runs."""
    nx, ny, fs = 2560, 2048, 1.0
    z = np.zeros((ny, nx))
    p = workloads.make_params("REASON_VHF", "planetary_ice", nr=1024, smpl=48e6, rst=2600.0,
                              aperture=60.0, fs=fs, ns=1, sx0=0.0, sdx=1.0, sy=0.0, sz=3000.0)
    par = parse_parameters(p)
    shots = make_shots(par)
    m = gpu_ctx.mesh_from_heights(z, fs, -fs * nx / 2, -fs * ny / 2)
    xyzn = np.zeros((6, 1), np.float32)
    xyzn[2] = -300.0
    xyzn[5] = 1.0
    t = gpu_ctx.upload_targets(xyzn, np.ones(1, np.uint8))
    d = DevOut(1, par.nr)
    try:
        with pytest.raises(ErpError) as ei:
            gpu_ctx.run_dev_noalloc(par.to_erp(), m, t, shots, d.out.value, d.nv.value)
        assert ei.value.code == ERP_E_NOMEM
        assert "allocation-free scratch" in str(ei.value), str(ei.value)
    finally:
        d.free()
        m.release()
        t.release()
    par, soa, tg, shots = _scene(32, 10.0, 1024, 3, ntg=2)
    (g, gv), (r, rv) = _both(gpu_ctx, par, soa, tg, shots)
    np.testing.assert_array_equal(gv, rv)
    compare_traces(g, r)
