"""ctypes binding of liberp.so (include/erp.h).

The library is loaded from this package directory.  There is no fallback:
if the HIP library is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

from ._abi import (ERP_E_NOTREADY, ERP_OK, PRISM_DTYPE, SHOT_DTYPE, ErpParams, ErpTerrain,
                   ErpTrack)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liberp.so")

_lib = None
_lock = threading.Lock()

EXPORTS = ("erp_ctx_create", "erp_ctx_destroy", "erp_last_error", "erp_mesh_upload",
           "erp_mesh_release", "erp_targets_upload", "erp_targets_release", "erp_run",
           "erp_run_dev", "erp_submit", "erp_poll", "erp_wait", "erp_run_footprint",
           "erp_profile_enable", "erp_profile_read", "erp_serve_slots",
           "erp_mesh_from_heights", "erp_mesh_download", "erp_focus", "erp_set_tuning",
           "erp_get_tuning", "erp_run_dev_noalloc", "erp_run_tracks", "erp_mesh_from_terrain")

STAGES = ("mask", "surface", "surface_conv", "target_select", "target_accum", "target_fft",
          "shot_final", "other", "target_sum")


class ErpError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"erp error {code}: {msg}")
        self.code = code


def lib() -> ctypes.CDLL:
    """Load liberp.so once (raises if it is not built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`"
                " or `make -C europaraypaths_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, i32, u32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_uint32
        P = ctypes.POINTER
        L.erp_ctx_create.argtypes = [i32, P(vp)]
        L.erp_ctx_create.restype = i32
        L.erp_ctx_destroy.argtypes = [vp]
        L.erp_ctx_destroy.restype = None
        L.erp_last_error.argtypes = [vp]
        L.erp_last_error.restype = ctypes.c_char_p
        L.erp_mesh_upload.argtypes = [vp, vp, i64, P(vp)]
        L.erp_mesh_upload.restype = i32
        L.erp_mesh_from_heights.argtypes = [vp, vp, i64, i64, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, P(vp)]
        L.erp_mesh_from_heights.restype = i32
        L.erp_mesh_from_terrain.argtypes = [vp, P(ErpTerrain), P(vp)]
        L.erp_mesh_from_terrain.restype = i32
        L.erp_mesh_download.argtypes = [vp, vp]
        L.erp_mesh_download.restype = i32
        L.erp_focus.argtypes = [vp, vp, i64, i64, i64, i64, vp, vp, vp, vp, vp, i64, vp]
        L.erp_focus.restype = i32
        L.erp_mesh_release.argtypes = [vp]
        L.erp_mesh_release.restype = None
        L.erp_targets_upload.argtypes = [vp, vp, vp, i64, P(vp)]
        L.erp_targets_upload.restype = i32
        L.erp_targets_release.argtypes = [vp]
        L.erp_targets_release.restype = None
        run_args = [vp, P(ErpParams), vp, vp, vp, i64, vp, i64, vp, vp, u32]
        L.erp_run.argtypes = run_args
        L.erp_run.restype = i32
        L.erp_run_dev.argtypes = run_args
        L.erp_run_dev.restype = i32
        L.erp_run_dev_noalloc.argtypes = run_args
        L.erp_run_dev_noalloc.restype = i32
        L.erp_run_tracks.argtypes = [vp, P(ErpTrack), i64, u32]
        L.erp_run_tracks.restype = i32
        L.erp_submit.argtypes = run_args + [P(vp)]
        L.erp_submit.restype = i32
        L.erp_poll.argtypes = [vp]
        L.erp_poll.restype = i32
        L.erp_wait.argtypes = [vp]
        L.erp_wait.restype = i32
        L.erp_serve_slots.argtypes = [vp, i32]
        L.erp_serve_slots.restype = i32
        L.erp_run_footprint.argtypes = [P(ErpParams), i64, i64, i64]
        L.erp_run_footprint.restype = i64
        L.erp_profile_enable.argtypes = [vp, i32]
        L.erp_profile_enable.restype = i32
        L.erp_profile_read.argtypes = [vp, i32, P(ctypes.c_double), P(i64), P(ctypes.c_double)]
        L.erp_profile_read.restype = i32
        L.erp_set_tuning.argtypes = [vp, i32, i64]
        L.erp_set_tuning.restype = i32
        L.erp_get_tuning.argtypes = [vp, i32, P(i64)]
        L.erp_get_tuning.restype = i32
        _lib = L
        return L


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


_hip = None


def _hiprt() -> ctypes.CDLL:
    """The HIP runtime, for the device buffers of DeviceTraces."""
    global _hip
    with _lock:
        if _hip is None:
            h = ctypes.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                    ctypes.c_int]
            _hip = h
        return _hip


class DeviceTraces:
    """ns x nr complex64 traces and ns int32 illuminated counts in device
    memory: the outputs of run_dev / run_tracks."""

    def __init__(self, ns: int, nr: int):
        self.ns, self.nr = ns, nr
        self.out, self.nvalid = ctypes.c_void_p(), ctypes.c_void_p()
        h = _hiprt()
        for p, n in ((self.out, ns * nr * 8), (self.nvalid, ns * 4)):
            rc = h.hipMalloc(ctypes.byref(p), max(1, n))
            if rc != 0:
                self.free()
                raise ErpError(rc, f"hipMalloc of {n} bytes failed")

    def read(self):
        """(traces (ns, nr) complex64, nvalid (ns,) int32) copied to the host."""
        o = np.zeros((self.ns, self.nr), np.complex64)
        v = np.zeros(self.ns, np.int32)
        h = _hiprt()
        for a, p in ((o, self.out), (v, self.nvalid)):
            if a.nbytes:
                rc = h.hipMemcpy(a.ctypes.data, p, a.nbytes, 2)  # device to host
                if rc != 0:
                    raise ErpError(rc, "hipMemcpy device to host failed")
        return o, v

    def free(self):
        for p in (self.out, self.nvalid):
            if p.value:
                _hiprt().hipFree(p)
                p.value = None


class Mesh:
    """A facet mesh resident in HBM (erp_mesh_upload / erp_mesh_from_heights)."""

    def __init__(self, ctx: "Context", soa: np.ndarray | None, heights=None, terrain=None):
        self.ctx = ctx
        h = ctypes.c_void_p()
        if terrain is not None:
            t, keep = terrain  # ErpTerrain and the host noise it points to
            self.F = int(t.nx * t.ny)
            ctx._check(lib().erp_mesh_from_terrain(ctx.handle, ctypes.byref(t), ctypes.byref(h)))
            del keep
        elif heights is None:
            soa = np.ascontiguousarray(soa, dtype=np.float32)
            if soa.ndim != 2 or soa.shape[0] != 12:
                raise ValueError("mesh must be a (12, F) float32 SoA array")
            self.F = soa.shape[1]
            ctx._check(lib().erp_mesh_upload(ctx.handle, _ptr(soa), self.F, ctypes.byref(h)))
        else:
            z, fs, ox, oy = heights
            z = np.ascontiguousarray(z, dtype=np.float64)
            if z.ndim != 2:
                raise ValueError("heights must be a (ny, nx) array")
            ny, nx = z.shape
            self.F = nx * ny
            ctx._check(lib().erp_mesh_from_heights(ctx.handle, _ptr(z), nx, ny, float(fs),
                                                   float(ox), float(oy), ctypes.byref(h)))
        self.handle = h

    def download(self) -> np.ndarray:
        """The resident mesh as the (12, F) float32 SoA (erp_mesh_download)."""
        out = np.zeros((12, self.F), np.float32)
        self.ctx._check(lib().erp_mesh_download(self.handle, _ptr(out)))
        return out

    def release(self):
        if getattr(self, "handle", None):
            lib().erp_mesh_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Targets:
    """Subsurface targets resident in HBM (erp_targets_upload)."""

    def __init__(self, ctx: "Context", xyzn: np.ndarray, ttype: np.ndarray):
        xyzn = np.ascontiguousarray(xyzn, dtype=np.float32)
        ttype = np.ascontiguousarray(ttype, dtype=np.uint8)
        if xyzn.ndim != 2 or xyzn.shape[0] != 6 or ttype.shape[0] != xyzn.shape[1]:
            raise ValueError("targets must be (6, T) float32 + (T,) uint8")
        self.ctx = ctx
        self.T = xyzn.shape[1]
        h = ctypes.c_void_p()
        ctx._check(lib().erp_targets_upload(ctx.handle, _ptr(xyzn), _ptr(ttype), self.T,
                                            ctypes.byref(h)))
        self.handle = h

    def release(self):
        if getattr(self, "handle", None):
            lib().erp_targets_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Job:
    """An asynchronous run (erp_submit); keeps its host buffers alive."""

    def __init__(self, ctx, handle, out, nvalid, keep):
        self.ctx, self.handle, self.out, self.nvalid, self._keep = ctx, handle, out, nvalid, keep
        self.done = False

    def poll(self) -> bool:
        if self.done:
            return True
        rc = lib().erp_poll(self.handle)
        if rc == ERP_E_NOTREADY:
            return False
        self.done = True
        self.ctx._check(rc)
        return True

    def wait(self):
        if not self.done:
            rc = lib().erp_wait(self.handle)
            self.done = True
            self.ctx._check(rc)
        return self.out, self.nvalid


class Context:
    """One HIP device context (erp_ctx_create)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        rc = lib().erp_ctx_create(device, ctypes.byref(h))
        if rc != ERP_OK:
            raise ErpError(rc, lib().erp_last_error(None).decode())
        self.handle = h
        self.device = device

    def _check(self, rc: int):
        if rc != ERP_OK:
            raise ErpError(rc, lib().erp_last_error(self.handle).decode())

    def upload_mesh(self, soa) -> Mesh:
        return Mesh(self, soa)

    def mesh_from_heights(self, heights, fs, ox, oy) -> Mesh:
        """Build a mesh on the device from a (ny, nx) height grid with spacing
        fs and origin (ox, oy): the mesh surfaces.terrain_soa would give."""
        return Mesh(self, None, heights=(heights, fs, ox, oy))

    def mesh_from_terrain(self, kind: int, nx: int, ny: int, fs: float, ox: float, oy: float,
                          rms: float, corr_len: float = 0.0, hurst: float = 0.0,
                          seed: int = 0, noise=None) -> Mesh:
        """erp_mesh_from_terrain: a Gaussian-field (ERP_TERRAIN_GAUSSIAN) or fBm
        (ERP_TERRAIN_FBM) terrain generated on the device, and its mesh.
        noise: the host draws (see surfaces.terrain_noise), or None for the
        device's own seeded draws."""
        if noise is not None:
            noise = np.ascontiguousarray(noise, dtype=np.float64)
        t = ErpTerrain(int(kind), 0, int(nx), int(ny), float(fs), float(ox), float(oy),
                       float(rms), float(corr_len), float(hurst), int(seed) & (2 ** 64 - 1),
                       noise.ctypes.data if noise is not None else None)
        return Mesh(self, None, terrain=(t, noise))

    def upload_targets(self, xyzn, ttype) -> Targets:
        return Targets(self, xyzn, ttype)

    @staticmethod
    def _args(params: ErpParams, shots, prisms):
        shots = np.ascontiguousarray(shots, dtype=SHOT_DTYPE)
        if prisms is None or len(prisms) == 0:
            prisms = None
            npr = 0
        else:
            prisms = np.ascontiguousarray(prisms, dtype=PRISM_DTYPE)
            npr = prisms.shape[0]
        return shots, prisms, npr

    def run(self, params: ErpParams, mesh: Mesh, targets: Targets | None, shots,
            prisms=None, flags: int = 0):
        """Synchronous erp_run -> (traces (ns, nr) complex64, nvalid (ns,) int32)."""
        shots, prisms, npr = self._args(params, shots, prisms)
        ns, nr = shots.shape[0], params.nr
        out = np.zeros((ns, nr), dtype=np.complex64)
        nvalid = np.zeros(ns, dtype=np.int32)
        self._check(lib().erp_run(self.handle, ctypes.byref(params), mesh.handle,
                                  targets.handle if targets is not None else None,
                                  _ptr(shots), ns, _ptr(prisms), npr, _ptr(out), _ptr(nvalid),
                                  flags))
        return out, nvalid

    def run_dev(self, params: ErpParams, mesh: Mesh, targets: Targets | None, shots,
                out_ptr: int, nvalid_ptr: int | None = None, prisms=None, flags: int = 0):
        """erp_run_dev with a caller-owned device output pointer."""
        shots, prisms, npr = self._args(params, shots, prisms)
        self._check(lib().erp_run_dev(self.handle, ctypes.byref(params), mesh.handle,
                                      targets.handle if targets is not None else None,
                                      _ptr(shots), shots.shape[0], _ptr(prisms), npr,
                                      ctypes.c_void_p(out_ptr),
                                      ctypes.c_void_p(nvalid_ptr) if nvalid_ptr else None,
                                      flags))

    def run_dev_noalloc(self, params: ErpParams, mesh: Mesh, targets: Targets | None, shots,
                        out_ptr: int, nvalid_ptr: int | None = None, prisms=None,
                        flags: int = 0):
        """erp_run_dev_noalloc: erp_run_dev with every work buffer carved from
        the library's per-device scratch array (no device allocation)."""
        shots, prisms, npr = self._args(params, shots, prisms)
        self._check(lib().erp_run_dev_noalloc(
            self.handle, ctypes.byref(params), mesh.handle,
            targets.handle if targets is not None else None, _ptr(shots), shots.shape[0],
            _ptr(prisms), npr, ctypes.c_void_p(out_ptr),
            ctypes.c_void_p(nvalid_ptr) if nvalid_ptr else None, flags))

    def run_tracks_dev(self, jobs, flags: int = 0):
        """erp_run_tracks over jobs = [(params, mesh, targets or None, shots,
        DeviceTraces)]: every job's radargram and counts into its buffers."""
        keep = []
        arr = (ErpTrack * max(1, len(jobs)))()
        for i, (params, mesh, targets, shots, dev) in enumerate(jobs):
            shots = np.ascontiguousarray(shots, dtype=SHOT_DTYPE)
            if dev.ns != shots.shape[0] or dev.nr != params.nr:
                raise ValueError(f"job {i}: output buffer shape does not match")
            keep.append((params, shots))
            arr[i] = ErpTrack(ctypes.pointer(params), mesh.handle,
                              targets.handle if targets is not None else None,
                              shots.ctypes.data, shots.shape[0], dev.out.value,
                              dev.nvalid.value)
        self._check(lib().erp_run_tracks(self.handle, arr, len(jobs), flags))

    def run_tracks(self, jobs, flags: int = 0):
        """erp_run_tracks over jobs = [(params, mesh, targets or None, shots)]
        -> [(traces (ns, nr) complex64, nvalid (ns,) int32)] on the host."""
        devs = []
        try:
            for params, mesh, targets, shots in jobs:
                devs.append(DeviceTraces(len(shots), params.nr))
            self.run_tracks_dev([j + (d,) for j, d in zip(jobs, devs)], flags)
            return [d.read() for d in devs]
        finally:
            for d in devs:
                d.free()

    def submit(self, params: ErpParams, mesh: Mesh, targets: Targets | None, shots,
               prisms=None, flags: int = 0) -> Job:
        shots, prisms, npr = self._args(params, shots, prisms)
        ns, nr = shots.shape[0], params.nr
        out = np.zeros((ns, nr), dtype=np.complex64)
        nvalid = np.zeros(ns, dtype=np.int32)
        h = ctypes.c_void_p()
        self._check(lib().erp_submit(self.handle, ctypes.byref(params), mesh.handle,
                                     targets.handle if targets is not None else None,
                                     _ptr(shots), ns, _ptr(prisms), npr, _ptr(out),
                                     _ptr(nvalid), flags, ctypes.byref(h)))
        return Job(self, h, out, nvalid, (shots, prisms, params))

    def focus(self, rdr, row0: int, nrows: int, hs, nsx, offs, rows, filt) -> np.ndarray:
        """erp_focus: focused rows row0 .. row0 + nrows - 1 of the complex64
        (Nr, Na) radargram -> (nrows, Na) complex64."""
        rdr = np.ascontiguousarray(rdr, dtype=np.complex64)
        Nr, Na = rdr.shape
        hs = np.ascontiguousarray(hs, dtype=np.int32)
        nsx = np.ascontiguousarray(nsx, dtype=np.int32)
        offs = np.ascontiguousarray(offs, dtype=np.int64)
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        filt = np.ascontiguousarray(filt, dtype=np.complex64)
        out = np.zeros((nrows, Na), np.complex64)
        self._check(lib().erp_focus(self.handle, _ptr(rdr), Nr, Na, row0, nrows, _ptr(hs),
                                    _ptr(nsx), _ptr(offs), _ptr(rows), _ptr(filt), rows.size,
                                    _ptr(out)))
        return out

    def set_tuning(self, key: int, value: int):
        """erp_set_tuning: cap a batch shape (ERP_TUNE_*) on this context; 0 = default."""
        self._check(lib().erp_set_tuning(self.handle, key, value))

    def get_tuning(self, key: int) -> int:
        v = ctypes.c_int64()
        self._check(lib().erp_get_tuning(self.handle, key, ctypes.byref(v)))
        return v.value

    def serve_slots(self, n: int):
        """Start worker slots for submit() up to n (each: thread, stream, buffers)."""
        self._check(lib().erp_serve_slots(self.handle, n))

    def profile(self, on: bool = True):
        """Enable (and clear) or disable the per-stage hipEvent timers."""
        self._check(lib().erp_profile_enable(self.handle, 1 if on else 0))

    def profile_read(self) -> dict:
        """{stage: (ms, launches, units)} accumulated since profile(True)."""
        out = {}
        for k, name in enumerate(STAGES):
            ms, n, u = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
            self._check(lib().erp_profile_read(self.handle, k, ctypes.byref(ms),
                                               ctypes.byref(n), ctypes.byref(u)))
            out[name] = (ms.value, n.value, u.value)
        return out

    @staticmethod
    def footprint(params: ErpParams, F: int, T: int, ns: int) -> int:
        return int(lib().erp_run_footprint(ctypes.byref(params), F, T, ns))

    def close(self):
        if getattr(self, "handle", None):
            lib().erp_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
