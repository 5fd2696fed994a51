"""ctypes wrapper of liberp_oracle.so — TEST INFRASTRUCTURE ONLY.

The CPU restatement of the reference path (erp_oracle.cpp) that tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg compare against.
The product package never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from europaraypaths_amd._abi import PRISM_DTYPE, SHOT_DTYPE, ErpParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liberp_oracle.so")
_lib = None


def build():
    """Compile the oracle with its Makefile (g++ only)."""
    subprocess.run(["make", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        L.erp_oracle_run.argtypes = [ctypes.POINTER(ErpParams), vp, i64, vp, vp, i64, vp, i64,
                                     vp, i64, i32, i32, vp, vp, vp]
        L.erp_oracle_run.restype = i32
        L.erp_oracle_facets.argtypes = [ctypes.POINTER(ErpParams), vp, i64, vp, i32, vp]
        L.erp_oracle_facets.restype = i32
        L.erp_oracle_target_pairs.argtypes = [ctypes.POINTER(ErpParams), vp, i64, vp, vp, i64,
                                              i64, vp, vp, i64, i32, vp]
        L.erp_oracle_target_pairs.restype = i32
        L.erp_oracle_dft.argtypes = [vp, i64, i32, vp]
        L.erp_oracle_dft.restype = i32
        d = ctypes.c_double
        L.erp_oracle_conv.argtypes = [i32, vp, vp, i32, d, d, d, d, d, i32, vp]
        L.erp_oracle_conv.restype = i32
        L.erp_oracle_version.restype = ctypes.c_char_p
        L.erp_oracle_set_mutation.argtypes = [i32]
        L.erp_oracle_set_mutation.restype = i32
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def run(params: ErpParams, soa, targets=None, shots=None, prisms=None, precision=64,
        nthreads=0, parts=False):
    """Oracle shot loop.  Returns (out (ns, nr) complex128, nvalid (ns,) [, parts])."""
    soa = np.ascontiguousarray(soa, dtype=np.float32)
    F = soa.shape[1]
    shots = np.ascontiguousarray(shots, dtype=SHOT_DTYPE)
    ns, nr = shots.shape[0], params.nr
    if targets is not None and targets[0].shape[1] > 0:
        xyzn = np.ascontiguousarray(targets[0], dtype=np.float32)
        tt = np.ascontiguousarray(targets[1], dtype=np.uint8)
        T = xyzn.shape[1]
    else:
        xyzn, tt, T = None, None, 0
    if prisms is not None and len(prisms):
        pr = np.ascontiguousarray(prisms, dtype=PRISM_DTYPE)
        npr = pr.shape[0]
    else:
        pr, npr = None, 0
    out = np.zeros((ns, nr, 2), dtype=np.float64)
    nvalid = np.zeros(ns, dtype=np.int32)
    prt = np.zeros((ns, 3, nr, 2), dtype=np.float64) if parts else None
    rc = lib().erp_oracle_run(ctypes.byref(params), _p(soa), F, _p(xyzn), _p(tt), T, _p(shots),
                              ns, _p(pr), npr, precision, nthreads, _p(out), _p(prt),
                              _p(nvalid))
    if rc != 0:
        raise RuntimeError(f"oracle run failed: {rc}")
    o = out[..., 0] + 1j * out[..., 1]
    if parts:
        return o, nvalid, prt[..., 0] + 1j * prt[..., 1]
    return o, nvalid


def facets(params: ErpParams, soa, shot, precision=64):
    """Per-facet intermediates (F, 12): illum, Itd, Ith, Iph, Rth, fRe, fRfrC,
    G, Psrc, bin, w, phase."""
    soa = np.ascontiguousarray(soa, dtype=np.float32)
    sh = np.ascontiguousarray(np.atleast_1d(shot), dtype=SHOT_DTYPE)
    out = np.zeros((soa.shape[1], 12))
    rc = lib().erp_oracle_facets(ctypes.byref(params), _p(soa), soa.shape[1], _p(sh), precision,
                                 _p(out))
    if rc:
        raise RuntimeError(rc)
    return out


def target_pairs(params: ErpParams, soa, targets, it, shot, prisms=None, precision=64):
    """Per (facet, target it) intermediates (F, 10): illum, Ttd, Tth,
    TargetTh, Tph, EI, EO, rngt, Pray_in, Pray_out."""
    soa = np.ascontiguousarray(soa, dtype=np.float32)
    xyzn = np.ascontiguousarray(targets[0], dtype=np.float32)
    tt = np.ascontiguousarray(targets[1], dtype=np.uint8)
    sh = np.ascontiguousarray(np.atleast_1d(shot), dtype=SHOT_DTYPE)
    pr = None if prisms is None or not len(prisms) else np.ascontiguousarray(prisms, PRISM_DTYPE)
    out = np.zeros((soa.shape[1], 10))
    rc = lib().erp_oracle_target_pairs(ctypes.byref(params), _p(soa), soa.shape[1], _p(xyzn),
                                       _p(tt), xyzn.shape[1], it, _p(sh), _p(pr),
                                       0 if pr is None else pr.shape[0], precision, _p(out))
    if rc:
        raise RuntimeError(rc)
    return out


def dft(x, inverse=False):
    x = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    a = np.stack([x.real, x.imag], axis=-1).copy()
    o = np.zeros_like(a)
    rc = lib().erp_oracle_dft(_p(a), x.shape[0], 1 if inverse else 0, _p(o))
    if rc:
        raise RuntimeError(rc)
    return o[:, 0] + 1j * o[:, 1]


def conv(kind, x, y=None, nr=None, dr=1.0, res=1.0, rst=0.0, smpl=1.0, B=1.0, precision=64):
    """kind 0: linear chirp convolution; 1: circular; 2: convolveComplex(x, y)."""
    x = np.asarray(x, dtype=np.complex128)
    nr = x.shape[0] if nr is None else nr
    a = np.stack([x.real, x.imag], -1).copy()
    b = None
    if y is not None:
        y = np.asarray(y, dtype=np.complex128)
        b = np.stack([y.real, y.imag], -1).copy()
    o = np.zeros((nr, 2))
    rc = lib().erp_oracle_conv(kind, _p(a), _p(b), nr, dr, res, rst, smpl, B, precision, _p(o))
    if rc:
        raise RuntimeError(rc)
    return o[:, 0] + 1j * o[:, 1]


# planted mistakes (erp_oracle.cpp MUT_*): for tests that show a pin catches them
MUT_FACET_G_PI_ONCE = 1
MUT_RADAREQ_NO_V = 2
MUT_RADAREQ_4PI_CUBED = 3
MUT_TPH_SWAPPED = 4
MUT_TARGETTH_FACET_N = 5
MUT_INCOHERENT_HALF = 6
MUT_FACETRERAD_PI_TWICE = 7
MUT_REFR_INDEX_DROPPED = 8
MUT_NO_TRANSMISSION = 9


class mutation:
    """Context manager: run the oracle with one planted mistake."""

    def __init__(self, m: int):
        self.m = m

    def __enter__(self):
        self.prev = lib().erp_oracle_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        lib().erp_oracle_set_mutation(self.prev)
        return False
