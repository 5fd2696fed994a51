"""Precision modes of the target accumulation at orbital range (400 km),
for the 9 MHz and the 60 MHz chirp, against the FP64 oracle; the study
itself is tests/precision_study.py.

At 4e5 m an FP32 range carries 0.03 m of rounding, most of a cycle at
60 MHz, so the reference's FP32 arithmetic (ERP_RUN_PHASE_F32) must miss
the 1e-3 rad phase bar there, while the FP64 product and the double-float
variant must meet it.  Summing the bin windows in plain or compensated FP32
instead of FP64 must keep the result within the same tolerance.
"""
import numpy as np
import pytest

from europaraypaths_amd._abi import (ERP_RUN_ACC_COMP, ERP_RUN_ACC_F32, ERP_RUN_PHASE_DF,
                                     ERP_RUN_PHASE_F32)
from oracle import oracle
from tests.helpers import compare_traces
from tests.precision_study import errors, orbital_scene

pytestmark = pytest.mark.gpu


# (platform, lower bound on the FP32 phase error, rad): FP32 ranges at
# 4e5 m carry 0.03 m of rounding, 1.1e-2 rad at 9 MHz and 7.5e-2 at 60 MHz
CHIRPS = [("REASON_HF", 5e-3), ("REASON_VHF", 2e-2)]
# the surface echo alone: 3.9e-3 rad at 9 MHz and about 0.1 rad at 60 MHz
# measured on MI355X (FP32 Itd and phasor(2 Itd) of surfacePT)
SURF_CHIRPS = [("REASON_HF", 2e-3), ("REASON_VHF", 2e-2)]


@pytest.mark.parametrize("platform,fp32_min", CHIRPS, ids=["9MHz", "60MHz"])
def test_precision_modes_at_400_km(gpu_ctx, platform, fp32_min):
    print(f"precision study {platform}", flush=True)
    par, soa, tg, shots = orbital_scene(platform, grid=128, ns=2)
    e = par.to_erp()
    ref, rv = oracle.run(e, soa, tg, shots)
    m = gpu_ctx.upload_mesh(soa)
    t = gpu_ctx.upload_targets(*tg)
    try:
        res = {}
        for name, flags in (("fp64", 0), ("df", ERP_RUN_PHASE_DF), ("fp32", ERP_RUN_PHASE_F32),
                            ("acc32", ERP_RUN_ACC_F32), ("comp", ERP_RUN_ACC_COMP)):
            print(f"precision mode {name}", flush=True)
            g, gv = gpu_ctx.run(e, m, t, shots, flags=flags)
            np.testing.assert_array_equal(gv, rv)
            res[name] = (g, errors(g, ref))
        for name in ("fp64", "df", "acc32", "comp"):
            compare_traces(res[name][0], ref)
        # FP32 ranges and phase: milliradians at 9 MHz, tens at 60 MHz
        assert res["fp32"][1]["phase_max"] > fp32_min, res["fp32"][1]
        assert res["df"][1]["phase_max"] < 1e-3, res["df"][1]
        # compensated FP32 sums reproduce the FP64 sums closely
        d = np.abs(res["comp"][0] - res["fp64"][0]).max() / np.abs(ref).max()
        assert d < 1e-6, d
    finally:
        m.release()
        t.release()


@pytest.mark.parametrize("platform,fp32_min", SURF_CHIRPS, ids=["9MHz", "60MHz"])
def test_precision_modes_surface_echo_at_400_km(gpu_ctx, platform, fp32_min):
    """The flags act on the surface echo as well (surface kernel: range, bin
    and phase arithmetic, LDS window sums): the orbital scene's surface echo
    alone, no targets."""
    print(f"precision study, surface echo, {platform}", flush=True)
    par, soa, _, shots = orbital_scene(platform, grid=128, ns=2, surface=True)
    e = par.to_erp()
    ref, rv = oracle.run(e, soa, None, shots)
    m = gpu_ctx.upload_mesh(soa)
    try:
        res = {}
        for name, flags in (("fp64", 0), ("df", ERP_RUN_PHASE_DF), ("fp32", ERP_RUN_PHASE_F32),
                            ("acc32", ERP_RUN_ACC_F32), ("comp", ERP_RUN_ACC_COMP)):
            g, gv = gpu_ctx.run(e, m, None, shots, flags=flags)
            np.testing.assert_array_equal(gv, rv)
            res[name] = (g, errors(g, ref))
            print(f"  {name}: {res[name][1]}", flush=True)
        for name in ("fp64", "df", "acc32", "comp"):
            compare_traces(res[name][0], ref)
        assert res["fp32"][1]["phase_max"] > fp32_min, res["fp32"][1]
        assert res["df"][1]["phase_max"] < 1e-3, res["df"][1]
        d = np.abs(res["comp"][0] - res["fp64"][0]).max() / np.abs(ref).max()
        assert d < 1e-5, d
    finally:
        m.release()


def test_precision_flags_rejected_where_they_do_not_apply(gpu_ctx):
    from europaraypaths_amd._abi import ERP_RUN_DETERMINISTIC
    from europaraypaths_amd._native import ErpError
    par, soa, tg, shots = orbital_scene("REASON_HF", grid=16, ns=1)
    m = gpu_ctx.upload_mesh(soa)
    t = gpu_ctx.upload_targets(*tg)
    try:
        for flags in (ERP_RUN_PHASE_F32 | ERP_RUN_PHASE_DF, ERP_RUN_ACC_F32 | ERP_RUN_ACC_COMP,
                      ERP_RUN_PHASE_DF | ERP_RUN_DETERMINISTIC, 0x40):
            with pytest.raises(ErpError):
                gpu_ctx.run(par.to_erp(), m, t, shots, flags=flags)
    finally:
        m.release()
        t.release()
