"""Batch-shape caps (erp_set_tuning): tiny shot batches and target launches
split shots across batches and one shot's (shot, target) pairs across
several launches, slices and groups.  The traces must not notice: each run
is compared with the FP64 oracle, and with the default shapes."""
import numpy as np
import pytest

from europaraypaths_amd._abi import (ERP_E_INTERNAL, ERP_E_INVALID, ERP_RUN_DETERMINISTIC,
                                    ERP_TUNE_PAIRS_PER_LAUNCH, ERP_TUNE_SHOTS_PER_BATCH,
                                    ERP_TUNE_SINGLE_SHOT, ERP_TUNE_WINDOW_SHIFT)
from europaraypaths_amd._native import ErpError
from oracle import oracle
from tests.helpers import compare_traces
from tests.test_gpu_parity import _targets, scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nr", [4096, 2048])
def test_small_batches_and_launches_give_the_same_traces(gpu_ctx, nr):
    # nr 4096: per-pair LDS chain; 2048: the fused multi-kernel chain
    par, soa, shots = scene(rough=0.3, nr=nr, h=3000.0, rst=2600.0, ns=5)
    tg = _targets(par, n=41)
    e = par.to_erp()
    m = gpu_ctx.upload_mesh(soa)
    t = gpu_ctx.upload_targets(*tg)
    o, ov = oracle.run(e, soa, tg, shots)
    g0, gv0 = gpu_ctx.run(e, m, t, shots)
    try:
        gpu_ctx.set_tuning(ERP_TUNE_SHOTS_PER_BATCH, 2)
        gpu_ctx.set_tuning(ERP_TUNE_PAIRS_PER_LAUNCH, 7)
        assert gpu_ctx.get_tuning(ERP_TUNE_PAIRS_PER_LAUNCH) == 7
        g1, gv1 = gpu_ctx.run(e, m, t, shots)
    finally:
        gpu_ctx.set_tuning(ERP_TUNE_SHOTS_PER_BATCH, 0)
        gpu_ctx.set_tuning(ERP_TUNE_PAIRS_PER_LAUNCH, 0)
    np.testing.assert_array_equal(gv0, ov)
    np.testing.assert_array_equal(gv1, ov)
    compare_traces(g0, o)
    compare_traces(g1, o)
    compare_traces(g1, g0)


def test_tuning_keys_are_checked(gpu_ctx):
    for key, value in ((0, 1), (7, 1), (ERP_TUNE_SHOTS_PER_BATCH, -1)):
        with pytest.raises(ErpError) as ex:
            gpu_ctx.set_tuning(key, value)
        assert ex.value.code == ERP_E_INVALID
    assert gpu_ctx.get_tuning(ERP_TUNE_SHOTS_PER_BATCH) == 0


def test_window_check_fails_a_run_whose_window_misses_terms(gpu_ctx):
    """A pair's bin window must hold every term its bounds allow.  Shifting
    every window up by 16 bins (ERP_TUNE_WINDOW_SHIFT, a test-only knob)
    leaves the lowest terms outside: the run must fail with ERP_E_INTERNAL
    (single track with the shot-pair and the single-shot kernel,
    deterministic mode, multi-track batch) instead of dropping them, and
    pass again once the shift is off."""
    par, soa, shots = scene(rough=0.3, nr=1024, h=3000.0, rst=2600.0, ns=3)
    tg = _targets(par, n=5)
    e = par.to_erp()
    m = gpu_ctx.upload_mesh(soa)
    t = gpu_ctx.upload_targets(*tg)
    try:
        gpu_ctx.set_tuning(ERP_TUNE_WINDOW_SHIFT, 16)
        try:
            for flags, single in ((0, 0), (0, 1), (ERP_RUN_DETERMINISTIC, 0)):
                gpu_ctx.set_tuning(ERP_TUNE_SINGLE_SHOT, single)
                with pytest.raises(ErpError) as ex:
                    gpu_ctx.run(e, m, t, shots, flags=flags)
                assert ex.value.code == ERP_E_INTERNAL, ex.value
                assert "window" in str(ex.value)
            gpu_ctx.set_tuning(ERP_TUNE_SINGLE_SHOT, 0)
            with pytest.raises(ErpError) as ex:
                gpu_ctx.run_tracks([(e, m, t, shots)])
            assert ex.value.code == ERP_E_INTERNAL, ex.value
        finally:
            gpu_ctx.set_tuning(ERP_TUNE_WINDOW_SHIFT, 0)
            gpu_ctx.set_tuning(ERP_TUNE_SINGLE_SHOT, 0)
        o, ov = oracle.run(e, soa, tg, shots)
        for g, gv in (gpu_ctx.run(e, m, t, shots), gpu_ctx.run_tracks([(e, m, t, shots)])[0]):
            np.testing.assert_array_equal(gv, ov)
            compare_traces(g, o)
    finally:
        m.release()
        t.release()
