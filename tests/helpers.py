"""Shared comparison helpers for the parity tests."""
import numpy as np


def compare_traces(gpu, ref, rel_amp=1e-4, phase=1e-3, floor_db=40.0, abs_rel=1e-4):
    """Per-bin parity of complex traces (..., nr).

    Bins whose oracle magnitude is within `floor_db` of the trace peak must
    agree within `phase` rad and `rel_amp` relative amplitude; every bin must
    agree within abs_rel * peak in absolute terms.  Returns a dict of the
    worst errors (for messages) and raises AssertionError on failure.
    """
    gpu = np.asarray(gpu, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    assert gpu.shape == ref.shape, (gpu.shape, ref.shape)
    g2 = gpu.reshape(-1, gpu.shape[-1])
    r2 = ref.reshape(-1, ref.shape[-1])
    worst = {"phase": 0.0, "amp": 0.0, "abs": 0.0, "nbins": 0}
    for g, r in zip(g2, r2):
        peak = np.max(np.abs(r))
        if peak == 0.0:
            assert np.max(np.abs(g)) == 0.0, "oracle trace is zero but GPU trace is not"
            continue
        sel = np.abs(r) >= peak * 10 ** (-floor_db / 20)
        ph = np.abs(np.angle(g[sel] / r[sel]))
        am = np.abs(np.abs(g[sel]) / np.abs(r[sel]) - 1)
        ab = np.max(np.abs(g - r)) / peak
        worst["phase"] = max(worst["phase"], float(ph.max(initial=0)))
        worst["amp"] = max(worst["amp"], float(am.max(initial=0)))
        worst["abs"] = max(worst["abs"], float(ab))
        worst["nbins"] += int(sel.sum())
    assert worst["phase"] <= phase, f"phase error {worst}"
    assert worst["amp"] <= rel_amp, f"amplitude error {worst}"
    assert worst["abs"] <= abs_rel, f"absolute error {worst}"
    return worst
