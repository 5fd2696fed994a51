"""SAR focusing on the GPU (erp_focus) against the oracle's restatement of
focus_rdrgrm (simple_focusing.py:310-404): per-pixel phase within 1e-3 rad
and amplitude within 1e-4 relative on pixels within 40 dB of each row's
peak, 1e-4 x peak absolute everywhere."""
import numpy as np
import pytest

from europaraypaths_amd import focusing
from oracle import focusing as ofo
from tests.helpers import compare_traces

pytestmark = pytest.mark.gpu

PAR = {"rx_window_m": 3000.0, "rx_window_offset_m": 4800.0, "rx_sample_rate": 48e6,
       "aperture": 10.0, "spacing": 50.0, "altitude": 5000.0, "eps_2": 3.15,
       "lam": 299792458 / 9e6}


@pytest.mark.parametrize("st,en", [(20, 60), (None, None)])
def test_focus_matches_oracle(gpu_ctx, st, en):
    rng = np.random.default_rng(7)
    Nr, Na = 96, 40
    rdr = (rng.standard_normal((Nr, Na)) + 1j * rng.standard_normal((Nr, Na))).astype(np.complex64)
    got = focusing.focus_rdrgrm(rdr, PAR, st=st, en=en, ctx=gpu_ctx)
    want = ofo.focus_rdrgrm(rdr.astype(np.complex128), PAR, st=st, en=en)
    rows = slice(st, en + 1) if st else slice(0, Nr)
    assert np.all(got[:rows.start] == 0) and np.all(got[rows.stop:] == 0)
    compare_traces(got[rows], want[rows])


def test_focus_point_target_gain(gpu_ctx):
    """A point target synthesised with the matched filter's own phase along
    its range hyperbola focuses to a peak at its trace with gain ~ aperture."""
    Nr, Na, i0, j0 = 128, 64, 40, 32
    n, hs, rb, mf = focusing.row_tables(PAR, i0, Nr)
    rdr = np.zeros((Nr, Na), np.complex64)
    L = 2 * n
    az_st = j0 - hs
    nn = (j0 - az_st) + (L - 2 * hs) // 2
    for m in range(2 * hs):
        t = (nn - m) % L
        if t < n:
            rdr[rb[m] % Nr, az_st + m] += np.conj(mf[t])
    foc = focusing.focus_rdrgrm(rdr, PAR, st=i0, en=i0, ctx=gpu_ctx)
    k = int(np.argmax(np.abs(foc[i0])))
    assert k == j0
    assert abs(foc[i0, j0]) > 0.9 * (2 * hs - 1)
