"""Edge cases of the HIP path through the C ABI: empty inputs, a single
facet, empty target sets, and argument errors (each reported as a status
code, with the context still usable afterwards)."""
import numpy as np
import pytest

from europaraypaths_amd import surfaces, workloads
from europaraypaths_amd._abi import (ERP_E_INVALID, ERP_E_UNSUPPORTED, ERP_RUN_DETERMINISTIC,
                                    ERP_RUN_PHASE_DF, ERP_RUN_PHASE_F32)
from europaraypaths_amd._native import Context, ErpError
from europaraypaths_amd.params import make_shots, parse_parameters
from oracle import oracle
from tests.helpers import compare_traces

pytestmark = pytest.mark.gpu


def _par(nr=256, ns=2, **over):
    p = workloads.make_params("REASON_VHF", "planetary_ice", nr=nr, smpl=48e6, rst=4600.0,
                              aperture=20.0, fs=20.0, ns=ns, sx0=-30.0, sdx=25.0, sy=7.0,
                              sz=5000.0, **over)
    return parse_parameters(p)


def test_empty_mesh_gives_zero_traces(gpu_ctx):
    par = _par()
    shots = make_shots(par)
    m = gpu_ctx.upload_mesh(np.zeros((12, 0), np.float32))
    g, gv = gpu_ctx.run(par.to_erp(), m, None, shots)
    o, ov = oracle.run(par.to_erp(), np.zeros((12, 0), np.float32), None, shots)
    assert g.shape == (2, par.nr)
    assert np.all(g == 0) and np.all(gv == 0)
    np.testing.assert_array_equal(gv, ov)
    assert np.all(o == 0)


def test_no_shots(gpu_ctx):
    par = _par()
    soa = surfaces.terrain_soa(8, 8, 20.0, -80.0, -80.0, np.zeros((8, 8)))
    m = gpu_ctx.upload_mesh(soa)
    g, gv = gpu_ctx.run(par.to_erp(), m, None, make_shots(par)[:0])
    assert g.shape == (0, par.nr) and gv.shape == (0,)


def test_single_facet_below_the_track(gpu_ctx):
    par = _par()
    shots = make_shots(par)
    soa = surfaces.terrain_soa(1, 1, 20.0, -10.0, -10.0, np.zeros((1, 1)))
    g, gv = gpu_ctx.run(par.to_erp(), gpu_ctx.upload_mesh(soa), None, shots)
    o, ov = oracle.run(par.to_erp(), soa, None, shots)
    np.testing.assert_array_equal(gv, ov)
    assert (gv == 1).all()
    compare_traces(g, o)


def test_empty_target_set_equals_no_targets(gpu_ctx):
    par = _par(nr=512)
    shots = make_shots(par)
    z = surfaces.gaussian_heights(24, 24, 20.0, 0.4, 60.0, 3)
    soa = surfaces.terrain_soa(24, 24, 20.0, -240.0, -240.0, z)
    m = gpu_ctx.upload_mesh(soa)
    t = gpu_ctx.upload_targets(np.zeros((6, 0), np.float32), np.zeros(0, np.uint8))
    a, av = gpu_ctx.run(par.to_erp(), m, None, shots, flags=ERP_RUN_DETERMINISTIC)
    b, bv = gpu_ctx.run(par.to_erp(), m, t, shots, flags=ERP_RUN_DETERMINISTIC)
    np.testing.assert_array_equal(av, bv)
    np.testing.assert_array_equal(a, b)


def test_argument_errors_leave_the_context_usable(gpu_ctx):
    par = _par()
    shots = make_shots(par)
    soa = surfaces.terrain_soa(4, 4, 20.0, -40.0, -40.0, np.zeros((4, 4)))
    m = gpu_ctx.upload_mesh(soa)
    e = par.to_erp()
    for flags, code in ((1 << 30, ERP_E_INVALID),
                        (ERP_RUN_PHASE_F32 | ERP_RUN_PHASE_DF, ERP_E_INVALID)):
        with pytest.raises(ErpError) as ex:
            gpu_ctx.run(e, m, None, shots, flags=flags)
        assert ex.value.code == code
    bad = par.to_erp()
    bad.nr = 0
    with pytest.raises(ErpError) as ex:
        gpu_ctx.run(bad, m, None, shots)
    assert ex.value.code == ERP_E_INVALID
    spec = _par(specular=True).to_erp()
    with pytest.raises(ErpError) as ex:
        gpu_ctx.run(spec, m, None, shots, flags=ERP_RUN_PHASE_DF)
    assert ex.value.code == ERP_E_UNSUPPORTED
    other = Context(0)
    try:
        with pytest.raises(ErpError) as ex:
            other.run(e, m, None, shots)
        assert ex.value.code == ERP_E_INVALID
    finally:
        other.close()
    g, gv = gpu_ctx.run(e, m, None, shots)
    o, ov = oracle.run(e, soa, None, shots)
    np.testing.assert_array_equal(gv, ov)
    compare_traces(g, o)
