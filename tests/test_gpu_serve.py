"""Served jobs give the same radargrams as the same jobs run alone.

A mixed-size job stream (several meshes, one of them shared by several
jobs, 2048-16384 range bins, 9 and 60 MHz, 1-3 target layers) is served
with every job arriving at once, so that up to four jobs run on the device
together in the library's worker slots.  Each job is then run alone on the
plain path.  Illuminated-facet counts must match exactly; traces must agree
within the path's stated tolerance (phase 1e-3 rad and amplitude 1e-4
relative on bins within 40 dB of each trace's peak, 1e-4 x peak absolute on
every bin): the only difference between the two runs is the order of the
floating-point atomic additions into range bins.
"""
import numpy as np
import pytest

from europaraypaths_amd import serve
from europaraypaths_amd._abi import ERP_MODE_CONVOLUTION, ERP_RUN_DETERMINISTIC
from tests.helpers import compare_traces

pytestmark = pytest.mark.gpu


def small_stream(n=10, seed=21):
    return serve.mixed_job_stream(n, None, seed=seed, reuse=0.4, ns_scale=1 / 200,
                                  f_range=(3e3, 6e4))


def test_each_job_is_the_same_alone_and_in_the_mix(gpu_ctx):
    jobs = small_stream()
    assert len({j.scene.mesh_id for j in jobs}) < len(jobs)    # a shared mesh
    srv = serve.Server(gpu_ctx, budget_bytes=8e9, slots=4)
    rep = srv.run(jobs)
    srv.clear_cache()
    assert rep.cache_hits > 0
    for j in jobs:
        r = rep.results[j.job_id]
        assert r.error is None, r.error
        out, nv = serve.run_plain(gpu_ctx, j)
        print(f"job {j.job_id}: F={j.scene.F} T={j.scene.T} ns={j.ns} nr={j.nr} "
              f"pairs={int(nv.sum())}", flush=True)
        np.testing.assert_array_equal(r.nvalid, nv)
        assert nv.sum() > 0
        compare_traces(r.out, out)


def test_deterministic_jobs_are_bitwise_the_same_alone_and_in_the_mix(gpu_ctx):
    """With ERP_RUN_DETERMINISTIC every bin is a sum of fixed-point values
    with integer atomics, so a job's radargram does not depend on what else
    runs beside it: served and alone must agree bit for bit."""
    jobs = small_stream(8, seed=23)
    for j in jobs:
        j.flags = ERP_RUN_DETERMINISTIC
    srv = serve.Server(gpu_ctx, budget_bytes=8e9, slots=4)
    rep = srv.run(jobs)
    srv.clear_cache()
    for j in jobs:
        r = rep.results[j.job_id]
        assert r.error is None, r.error
        out, nv = serve.run_plain(gpu_ctx, j)
        print(f"det job {j.job_id}: F={j.scene.F} ns={j.ns} nr={j.nr}", flush=True)
        np.testing.assert_array_equal(r.nvalid, nv)
        assert np.array_equal(r.out.view(np.uint32), out.view(np.uint32))


def test_submit_reports_job_errors_and_keeps_serving(gpu_ctx):
    jobs = small_stream(3, seed=5)
    j = jobs[0]
    m = gpu_ctx.upload_mesh(j.scene.soa)
    t = gpu_ctx.upload_targets(*j.scene.targets)
    try:
        # a non-specular run without convolution is a mode the reference
        # cannot run either; it fails inside the worker before any kernel is
        # queued and the error comes back through erp_wait
        e = j.par.to_erp()
        e.mode &= ~ERP_MODE_CONVOLUTION
        bad = gpu_ctx.submit(e, m, t, j.shots)
        good = gpu_ctx.submit(j.par.to_erp(), m, t, j.shots)
        with pytest.raises(Exception, match="convolution"):
            bad.wait()
        out, nv = good.wait()
        ref, rnv = gpu_ctx.run(j.par.to_erp(), m, t, j.shots)
        np.testing.assert_array_equal(nv, rnv)
        compare_traces(out, ref)
    finally:
        m.release()
        t.release()


def test_stream_at_full_size_with_poisson_arrivals(gpu_ctx):
    """The serving workload at its stated sizes: the first four jobs of the
    job stream (SeedSequence(4): F log-uniform in [1e4, 4e6], U{100..2000}
    shots at full count, 2048-16384 bins, 1-3 layers in one or two media)
    arriving as a Poisson stream at 20 jobs/s, in the deterministic mode.
    Every served radargram equals its job run alone bit for bit, and each
    latency runs from the job's arrival to its result."""
    jobs = serve.with_arrivals(serve.mixed_job_stream(4, None, seed=4), 20.0, seed=4)
    for j in jobs:
        j.flags = ERP_RUN_DETERMINISTIC
    srv = serve.Server(gpu_ctx, budget_bytes=250e9, slots=4)
    rep = srv.run(jobs)
    srv.clear_cache()
    print(f"served 4 full-size jobs: {rep.summary()}", flush=True)
    for j in jobs:
        r = rep.results[j.job_id]
        assert r.error is None, r.error
        assert r.finished >= r.admitted >= r.arrival - 1e-9
        out, nv = serve.run_plain(gpu_ctx, j)
        print(f"job {j.job_id}: F={j.scene.F} ns={j.ns} nr={j.nr} pairs={int(nv.sum())}",
              flush=True)
        np.testing.assert_array_equal(r.nvalid, nv)
        assert np.array_equal(r.out.view(np.uint32), out.view(np.uint32))
