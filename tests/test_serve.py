"""Host logic of the job server: admission by footprint, the resident-mesh
cache, rejection of oversized jobs and the job-stream generator.  A fake
device context stands in for the library so these run without a GPU."""
import time

import numpy as np
import pytest

from europaraypaths_amd import serve, workloads
from europaraypaths_amd.params import make_shots, parse_parameters


class FakeBuf:
    def __init__(self, ctx, kind, key):
        self.ctx, self.kind, self.key, self.live = ctx, kind, key, True
        ctx.uploads.append((kind, key))

    def release(self):
        # a mesh must never be released while a job that reads it is running
        assert not any(j.mesh is self and not j.finished for j in self.ctx.jobs), \
            "mesh released while in use"
        self.live = False


class FakeJob:
    def __init__(self, ctx, mesh, ns, nr, duration, fail=False):
        self.ctx, self.mesh = ctx, mesh
        self.t_end = time.perf_counter() + duration
        self.out = np.zeros((ns, nr), np.complex64)
        self.nvalid = np.full(ns, 7, np.int32)
        self.fail, self.finished = fail, False

    def poll(self):
        if time.perf_counter() < self.t_end:
            return False
        self.finished = True
        if self.fail:
            raise RuntimeError("injected job failure")
        return True


class FakeCtx:
    """upload/submit/footprint with the library's call shapes; jobs take
    `duration` seconds each and at most `slots` run at a time."""

    def __init__(self, duration=0.01, per_shot_bytes=1000):
        self.duration, self.per_shot = duration, per_shot_bytes
        self.uploads, self.jobs, self.slots = [], [], None
        self.max_running = 0

    def serve_slots(self, n):
        self.slots = n

    def footprint(self, e, F, T, ns):
        return self.per_shot * ns

    def upload_mesh(self, soa):
        return FakeBuf(self, "mesh", soa.shape[1])

    def upload_targets(self, xyzn, ttype):
        return FakeBuf(self, "targets", xyzn.shape[1])

    def submit(self, e, mesh, targets, shots, flags=0):
        assert mesh.live and (targets is None or targets.live)
        running = sum(not j.finished and time.perf_counter() < j.t_end for j in self.jobs)
        self.max_running = max(self.max_running, running + 1)
        j = FakeJob(self, mesh, len(shots), e.nr, self.duration, fail=bool(flags & 1))
        self.jobs.append(j)
        return j


def tiny_par(ns):
    p = workloads.make_params("REASON_VHF", "planetary_ice", nr=256, smpl=48e6, rst=4600.0,
                              fs=20.0, ns=ns, sz=5000.0)
    return parse_parameters(p)


def tiny_scene(mid, F=100, T=10):
    soa = np.zeros((12, F), np.float32)
    tg = (np.zeros((6, T), np.float32), np.ones(T, np.uint8))
    return serve.Scene(mid, soa, tg)


def tiny_job(k, arrival, scene, ns=3, flags=0):
    par = tiny_par(ns)
    return serve.SimJob(k, arrival, scene, par, make_shots(par), flags=flags)


def test_every_job_completes_with_latency_from_arrival():
    ctx = FakeCtx(duration=0.005)
    srv = serve.Server(ctx, budget_bytes=1e9, slots=2)
    scenes = [tiny_scene(i) for i in range(3)]
    jobs = [tiny_job(k, 0.002 * k, scenes[k % 3]) for k in range(9)]
    rep = srv.run(jobs)
    assert ctx.slots == 2
    assert sorted(rep.results) == list(range(9))
    for r in rep.results.values():
        assert r.error is None
        assert r.finished >= r.admitted >= r.arrival - 1e-9
        assert r.latency >= 0.005 - 1e-4
        assert r.out.shape == (3, 256)
    assert rep.pairs == 9 * 3 * 7
    s = rep.summary()
    assert s["completed"] == 9 and s["latency_ms"]["p99"] >= s["latency_ms"]["p50"]
    assert ctx.max_running <= srv.max_inflight


def test_mesh_cache_uploads_each_scene_once_under_a_large_budget():
    ctx = FakeCtx(duration=0.001)
    srv = serve.Server(ctx, budget_bytes=1e9, slots=4)
    scenes = [tiny_scene(i) for i in range(4)]
    order = [0, 1, 0, 2, 1, 3, 0, 3]
    jobs = [tiny_job(k, 0.0, scenes[s]) for k, s in enumerate(order)]
    rep = srv.run(jobs)
    meshes = [k for kind, k in ctx.uploads if kind == "mesh"]
    assert len(meshes) == 4
    assert rep.cache_misses == 4 and rep.cache_hits == 4
    assert rep.evictions == 0


def test_admission_respects_the_budget_and_evicts_unused_meshes():
    ctx = FakeCtx(duration=0.003, per_shot_bytes=1000)
    sc = [tiny_scene(i, F=1000, T=0) for i in range(3)]
    for s in sc:
        s.targets = None
    mesh_b = sc[0].device_bytes
    # room for one job (3 shots = 3000 bytes) and one or two meshes at a time
    budget = 3000 + 2 * mesh_b + 10
    srv = serve.Server(ctx, budget_bytes=budget, slots=4)
    jobs = [tiny_job(k, 0.0, sc[k % 3]) for k in range(9)]
    rep = srv.run(jobs)
    assert all(r.error is None for r in rep.results.values())
    assert rep.peak_reserved <= budget
    assert rep.evictions > 0
    # one job at a time fits: nothing overlapped
    assert ctx.max_running == 1


def test_oversized_job_is_rejected_and_the_rest_complete():
    ctx = FakeCtx(duration=0.001, per_shot_bytes=10 ** 6)
    srv = serve.Server(ctx, budget_bytes=5.5e6, slots=2)
    s = tiny_scene(0)
    jobs = [tiny_job(0, 0.0, s, ns=2), tiny_job(1, 0.0, s, ns=9), tiny_job(2, 0.0, s, ns=3)]
    rep = srv.run(jobs)
    assert rep.rejected == 1
    assert rep.results[1].error and "budget" in rep.results[1].error
    assert rep.results[0].error is None and rep.results[2].error is None


def test_failed_job_is_reported_without_stopping_the_server():
    ctx = FakeCtx(duration=0.001)
    srv = serve.Server(ctx, budget_bytes=1e9, slots=2)
    s = tiny_scene(0)
    jobs = [tiny_job(0, 0.0, s), tiny_job(1, 0.0, s, flags=1), tiny_job(2, 0.001, s)]
    rep = srv.run(jobs)
    assert "injected" in rep.results[1].error
    assert rep.results[0].error is None and rep.results[2].error is None
    # the failed job's mesh reference was returned: the cache can be cleared
    srv.clear_cache()
    assert not srv.cache and srv.cached_bytes == 0


def test_poisson_arrivals_are_honoured():
    ctx = FakeCtx(duration=0.0005)
    srv = serve.Server(ctx, budget_bytes=1e9, slots=2)
    s = tiny_scene(0)
    jobs = [tiny_job(k, 0.01 * k, s) for k in range(5)]
    t0 = time.perf_counter()
    rep = srv.run(jobs)
    assert time.perf_counter() - t0 >= 0.04
    for r in rep.results.values():
        assert r.admitted >= r.arrival - 1e-9


def test_job_stream_is_seeded_and_reuses_meshes():
    kw = dict(ns_scale=0.01, f_range=(1e3, 4e3))
    a = serve.mixed_job_stream(40, 50.0, seed=11, **kw)
    b = serve.mixed_job_stream(40, 50.0, seed=11, **kw)
    c = serve.mixed_job_stream(40, 50.0, seed=12, **kw)
    assert [j.arrival for j in a] == [j.arrival for j in b]
    assert [j.scene.mesh_id for j in a] == [j.scene.mesh_id for j in b]
    assert all(np.array_equal(x.scene.soa, y.scene.soa) for x, y in zip(a, b))
    assert [j.arrival for j in a] != [j.arrival for j in c]
    nscenes = len({j.scene.mesh_id for j in a})
    reused = 40 - nscenes
    assert 4 <= reused <= 22            # 30 % of 39 draws, loosely
    # exponential inter-arrivals at 50 jobs/s
    gaps = np.diff([j.arrival for j in a])
    assert (gaps >= 0).all() and 0.005 < gaps.mean() < 0.05
    for j in a:
        assert j.nr in serve.NR_CHOICES
        assert 1 <= j.ns <= 20
        assert 1e3 * 0.8 <= j.scene.F <= 4e3 * 1.2
        assert 25e3 <= float(j.shots["sz"][0]) <= 50e3
        depths = np.unique(-j.scene.targets[0][2])
        nl = len(depths)
        assert 1 <= nl <= 3
        assert 3.0 <= j.par.eps_2 <= 3.4
        if nl == 1:
            assert j.scene.eps_3 == 0.0 and float(j.par.eps_3) == 0.0
        else:
            # the layers below the first lie in their own medium, below an
            # interface halfway between the first two layers
            assert 3.0 <= j.scene.eps_3 <= 3.4
            assert float(j.par.eps_3) == pytest.approx(j.scene.eps_3, rel=1e-6)
            assert float(j.par.z_int) == pytest.approx(-0.5 * (depths[0] + depths[1]), rel=1e-5)


def test_all_at_once_stream():
    jobs = serve.mixed_job_stream(3, None, seed=1, ns_scale=0.01, f_range=(1e3, 2e3))
    assert [j.arrival for j in jobs] == [0.0, 0.0, 0.0]


def test_job_rx_window_holds_surface_and_deepest_layer():
    jobs = serve.mixed_job_stream(12, None, seed=3, ns_scale=0.01, f_range=(1e3, 2e3))
    assert {j.nr for j in jobs} >= {2048}
    for j in jobs:
        h = float(j.shots["sz"][0])
        rst = float(j.shots["rst"][0])
        depth = -float(j.scene.targets[0][2].min())
        n = np.sqrt(max(j.par.eps_2, j.par.eps_3))
        far = h + n * depth
        assert rst < h and far < rst + j.par.nr * float(j.par.dr)
