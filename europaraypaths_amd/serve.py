"""Continuous-batching job server for the path on one GPU.

The reference runs one simulation per process (`sim <params> <facets>
<targets> <outdir>`, sim.cu:92-1008) and keeps nothing resident between
runs.  This module serves a stream of independent simulation jobs (mesh,
target layers, track, chirp) on one device:

* Jobs arrive over time.  A job is admitted when its device footprint fits
  the memory budget beside what is already reserved: the footprints of the
  jobs in flight (erp_run_footprint) plus the meshes held in the cache.
* Admitted jobs go to the library's worker slots (erp_submit), each with
  its own stream and persistent work buffers, so jobs that arrive while
  others are on the GPU overlap with them on the device.
* Meshes and their target layers stay resident in HBM, keyed by mesh id,
  and are evicted least-recently-used when a new job needs the room and no
  job in flight holds them.
* Each job's latency runs from its arrival time to the moment its radargram
  is visible in host memory.

`mixed_job_stream` draws the mixed-size job stream (facet counts
log-uniform, 1-3 target layers, 2048-16384 range bins, 9 or 60 MHz) from one
seed, with a fraction of jobs reusing an earlier job's mesh.
"""
from __future__ import annotations

import time
from collections import OrderedDict, deque
from dataclasses import dataclass, field

import numpy as np

from .params import SimulationParameters, make_shots, parse_parameters
from .surfaces import fbm_heights, target_layer, terrain_soa

MESH_BYTES_PER_FACET = 48   # 12 float32 attributes
TARGET_BYTES_PER_POINT = 25  # 6 float32 + 1 type byte


@dataclass
class Scene:
    """A mesh and its target layers (what the mesh cache keeps resident)."""
    mesh_id: int
    soa: np.ndarray                      # (12, F) float32
    targets: tuple | None                # ((6, T) float32, (T,) uint8)
    eps_2: float = 3.15
    # the medium below z_int (0: one medium for every layer)
    eps_3: float = 0.0
    z_int: float = 0.0

    @property
    def F(self) -> int:
        return int(self.soa.shape[1])

    @property
    def T(self) -> int:
        return 0 if self.targets is None else int(self.targets[0].shape[1])

    @property
    def device_bytes(self) -> int:
        fpad = (self.F + 63) // 64 * 64
        tpad = (self.T + 63) // 64 * 64
        return MESH_BYTES_PER_FACET * fpad + TARGET_BYTES_PER_POINT * tpad


@dataclass
class SimJob:
    """One simulation request."""
    job_id: int
    arrival: float                       # seconds after the stream starts
    scene: Scene
    par: SimulationParameters
    shots: np.ndarray
    flags: int = 0

    @property
    def ns(self) -> int:
        return int(len(self.shots))

    @property
    def nr(self) -> int:
        return int(self.par.nr)


@dataclass
class JobResult:
    job_id: int
    out: np.ndarray | None               # (ns, nr) complex64
    nvalid: np.ndarray | None            # (ns,) int32 illuminated facets per shot
    arrival: float
    admitted: float
    finished: float
    cache_hit: bool
    error: str | None = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def pairs(self) -> int:
        return 0 if self.nvalid is None else int(self.nvalid.sum())


@dataclass
class ServeReport:
    results: dict = field(default_factory=dict)   # job_id -> JobResult
    makespan: float = 0.0                         # first arrival -> last completion
    cache_hits: int = 0
    cache_misses: int = 0
    evictions: int = 0
    rejected: int = 0
    peak_reserved: int = 0

    @property
    def pairs(self) -> int:
        return sum(r.pairs for r in self.results.values())

    def latency_ms(self) -> dict:
        lat = [r.latency for r in self.results.values() if r.error is None]
        if not lat:
            return {"p50": None, "p99": None, "mean": None, "max": None}
        lat = 1e3 * np.asarray(lat)
        return {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                "mean": float(lat.mean()), "max": float(lat.max())}

    def summary(self) -> dict:
        ok = [r for r in self.results.values() if r.error is None]
        return {"jobs": len(self.results), "completed": len(ok), "rejected": self.rejected,
                "pairs": self.pairs, "makespan_s": self.makespan,
                "pairs_per_s": self.pairs / self.makespan if self.makespan > 0 else 0.0,
                "latency_ms": self.latency_ms(), "cache_hits": self.cache_hits,
                "cache_misses": self.cache_misses, "evictions": self.evictions,
                "peak_reserved_bytes": self.peak_reserved}


class _Resident:
    __slots__ = ("mesh", "targets", "bytes", "refs")

    def __init__(self, mesh, targets, nbytes):
        self.mesh, self.targets, self.bytes, self.refs = mesh, targets, nbytes, 0


class Server:
    """Admission by device footprint, a device-resident mesh cache and
    overlapped execution in the library's worker slots.

    ctx: a `_native.Context` (or an object with the same upload_mesh /
    upload_targets / submit / footprint / serve_slots methods).
    """

    def __init__(self, ctx, budget_bytes: float = 250e9, slots: int = 4,
                 max_inflight: int | None = None, poll_interval: float = 2e-4):
        self.ctx = ctx
        self.budget = int(budget_bytes)
        self.slots = slots
        self.max_inflight = max_inflight or 2 * slots
        self.poll_interval = poll_interval
        ctx.serve_slots(slots)
        self.cache: OrderedDict[int, _Resident] = OrderedDict()
        self.cached_bytes = 0
        self.inflight_bytes = 0

    # -- memory accounting -------------------------------------------------
    def run_footprint(self, job: SimJob) -> int:
        return int(self.ctx.footprint(job.par.to_erp(), job.scene.F, job.scene.T, job.ns))

    def reserved(self) -> int:
        return self.cached_bytes + self.inflight_bytes

    def _evict_for(self, need: int, report: ServeReport, keep: int) -> None:
        """Drop unreferenced cached meshes (oldest first) until `need` fits."""
        for mid in list(self.cache.keys()):
            if self.reserved() + need <= self.budget:
                return
            r = self.cache[mid]
            if r.refs or mid == keep:
                continue
            self._drop(mid)
            report.evictions += 1

    def _drop(self, mid: int) -> None:
        r = self.cache.pop(mid)
        r.mesh.release()
        if r.targets is not None:
            r.targets.release()
        self.cached_bytes -= r.bytes

    def clear_cache(self) -> None:
        for mid in [m for m, r in self.cache.items() if not r.refs]:
            self._drop(mid)

    def _try_admit(self, job: SimJob, report: ServeReport):
        """Reserve memory for `job` and return (resident, hit, bytes) or None."""
        sc = job.scene
        need = self.run_footprint(job)
        hit = sc.mesh_id in self.cache
        mesh_bytes = 0 if hit else sc.device_bytes
        if need + mesh_bytes > self.budget:
            raise MemoryError(f"job {job.job_id} needs {need + mesh_bytes} bytes, "
                              f"more than the {self.budget} byte budget")
        if self.reserved() + need + mesh_bytes > self.budget:
            self._evict_for(need + mesh_bytes, report, sc.mesh_id)
            if self.reserved() + need + mesh_bytes > self.budget:
                return None
        if hit:
            res = self.cache[sc.mesh_id]
            self.cache.move_to_end(sc.mesh_id)
        else:
            m = self.ctx.upload_mesh(sc.soa)
            t = self.ctx.upload_targets(*sc.targets) if sc.targets is not None else None
            res = _Resident(m, t, sc.device_bytes)
            self.cache[sc.mesh_id] = res
            self.cached_bytes += res.bytes
        res.refs += 1
        self.inflight_bytes += need
        report.peak_reserved = max(report.peak_reserved, self.reserved())
        return res, hit, need

    # -- the request loop --------------------------------------------------
    def run(self, jobs, keep_outputs: bool = True) -> ServeReport:
        """Serve `jobs` (arrival times relative to the call) to completion."""
        jobs = sorted(jobs, key=lambda j: (j.arrival, j.job_id))
        report = ServeReport()
        waiting: deque = deque()
        inflight = []   # (job, handle, resident, need, admitted, hit)
        i = 0
        t0 = time.perf_counter()
        while i < len(jobs) or waiting or inflight:
            now = time.perf_counter() - t0
            while i < len(jobs) and jobs[i].arrival <= now:
                waiting.append(jobs[i])
                i += 1
            progressed = False
            # admission, first come first served: a job that does not fit
            # holds back the ones behind it so that large jobs cannot starve
            while waiting and len(inflight) < self.max_inflight:
                job = waiting[0]
                try:
                    got = self._try_admit(job, report)
                except MemoryError as ex:
                    waiting.popleft()
                    report.rejected += 1
                    t = time.perf_counter() - t0
                    report.results[job.job_id] = JobResult(job.job_id, None, None, job.arrival,
                                                           t, t, False, error=str(ex))
                    progressed = True
                    continue
                if got is None:
                    break
                res, hit, need = got
                waiting.popleft()
                report.cache_hits += hit
                report.cache_misses += not hit
                h = self.ctx.submit(job.par.to_erp(), res.mesh, res.targets, job.shots,
                                    flags=job.flags)
                inflight.append((job, h, res, need, time.perf_counter() - t0, hit))
                progressed = True
            # completions
            still = []
            for rec in inflight:
                job, h, res, need, adm, hit = rec
                err = None
                try:
                    done = h.poll()
                except Exception as ex:  # noqa: BLE001 - a failed job is reported, not fatal
                    done, err = True, str(ex)
                if not done:
                    still.append(rec)
                    continue
                t = time.perf_counter() - t0
                res.refs -= 1
                self.inflight_bytes -= need
                out, nv = (h.out, h.nvalid) if err is None else (None, None)
                report.results[job.job_id] = JobResult(
                    job.job_id, out if keep_outputs else None, nv, job.arrival, adm, t, hit,
                    error=err)
                progressed = True
            inflight = still
            if not progressed:
                nxt = jobs[i].arrival - (time.perf_counter() - t0) if i < len(jobs) else None
                if inflight or waiting or nxt is None:
                    time.sleep(self.poll_interval)
                else:
                    time.sleep(max(0.0, min(nxt, 0.01)))
        ends = [r.finished for r in report.results.values()]
        starts = [j.arrival for j in jobs]
        report.makespan = (max(ends) - min(starts)) if ends else 0.0
        return report


def run_plain(ctx, job: SimJob):
    """The plain single-job path for one job, as `simulate` does it: upload the
    mesh and targets, run every shot synchronously, copy the radargram back,
    release the device copies.  Returns (out (ns, nr), nvalid (ns,))."""
    m = ctx.upload_mesh(job.scene.soa)
    t = ctx.upload_targets(*job.scene.targets) if job.scene.targets is not None else None
    try:
        return ctx.run(job.par.to_erp(), m, t, job.shots, flags=job.flags)
    finally:
        m.release()
        if t is not None:
            t.release()


# ---------------------------------------------------------------------------
# the mixed-size job stream
# ---------------------------------------------------------------------------
NR_CHOICES = (2048, 4096, 8192, 16384)
FACET_SPACING = 50.0
LAYER_SPACING = 500.0


def make_scene(mesh_id: int, rng: np.random.Generator, f_range=(1e4, 4e6)) -> Scene:
    """fBm mesh (log-uniform facet count, Hurst U(0.6, 0.9), 10 m RMS, 50 m
    facets) with 1-3 flat target layers at depths U(0.5, 6) km on a 500 m
    grid over the same extent, each layer's medium eps ~ U(3.0, 3.4).  The
    reference models one subsurface medium (file_io.cu:81); this build adds
    a second one (erp_params eps_3 below z_int), so the first layer lies in
    eps_2 and, with two or three layers, the deeper ones in eps_3 below an
    interface halfway between the first and the second layer."""
    F = float(np.exp(rng.uniform(np.log(f_range[0]), np.log(f_range[1]))))
    n = max(8, int(round(np.sqrt(F))))
    hurst = float(rng.uniform(0.6, 0.9))
    fs = FACET_SPACING
    ox = oy = -fs * n / 2
    z = fbm_heights(n, n, fs, 10.0, hurst, seed=int(rng.integers(2 ** 31)))
    soa = terrain_soa(n, n, fs, ox, oy, z)
    nl = int(rng.integers(1, 4))
    depths = np.sort(rng.uniform(500.0, 6000.0, nl))
    m = max(2, int(n * fs / LAYER_SPACING))
    lo = -LAYER_SPACING * (m - 1) / 2
    xs, ts = [], []
    for d in depths:
        x, t = target_layer(m, m, LAYER_SPACING, lo, lo, -float(d))
        xs.append(x)
        ts.append(t)
    tg = (np.concatenate(xs, axis=1), np.concatenate(ts))
    eps = rng.uniform(3.0, 3.4, 2)
    if nl == 1:
        return Scene(mesh_id, soa, tg, eps_2=float(eps[0]))
    return Scene(mesh_id, soa, tg, eps_2=float(eps[0]), eps_3=float(eps[1]),
                 z_int=-0.5 * float(depths[0] + depths[1]))


def make_job(job_id: int, arrival: float, scene: Scene, rng: np.random.Generator,
             ns_range=(100, 2000), ns_scale: float = 1.0) -> SimJob:
    """Track, chirp and receiver of one job over `scene`: ns ~ U{ns_range}
    (times ns_scale), nr from {2048, ..., 16384}, 9 or 60 MHz, altitude
    U(25, 50) km, 10 deg aperture, a straight track across the mesh."""
    from .workloads import make_params
    ns = max(1, int(round(rng.integers(ns_range[0], ns_range[1] + 1) * ns_scale)))
    nr = int(rng.choice(NR_CHOICES))
    platform = "REASON_HF" if rng.integers(2) == 0 else "REASON_VHF"
    h = float(rng.uniform(25e3, 50e3))
    n = int(round(np.sqrt(scene.F)))
    extent = n * FACET_SPACING
    sdx = min(FACET_SPACING, 0.8 * extent / max(ns - 1, 1))
    sx0 = -0.5 * sdx * (ns - 1)
    media = {"eps_3": scene.eps_3, "z_interface": scene.z_int} if scene.eps_3 > 0 else {}
    p = make_params(platform, "planetary_ice", nr=nr, smpl=48e6, rst=h - 1000.0,
                    aperture=10.0, fs=FACET_SPACING, ns=ns, sx0=sx0, sdx=sdx, sy=0.0,
                    sz=h, eps_2=scene.eps_2, **media)
    par = parse_parameters(p)
    return SimJob(job_id, arrival, scene, par, make_shots(par))


def poisson_arrivals(n_jobs: int, rate: float | None, seed: int = 4) -> np.ndarray:
    """Arrival times (s) of `n_jobs` jobs at `rate` jobs/s, first at 0
    (None: every job at 0).  Drawn from the last child of SeedSequence(seed),
    the same draw mixed_job_stream uses."""
    if rate is None:
        return np.zeros(n_jobs)
    rng = np.random.default_rng(np.random.SeedSequence(seed).spawn(n_jobs + 1)[-1])
    t = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    return t - t[0]


def with_arrivals(jobs, rate: float | None, seed: int = 4):
    """The same jobs with the arrival times of another rate."""
    import dataclasses
    t = poisson_arrivals(len(jobs), rate, seed)
    return [dataclasses.replace(j, arrival=float(t[k])) for k, j in enumerate(jobs)]


def mixed_job_stream(n_jobs: int, rate: float | None, seed: int = 4, reuse: float = 0.3,
                     ns_scale: float = 1.0, ns_range=(100, 2000), f_range=(1e4, 4e6)):
    """`n_jobs` jobs with Poisson arrivals at `rate` jobs/s (None: all at t=0).

    Every draw comes from numpy SeedSequence(seed): one child per job for its
    scene and track, one for the arrival times.  With probability `reuse` a
    job (after the first) takes the scene of a uniformly chosen earlier job,
    which is what exercises the resident-mesh cache.
    """
    kids = np.random.SeedSequence(seed).spawn(n_jobs + 1)
    arrivals = poisson_arrivals(n_jobs, rate, seed)
    scenes: list[Scene] = []
    jobs = []
    for k in range(n_jobs):
        rng = np.random.default_rng(kids[k])
        if scenes and rng.uniform() < reuse:
            sc = scenes[int(rng.integers(len(scenes)))]
        else:
            sc = make_scene(len(scenes), rng, f_range)
            scenes.append(sc)
        jobs.append(make_job(k, float(arrivals[k]), sc, rng, ns_range, ns_scale))
    return jobs
