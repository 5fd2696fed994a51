// erp_api.hip — the extern "C" boundary of include/erp.h.
//
// Replaces the process boundary of the reference (sim.cu:92-613: parameter
// parse, file loads, device allocation and upload) and its exit()-on-error
// policy (sim.cu:60-70) with status codes and a per-context error string.
#include "erp_internal.h"

#include <algorithm>
#include <cmath>

#include <climits>
#include <cstring>
#include <exception>
#include <new>

// The scratch of erp_run_dev_noalloc (see scratch_init below).
static __device__ __align__(256) unsigned char scratch[ERP_SCRATCH_BYTES];

namespace {

thread_local std::string g_null_ctx_error;

template <class Fn> int guarded(erp_ctx* ctx, Fn&& fn) {
    try {
        fn();
        return ERP_OK;
    } catch (const ErpError& e) {
        if (ctx) ctx->last_error = e.msg; else g_null_ctx_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->last_error = "host allocation failed"; else g_null_ctx_error = "host allocation failed";
        return ERP_E_NOMEM;
    } catch (const std::exception& e) {
        if (ctx) ctx->last_error = e.what(); else g_null_ctx_error = e.what();
        return ERP_E_INVALID;
    }
}

void set_device(erp_ctx* ctx) { ERP_HIP_TRY(hipSetDevice(ctx->device)); }

void validate_params(const erp_params* p) {
    if (!p) throw ErpError{ERP_E_INVALID, "params is NULL"};
    if (p->nr <= 0 || p->nr > 32767) throw ErpError{ERP_E_INVALID, "nr must be in [1, 32767]"};
    if (!(p->dr > 0.f)) throw ErpError{ERP_E_INVALID, "dr must be > 0"};
    if (!(p->lam > 0.f)) throw ErpError{ERP_E_INVALID, "lam must be > 0"};
    if (2 * (int64_t)p->nr - 1 > (1 << 16))
        throw ErpError{ERP_E_INVALID, "nr too large for the FFT tables"};
}

}  // namespace

hipEvent_t StageTimer::get() {
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e;
    ERP_HIP_TRY(hipEventCreate(&e));
    return e;
}

void StageTimer::begin(int stage, hipStream_t s, double u) {
    if (!on) return;
    Pending p{stage, get(), get(), u};
    ERP_HIP_TRY(hipEventRecord(p.a, s));
    pending.push_back(p);
    cur = (int)pending.size() - 1;
}

void StageTimer::end(hipStream_t s) {
    if (!on || cur < 0) return;
    ERP_HIP_TRY(hipEventRecord(pending[cur].b, s));
    cur = -1;
}

void StageTimer::resolve() {
    for (Pending& p : pending) {
        float ms_ = 0.f;
        if (hipEventElapsedTime(&ms_, p.a, p.b) == hipSuccess) {
            ms[p.stage] += ms_;
            launches[p.stage] += 1;
            units[p.stage] += p.units;
        }
        pool.push_back(p.a);
        pool.push_back(p.b);
    }
    pending.clear();
    cur = -1;
}

void StageTimer::reset() {
    for (int k = 0; k < ERP_NSTAGES; ++k) {
        ms[k] = 0;
        launches[k] = 0;
        units[k] = 0;
    }
}

// One asynchronous job: a copy of the call's host arguments and its status.
struct erp_job {
    erp_ctx* ctx = nullptr;
    erp_params par{};
    const erp_mesh* mesh = nullptr;
    const erp_targets* targets = nullptr;
    std::vector<erp_shot> shots;
    std::vector<erp_prism> prisms;
    uint32_t flags = 0;
    float* host_out = nullptr;
    int32_t* host_nvalid = nullptr;
    std::mutex mu;
    std::condition_variable cv;
    bool done = false;
    int status = ERP_OK;
    std::string err;
};

// A worker slot: one host thread, one stream and work buffers that persist
// from job to job, so that a warm slot allocates nothing.
struct ErpSlot {
    std::thread th;
    hipStream_t stream = nullptr;
    WorkBufs work;
    DevBuf out, nvalid;
};

namespace {

void release_work(WorkBufs& w) {
    DevBuf* bufs[] = {&w.cnt, &w.off, &w.V, &w.Vd, &w.tcnt, &w.tlist, &w.tcount, &w.rec,
                      &w.bounds, &w.psurf, &w.srow, &w.refl, &w.shotS, &w.blu, &w.tA, &w.tB, &w.tC,
                      &w.pairs, &w.t0trace, &w.shots, &w.prisms, &w.out, &w.nvalid, &w.oow,
                      &w.umask, &w.ucnt, &w.uoff, &w.U, &w.ubase, &w.brec, &w.items, &w.bbox};
    for (DevBuf* b : bufs) b->release();
}

// Runs one job on a slot: the pipeline on the slot's stream, the copies back
// to the caller's host buffers, then a wait on the stream.
void run_job(erp_ctx* ctx, ErpSlot* s, erp_job* j) {
    const size_t nout = (size_t)j->shots.size() * j->par.nr;
    const int64_t ns = (int64_t)j->shots.size();
    s->out.ensure(sizeof(float2) * std::max<size_t>(nout, 1));
    s->nvalid.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(ns, 1));
    erp_run_impl(ctx, &j->par, j->mesh, j->targets, j->shots.data(), ns,
                 j->prisms.empty() ? nullptr : j->prisms.data(), (int64_t)j->prisms.size(),
                 s->out.as<float2>(), s->nvalid.as<int32_t>(), j->flags, s->stream, s->work,
                 nullptr);
    if (nout)
        ERP_HIP_TRY(hipMemcpyAsync(j->host_out, s->out.p, sizeof(float2) * nout,
                                   hipMemcpyDeviceToHost, s->stream));
    if (j->host_nvalid && ns)
        ERP_HIP_TRY(hipMemcpyAsync(j->host_nvalid, s->nvalid.p, sizeof(int32_t) * ns,
                                   hipMemcpyDeviceToHost, s->stream));
    ERP_HIP_TRY(hipStreamSynchronize(s->stream));
}

void slot_main(erp_ctx* ctx, ErpSlot* s) {
    hipSetDevice(ctx->device);
    for (;;) {
        erp_job* j = nullptr;
        {
            std::unique_lock<std::mutex> lk(ctx->qmu);
            ctx->qcv.wait(lk, [&] { return ctx->stopping || !ctx->queue.empty(); });
            if (ctx->queue.empty()) return;  // stopping with nothing queued
            j = ctx->queue.front();
            ctx->queue.pop_front();
            // counted before the job can take pointers into the table caches
            ctx->active.fetch_add(1);
        }
        int rc = ERP_OK;
        std::string err;
        try {
            run_job(ctx, s, j);
        } catch (const ErpError& e) {
            rc = e.code;
            err = e.msg;
        } catch (const std::bad_alloc&) {
            rc = ERP_E_NOMEM;
            err = "host allocation failed";
        } catch (const std::exception& e) {
            rc = ERP_E_INVALID;
            err = e.what();
        }
        if (rc != ERP_OK) hipStreamSynchronize(s->stream);  // drain what was queued
        ctx->active.fetch_sub(1);
        {
            // notify under the lock: once it is released the caller may
            // free the job (erp_wait / erp_poll)
            std::lock_guard<std::mutex> lk(j->mu);
            j->status = rc;
            j->err = err;
            j->done = true;
            j->cv.notify_all();
        }
    }
}

// Starts worker slots up to n (never fewer than exist).
void start_slots(erp_ctx* ctx, int n) {
    std::lock_guard<std::mutex> lk(ctx->qmu);
    while ((int)ctx->slots.size() < n) {
        ErpSlot* s = new ErpSlot();
        hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete s;
            throw ErpError{ERP_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)};
        }
        s->th = std::thread(slot_main, ctx, s);
        ctx->slots.push_back(s);
    }
}

void stop_slots(erp_ctx* ctx) {
    {
        std::lock_guard<std::mutex> lk(ctx->qmu);
        ctx->stopping = true;
        // jobs still queued at destruction never run
        for (erp_job* j : ctx->queue) {
            std::lock_guard<std::mutex> jl(j->mu);
            j->status = ERP_E_INVALID;
            j->err = "context destroyed before the job ran";
            j->ctx = nullptr;
            j->done = true;
            j->cv.notify_all();
        }
        ctx->queue.clear();
    }
    ctx->qcv.notify_all();
    for (ErpSlot* s : ctx->slots) {
        if (s->th.joinable()) s->th.join();
        release_work(s->work);
        s->out.release();
        s->nvalid.release();
        if (s->stream) hipStreamDestroy(s->stream);
        delete s;
    }
    ctx->slots.clear();
}

constexpr int kDefaultSlots = 4;
constexpr int kMaxSlots = 16;

// ---------------------------------------------------------------------------
// scratch of erp_run_dev_noalloc
// ---------------------------------------------------------------------------
// Each device holds its own copy of `scratch` from the moment the library's
// code object is loaded on it until the process ends.  In host code its name
// is not a device address: the address of the current device's copy is
// looked up once per device with hipGetSymbolAddress (at the first context
// creation on the device, that device current), and only that pointer is
// handed to kernels and copies.
constexpr int kScratchDevices = 64;
struct ScratchDev {
    void* ptr = nullptr;  // this device's copy of `scratch`
    std::mutex run_mu;    // one allocation-free run at a time per device
};
ScratchDev g_scratch[kScratchDevices];
std::mutex g_scratch_init_mu;

// With `device` current: look up its copy of the array (once).
void scratch_init(int device) {
    if (device < 0 || device >= kScratchDevices) return;  // the entry point reports it
    std::lock_guard<std::mutex> lk(g_scratch_init_mu);
    if (g_scratch[device].ptr) return;
    void* p = nullptr;
    ERP_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw ErpError{ERP_E_HIP, "hipGetSymbolAddress(scratch) returned NULL"};
    g_scratch[device].ptr = p;
}

}  // namespace


extern "C" {

int erp_ctx_create(int device, erp_ctx** out) {
    if (!out) return ERP_E_INVALID;
    *out = nullptr;
    erp_ctx* ctx = new (std::nothrow) erp_ctx();
    if (!ctx) return ERP_E_NOMEM;
    ctx->device = device;
    int rc = guarded(nullptr, [&] {
        int n = 0;
        ERP_HIP_TRY(hipGetDeviceCount(&n));
        if (device < 0 || device >= n)
            throw ErpError{ERP_E_INVALID, "no HIP device " + std::to_string(device)};
        ERP_HIP_TRY(hipSetDevice(device));
        ERP_HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        ERP_HIP_TRY(hipGetDeviceProperties(&prop, device));
        ctx->num_cu = prop.multiProcessorCount;
        fft_init_twiddles(ctx);
        scratch_init(device);
        ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    });
    if (rc != ERP_OK) {
        if (ctx->stream) hipStreamDestroy(ctx->stream);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return ERP_OK;
}

void erp_ctx_destroy(erp_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    stop_slots(ctx);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (auto* c : ctx->chirps) delete c;
    for (auto* b : ctx->blues) delete b;
    ctx->chirps.clear();
    ctx->blues.clear();
    release_work(ctx->work);
    mt_state_free(ctx->mt);
    ctx->mt = nullptr;
    ctx->twiddle.release();
    for (auto& t : ctx->inter) {
        t.second->release();
        delete t.second;
    }
    ctx->inter.clear();
    ctx->prof.resolve();
    for (hipEvent_t e : ctx->prof.pool) hipEventDestroy(e);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* erp_last_error(const erp_ctx* ctx) {
    if (!ctx) return g_null_ctx_error.c_str();
    return ctx->last_error.c_str();
}

// Row length of a mesh whose facets come in rows of equal y (the first
// facet whose y differs from the first one's starts the second row), when
// it suits the surface kernels' 64 x 32 tiles; else 0.  Only the order in
// which the kernels visit facets depends on it, never a result.
static int tile_width(const float* y, int64_t F) {
    int64_t k = 1;
    while (k < F && y[k] == y[0]) ++k;
    if (k >= F || k % 64 != 0 || k > (int64_t)1 << 20 || F % (k * 32) != 0) return 0;
    return (int)k;
}

int erp_mesh_upload(erp_ctx* ctx, const float* soa, int64_t F, erp_mesh** out) {
    if (!ctx || !out || (F > 0 && !soa) || F < 0) return ERP_E_INVALID;
    *out = nullptr;
    erp_mesh* m = new (std::nothrow) erp_mesh();
    if (!m) return ERP_E_NOMEM;
    int rc = guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        set_device(ctx);
        m->ctx = ctx;
        m->device = ctx->device;
        m->F = F;
        m->Fpad = ((F + 63) / 64) * 64;
        m->soa.ensure(sizeof(float) * 12 * (size_t)std::max<int64_t>(m->Fpad, 64));
        for (int a = 0; a < 12; ++a) {
            if (F > 0)
                ERP_HIP_TRY(hipMemcpyAsync(m->soa.as<float>() + (size_t)a * m->Fpad,
                                           soa + (size_t)a * F, sizeof(float) * F,
                                           hipMemcpyHostToDevice, ctx->stream));
        }
        ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
        bool ok = true;
        for (int64_t i = 0; i < F && ok; ++i) {
            float g[12];
            for (int a = 0; a < 12; ++a) g[a] = soa[(size_t)a * F + i];
            const double wx = (double)g[7] * g[11] - (double)g[8] * g[10];
            const double wy = (double)g[8] * g[9] - (double)g[6] * g[11];
            const double wz = (double)g[6] * g[10] - (double)g[7] * g[9];
            const double ep = std::max(std::fabs(g[3] - wx), std::max(std::fabs(g[4] - wy),
                                                                      std::fabs(g[5] - wz)));
            const double em = std::max(std::fabs(g[3] + wx), std::max(std::fabs(g[4] + wy),
                                                                      std::fabs(g[5] + wz)));
            ok = std::min(ep, em) <= 1e-5;
        }
        m->frame_ok = ok;
        m->tile_nx = tile_width(soa + F, F);
    });
    if (rc != ERP_OK) {
        m->soa.release();
        delete m;
        return rc;
    }
    *out = m;
    return ERP_OK;
}

// A mesh built on the device from an ny x nx height grid that fill() writes
// into device memory (FP64, row-major, x fastest).
extern "C++" {
template <class Fill>
static int mesh_from_device_heights(erp_ctx* ctx, int64_t nx, int64_t ny, double fs, double ox,
                                    double oy, erp_mesh** out, Fill&& fill) {
    *out = nullptr;
    erp_mesh* m = new (std::nothrow) erp_mesh();
    if (!m) return ERP_E_NOMEM;
    DevBuf z;
    int rc = guarded(ctx, [&] {
        if (!(fs > 0.0)) throw ErpError{ERP_E_INVALID, "fs must be > 0"};
        if (nx > (int64_t)1 << 20 || ny > (int64_t)1 << 20 || nx * ny > (int64_t)1 << 31)
            throw ErpError{ERP_E_INVALID, "height grid too large"};
        std::lock_guard<std::mutex> lk(ctx->mu);
        set_device(ctx);
        const int64_t F = nx * ny;
        m->ctx = ctx;
        m->device = ctx->device;
        m->F = F;
        m->Fpad = ((F + 63) / 64) * 64;
        m->soa.ensure(sizeof(float) * 12 * (size_t)std::max<int64_t>(m->Fpad, 64));
        z.ensure(sizeof(double) * (size_t)std::max<int64_t>(F, 1));
        if (F > 0) {
            fill(z.as<double>());
            mesh_from_heights(z.as<double>(), nx, ny, fs, ox, oy, m->soa.as<float>(), m->Fpad,
                              ctx->stream);
        }
        ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
        // tangents and normal form an orthonormal frame by construction
        m->frame_ok = true;
        m->tile_nx = (nx % 64 == 0 && ny % 32 == 0 && nx <= ((int64_t)1 << 20)) ? (int)nx : 0;
    });
    z.release();
    if (rc != ERP_OK) {
        m->soa.release();
        delete m;
        return rc;
    }
    *out = m;
    return ERP_OK;
}
}  // extern "C++"

int erp_mesh_from_heights(erp_ctx* ctx, const double* heights, int64_t nx, int64_t ny,
                          double fs, double ox, double oy, erp_mesh** out) {
    if (!ctx || !out || nx < 0 || ny < 0 || (nx * ny > 0 && !heights)) return ERP_E_INVALID;
    return mesh_from_device_heights(ctx, nx, ny, fs, ox, oy, out, [&](double* z) {
        ERP_HIP_TRY(hipMemcpyAsync(z, heights, sizeof(double) * nx * ny, hipMemcpyHostToDevice,
                                   ctx->stream));
    });
}

int erp_mesh_from_terrain(erp_ctx* ctx, const erp_terrain* t, erp_mesh** out) {
    if (!ctx || !out || !t) return ERP_E_INVALID;
    if (t->kind != ERP_TERRAIN_GAUSSIAN && t->kind != ERP_TERRAIN_FBM) {
        ctx->last_error = "unknown terrain kind";
        return ERP_E_INVALID;
    }
    if (t->nx < 1 || t->ny < 1 || t->nx > (1 << 16) || t->ny > (1 << 16) ||
        !(t->rms >= 0.0) || (t->kind == ERP_TERRAIN_GAUSSIAN && !(t->corr_len >= 0.0)) ||
        (t->kind == ERP_TERRAIN_FBM && !std::isfinite(t->hurst))) {
        ctx->last_error = "terrain: 1 <= nx, ny <= 65536, rms >= 0, corr_len >= 0, finite hurst";
        return ERP_E_INVALID;
    }
    return mesh_from_device_heights(ctx, t->nx, t->ny, t->fs, t->ox, t->oy, out,
                                    [&](double* z) { terrain_heights(ctx, *t, z, ctx->stream); });
}

int erp_mesh_download(const erp_mesh* m, float* soa) {
    if (!m || (m->F > 0 && !soa)) return ERP_E_INVALID;
    erp_ctx* ctx = m->ctx;
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        set_device(ctx);
        for (int a = 0; a < 12; ++a)
            if (m->F > 0)
                ERP_HIP_TRY(hipMemcpyAsync(soa + (size_t)a * m->F,
                                           m->soa.as<float>() + (size_t)a * m->Fpad,
                                           sizeof(float) * m->F, hipMemcpyDeviceToHost,
                                           ctx->stream));
        ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    });
}

void erp_mesh_release(erp_mesh* m) {
    if (!m) return;
    hipSetDevice(m->device);
    m->soa.release();
    delete m;
}

int erp_targets_upload(erp_ctx* ctx, const float* xyzn, const uint8_t* type, int64_t T,
                       erp_targets** out) {
    if (!ctx || !out || T < 0 || (T > 0 && (!xyzn || !type))) return ERP_E_INVALID;
    *out = nullptr;
    erp_targets* t = new (std::nothrow) erp_targets();
    if (!t) return ERP_E_NOMEM;
    int rc = guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        set_device(ctx);
        if (T > (int64_t)1 << 28) throw ErpError{ERP_E_INVALID, "too many targets"};
        t->ctx = ctx;
        t->device = ctx->device;
        t->T = T;
        t->Tpad = ((T + 63) / 64) * 64;
        t->xyzn.ensure(sizeof(float) * 6 * (size_t)std::max<int64_t>(t->Tpad, 64));
        t->type.ensure((size_t)std::max<int64_t>(T, 64));
        for (int a = 0; a < 6; ++a)
            if (T > 0)
                ERP_HIP_TRY(hipMemcpyAsync(t->xyzn.as<float>() + (size_t)a * t->Tpad,
                                           xyzn + (size_t)a * T, sizeof(float) * T,
                                           hipMemcpyHostToDevice, ctx->stream));
        if (T > 0)
            ERP_HIP_TRY(hipMemcpyAsync(t->type.p, type, (size_t)T, hipMemcpyHostToDevice,
                                       ctx->stream));
        t->n0 = std::count(type, type + T, (uint8_t)0);
        ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    });
    if (rc != ERP_OK) {
        t->xyzn.release();
        t->type.release();
        delete t;
        return rc;
    }
    *out = t;
    return ERP_OK;
}

void erp_targets_release(erp_targets* t) {
    if (!t) return;
    hipSetDevice(t->device);
    t->xyzn.release();
    t->type.release();
    delete t;
}

int erp_run_dev(erp_ctx* ctx, const erp_params* par, const erp_mesh* mesh,
                const erp_targets* targets, const erp_shot* shots, int64_t ns,
                const erp_prism* prisms, int64_t nprisms, void* out_dev, void* nvalid_dev,
                uint32_t flags) {
    if (!ctx) return ERP_E_INVALID;
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        validate_params(par);
        if (!mesh || !shots || ns < 0 || !out_dev || nprisms < 0 || (nprisms > 0 && !prisms))
            throw ErpError{ERP_E_INVALID, "invalid argument"};
        if (mesh->ctx != ctx || (targets && targets->ctx != ctx))
            throw ErpError{ERP_E_INVALID, "mesh/targets belong to another context"};
        set_device(ctx);
        erp_run_impl(ctx, par, mesh, targets, shots, ns, prisms, nprisms, (float2*)out_dev,
                     (int32_t*)nvalid_dev, flags, ctx->stream, ctx->work, &ctx->prof);
        ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
        ctx->prof.resolve();
    });
}

int erp_run_tracks(erp_ctx* ctx, const erp_track* tracks, int64_t ntracks, uint32_t flags) {
    if (!ctx) return ERP_E_INVALID;
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (ntracks < 0 || (ntracks > 0 && !tracks))
            throw ErpError{ERP_E_INVALID, "invalid track list"};
        for (int64_t i = 0; i < ntracks; ++i) {
            const erp_track& t = tracks[i];
            validate_params(t.par);
            if (!t.mesh || t.ns < 0 || (t.ns > 0 && (!t.shots || !t.out_dev)))
                throw ErpError{ERP_E_INVALID, "invalid argument in track " + std::to_string(i)};
            if (t.mesh->ctx != ctx || (t.targets && t.targets->ctx != ctx))
                throw ErpError{ERP_E_INVALID, "mesh/targets of track " + std::to_string(i) +
                                                  " belong to another context"};
        }
        set_device(ctx);
        // the cached range kernel spectra and Bluestein tables the batches
        // point to stay in the cache while the run holds them
        ctx->active.fetch_add(1);
        struct Hold {
            erp_ctx* c;
            ~Hold() { c->active.fetch_sub(1); }
        } hold{ctx};
        erp_run_tracks_impl(ctx, tracks, ntracks, flags, ctx->stream);
        ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
        ctx->prof.resolve();
    });
}

int erp_run_dev_noalloc(erp_ctx* ctx, const erp_params* par, const erp_mesh* mesh,
                        const erp_targets* targets, const erp_shot* shots, int64_t ns,
                        const erp_prism* prisms, int64_t nprisms, void* out_dev,
                        void* nvalid_dev, uint32_t flags) {
    if (!ctx) return ERP_E_INVALID;
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        validate_params(par);
        if (!mesh || !shots || ns < 0 || !out_dev || nprisms < 0 || (nprisms > 0 && !prisms))
            throw ErpError{ERP_E_INVALID, "invalid argument"};
        if (mesh->ctx != ctx || (targets && targets->ctx != ctx))
            throw ErpError{ERP_E_INVALID, "mesh/targets belong to another context"};
        if (ctx->device >= kScratchDevices || !g_scratch[ctx->device].ptr)
            throw ErpError{ERP_E_INVALID, "no allocation-free scratch on device " +
                                              std::to_string(ctx->device)};
        set_device(ctx);
        std::lock_guard<std::mutex> sl(g_scratch[ctx->device].run_mu);
        Arena a;
        a.base = static_cast<char*>(g_scratch[ctx->device].ptr);
        a.cap = (size_t)ERP_SCRATCH_BYTES;
        WorkBufs w;  // every buffer is lent from the arena; nothing to release
        try {
            erp_run_impl(ctx, par, mesh, targets, shots, ns, prisms, nprisms, (float2*)out_dev,
                         (int32_t*)nvalid_dev, flags, ctx->stream, w, nullptr, &a);
            ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            // the scratch is free for the next call only once queued work ended
            hipStreamSynchronize(ctx->stream);
            throw;
        }
    });
}

int erp_run(erp_ctx* ctx, const erp_params* par, const erp_mesh* mesh,
            const erp_targets* targets, const erp_shot* shots, int64_t ns,
            const erp_prism* prisms, int64_t nprisms, float* out, int32_t* nvalid,
            uint32_t flags) {
    if (!ctx) return ERP_E_INVALID;
    if (!out || !par) {
        ctx->last_error = "out/params is NULL";
        return ERP_E_INVALID;
    }
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        validate_params(par);
        if (!mesh || !shots || ns < 0 || nprisms < 0 || (nprisms > 0 && !prisms))
            throw ErpError{ERP_E_INVALID, "invalid argument"};
        if (mesh->ctx != ctx || (targets && targets->ctx != ctx))
            throw ErpError{ERP_E_INVALID, "mesh/targets belong to another context"};
        set_device(ctx);
        const size_t nout = (size_t)ns * par->nr;
        ctx->work.nvalid.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(ns, 1));
        DevBuf& ob = ctx->work.tB;  // output staging
        ob.ensure(sizeof(float2) * std::max<size_t>(nout, 1));
        erp_run_impl(ctx, par, mesh, targets, shots, ns, prisms, nprisms, ob.as<float2>(),
                     ctx->work.nvalid.as<int32_t>(), flags, ctx->stream, ctx->work, &ctx->prof);
        if (nout)
            ERP_HIP_TRY(hipMemcpyAsync(out, ob.p, sizeof(float2) * nout, hipMemcpyDeviceToHost,
                                       ctx->stream));
        if (nvalid && ns)
            ERP_HIP_TRY(hipMemcpyAsync(nvalid, ctx->work.nvalid.p, sizeof(int32_t) * ns,
                                       hipMemcpyDeviceToHost, ctx->stream));
        ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
        ctx->prof.resolve();
    });
}

int erp_serve_slots(erp_ctx* ctx, int n) {
    if (!ctx || n < 1 || n > kMaxSlots) return ERP_E_INVALID;
    return guarded(ctx, [&] {
        set_device(ctx);
        start_slots(ctx, n);
    });
}

int erp_submit(erp_ctx* ctx, const erp_params* par, const erp_mesh* mesh,
               const erp_targets* targets, const erp_shot* shots, int64_t ns,
               const erp_prism* prisms, int64_t nprisms, float* out, int32_t* nvalid,
               uint32_t flags, erp_job** job) {
    if (!ctx || !job) return ERP_E_INVALID;
    *job = nullptr;
    erp_job* j = nullptr;
    int rc = guarded(ctx, [&] {
        validate_params(par);
        if (!mesh || !shots || ns < 0 || !out || nprisms < 0 || (nprisms > 0 && !prisms))
            throw ErpError{ERP_E_INVALID, "invalid argument"};
        if (mesh->ctx != ctx || (targets && targets->ctx != ctx))
            throw ErpError{ERP_E_INVALID, "mesh/targets belong to another context"};
        set_device(ctx);
        if (ctx->slots.empty()) start_slots(ctx, kDefaultSlots);
        j = new erp_job();
        j->ctx = ctx;
        j->par = *par;
        j->mesh = mesh;
        j->targets = targets;
        j->shots.assign(shots, shots + ns);
        if (nprisms) j->prisms.assign(prisms, prisms + nprisms);
        j->flags = flags;
        j->host_out = out;
        j->host_nvalid = nvalid;
        {
            std::lock_guard<std::mutex> lk(ctx->qmu);
            ctx->queue.push_back(j);
        }
        ctx->qcv.notify_one();
    });
    if (rc != ERP_OK) {
        delete j;
        return rc;
    }
    *job = j;
    return ERP_OK;
}

static int job_finish(erp_job* j) {
    const int rc = j->status;
    if (rc != ERP_OK && j->ctx) j->ctx->last_error = j->err;
    delete j;
    return rc;
}

int erp_poll(erp_job* j) {
    if (!j) return ERP_E_INVALID;
    {
        std::lock_guard<std::mutex> lk(j->mu);
        if (!j->done) return ERP_E_NOTREADY;
    }
    return job_finish(j);
}

int erp_wait(erp_job* j) {
    if (!j) return ERP_E_INVALID;
    {
        std::unique_lock<std::mutex> lk(j->mu);
        j->cv.wait(lk, [&] { return j->done; });
    }
    return job_finish(j);
}

int erp_set_tuning(erp_ctx* ctx, int key, int64_t value) {
    if (!ctx || value < 0) return ERP_E_INVALID;
    switch (key) {
        case ERP_TUNE_SHOTS_PER_BATCH: ctx->tune_shots = value; return ERP_OK;
        case ERP_TUNE_PAIRS_PER_LAUNCH: ctx->tune_pairs = value; return ERP_OK;
        case ERP_TUNE_TRACK_ROWS: ctx->tune_rows = value; return ERP_OK;
        case ERP_TUNE_DIRECT_LAUNCH: ctx->tune_direct = value; return ERP_OK;
        case ERP_TUNE_WINDOW_SHIFT: ctx->tune_shift = value; return ERP_OK;
        case ERP_TUNE_SINGLE_SHOT: ctx->tune_single = value; return ERP_OK;
        default: return ERP_E_INVALID;
    }
}

int erp_get_tuning(erp_ctx* ctx, int key, int64_t* value) {
    if (!ctx || !value) return ERP_E_INVALID;
    switch (key) {
        case ERP_TUNE_SHOTS_PER_BATCH: *value = ctx->tune_shots; return ERP_OK;
        case ERP_TUNE_PAIRS_PER_LAUNCH: *value = ctx->tune_pairs; return ERP_OK;
        case ERP_TUNE_TRACK_ROWS: *value = ctx->tune_rows; return ERP_OK;
        case ERP_TUNE_DIRECT_LAUNCH: *value = ctx->tune_direct; return ERP_OK;
        case ERP_TUNE_WINDOW_SHIFT: *value = ctx->tune_shift; return ERP_OK;
        case ERP_TUNE_SINGLE_SHOT: *value = ctx->tune_single; return ERP_OK;
        default: return ERP_E_INVALID;
    }
}

int erp_profile_enable(erp_ctx* ctx, int on) {
    if (!ctx) return ERP_E_INVALID;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->prof.on = on != 0;
    ctx->prof.reset();
    return ERP_OK;
}

int erp_profile_read(erp_ctx* ctx, int stage, double* ms, int64_t* launches, double* units) {
    if (!ctx || stage < 0 || stage >= ERP_NSTAGES) return ERP_E_INVALID;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ms) *ms = ctx->prof.ms[stage];
    if (launches) *launches = ctx->prof.launches[stage];
    if (units) *units = ctx->prof.units[stage];
    return ERP_OK;
}

int erp_focus(erp_ctx* ctx, const float* rdr, int64_t Nr, int64_t Na, int64_t row0,
              int64_t nrows, const int32_t* row_hs, const int32_t* row_nsx,
              const int64_t* row_off, const int32_t* rows, const float* filt, int64_t ntab,
              float* out) {
    if (!ctx) return ERP_E_INVALID;
    return guarded(ctx, [&] {
        if (Nr <= 0 || Na <= 0 || nrows < 0 || row0 < 0 || row0 + nrows > Nr || ntab < 0 ||
            Nr > INT32_MAX || Na > INT32_MAX || Nr * Na > ((int64_t)1 << 34))
            throw ErpError{ERP_E_INVALID, "invalid radargram or row range"};
        if (nrows == 0) return;
        if (!rdr || !out || !row_hs || !row_nsx || !row_off || (ntab > 0 && (!rows || !filt)))
            throw ErpError{ERP_E_INVALID, "NULL argument"};
        for (int64_t q = 0; q < nrows; ++q) {
            const int64_t hs = row_hs[q], nsx = row_nsx[q], off = row_off[q];
            if (hs < 0 || nsx < 0 || 2 * hs > nsx || off < 0 || off + nsx > ntab ||
                nsx > INT32_MAX / 2)
                throw ErpError{ERP_E_INVALID, "row " + std::to_string(row0 + q) +
                                                  ": aperture does not fit the tables"};
        }
        for (int64_t t = 0; t < ntab; ++t)
            if (rows[t] < 0 || rows[t] >= Nr)
                throw ErpError{ERP_E_INVALID, "aperture row index out of the radargram"};
        std::lock_guard<std::mutex> lk(ctx->mu);
        set_device(ctx);
        DevBuf d_rdr, d_hs, d_nsx, d_off, d_rows, d_filt, d_out;
        d_rdr.ensure(sizeof(float2) * (size_t)(Nr * Na));
        d_hs.ensure(sizeof(int) * (size_t)nrows);
        d_nsx.ensure(sizeof(int) * (size_t)nrows);
        d_off.ensure(sizeof(int64_t) * (size_t)nrows);
        d_rows.ensure(sizeof(int) * (size_t)std::max<int64_t>(ntab, 1));
        d_filt.ensure(sizeof(float2) * (size_t)std::max<int64_t>(ntab, 1));
        d_out.ensure(sizeof(float2) * (size_t)(nrows * Na));
        hipStream_t st = ctx->stream;
        ERP_HIP_TRY(hipMemcpyAsync(d_rdr.p, rdr, sizeof(float2) * Nr * Na,
                                   hipMemcpyHostToDevice, st));
        ERP_HIP_TRY(hipMemcpyAsync(d_hs.p, row_hs, sizeof(int) * nrows, hipMemcpyHostToDevice, st));
        ERP_HIP_TRY(hipMemcpyAsync(d_nsx.p, row_nsx, sizeof(int) * nrows, hipMemcpyHostToDevice,
                                   st));
        ERP_HIP_TRY(hipMemcpyAsync(d_off.p, row_off, sizeof(int64_t) * nrows,
                                   hipMemcpyHostToDevice, st));
        if (ntab > 0) {
            ERP_HIP_TRY(hipMemcpyAsync(d_rows.p, rows, sizeof(int) * ntab, hipMemcpyHostToDevice,
                                       st));
            ERP_HIP_TRY(hipMemcpyAsync(d_filt.p, filt, sizeof(float2) * ntab,
                                       hipMemcpyHostToDevice, st));
        }
        focus_rows(d_rdr.as<float2>(), (int)Nr, (int)Na, (int)row0, (int)nrows, d_hs.as<int>(),
                   d_nsx.as<int>(), d_off.as<int64_t>(), d_rows.as<int>(), d_filt.as<float2>(),
                   d_out.as<float2>(), st);
        ERP_HIP_TRY(hipMemcpyAsync(out, d_out.p, sizeof(float2) * nrows * Na,
                                   hipMemcpyDeviceToHost, st));
        ERP_HIP_TRY(hipStreamSynchronize(st));
    });
}

int64_t erp_run_footprint(const erp_params* par, int64_t F, int64_t T, int64_t ns) {
    if (!par) return -1;
    return erp_footprint_impl(par, F, T, ns);
}

}  // extern "C"
