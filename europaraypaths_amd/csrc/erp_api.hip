// erp_api.hip — the extern "C" boundary of include/erp.h.
//
// Replaces the process boundary of the reference (sim.cu:92-613: parameter
// parse, file loads, device allocation and upload) and its exit()-on-error
// policy (sim.cu:60-70) with status codes and a per-context error string.
#include "erp_internal.h"

#include <algorithm>
#include <cmath>

#include <cstring>
#include <exception>
#include <new>

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

struct erp_job {
    erp_ctx* ctx;
    hipStream_t stream;
    hipEvent_t done;
    WorkBufs work;
    DevBuf out;
    DevBuf nvalid;
    float* host_out;
    int32_t* host_nvalid;
    int64_t n_out;
    int64_t ns;
    int status;
    std::string err;
};

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
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (auto* c : ctx->chirps) delete c;
    for (auto* b : ctx->blues) delete b;
    ctx->chirps.clear();
    ctx->blues.clear();
    WorkBufs& w = ctx->work;
    DevBuf* bufs[] = {&w.cnt, &w.off, &w.V, &w.Vd, &w.tcnt, &w.tlist, &w.tcount, &w.rec,
                      &w.bounds, &w.psurf, &w.refl, &w.shotS, &w.blu, &w.tA, &w.tB, &w.tC,
                      &w.pairs, &w.t0trace, &w.shots, &w.prisms, &w.out, &w.nvalid};
    for (DevBuf* b : bufs) b->release();
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
    });
    if (rc != ERP_OK) {
        m->soa.release();
        delete m;
        return rc;
    }
    *out = m;
    return ERP_OK;
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
                     (int32_t*)nvalid_dev, flags, ctx->stream, ctx->work);
        ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
        ctx->prof.resolve();
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
                     ctx->work.nvalid.as<int32_t>(), flags, ctx->stream, ctx->work);
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

int erp_submit(erp_ctx* ctx, const erp_params* par, const erp_mesh* mesh,
               const erp_targets* targets, const erp_shot* shots, int64_t ns,
               const erp_prism* prisms, int64_t nprisms, float* out, int32_t* nvalid,
               uint32_t flags, erp_job** job) {
    if (!ctx || !job) return ERP_E_INVALID;
    *job = nullptr;
    erp_job* j = new (std::nothrow) erp_job();
    if (!j) return ERP_E_NOMEM;
    j->ctx = ctx;
    j->stream = nullptr;
    j->done = nullptr;
    int rc = guarded(ctx, [&] {
        validate_params(par);
        if (!mesh || !shots || ns < 0 || !out || nprisms < 0 || (nprisms > 0 && !prisms))
            throw ErpError{ERP_E_INVALID, "invalid argument"};
        set_device(ctx);
        ERP_HIP_TRY(hipStreamCreateWithFlags(&j->stream, hipStreamNonBlocking));
        ERP_HIP_TRY(hipEventCreateWithFlags(&j->done, hipEventDisableTiming));
        {
            // chirp/Bluestein caches are shared: build them under the lock
            std::lock_guard<std::mutex> lk(ctx->mu);
            const size_t nout = (size_t)ns * par->nr;
            j->out.ensure(sizeof(float2) * std::max<size_t>(nout, 1));
            j->nvalid.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(ns, 1));
            // make sure context-level tables exist before the job's stream uses them
            ERP_HIP_TRY(hipStreamSynchronize(ctx->stream));
            const bool prof_on = ctx->prof.on;
            ctx->prof.on = false;  // stage timers cover synchronous runs only
            try {
                erp_run_impl(ctx, par, mesh, targets, shots, ns, prisms, nprisms,
                             j->out.as<float2>(), j->nvalid.as<int32_t>(), flags, j->stream,
                             j->work);
            } catch (...) {
                ctx->prof.on = prof_on;
                throw;
            }
            ctx->prof.on = prof_on;
            j->host_out = out;
            j->host_nvalid = nvalid;
            j->n_out = (int64_t)nout;
            j->ns = ns;
            if (nout)
                ERP_HIP_TRY(hipMemcpyAsync(out, j->out.p, sizeof(float2) * nout,
                                           hipMemcpyDeviceToHost, j->stream));
            if (nvalid && ns)
                ERP_HIP_TRY(hipMemcpyAsync(nvalid, j->nvalid.p, sizeof(int32_t) * ns,
                                           hipMemcpyDeviceToHost, j->stream));
            ERP_HIP_TRY(hipEventRecord(j->done, j->stream));
        }
    });
    if (rc != ERP_OK) {
        if (j->done) hipEventDestroy(j->done);
        if (j->stream) hipStreamDestroy(j->stream);
        delete j;
        return rc;
    }
    *job = j;
    return ERP_OK;
}

static void job_free(erp_job* j) {
    hipSetDevice(j->ctx->device);
    WorkBufs& w = j->work;
    DevBuf* bufs[] = {&w.cnt, &w.off, &w.V, &w.Vd, &w.tcnt, &w.tlist, &w.tcount, &w.rec,
                      &w.bounds, &w.psurf, &w.refl, &w.shotS, &w.blu, &w.tA, &w.tB, &w.tC,
                      &w.pairs, &w.t0trace, &w.shots, &w.prisms, &w.out, &w.nvalid};
    for (DevBuf* b : bufs) b->release();
    j->out.release();
    j->nvalid.release();
    if (j->done) hipEventDestroy(j->done);
    if (j->stream) hipStreamDestroy(j->stream);
    delete j;
}

int erp_poll(erp_job* j) {
    if (!j) return ERP_E_INVALID;
    hipError_t e = hipEventQuery(j->done);
    if (e == hipErrorNotReady) return ERP_E_NOTREADY;
    int rc = (e == hipSuccess) ? ERP_OK : ERP_E_HIP;
    if (rc != ERP_OK) j->ctx->last_error = hipGetErrorString(e);
    job_free(j);
    return rc;
}

int erp_wait(erp_job* j) {
    if (!j) return ERP_E_INVALID;
    hipError_t e = hipEventSynchronize(j->done);
    int rc = (e == hipSuccess) ? ERP_OK : ERP_E_HIP;
    if (rc != ERP_OK) j->ctx->last_error = hipGetErrorString(e);
    job_free(j);
    return rc;
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

int64_t erp_run_footprint(const erp_params* par, int64_t F, int64_t T, int64_t ns) {
    if (!par) return -1;
    return erp_footprint_impl(par, F, T, ns);
}

}  // extern "C"
