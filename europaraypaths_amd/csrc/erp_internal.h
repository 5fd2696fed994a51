// erp_internal.h — shared declarations of the gfx950 HIP implementation of
// the per-shot raypath / phasor / range-compression path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include <string>
#include <vector>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

#include "../../include/erp.h"

// ---------------------------------------------------------------------------
// error plumbing: every failure is returned as a status, never exit()
// ---------------------------------------------------------------------------
struct ErpError {
    int code;
    std::string msg;
};

#define ERP_HIP_TRY(expr)                                                     \
    do {                                                                      \
        hipError_t _e = (expr);                                               \
        if (_e != hipSuccess)                                                 \
            throw ErpError{ERP_E_HIP, std::string(#expr) + ": " +             \
                                          hipGetErrorString(_e)};             \
    } while (0)

#define ERP_CHECK_LAUNCH(name)                                                \
    do {                                                                      \
        hipError_t _e = hipGetLastError();                                    \
        if (_e != hipSuccess)                                                 \
            throw ErpError{ERP_E_HIP, std::string("launch ") + (name) +       \
                                          ": " + hipGetErrorString(_e)};      \
    } while (0)

// ---------------------------------------------------------------------------
// constants shared by every kernel (the reference's literals)
// ---------------------------------------------------------------------------
// geometry_kernels.cu:45 `const float pi = 3.14159265f`
#define ERP_PI_F 3.14159265f
// geometry_kernels.cu:388,411 `pow(4 * 3.14159, 2)` evaluated in double
#define ERP_FOURPI5_SQ (4.0 * 3.14159 * 4.0 * 3.14159)

// FFT twiddle table size (largest supported transform)
#define ERP_TWN (1 << 18)

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    // a region of memory the buffer does not own (see Arena): never freed,
    // never grown; asking for more than it holds is an error
    bool borrowed = false;
    void borrow(void* q, size_t n) {
        p = q;
        bytes = n;
        borrowed = true;
    }
    void ensure(size_t n) {
        if (n <= bytes) return;
        if (borrowed)
            throw ErpError{ERP_E_NOMEM, "a work buffer needs " + std::to_string(n) +
                                            " bytes, its scratch region holds " +
                                            std::to_string(bytes)};
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
        if (n == 0) return;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            throw ErpError{ERP_E_NOMEM, "hipMalloc of " + std::to_string(n) +
                                            " bytes failed: " + hipGetErrorString(e)};
        }
        bytes = n;
    }
    // for sizes that vary from call to call: grow by up to a quarter more
    // than asked (never beyond cap, the caller's bound on n), so that a
    // slowly rising size does not free and reallocate (hipFree waits for the
    // device) on every call
    void ensure_grow(size_t n, size_t cap) {
        if (n > bytes) ensure(std::max(n, std::min(n + n / 4, cap)));
    }
    void release() {
        if (p && !borrowed) hipFree(p);
        p = nullptr;
        bytes = 0;
        borrowed = false;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Device memory of one run carved from a region the caller provides (the
// allocation-free entry point): 256-byte aligned bump allocation, nothing is
// freed before the run ends.
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    static size_t round(size_t n) { return (n + 255) & ~(size_t)255; }
    size_t left() const { return cap - used; }
    void* take(size_t n, const char* what) {
        const size_t r = round(std::max<size_t>(n, 1));
        if (r > left())
            throw ErpError{ERP_E_NOMEM, std::string("scratch exhausted: ") + what + " needs " +
                                            std::to_string(r) + " bytes, " +
                                            std::to_string(left()) + " of " +
                                            std::to_string(cap) + " left"};
        void* p = base + used;
        used += r;
        return p;
    }
    template <class T> T* take_n(size_t n, const char* what) {
        return static_cast<T*>(take(sizeof(T) * n, what));
    }
    void lend(DevBuf& b, size_t n, const char* what) { b.borrow(take(n, what), round(std::max<size_t>(n, 1))); }
};

struct erp_mesh {
    erp_ctx* ctx;
    int device;
    int64_t F;
    DevBuf soa;  // 12 x F float32 (stride Fpad)
    int64_t Fpad;
    // every facet normal equals +-(u x v) to 1e-5: kernels may derive n from
    // the tangents instead of loading it
    bool frame_ok = false;
    // facets in rows of tile_nx (a multiple of 64, row count a multiple of
    // 32): the surface kernels then take them in 64 x 32 tiles; 0 = in order
    int tile_nx = 0;
};

struct erp_targets {
    erp_ctx* ctx;
    int device;
    int64_t T;
    DevBuf xyzn;  // 6 x Tpad float32
    DevBuf type;  // T uint8
    int64_t Tpad;
    int64_t n0 = 0;  // targets of type 0
};

// FFT plan for one power-of-two size N (see erp_fft.hip)
struct FftPlan {
    int N = 0, log2N = 0;
    int N1 = 0, N2 = 0;  // two-pass split: N = N1 (row, contiguous) * N2 (rows)
    bool two_pass = false;
};

// Per-(nr, dr, res) chirp spectra and per-L Bluestein tables
struct ChirpKey {
    int nr;
    float dr, res;
    int kind;  // 0 = padded (linear), 1 = circular (genChirp), rst in key2
    float rst;
    bool operator==(const ChirpKey& o) const {
        return nr == o.nr && dr == o.dr && res == o.res && kind == o.kind && rst == o.rst;
    }
};

struct ChirpSpec {
    ChirpKey key;
    DevBuf spec;  // P complex64, permuted forward order, scaled by 1/P
    DevBuf nat;   // the same spectrum in natural order (filled on first use)
    DevBuf wl;    // the natural one in the layout of k_tpair_wl (first use)
};

struct BlueTables {
    int L = 0, M = 0;
    float smpl = 0, B = 0;
    DevBuf w;     // L complex64: exp(-i pi n^2 / L)
    DevBuf Bw;    // M complex64: FFT_M of the conjugate chirp, scaled 1/M
    DevBuf mask;  // L uint8: crop mask (1 = keep)
};

struct WorkBufs {
    DevBuf cnt, off, V, Vd, tcnt, tlist, tcount;  // counts / offsets
    DevBuf rec;        // compacted facet records
    DevBuf bounds;     // per-shot bounds of illuminated facets
    DevBuf psurf;      // S x P complex: surface phasor trace (padded)
    DevBuf srow;       // S x nr FP64 / fixed-point surface bins + S per-shot maxima
    DevBuf refl;       // S x nr complex
    DevBuf shotS;      // S x P complex: per-shot spectral accumulator
    DevBuf blu;        // S x M complex: Bluestein workspace
    DevBuf tA, tB, tC; // per-target chunk buffers, C x P complex
    DevBuf pairs;      // chunk pair list
    DevBuf t0trace;    // S x P complex: summed type-0 target histograms
    DevBuf shots;      // S erp_shot
    DevBuf prisms;
    DevBuf out;        // S x nr complex (device output staging)
    DevBuf nvalid;     // S int32
    DevBuf oow;        // count of target terms outside their pair's bin window
    // shot-pair target kernel: union masks, counts, offsets and sizes per
    // block of two shots; block records; work items
    DevBuf umask, ucnt, uoff, U, ubase, brec, items;
    DevBuf bbox;       // per (shot, facet block) box of the lit facets
};

// stage timers: events recorded on the run stream, resolved after it syncs
struct StageTimer {
    bool on = false;
    double ms[ERP_NSTAGES] = {};
    int64_t launches[ERP_NSTAGES] = {};
    double units[ERP_NSTAGES] = {};
    struct Pending {
        int stage;
        hipEvent_t a, b;
        double units;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    hipEvent_t get();
    void begin(int stage, hipStream_t s, double units);
    void end(hipStream_t s);
    void resolve();  // after a stream sync
    void reset();
    int cur = -1;
};

struct ErpSlot;  // asynchronous job worker (erp_api.hip)

struct MtState;
void mt_state_free(MtState* m);

struct erp_ctx {
    int device = 0;
    StageTimer prof;
    hipStream_t stream = nullptr;
    std::string last_error;
    DevBuf twiddle;  // ladder table of the LDS FFT stages (65536 complex64)
    std::vector<std::pair<int, DevBuf*>> inter;  // inter-pass twiddles per size
    std::vector<ChirpSpec*> chirps;
    std::vector<BlueTables*> blues;
    WorkBufs work;
    std::mutex mu;        // serialises the synchronous entry points
    // guards inter / chirps / blues (shared by all streams); recursive because
    // building a chirp or Bluestein table runs FFTs that look up inter tables
    std::recursive_mutex cache_mu;
    int num_cu = 0;
    // asynchronous jobs: worker slots, each with its own thread, stream and
    // work buffers, fed from one FIFO
    std::vector<ErpSlot*> slots;
    std::deque<erp_job*> queue;
    std::mutex qmu;
    std::condition_variable qcv;
    bool stopping = false;
    std::atomic<int> active{0};  // jobs running in slots
    // caps on the batch shapes (erp_set_tuning; 0 = default)
    std::atomic<int64_t> tune_shots{0}, tune_pairs{0};
    // multi-track batches: rows per batch (0 = default), launch directly
    // instead of replaying captured graphs (> 0)
    std::atomic<int64_t> tune_rows{0}, tune_direct{0};
    // > 0 only in tests: shift every pair's bin window up by this many bins
    std::atomic<int64_t> tune_shift{0};
    // > 0: the plain target kernel one (shot, target) pair at a time
    std::atomic<int64_t> tune_single{0};
    MtState* mt = nullptr;  // buffers and graphs of erp_run_tracks
};

// ---------------------------------------------------------------------------
// FFT interface (erp_fft.hip)
// ---------------------------------------------------------------------------
FftPlan fft_plan(int N);
// Forward transform of `batch` rows of length N (row stride N), in place.
// Output in the plan's permuted order (natural order for single-pass sizes).
void fft_forward(erp_ctx* ctx, const FftPlan& p, float2* data, int batch, hipStream_t s);
// Inverse (unnormalised) of a permuted-order spectrum, in place, natural output.
void fft_inverse(erp_ctx* ctx, const FftPlan& p, float2* data, int batch, hipStream_t s);
void fft_init_twiddles(erp_ctx* ctx);
// Fused per-target spectral chain (erp_fft.hip): for n pairs with windowed
// inward traces in A (rows of P), accumulates scale * sum_p FFT(A_p) *
// FFT(window(IFFT(FFT(A_p) G))) into S[shot] for each shot segment.  Groups
// [gbeg, gend) partition the pairs without crossing shots; segment k sums
// groups (two-pass sizes) or pairs (single-pass sizes) [gseg_beg, gseg_end).
// one workgroup per (shot, target) pair, whole chain in LDS: P = 2n, n the
// next power of two of nr, from 1024 to 16384 (nr < n: the cropped variant)
bool target_half_supported(const FftPlan& p, int nr);
int target_half_slice();
// the chain runs with the wave-local transforms (k_tpair_wl): its range
// kernel spectra are then in that kernel's layout (wl_perm_table) and its
// rows are summed with target_sum_half(..., wl = true)
bool target_wl(const FftPlan& p);
// Gwl = Gnat in the layout of k_tpair_wl (P complex64 each)
void wl_perm_table(const float2* Gnat, float2* Gwl, int P, hipStream_t st);
// Gpair / npairs_dev (multi-track batches): per-pair range kernel spectra
// and the pair count on the device, n then being the grid's bound
void target_pairs_half(erp_ctx* ctx, const FftPlan& p, const float2* A, const int* pbase,
                       const int* pwin, int nr, int n, const float2* Gnat, float2* work,
                       hipStream_t st, const float4* const* Gpair = nullptr,
                       const int* npairs_dev = nullptr);
// seg_scale (multi-track batches): per-segment scale instead of `scale`;
// seq: each slice added to S on its own, in order, so that S does not
// depend on how launches group the slices (deterministic mode)
void target_sum_half(const FftPlan& p, const float2* work, float2* partial, const int* sl_beg,
                     const int* sl_end, int nslices, const int* seg_shot, const int* seg_beg,
                     const int* seg_end, int nseg, float2* S, float scale, hipStream_t st,
                     const float* seg_scale = nullptr, bool seq = false, bool wl = false);
// natural-order copy of a forward spectrum produced by fft_forward
void fft_natural_order(const FftPlan& p, const float2* src, float2* dst, hipStream_t st);
void target_spectra_fused(erp_ctx* ctx, const FftPlan& p, float2* A, float2* Y,
                          const int* pbase, const int* pwin, int nr, int n, const float2* G,
                          float2* work, int ngroups, const int* gbeg, const int* gend,
                          const int* gseg_shot, const int* gseg_beg, const int* gseg_end,
                          int nseg, float2* S, float scale, hipStream_t st);

// ---------------------------------------------------------------------------
// pipeline (erp_pipeline.hip)
// ---------------------------------------------------------------------------
// timer: stage timers to record into, or nullptr (asynchronous jobs).
// arena: nullptr, or the region every work buffer and table of the run is
// carved from (then `work` must be empty and nothing is allocated; batch
// shapes shrink to fit and a job that cannot fit fails with ERP_E_NOMEM).
void erp_run_impl(erp_ctx* ctx, const erp_params* par, const erp_mesh* mesh,
                  const erp_targets* targets, const erp_shot* shots, int64_t ns,
                  const erp_prism* prisms, int64_t nprisms, float2* out_dev,
                  int32_t* nvalid_dev, uint32_t flags, hipStream_t stream,
                  WorkBufs& work, StageTimer* timer, Arena* arena = nullptr);
int64_t erp_footprint_impl(const erp_params* par, int64_t F, int64_t T, int64_t ns);
// multi-track batches of the eligible tracks, every other track alone
void erp_run_tracks_impl(erp_ctx* ctx, const erp_track* tracks, int64_t ntracks, uint32_t flags,
                         hipStream_t stream);
// focused rows row0 .. row0 + nrows - 1 of an Nr x Na radargram (erp_focus.hip)
void focus_rows(const float2* rdr, int Nr, int Na, int row0, int nrows, const int* row_hs,
                const int* row_nsx, const int64_t* row_off, const int* rows, const float2* filt,
                float2* out, hipStream_t st);
// 12 x Fpad SoA of a height grid (see erp_mesh_from_heights)
void mesh_from_heights(const double* heights, int64_t nx, int64_t ny, double fs, double ox,
                       double oy, float* soa, int64_t Fpad, hipStream_t st);
// ny x nx FP64 heights of a synthetic terrain into device memory h (see
// erp_mesh_from_terrain; erp_terrain.hip); synchronous on st
void terrain_heights(erp_ctx* ctx, const erp_terrain& t, double* h, hipStream_t st);

inline int next_pow2(int64_t n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}
inline int ilog2(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}
