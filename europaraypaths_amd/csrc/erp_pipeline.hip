// erp_pipeline.hip — the per-shot path of sim.cu:615-1008 as gfx950 kernels.
//
// Per batch of S shots (all shots of a batch run together):
//   k_illum_mask   aperture predicate of every (shot, facet): FP32 screen,
//                  exact FP64 decision near the threshold; one 64-bit ballot
//                  word per wave -> mask, per-block counts.
//                  (compSourceInclination + cropByAperture,
//                  geometry_kernels.cu:75-100, 143-204)
//   k_scan_blocks  per-shot exclusive scan of the block counts -> stable
//                  record offsets, V (illuminated facets) per shot.
//   k_surface      fused compIncidentRays + compRefractedRays +
//                  compReflectedEnergy + surfacePT / genReflPhasor
//                  (geometry_kernels.cu:103-140, 426-517; radar_kernels.cu:
//                  314-348, 714-764): one thread per illuminated facet, FP64
//                  range and phase, two-bin split accumulated in an LDS window,
//                  compacted facet records for the target stage.
//   surface range compression: FFT -> x chirp spectrum -> IFFT, batched over
//                  shots (convolvePhasorChirpLinear, radar_kernels.cu:554-649).
//   k_target_select per-shot in-aperture target lists (sim.cu:808-826).
//   k_target_accum one workgroup per (shot, target): compTargetRays +
//                  compRefrEnergyIn/Out + accumulateTarget/radiateTarget
//                  (geometry_kernels.cu:272-350, 520-584; radar_kernels.cu:
//                  767-892) over all illuminated facets, LDS window histogram.
//   target spectra: FFT, x chirp(2 res), IFFT, window, FFT, and the product
//                  with the outward trace's spectrum summed over the shot's
//                  targets in the frequency domain; one inverse + Bluestein
//                  DFT_L / crop / IDFT_L per shot replaces the reference's
//                  per-target convolveComplex (elementary_kernels.cu:258-353).
//   k_finalize     refl + refr -> output (combineRadarSignals, radar_kernels.cu:
//                  302-311).
//
// Every convolution of the reference (2nr circular + window, nr circular) is
// evaluated as an exact linear convolution with a rearranged kernel g of
// length 2nr-1 through a power-of-two FFT of size P = nextpow2(2nr-1):
//   out[j] = sum_m x[m] g[j + nr - 1 - m],  j < nr.
#include "erp_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

constexpr int NT = 256;        // threads per workgroup
constexpr int FB = 2048;       // facets per block in the surface kernels
constexpr int FPT = FB / NT;   // facets per thread (8)
constexpr int WPB = NT / 64;   // waves per block
constexpr int MASKW = FB / 64; // mask words per block (32)
constexpr int SWIN = 2048;     // surface LDS window (bins)
constexpr int TWIN = 4096;     // target LDS window (bins)
constexpr int NTT = 512;       // threads per workgroup of the target kernels
// (shot, target) pairs per target launch: their P-point rows take at most
// kPairRowBytes per buffer set.  Large launches keep the last round of
// workgroups a small part of each launch and need fewer host round trips.
constexpr int64_t kPairRowBytes = (int64_t)6144 << 20;
constexpr int kMaxPairsPerLaunch = 16384;
// Shots per batch (default and maximum).  Each batch waits on the host a
// few times (illuminated counts, target counts and lists), so a 50-shot
// step runs as one batch rather than two.
constexpr int kShotsPerBatch = 64;
// bin window of k_target_accum_fast: FP64, 79 KB of LDS, so two workgroups
// fill a CU's 160 KB; the shot bounds of the 2048^2 fBm workload allow at
// most about 4.7e3 bins per pair
constexpr int TWF = 5056;
constexpr int NTF = 512;       // threads of k_target_accum_fast (two workgroups per CU)
// Bin windows accumulate in FP64: LDS atomic adds of doubles (ds_add_f64) run
// at more than twice the rate of FP32 ones (ds_add_f32) on gfx950 and remove
// FP32 accumulation error from the bins.

// Kernel-side parameter block (derived once per run on the host).
struct KPar {
    float lam, fs, P, ks, eps1, eps2, nu1, nu2, alpha1, alpha2, rng_res, dr;
    float c, c1, c2;
    float piLlam;      // PI_F*fs/lam (facet_G's sinc argument scale)
    float Dmax;        // 4*PI_F*fs^2/lam^2
    float fs_lam;      // fs/lam (facetReradiation)
    float fs2_lam;     // fs*fs/lam
    float nf;          // sqrt(eps_2) in float (accumulateTarget's n)
    int ortho;         // every facet frame (u, v, n) orthonormal (see sc_angles)
    float sqrtP;
    float invK;        // 1/(4*3.14159)^2
    float sqrt_e1e2;   // sqrt(eps_1/eps_2)
    float sqrt_e2e1;   // sqrt(eps_2/eps_1)
    double drd, inv_dr, cos_ap, n_d, inv_cyc, lam_d;
    // c1/c2 and c/c2 (genRefrPhasor, refrRadarSignal) divided in FP64: a
    // float quotient is off by up to 6e-8, i.e. 5e-5 rad of phase over a
    // 300 m refracted leg at 60 MHz
    double c1c2, cc2;
    // second subsurface medium below z = zi (has3 = 0: none): dn = n3 - n2,
    // tsq = sqrt(1 - rho23^2) (amplitude factor of a crossing path),
    // n_lo / n_hi = min / max(n2, n3) (window bounds)
    int has3;
    double zi, dn, n_lo, n_hi;
    float tsq;
    double B, smpl;
    int nr, P_fft, rerad;
    int row_live;      // entries of a pair row the spectral chain reads: P_fft / 2
                       // for the per-pair LDS chain, P_fft otherwise
    uint32_t mode;
    uint64_t seed;
    int win_shift;     // 0; > 0 only to test the target kernel's window check
};

struct ShotK {
    double sx, sy, sz;
    float snx, sny, snz, rst, grefl, grefr;
};

// Multi-track batches (erp_run_tracks): the rows of a batch are shots of
// several tracks of one range length; every kernel that reads a track's
// data takes it from that track's descriptor instead of its arguments.
struct TrackK {
    KPar kp;
    const float* mesh;             // 12 x Fpad SoA
    int64_t Fpad, F;
    int nblk, tile_nx, frame_ok;
    const float* tg;               // 6 x Tpad targets
    int64_t Tpad;
    int T;
    const unsigned char* ttype;
    const float2* g_surf;          // surface range kernel spectrum (fft order)
    const float2* g_targ_nat;      // target range kernel spectrum (natural order, or
                                   // in the layout of k_tpair_wl where it runs)
    const float2* blue_w;          // Bluestein tables of the track's band
    const float2* blue_Bw;
    const unsigned char* blue_mask;
    float2* out;                   // the track's radargram, ns x nr (device)
    int32_t* nvalid;               // its illuminated counts (device) or null
    float tscale;                  // -1 / sqrt(P): outward trace of a target
};
struct RowK {
    int trk;   // track of the row
    int shot;  // the row's shot index within its track
};

__device__ __forceinline__ ShotK load_shot(const erp_shot* __restrict__ shots, int s) {
    const erp_shot& h = shots[s];
    ShotK k;
    k.sx = h.sx; k.sy = h.sy; k.sz = h.sz;
    k.snx = h.snx; k.sny = h.sny; k.snz = h.snz;
    k.rst = h.rst; k.grefl = h.grefl_lin; k.grefr = h.grefr_lin;
    return k;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ---------------------------------------------------------------------------
// exact FP64 pieces (no contraction: identical operation sequence to the
// oracle's -ffp-contract=off build, so bins compare bit for bit)
// ---------------------------------------------------------------------------
__device__ __noinline__ bool illum_exact(float fx, float fy, float fz, const ShotK& s,
                                         double cos_ap) {
#pragma clang fp contract(off)
    double dx = (double)fx - s.sx, dy = (double)fy - s.sy, dz = (double)fz - s.sz;
    double Itd = sqrt(dx * dx + dy * dy + dz * dz);
    double Ix = dx / Itd, Iy = dy / Itd, Iz = dz / Itd;
    double c = (-Ix) * (double)s.snx + (-Iy) * (double)s.sny + (-Iz) * (double)s.snz;
    return c > cos_ap && c <= 1.0;
}

__device__ __forceinline__ double dist3(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    return sqrt(dx * dx + dy * dy + dz * dz);
}

__device__ __forceinline__ double madd_nc(double a, double b, double n) {
#pragma clang fp contract(off)
    return a + b * n;  // a + b*n, two roundings
}

// Range -> (bin, weight) of radar_kernels.cu:736-737 in FP64; valid iff
// -1 < x < nr (the CUDA (short) conversion saturates).
__device__ __forceinline__ bool bin_of(double rng, double rst, double dr, int nr, int* b,
                                       float* w) {
#pragma clang fp contract(off)
    double x = (rng - rst) / dr;
    if (!(x > -1.0) || !(x < (double)nr)) return false;
    int bi = (int)x;
    *b = bi;
    *w = (float)(x - (double)bi);
    return true;
}
__device__ __forceinline__ bool bin_of(double rng, double rst, double dr, int nr, int* b,
                                       double* w) {
    float wf;
    const bool ok = bin_of(rng, rst, dr, nr, b, &wf);
    *w = wf;
    return ok;
}

// exp(-i 2 pi cyc): phase from FP64 cycles, reduced before FP32 sincospi.
__device__ __forceinline__ float2 phasor_cycles(double cyc) {
    double f = cyc - floor(cyc);
    float s, c;
    sincospif(-2.0f * (float)f, &s, &c);
    return make_float2(c, s);
}

// Cosine and sine of the angle between -I and a stored normal m as the
// reference forms them (geometry_kernels.cu:103-140): c = clamp(-I.m, -1, 1),
// s = sin(acos(c)) = sqrt((1 - c)(1 + c)), with I = (ix, iy, iz) in FP64.
// Near c = 1 the sine depends on |m| (stored normals are unit length only to
// float rounding) at the 1e-4 level, which FP32 products cannot resolve.
__device__ __forceinline__ float2 cos_sin_ref(double ix, double iy, double iz, float mx, float my,
                                              float mz) {
    double c = -(ix * mx + iy * my + iz * mz);
    c = fmin(1.0, fmax(-1.0, c));
    return make_float2((float)c, sqrtf((float)((1.0 - c) * (1.0 + c))));
}

// cos_sin_ref with the hardware square root (about 1 ulp, no correctly
// rounded expansion): sin = sqrt((1 - c)(1 + c)) with c in FP64, so the
// cancellation near normal incidence stays exact and only the final FP32
// root is approximate
__device__ __forceinline__ float2 cos_sin_hw(double ix, double iy, double iz, float mx, float my,
                                             float mz) {
    double c = -(ix * mx + iy * my + iz * mz);
    c = fmin(1.0, fmax(-1.0, c));
    return make_float2((float)c, __builtin_amdgcn_sqrtf((float)((1.0 - c) * (1.0 + c))));
}

// (I.u, I.v) in FP64.  Needed for a general (non-orthonormal) facet frame,
// where the azimuth atan2(I.u, I.v) is ill-conditioned near u x v and FP32
// dot products lose it.
__device__ __forceinline__ float2 uv_dots_exact(double ix, double iy, double iz, float ux,
                                                float uy, float uz, float vx, float vy,
                                                float vz) {
    return make_float2((float)(ix * ux + iy * uy + iz * uz), (float)(ix * vx + iy * vy + iz * vz));
}

// sin(th) cos(ph) and sin(th) sin(ph) for ph = atan2(I.u, I.v)
// (compIncidentRays), given s = sin(th); ph = 0 when both projections
// vanish, as atan2(0, 0) = 0.
__device__ __forceinline__ float2 sc_angles(float Iu, float Iv, float s) {
    const float r2 = Iu * Iu + Iv * Iv;
    if (!(r2 > 0.f)) return make_float2(s, 0.f);
    const float ri = rsqrtf(r2);
    return make_float2(s * Iv * ri, s * Iu * ri);
}

// elementary_kernels.cu:50-53
__device__ __forceinline__ float sinc_f(float x) {
    if (fabsf(x) < 1e-6f) return 1.0f;
    return sinpif(x) / (3.14159265358979f * x);
}

// sin(pi r) and cos(pi r) for |r| <= 0.5 from their Taylor series (degree 13
// and 14); truncation error < 7e-10, FP32 evaluation error < 2e-7.
__device__ __forceinline__ void sincospi_red(float r, float* s, float* c) {
    const float r2 = r * r;
    float ps = 0.00046630280576761234f;
    ps = fmaf(ps, r2, -0.0073704309457143478f);
    ps = fmaf(ps, r2, 0.082145886611128191f);
    ps = fmaf(ps, r2, -0.59926452932079188f);
    ps = fmaf(ps, r2, 2.5501640398773451f);
    ps = fmaf(ps, r2, -5.1677127800499694f);
    ps = fmaf(ps, r2, 3.1415926535897931f);
    *s = r * ps;
    float pc = -0.00010463810492484565f;
    pc = fmaf(pc, r2, 0.0019295743094039221f);
    pc = fmaf(pc, r2, -0.025806891390014051f);
    pc = fmaf(pc, r2, 0.23533063035889312f);
    pc = fmaf(pc, r2, -1.3352627688545893f);
    pc = fmaf(pc, r2, 4.0587121264167676f);
    pc = fmaf(pc, r2, -4.934802200544679f);
    *c = fmaf(pc, r2, 1.0f);
}

// sin(pi r) only, |r| <= 0.5
__device__ __forceinline__ float sinpi_red(float r) {
    const float r2 = r * r;
    float ps = 0.00046630280576761234f;
    ps = fmaf(ps, r2, -0.0073704309457143478f);
    ps = fmaf(ps, r2, 0.082145886611128191f);
    ps = fmaf(ps, r2, -0.59926452932079188f);
    ps = fmaf(ps, r2, 2.5501640398773451f);
    ps = fmaf(ps, r2, -5.1677127800499694f);
    ps = fmaf(ps, r2, 3.1415926535897931f);
    return r * ps;
}

// sinc(x) of elementary_kernels.cu:50-53 as numerator / denominator so that a
// product of sincs needs one reciprocal: sinc = num / den.
__device__ __forceinline__ void sinc_nd(float x, float* num, float* den) {
    const bool tiny = fabsf(x) < 1e-6f;
    const float n = rintf(x);
    float s = sinpi_red(x - n);
    s = ((int)n & 1) ? -s : s;
    *num = tiny ? 1.0f : s;
    *den = tiny ? 1.0f : 3.14159265358979f * x;
}

// sin(pi x) from the hardware sine, whose argument is in revolutions:
// x/2 - rint(x/2) is exact and lies in [-1/2, 1/2], where v_sin_f32 is
// accurate to 1.3e-7 absolute (checked against double on gfx950).
__device__ __forceinline__ float sinpi_hw(float x) {
    const float h = 0.5f * x;
    return __builtin_amdgcn_sinf(h - rintf(h));
}

// sinc as numerator / denominator (see sinc_nd) with the hardware sine
__device__ __forceinline__ void sinc_nd_hw(float x, float* num, float* den) {
    const bool tiny = fabsf(x) < 1e-6f;
    *num = tiny ? 1.0f : sinpi_hw(x);
    *den = tiny ? 1.0f : 3.14159265358979f * x;
}

// pi sinc(x) as numerator / denominator: sin(pi |x|) / |x| (an even
// function), with |x| raised to at least 1e-6: below that sin(pi x) / x
// equals pi to within 2e-12 relative, so x = 0 needs no select.  A product
// of four needs one reciprocal and one factor 1/pi^4, instead of four
// multiplications by pi.
__device__ __forceinline__ void pisinc_nd_hw(float x, float* num, float* den) {
    const float xa = fmaxf(fabsf(x), 1e-6f);
    *num = sinpi_hw(xa);
    *den = xa;
}

// exp(-i pi t) for t = 2 r c given r and c (FP64): the half-cycle count
// h = t / 2 = r c is formed directly, which is the same double as 0.5 (r 2c)
__device__ __forceinline__ float2 phasor_rc_hw(double r, double c) {
    const double h = r * c;
    const float f = (float)(h - rint(h));
    return make_float2(__builtin_amdgcn_cosf(f), -__builtin_amdgcn_sinf(f));
}

// exp(-i pi t) for FP64 t: exact reduction to revolutions in double, then
// the hardware sine and cosine
__device__ __forceinline__ float2 phasor_half_cycles_hw(double t) {
    const double h = 0.5 * t;
    const float r = (float)(h - rint(h));
    return make_float2(__builtin_amdgcn_cosf(r), -__builtin_amdgcn_sinf(r));
}

// exp(-i pi t) for FP64 t: exact reduction in double, FP32 polynomial
__device__ __forceinline__ float2 phasor_half_cycles(double t) {
    const double n = rint(t);
    const float r = (float)(t - n);
    float s, c;
    sincospi_red(r, &s, &c);
    if (((long long)n) & 1) {
        s = -s;
        c = -c;
    }
    return make_float2(c, -s);
}

// splitmix64 hash of (seed, shot, facet): the incoherent phase u in [0,1)
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float incoherent_u(uint64_t seed, int64_t shot, int64_t facet) {
    uint64_t h = mix64(seed ^ mix64((uint64_t)shot * 0x100000001B3ull ^ mix64((uint64_t)facet)));
    return (float)((double)(h >> 40) * (1.0 / 16777216.0));
}

// order-preserving float <-> uint for atomic min/max
__device__ __forceinline__ unsigned f2o(float f) {
    unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(unsigned u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// per-shot bounds of the illuminated facets (for target windows)
struct Bounds {
    unsigned long long itd_min, itd_max;  // double bits (positive)
    unsigned xmin, ymin, zmin, xmax, ymax, zmax;  // f2o-encoded
    // a sphere holding every lit facet: centre (the box's) and radius, or
    // rad < 0 where k_shot_sphere has not run
    float cx, cy, cz, rad;
};

// Arithmetic of the range, bin and phase of a triple (precision study):
// FP64 (the product), FP32 (the reference's own arithmetic,
// radar_kernels.cu:681-686,811-813) or double-float (pairs of floats).
constexpr int PH_F64 = 0, PH_F32 = 1, PH_DF = 2;
// Accumulation of the bin window: FP64 (the product), plain FP32, or FP32
// with the rounding error of every addition recovered (TwoSum on the value
// an atomic returns) and accumulated beside it.
constexpr int ACC_F64 = 0, ACC_F32 = 1, ACC_COMP = 2;

// double-float: an unevaluated sum hi + lo of two floats (Dekker, Knuth).
// Contraction is off inside: an FMA formed from a product and a later sum
// would use the exact product where the error term assumes the rounded one.
struct DF {
    float hi, lo;
};
__device__ __forceinline__ DF df_two_sum(float a, float b) {
#pragma clang fp contract(off)
    const float s = a + b;
    const float bb = s - a;
    return DF{s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ DF df_fast(float a, float b) {
#pragma clang fp contract(off)
    const float s = a + b;
    return DF{s, b - (s - a)};
}
__device__ __forceinline__ DF df_of(double x) {
    const float h = (float)x;
    return DF{h, (float)(x - (double)h)};
}
__device__ __forceinline__ DF df_add(DF a, DF b) {
#pragma clang fp contract(off)
    const DF s = df_two_sum(a.hi, b.hi);
    return df_fast(s.hi, s.lo + (a.lo + b.lo));
}
__device__ __forceinline__ DF df_neg(DF a) { return DF{-a.hi, -a.lo}; }
__device__ __forceinline__ DF df_mul(DF a, DF b) {
#pragma clang fp contract(off)
    const float p = a.hi * b.hi;
    const float e = fmaf(a.hi, b.hi, -p);
    return df_fast(p, e + (a.hi * b.lo + a.lo * b.hi));
}
__device__ __forceinline__ DF df_sqrt(DF a) {
#pragma clang fp contract(off)
    if (!(a.hi > 0.f)) return DF{0.f, 0.f};
    const float x = sqrtf(a.hi);
    const float p = x * x;
    const float e = fmaf(x, x, -p);
    return df_fast(x, ((a.hi - p) - e + a.lo) / (2.f * x));
}
// |a - b| for float a and double-float b components
__device__ __forceinline__ DF df_dist(float ax, float ay, float az, DF bx, DF by, DF bz) {
    const DF dx = df_add(DF{ax, 0.f}, df_neg(bx));
    const DF dy = df_add(DF{ay, 0.f}, df_neg(by));
    const DF dz = df_add(DF{az, 0.f}, df_neg(bz));
    return df_sqrt(df_add(df_add(df_mul(dx, dx), df_mul(dy, dy)), df_mul(dz, dz)));
}

// ---------------------------------------------------------------------------
// generic small kernels
// ---------------------------------------------------------------------------
__global__ void k_zero(float2* __restrict__ p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = make_float2(0.f, 0.f);
}

// pair i of a chunk: the host uploads its shot and its index in that
// shot's target list; the target id is read here from the list the select
// kernel wrote, so the lists never travel to the host
__global__ void k_pair_targets(int* __restrict__ ptgt, const int* __restrict__ pshot,
                               const int* __restrict__ list1, int Tcap, int n,
                               int* __restrict__ next_pair) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ptgt[i] = list1[(size_t)pshot[i] * Tcap + ptgt[i]];
    if (i == 0) *next_pair = 0;  // the target kernel's pair counter
}

__global__ void k_init_bounds(Bounds* b, int S) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) {
        b[s].itd_min = ~0ull; b[s].itd_max = 0ull;
        b[s].xmin = b[s].ymin = b[s].zmin = 0xFFFFFFFFu;
        b[s].xmax = b[s].ymax = b[s].zmax = 0u;
        b[s].cx = b[s].cy = b[s].cz = 0.f;
        b[s].rad = -1.f;
    }
}

// Bounding sphere of each shot's lit facets: centre c = the centre of the
// shot's box, radius = the largest distance from c to a corner of a lit
// facet block's box (every lit facet lies in its block's box).  Each facet
// block (a mesh row or tile) is a thin box, so the sphere is close to the
// lit set; a pair's window takes the nearer of the box and sphere bounds.
// Rows of a multi-track batch: a padding row and blocks past a row's mesh
// have no boxes.
__global__ void __launch_bounds__(256) k_shot_sphere(Bounds* __restrict__ bounds,
                                                     const float4* __restrict__ bbox,
                                                     const int* __restrict__ cnt, int nblk,
                                                     const TrackK* __restrict__ trk,
                                                     const RowK* __restrict__ rows) {
    const int s = blockIdx.x;
    int n = nblk;
    if (trk) {
        if (rows[s].shot < 0) return;
        n = trk[rows[s].trk].nblk;
    }
    Bounds& bd = bounds[s];
    const float cx = 0.5f * (o2f(bd.xmin) + o2f(bd.xmax));
    const float cy = 0.5f * (o2f(bd.ymin) + o2f(bd.ymax));
    const float cz = 0.5f * (o2f(bd.zmin) + o2f(bd.zmax));
    double m = 0.0;
    for (int b = threadIdx.x; b < n; b += 256) {
        const size_t i = (size_t)s * nblk + b;
        if (cnt[i] == 0) continue;
        const float4 lo = bbox[2 * i], hi = bbox[2 * i + 1];
        const double ex = fmax(fabs((double)lo.x - cx), fabs((double)hi.x - cx));
        const double ey = fmax(fabs((double)lo.y - cy), fabs((double)hi.y - cy));
        const double ez = fmax(fabs((double)lo.z - cz), fabs((double)hi.z - cz));
        m = fmax(m, ex * ex + ey * ey + ez * ez);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        bd.cx = cx;
        bd.cy = cy;
        bd.cz = cz;
        // rounded up: the float radius is at least the FP64 one
        bd.rad = (float)sqrt(m) * (1.f + 1e-6f);
    }
}

// ---------------------------------------------------------------------------
// 1. aperture mask + block counts
// ---------------------------------------------------------------------------
// Facet of slot j (j = q NT + threadIdx.x) of block blk.  Blocks are FB
// consecutive facets, or, for a mesh laid out as rows of tile_nx facets
// (tile_nx a multiple of 64 and the row count a multiple of 32), tiles of
// 64 columns x 32 rows: a tile's facets lie close together, so its ranges
// span few bins and the surface window catches them.  Either way a
// bijection onto [0, F); results do not depend on it.
__device__ __forceinline__ int64_t facet_of(int blk, int j, int tile_nx) {
    if (tile_nx == 0) return (int64_t)blk * FB + j;
    const int tpr = tile_nx / 64;
    const int trow = blk / tpr, tcol = blk - trow * tpr;
    return ((int64_t)trow * 32 + (j >> 6)) * tile_nx + (int64_t)tcol * 64 + (j & 63);
}

// sb shots per workgroup (blockIdx.y: shots [sb y, sb y + sb) of the batch):
// each facet is loaded once for all of them.  Multi-track rows each have
// their own mesh and take sb = 1.
constexpr int MASK_SB = 8;
__global__ void __launch_bounds__(NT) k_illum_mask(
    const float* __restrict__ mesh, int64_t Fpad, int64_t F,
    const erp_shot* __restrict__ shots, double cos_ap,
    unsigned long long* __restrict__ mask, int* __restrict__ cnt, int nblk, int tile_nx,
    const TrackK* __restrict__ trk, const RowK* __restrict__ rows, int S, int sb) {
    const int blk = blockIdx.x;
    const int s0 = blockIdx.y * sb, s1 = min(s0 + sb, S);
    if (trk) {
        // multi-track: the row's mesh; blocks past it and padding rows
        // (shot -1) hold nothing
        const TrackK& t = trk[rows[s0].trk];
        if (blk >= t.nblk || rows[s0].shot < 0) return;
        mesh = t.mesh;
        Fpad = t.Fpad;
        F = t.F;
        cos_ap = t.kp.cos_ap;
        tile_nx = t.tile_nx;
    }
    const float cosf_ap = (float)cos_ap;
    const float* X = mesh;
    const float* Y = mesh + Fpad;
    const float* Z = mesh + 2 * Fpad;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __shared__ int wc[MASK_SB][WPB];
    float fx[FPT], fy[FPT], fz[FPT];
    bool in[FPT];
#pragma unroll
    for (int q = 0; q < FPT; ++q) {
        const int64_t i = facet_of(blk, q * NT + threadIdx.x, tile_nx);
        in[q] = i < F;
        fx[q] = in[q] ? X[i] : 0.f;
        fy[q] = in[q] ? Y[i] : 0.f;
        fz[q] = in[q] ? Z[i] : 0.f;
    }
    for (int s = s0; s < s1; ++s) {
        const ShotK sh = load_shot(shots, s);
        const float sxf = (float)sh.sx, syf = (float)sh.sy, szf = (float)sh.sz;
        const float srcmag = fabsf(sxf) + fabsf(syf) + fabsf(szf);
        int total = 0;
#pragma unroll
        for (int q = 0; q < FPT; ++q) {
            bool p = false;
            if (in[q]) {
                const float dx = fx[q] - sxf, dy = fy[q] - syf, dz = fz[q] - szf;
                const float d2 = dx * dx + dy * dy + dz * dz;
                const float rinv = rsqrtf(d2);
                const float c = -(dx * sh.snx + dy * sh.sny + dz * sh.snz) * rinv;
                const float margin =
                    1e-5f + 4e-7f * (srcmag + fabsf(fx[q]) + fabsf(fy[q]) + fabsf(fz[q])) * rinv;
                if (c < cosf_ap - margin) {
                    p = false;
                } else if (c > cosf_ap + margin && c < 1.0f - margin) {
                    p = true;
                } else {
                    p = illum_exact(fx[q], fy[q], fz[q], sh, cos_ap);
                }
            }
            const unsigned long long b = __ballot(p);
            if (lane == 0) mask[((size_t)s * nblk + blk) * MASKW + q * WPB + wv] = b;
            total += __popcll(b);
        }
        if (lane == 0) wc[s - s0][wv] = total;
    }
    __syncthreads();
    if (threadIdx.x < s1 - s0) {
        int t = 0;
        for (int w = 0; w < WPB; ++w) t += wc[threadIdx.x][w];
        cnt[(size_t)(s0 + threadIdx.x) * nblk + blk] = t;
    }
}

// ---------------------------------------------------------------------------
// 2. per-shot exclusive scan of block counts
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_scan_blocks(const int* __restrict__ cnt,
                                                      int* __restrict__ off,
                                                      int* __restrict__ V, int nblk,
                                                      const TrackK* __restrict__ trk,
                                                      const RowK* __restrict__ rows) {
    const int s = blockIdx.x;
    // rows of a multi-track batch: stride nblk, the row's mesh's own count;
    // a padding row (shot -1) illuminates nothing
    if (trk && rows[s].shot < 0) {
        if (threadIdx.x == 0) V[s] = 0;
        return;
    }
    const int n = trk ? trk[rows[s].trk].nblk : nblk;
    __shared__ int buf[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = (i < n) ? cnt[(size_t)s * nblk + i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            int t = (threadIdx.x >= o) ? buf[threadIdx.x - o] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < n) off[(size_t)s * nblk + i] = carry + buf[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) V[s] = carry;
}

// ---------------------------------------------------------------------------
// 3. fused surface kernel
// ---------------------------------------------------------------------------
// Compacted facet record (SoA over Vtot records of the batch).
struct Rec {
    float* geo;     // 12 x cap: x,y,z,nx,ny,nz,ux,uy,uz,vx,vy,vz
    double* itd;    // cap
    float* F;       // cap: sqrt(P n Grefr G1 [fRfrC]) lam^2 / (K Itd)
    float* fRfrC;   // cap
    float* ang;     // 3 x cap: Ith, Iph, Rth
    // the facet record packed for k_target_accum_fast (52 bytes in four
    // loads, plus Itd from itd above): qa = (x, y, z, F),
    // qb = (nx, ny, nz, ux), qc = (uy, uz, vx, vy), qd = vz
    float4* qa;
    float4* qb;
    float4* qc;
    float* qd;
    int64_t cap;
};

// Facet records of a block of two consecutive shots for the shot-pair target
// kernel: the facets that either shot illuminates (the union of the two
// masks), laid out per block as rec_slot lays out one shot's, with the
// geometry once and the shot-dependent source leg per shot: Itd, and the
// gain F Dmax / pi^4 times the source leg's phasor exp(-i 2 pi Itd / lam')
// (lam' = 1 / inv_cyc), so that a term's phasor is this one times the
// target leg's.
struct BRec {
    float4* qa;    // (x, y, z, 0)
    float4* qb;    // (nx, ny, nz, ux)
    float4* qc;    // (uy, uz, vx, vy)
    float* qd;     // vz
    double2* itd;  // Itd of the block's first and second shot; -1: not lit by it
    float4* src;   // gain x source phasor of the first shot (x, y), the second (z, w)
};
// Where k_surface writes block records: rec.qa null = none.  umask / uoff:
// the union masks and their exclusive block offsets per block of shots,
// U its size, ubase its first record.
struct BlkRec {
    BRec rec;
    const unsigned long long* umask;
    const int* uoff;
    const int* U;
    const int64_t* ubase;
};

__device__ __forceinline__ void lds_add2(double2* win, int idx, float2 v) {
    atomicAdd(&win[idx].x, (double)v.x);
    atomicAdd(&win[idx].y, (double)v.y);
}
__device__ __forceinline__ void glb_add2(float2* row, int idx, float2 v) {
    atomicAdd(&row[idx].x, v.x);
    atomicAdd(&row[idx].y, v.y);
}
__device__ __forceinline__ void glb_add2d(double2* row, int idx, double x, double y) {
    atomicAdd(&row[idx].x, x);
    atomicAdd(&row[idx].y, y);
}

// Deterministic accumulation: every term is rounded once to a signed 64-bit
// fixed-point value v * 2^E and summed with integer atomics, which are exact
// and so independent of the order in which threads and workgroups arrive.
// E is chosen from a bound M >= |component| of every term and the number of
// terms n that can reach one bin: |sum| * 2^E <= 2 n M 2^E <= 2^61.
__device__ __forceinline__ double det_scale(double n, double M) {
    if (!(M > 0.0)) return 0.0;
    int e;
    frexp(2.0 * n * M, &e);  // 2 n M < 2^e
    return ldexp(1.0, 61 - e);
}
__device__ __forceinline__ unsigned long long det_fix(float v, double scale) {
    return (unsigned long long)__double2ll_rn((double)v * scale);
}
// LDS windows keep real and imaginary parts in separate arrays: 8-byte
// elements at an 8-byte stride spread a wave's bins over every bank pair
// (an interleaved 16-byte complex uses only every other pair)
struct WinS {
    double* x;
    double* y;
};
__device__ __forceinline__ void lds_add2s(WinS w, int idx, float2 v) {
    atomicAdd(&w.x[idx], (double)v.x);
    atomicAdd(&w.y[idx], (double)v.y);
}
__device__ __forceinline__ void lds_add2_fix(unsigned long long* wx, unsigned long long* wy,
                                             int idx, float2 v, double scale) {
    atomicAdd(&wx[idx], det_fix(v.x, scale));
    atomicAdd(&wy[idx], det_fix(v.y, scale));
}
// det_fix for |v scale| < 2^51: v scale is exact (scale is a power of two),
// and adding 1.5 * 2^52 rounds it to the nearest integer, ties to even (as
// __double2ll_rn), held in the low mantissa bits: the same value in three
// instructions instead of a full double -> int64 conversion
__device__ __forceinline__ unsigned long long det_fix51(float v, double scale) {
    const double magic = 6755399441055744.0;  // 1.5 * 2^52
    const double x = (double)v * scale + magic;
    return (unsigned long long)(__double_as_longlong(x) - __double_as_longlong(magic));
}
__device__ __forceinline__ void lds_add2_fix51(unsigned long long* wx, unsigned long long* wy,
                                               int idx, float2 v, double scale) {
    atomicAdd(&wx[idx], det_fix51(v.x, scale));
    atomicAdd(&wy[idx], det_fix51(v.y, scale));
}
// FP32 window add; ACC_COMP also accumulates the rounding error of the
// addition the atomic performed (TwoSum of the returned old value and v)
template <int ACC>
__device__ __forceinline__ void lds_add_f32(float* w, int idx, float v, int stride2) {
    if constexpr (ACC == ACC_COMP) {
        const float old = atomicAdd(&w[idx], v);
        const float s = old + v;
        const float bb = s - old;
        const float e = (old - (s - bb)) + (v - bb);
        atomicAdd(&w[stride2 + idx], e);
    } else {
        atomicAdd(&w[idx], v);
    }
}
template <int ACC>
__device__ __forceinline__ void win_add(WinS w, int idx, float2 v) {
    if constexpr (ACC == ACC_F64) {
        lds_add2s(w, idx, v);
    } else {
        // FP32 layout inside the same LDS: x at [0, n), y at [n, 2n),
        // compensations at [2n, 4n) (n = the window's FP64 capacity)
        float* f = reinterpret_cast<float*>(w.x);
        const int n = (int)(w.y - w.x);
        lds_add_f32<ACC>(f, idx, v.x, 2 * n);
        lds_add_f32<ACC>(f + n, idx, v.y, 2 * n);
    }
}
template <int ACC>
__device__ __forceinline__ void win_zero(WinS w, int idx) {
    if constexpr (ACC == ACC_F64) {
        w.x[idx] = 0.0;
        w.y[idx] = 0.0;
    } else {
        float* f = reinterpret_cast<float*>(w.x);
        const int n = (int)(w.y - w.x);
        f[idx] = 0.f;
        f[n + idx] = 0.f;
        if constexpr (ACC == ACC_COMP) {
            f[2 * n + idx] = 0.f;
            f[3 * n + idx] = 0.f;
        }
    }
}
template <int ACC>
__device__ __forceinline__ float2 win_get(WinS w, int idx) {
    if constexpr (ACC == ACC_F64) {
        return make_float2((float)w.x[idx], (float)w.y[idx]);
    } else {
        const float* f = reinterpret_cast<const float*>(w.x);
        const int n = (int)(w.y - w.x);
        if constexpr (ACC == ACC_COMP)
            return make_float2(f[idx] + f[2 * n + idx], f[n + idx] + f[3 * n + idx]);
        return make_float2(f[idx], f[n + idx]);
    }
}

__device__ __forceinline__ void glb_add2_fix(unsigned long long* row, int idx, float2 v,
                                             double scale) {
    atomicAdd(&row[2 * idx], det_fix(v.x, scale));
    atomicAdd(&row[2 * idx + 1], det_fix(v.y, scale));
}
__device__ __forceinline__ double det_unfix(unsigned long long u, double inv_scale) {
    return (double)(long long)u * inv_scale;
}

// Scale of a shot's fixed-point type-0 row: V facets times n0 type-0
// targets terms, each at most the largest component of any pair (bits m),
// with 1e-4 of margin
__device__ __forceinline__ double t0_det_scale(int V, int n0, unsigned m) {
    return det_scale((double)V * (double)max(n0, 1), 1.0001 * (double)__uint_as_float(m));
}

// Record slot of the illuminated facet of rank i (facet order) among the V
// of its shot.  The ranks are laid out as 64 rows of equal i mod 64, so
// consecutive slots hold facets 64 ranks apart.  Neighbouring facets have
// nearly equal ranges, and the target kernels, whose lanes read consecutive
// slots, would otherwise add into the same few LDS bins in one instruction.
// Every consumer sums over all of a shot's records, so the order is free.
__device__ __forceinline__ int rec_slot(int i, int V) {
    constexpr int G = 64;
    const int q = V / G, rem = V - q * G;
    const int a = i % G, c = i / G;
    return (a < rem ? a * (q + 1) : rem * (q + 1) + (a - rem) * q) + c;
}

__global__ void __launch_bounds__(NT) k_surface(
    const float* __restrict__ mesh, int64_t Fpad, int64_t F, const erp_shot* __restrict__ shots,
    const unsigned long long* __restrict__ mask, const int* __restrict__ off,
    const int* __restrict__ Vs, const int64_t* __restrict__ rec_base, int nblk, KPar kp,
    Rec rec, int write_rec, int do_trace, double2* __restrict__ srow, int64_t shot0,
    Bounds* __restrict__ bounds, int det, unsigned* __restrict__ smax, int tile_nx,
    int prec_ph, int prec_acc, const TrackK* __restrict__ trk, const RowK* __restrict__ rows,
    BlkRec brk, float4* __restrict__ bbox) {
    // det 0: FP64 bins (LDS window, FP64 atomics into srow)
    // det 1: first deterministic pass, largest |component| per shot into smax
    // det 2: second deterministic pass, fixed-point bins (srow holds int64)
    // prec_ph, prec_acc (precision study, det 0, non-specular): the trace's range, bin
    // and phase in FP64 / FP32 (the reference's surfacePT arithmetic) /
    // double-float, and its LDS window sums in FP64 / FP32 / compensated
    // FP32 (the window partial sums enter the FP64 row as they are); the
    // facet records for the target kernels always keep the FP64 range
    const int blk = blockIdx.x, s = blockIdx.y;
    // the shot's index within its track (the incoherent phase hash)
    int64_t shot_id = shot0 + s;
    if (trk) {
        // multi-track: the row's mesh and parameters; blocks past its mesh
        // and padding rows (shot -1) hold nothing; a track without targets
        // keeps no facet records
        const TrackK& t = trk[rows[s].trk];
        if (blk >= t.nblk || rows[s].shot < 0) return;
        mesh = t.mesh;
        Fpad = t.Fpad;
        F = t.F;
        kp = t.kp;
        tile_nx = t.tile_nx;
        shot_id = rows[s].shot;
        if (t.T == 0) write_rec = 0;
    }
    const unsigned long long* mw = mask + ((size_t)s * nblk + blk) * MASKW;
    __shared__ unsigned long long m_s[MASKW];
    __shared__ int pre_s[MASKW + 1];
    __shared__ double swin[2 * SWIN];
    __shared__ int wmin_s;
    __shared__ float red_f[6][WPB];
    __shared__ double red_d[2][WPB];
    // block records: the union mask of the shot's block of two and the
    // prefix of its words
    __shared__ unsigned long long um_s[MASKW];
    __shared__ int upre_s[MASKW];
    const bool blk_rec = write_rec && brk.rec.qa;
    if (threadIdx.x < MASKW) {
        m_s[threadIdx.x] = mw[threadIdx.x];
        if (blk_rec)
            um_s[threadIdx.x] = brk.umask[((size_t)(s >> 1) * nblk + blk) * MASKW + threadIdx.x];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < MASKW; ++w) {
            pre_s[w] = t;
            t += __popcll(m_s[w]);
        }
        pre_s[MASKW] = t;
        wmin_s = 0x7FFFFFFF;
        if (blk_rec) {
            int u = 0;
            for (int w = 0; w < MASKW; ++w) {
                upre_s[w] = u;
                u += __popcll(um_s[w]);
            }
        }
    }
    __syncthreads();
    if (pre_s[MASKW] == 0) return;  // dark block (uniform exit)

    const ShotK sh = load_shot(shots, s);
    const int V = Vs[s];
    const int64_t rbase = rec_base[s];
    const int rank0 = off[(size_t)s * nblk + blk];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const bool spec = kp.mode & ERP_MODE_SPECULAR;
    const bool lossless = kp.mode & ERP_MODE_LOSSLESS;
    const bool incoh = kp.mode & ERP_MODE_INCOHERENT;
    const double rst = (double)sh.rst;
    const float Grefr = sh.grefr;

    // per-thread contributions kept for the windowed scatter
    int bq[FPT];
    float wq[FPT];
    float2 cq[FPT];
    float lxmin = INFINITY, lymin = INFINITY, lzmin = INFINITY;
    float lxmax = -INFINITY, lymax = -INFINITY, lzmax = -INFINITY;
    double litmin = INFINITY, litmax = 0.0;
    int bmin = 0x7FFFFFFF;

#pragma unroll
    for (int q = 0; q < FPT; ++q) {
        bq[q] = -1;
        const int w = q * WPB + wv;
        const unsigned long long m = m_s[w];
        if (!((m >> lane) & 1ull)) continue;
        const int64_t i = facet_of(blk, q * NT + threadIdx.x, tile_nx);
        const int rank = pre_s[w] + __popcll(m & ((1ull << lane) - 1ull));
        float g[12];
#pragma unroll
        for (int a = 0; a < 12; ++a) g[a] = mesh[(size_t)a * Fpad + i];
        // compIncidentRays (FP64 range, FP32 directions)
        const double dx = (double)g[0] - sh.sx, dy = (double)g[1] - sh.sy,
                     dz = (double)g[2] - sh.sz;
        const double Itd = dist3(dx, dy, dz);
        const double invd = 1.0 / Itd;
        const double ix = dx * invd, iy = dy * invd, iz = dz * invd;
        const float Ix = (float)ix, Iy = (float)iy, Iz = (float)iz;
        const float2 csI = cos_sin_ref(ix, iy, iz, g[3], g[4], g[5]);
        const float cI = csI.x, sI = csI.y;
        float Iu = Ix * g[6] + Iy * g[7] + Iz * g[8];
        float Iv = Ix * g[9] + Iy * g[10] + Iz * g[11];
        if (!kp.ortho) {
            const float2 e = uv_dots_exact(ix, iy, iz, g[6], g[7], g[8], g[9], g[10], g[11]);
            Iu = e.x;
            Iv = e.y;
        }
        // compRefractedRays (snellsLaw): past the critical angle the
        // reference sets Rth = 1e4 and uses cos(1e4); Rth itself (and the
        // angles Ith, Iph) are only stored for the kernels that read them
        const float sR = kp.sqrt_e1e2 * sI;
        const bool tir = sR > 1.0f;
        const float cR = tir ? cosf(1e4f) : sqrtf(fmaxf(0.f, 1.f - sR * sR));
        // compReflectedEnergy (pol 0, H)
        const float rho = (kp.nu2 * cI - kp.nu1 * cR) / (kp.nu2 * cI + kp.nu1 * cR);
        const float Itdf = (float)Itd;
        const float atm = expf(-4.0f * kp.alpha1 * Itdf);
        float fRfrC = tir ? 0.f : (1.f - rho * rho);
        fRfrC *= atm;
        const float kc = kp.ks * cI;
        const float rough = expf(-4.f * (kc * kc));
        float fRe = rho * rho * atm * rough;
        // facet_G(Ith, Iph) (radar_kernels.cu:658-677)
        // facet_G(Ith, Iph): sin(Ith)cos(Iph) = I.v and sin(Ith)sin(Iph) = I.u for the
        // orthonormal facet basis (u, v, n)
        const float2 scI = sc_angles(Iu, Iv, sI);
        const float s1 = sinc_f(kp.piLlam * scI.x), s2 = sinc_f(kp.piLlam * scI.y);
        const float G1 = (s1 * s2) * (s1 * s2) * kp.Dmax;
        if (spec) {
            // facetReradiation(Itd, 2 Ith, -Iph) * radarEq(P, Grefl, fs, lam, Itd, V);
            // sin(2 Ith) cos(-Iph) = 2 cos(Ith) I.v, sin(2 Ith) sin(-Iph) = -2 cos(Ith) I.u
            const float r1 = sinc_f(kp.fs_lam * 2.f * cI * scI.x),
                        r2 = sinc_f(kp.fs_lam * -2.f * cI * scI.y);
            const float rr = kp.fs2_lam * r1 * r2;
            const float Gl = sh.grefl;
            const double Itd2 = Itd * Itd;
            const float req = (float)((double)V *
                                      ((double)(kp.P * Gl * Gl * kp.fs * kp.fs * kp.fs * kp.fs) /
                                       (ERP_FOURPI5_SQ * Itd2 * Itd2)));
            fRe *= rr * rr * req;
        }
        // bounds for the target windows
        lxmin = fminf(lxmin, g[0]); lxmax = fmaxf(lxmax, g[0]);
        lymin = fminf(lymin, g[1]); lymax = fmaxf(lymax, g[1]);
        lzmin = fminf(lzmin, g[2]); lzmax = fmaxf(lzmax, g[2]);
        litmin = fmin(litmin, Itd); litmax = fmax(litmax, Itd);

        if (write_rec) {
            const int64_t r = rbase + rec_slot(rank0 + rank, V);
            if (rec.itd) rec.itd[r] = Itd;
            // target amplitude factor: sqrt(P n Grefr G1 fRfrC) lam^2/(K Itd)
            const float fac = lossless ? 1.f : fRfrC;
            // NoConv specular keeps fRe here for reflRadarSignal's direct sum
            const float Fr = (spec && !(kp.mode & ERP_MODE_CONVOLUTION))
                                 ? fRe
                                 : sqrtf(kp.P * kp.nf * Grefr * G1 * fac) * kp.lam * kp.lam *
                                       kp.invK / Itdf;
            if (blk_rec) {
                // the facet's slot among its block's union (both shots of the
                // block that light it write the same geometry)
                const int j = s >> 1, k = s & 1;
                const int ur = brk.uoff[(size_t)j * nblk + blk] + upre_s[w] +
                               __popcll(um_s[w] & ((1ull << lane) - 1ull));
                const int64_t u = brk.ubase[j] + rec_slot(ur, brk.U[j]);
                brk.rec.qa[u] = make_float4(g[0], g[1], g[2], 0.f);
                brk.rec.qb[u] = make_float4(g[3], g[4], g[5], g[6]);
                brk.rec.qc[u] = make_float4(g[7], g[8], g[9], g[10]);
                brk.rec.qd[u] = g[11];
                reinterpret_cast<double*>(brk.rec.itd + u)[k] = Itd;
                const float gain = Fr * (kp.Dmax * 0.0102659823f);
                const float2 ph = phasor_rc_hw(Itd, kp.inv_cyc);
                reinterpret_cast<float2*>(brk.rec.src + u)[k] = make_float2(ph.x * gain,
                                                                            ph.y * gain);
            }
            if (rec.geo) {
#pragma unroll
                for (int a = 0; a < 12; ++a) rec.geo[(size_t)a * rec.cap + r] = g[a];
                rec.F[r] = Fr;
                rec.fRfrC[r] = fRfrC;
                rec.ang[r] = atan2f(sI, cI);                 // Ith
                rec.ang[rec.cap + r] = atan2f(Iu, Iv);       // Iph
                rec.ang[2 * rec.cap + r] = tir ? 1e4f : asinf(sR);  // Rth
            }
            if (rec.qa) {
                // F Dmax / pi^4: the gain factor of the target kernel's
                // amplitude, with 1/pi^4 taking out the four factors pi of
                // its pi-sincs
                rec.qa[r] = make_float4(g[0], g[1], g[2], Fr * (kp.Dmax * 0.0102659823f));
                rec.qb[r] = make_float4(g[3], g[4], g[5], g[6]);
                rec.qc[r] = make_float4(g[7], g[8], g[9], g[10]);
                rec.qd[r] = g[11];
            }
        }
        if (!do_trace) continue;
        int b;
        float wgt;
        if (!spec) {
            // surfacePT: sqrt(friis(friis(P,Gr,G,lam,R),G,Gr,lam,R) [* fRe])
            float2 pst = make_float2(1.f, 0.f);
            if (prec_ph == PH_F32) {
                // Itd, the bin coordinate and phasor(2 Itd) in FP32
                // (compIncidentRays, radar_kernels.cu:681-686, 733-737)
                const float ax = g[0] - (float)sh.sx, ay = g[1] - (float)sh.sy,
                            az = g[2] - (float)sh.sz;
                const float I32 = sqrtf(ax * ax + ay * ay + az * az);
                const float x32 = (I32 - sh.rst) / kp.dr;
                if (!(x32 > -1.f) || !(x32 < (float)kp.nr)) continue;
                b = (int)x32;
                wgt = x32 - (float)b;
                float sn, cs;
                sincosf(-(2.0f * ERP_PI_F / kp.lam) * (2.f * I32), &sn, &cs);
                pst = make_float2(cs, sn);
            } else if (prec_ph == PH_DF) {
                const DF Id = df_dist(g[0], g[1], g[2], df_of(sh.sx), df_of(sh.sy), df_of(sh.sz));
                const DF xd = df_mul(df_add(Id, df_neg(df_of(rst))), df_of(kp.inv_dr));
                const double xq = (double)xd.hi + (double)xd.lo;
                if (!(xq > -1.0) || !(xq < (double)kp.nr)) continue;
                b = (int)xq;
                wgt = (float)(xq - (double)b);
                // exp(-i pi t), t = 2 Itd 2 inv_cyc half cycles, reduced in double-float
                const DF t = df_mul(Id, df_of(2.0 * kp.inv_cyc));
                const float rr = rintf(t.hi);
                const float fr = (t.hi - rr) + t.lo;
                pst = make_float2(__builtin_amdgcn_cosf(fr), -__builtin_amdgcn_sinf(fr));
            } else if (!bin_of(Itd, rst, kp.drd, kp.nr, &b, &wgt)) {
                continue;
            }
            float amp = kp.sqrtP * kp.lam * kp.lam * Grefr * G1 * kp.invK / (Itdf * Itdf);
            if (!lossless) amp *= sqrtf(fRe);
            float2 ph;
            if (incoh) {
                float sn, cs;
                sincospif(2.0f * incoherent_u(kp.seed, shot_id, i), &sn, &cs);
                ph = make_float2(cs, sn);
            } else if (prec_ph != PH_F64) {
                ph = pst;
            } else {
                ph = phasor_cycles(2.0 * Itd * kp.inv_cyc);
            }
            bq[q] = b;
            wq[q] = wgt;
            cq[q] = make_float2(ph.x * amp, ph.y * amp);
        } else {
            // genReflPhasor: fRe * exp(+i 4 pi/lam Itd), one bin, no split
            if (!bin_of(Itd, rst, kp.drd, kp.nr, &b, &wgt)) continue;
            const float2 ph = phasor_cycles(-2.0 * Itd * kp.inv_cyc);
            bq[q] = b;
            wq[q] = 0.f;
            cq[q] = make_float2(ph.x * fRe, ph.y * fRe);
        }
        bmin = min(bmin, b);
    }

    if (det == 1 && do_trace) {
        float m = 0.f;
#pragma unroll
        for (int q = 0; q < FPT; ++q)
            if (bq[q] >= 0) m = fmaxf(m, fmaxf(fabsf(cq[q].x), fabsf(cq[q].y)));
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        if (lane == 0 && m > 0.f) atomicMax(&smax[s], __float_as_uint(m));
    }
    // block reductions: window base and bounds
    for (int o = 32; o > 0; o >>= 1) {
        bmin = min(bmin, __shfl_xor(bmin, o));
        lxmin = fminf(lxmin, __shfl_xor(lxmin, o)); lxmax = fmaxf(lxmax, __shfl_xor(lxmax, o));
        lymin = fminf(lymin, __shfl_xor(lymin, o)); lymax = fmaxf(lymax, __shfl_xor(lymax, o));
        lzmin = fminf(lzmin, __shfl_xor(lzmin, o)); lzmax = fmaxf(lzmax, __shfl_xor(lzmax, o));
        litmin = fmin(litmin, __shfl_xor(litmin, o)); litmax = fmax(litmax, __shfl_xor(litmax, o));
    }
    if (lane == 0) {
        atomicMin(&wmin_s, bmin);
        red_f[0][wv] = lxmin; red_f[1][wv] = lymin; red_f[2][wv] = lzmin;
        red_f[3][wv] = lxmax; red_f[4][wv] = lymax; red_f[5][wv] = lzmax;
        red_d[0][wv] = litmin; red_d[1][wv] = litmax;
    }
    // (FP64 window, or the FP32 layout of the study's sums in the same LDS)
    for (int k = threadIdx.x; k < 2 * SWIN; k += NT) swin[k] = 0.0;
    __syncthreads();
    if (threadIdx.x == 0 && bounds) {
        float a[6];
        double d0 = red_d[0][0], d1 = red_d[1][0];
        for (int k = 0; k < 6; ++k) a[k] = red_f[k][0];
        for (int w = 1; w < WPB; ++w) {
            a[0] = fminf(a[0], red_f[0][w]); a[1] = fminf(a[1], red_f[1][w]);
            a[2] = fminf(a[2], red_f[2][w]); a[3] = fmaxf(a[3], red_f[3][w]);
            a[4] = fmaxf(a[4], red_f[4][w]); a[5] = fmaxf(a[5], red_f[5][w]);
            d0 = fmin(d0, red_d[0][w]); d1 = fmax(d1, red_d[1][w]);
        }
        Bounds* bd = bounds + s;
        if (bbox) {
            // this block's box, for the shot's bounding sphere (k_shot_sphere)
            float4* q = bbox + 2 * ((size_t)s * nblk + blk);
            q[0] = make_float4(a[0], a[1], a[2], 0.f);
            q[1] = make_float4(a[3], a[4], a[5], 0.f);
        }
        atomicMin(&bd->xmin, f2o(a[0])); atomicMin(&bd->ymin, f2o(a[1]));
        atomicMin(&bd->zmin, f2o(a[2])); atomicMax(&bd->xmax, f2o(a[3]));
        atomicMax(&bd->ymax, f2o(a[4])); atomicMax(&bd->zmax, f2o(a[5]));
        atomicMin(&bd->itd_min, (unsigned long long)__double_as_longlong(d0));
        atomicMax(&bd->itd_max, (unsigned long long)__double_as_longlong(d1));
    }
    if (!do_trace || det == 1) return;
    const int base = wmin_s;
    if (base == 0x7FFFFFFF) return;
    const int nr = kp.nr;
    double2* row = srow + (size_t)s * nr;
    if (det == 2) {
        // the bound covers every term of the shot: 2 terms per facet, V facets
        const double scale = det_scale((double)V, (double)__uint_as_float(smax[s]));
        unsigned long long* wux = reinterpret_cast<unsigned long long*>(swin);
        unsigned long long* wuy = wux + SWIN;
        unsigned long long* ru = reinterpret_cast<unsigned long long*>(row);
#pragma unroll
        for (int q = 0; q < FPT; ++q) {
            const int b = bq[q];
            if (b < 0) continue;
            const float w = wq[q];
            const float2 c = cq[q];
            const float2 c0 = make_float2(c.x * (1.f - w), c.y * (1.f - w));
            const int k0 = b - base;
            if (k0 < SWIN) lds_add2_fix(wux, wuy, k0, c0, scale);
            else glb_add2_fix(ru, b, c0, scale);
            if (!spec && b + 1 < nr) {
                const float2 c1 = make_float2(c.x * w, c.y * w);
                if (k0 + 1 < SWIN) lds_add2_fix(wux, wuy, k0 + 1, c1, scale);
                else glb_add2_fix(ru, b + 1, c1, scale);
            }
        }
        __syncthreads();
        for (int k = threadIdx.x; k < SWIN; k += NT) {
            const int b = base + k;
            const unsigned long long x = wux[k], y = wuy[k];
            if (b < nr && (x | y)) {
                atomicAdd(&ru[2 * b], x);
                atomicAdd(&ru[2 * b + 1], y);
            }
        }
        return;
    }
    const WinS W{swin, swin + SWIN};
    // one window add of the study's arithmetic (a uniform choice)
    auto wadd = [&](int k, float2 v) {
        if (prec_acc == ACC_F32) win_add<ACC_F32>(W, k, v);
        else if (prec_acc == ACC_COMP) win_add<ACC_COMP>(W, k, v);
        else lds_add2s(W, k, v);
    };
#pragma unroll
    for (int q = 0; q < FPT; ++q) {
        const int b = bq[q];
        if (b < 0) continue;
        const float w = wq[q];
        const float2 c = cq[q];
        const float2 c0 = make_float2(c.x * (1.f - w), c.y * (1.f - w));
        const int k0 = b - base;
        if (k0 < SWIN) wadd(k0, c0); else glb_add2d(row, b, c0.x, c0.y);
        if (!spec && b + 1 < nr) {
            const float2 c1 = make_float2(c.x * w, c.y * w);
            if (k0 + 1 < SWIN) wadd(k0 + 1, c1); else glb_add2d(row, b + 1, c1.x, c1.y);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < SWIN; k += NT) {
        double vx, vy;
        if (prec_acc == ACC_F64) {
            vx = swin[k];
            vy = swin[SWIN + k];
        } else {
            const float2 v = prec_acc == ACC_F32 ? win_get<ACC_F32>(W, k) : win_get<ACC_COMP>(W, k);
            vx = v.x;
            vy = v.y;
        }
        const int b = base + k;
        if (b < nr && (vx != 0.0 || vy != 0.0)) glb_add2d(row, b, vx, vy);
    }
}

// deterministic type-0 rows: the fixed-point (int64) pairs of each shot's
// row (the shot's scale, t0_det_scale) -> FP32 rows of P, zero from nr on
__global__ void k_fix_rows(float2* __restrict__ dst, const unsigned long long* __restrict__ src,
                           int P, int nr, int rows, const int* __restrict__ Vs,
                           const int* __restrict__ tcnt, const unsigned* __restrict__ tmax) {
    const size_t n = (size_t)P * rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / P), k = (int)(i - (size_t)r * P);
        float2 v = make_float2(0.f, 0.f);
        if (k < nr) {
            const double sc = t0_det_scale(Vs[r], tcnt[2 * r + 1], tmax[r]);
            const double inv = sc > 0.0 ? 1.0 / sc : 0.0;
            const unsigned long long* q = src + (size_t)r * P;
            v = make_float2((float)det_unfix(q[2 * k], inv), (float)det_unfix(q[2 * k + 1], inv));
        }
        dst[i] = v;
    }
}

// surface rows (FP64 or fixed point, stride nr) -> FP32 rows of the FFT
// buffer (stride P); bins nr..P-1 are left as they are (zero)
__global__ void k_surf_rows(float2* __restrict__ psurf, int P, const double2* __restrict__ srow,
                            int nr, int S, int det, const unsigned* __restrict__ smax,
                            const int* __restrict__ Vs) {
    const int s = blockIdx.y;
    if (s >= S) return;
    const double2* r = srow + (size_t)s * nr;
    double inv = 0.0;
    if (det) {
        const double sc = det_scale((double)Vs[s], (double)__uint_as_float(smax[s]));
        inv = sc > 0.0 ? 1.0 / sc : 0.0;
    }
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nr; b += gridDim.x * blockDim.x) {
        const double2 v = r[b];
        float2 o;
        if (det) {
            o = make_float2((float)det_unfix(__double_as_longlong(v.x), inv),
                            (float)det_unfix(__double_as_longlong(v.y), inv));
        } else {
            o = make_float2((float)v.x, (float)v.y);
        }
        psurf[(size_t)s * P + b] = o;
    }
}

// ---------------------------------------------------------------------------
// 4. pointwise spectral kernels
// ---------------------------------------------------------------------------
// dst[r][k] = src[r][k] * spec[k]
__global__ void k_mul_spec(float2* __restrict__ dst, const float2* __restrict__ src,
                           const float2* __restrict__ spec, int P, int rows) {
    const size_t n = (size_t)P * rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        dst[i] = cmul(src[i], spec[i & (size_t)(P - 1)]);
    }
}

// out[r][j] = scale * y[r][j + nr] for j < nr (window of the exact linear
// convolution); if pad, also zero out[r][nr .. P).
__global__ void k_window(float2* __restrict__ out, int ostride, const float2* __restrict__ y,
                         int P, int nr, int rows, float scale, int pad) {
    const int lim = pad ? P : nr;
    const size_t n = (size_t)lim * rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / lim), j = (int)(i - (size_t)r * lim);
        float2 v = make_float2(0.f, 0.f);
        if (j < nr) {
            v = y[(size_t)r * P + j + nr];
            v.x *= scale;
            v.y *= scale;
        }
        out[(size_t)r * ostride + j] = v;
    }
}

// out[r][j] = in[r][j], j < ncols
__global__ void k_copy_rows(float2* __restrict__ out, int ostride, const float2* __restrict__ in,
                            int istride, int ncols, int rows) {
    const size_t n = (size_t)ncols * rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / ncols), j = (int)(i - (size_t)r * ncols);
        out[(size_t)r * ostride + j] = in[(size_t)r * istride + j];
    }
}

// ---------------------------------------------------------------------------
// 5. chirp tables
// ---------------------------------------------------------------------------
// Rearranged kernel g of the reference convolutions, placed at e = 1 .. 2nr-1
// (e' = e - 1) so that the exact linear convolution's range window starts at
// index nr (a row boundary of the two-pass FFT layout):
//  kind 0 (convolvePhasorChirpLinear with genCenteredChirpPadded):
//     g[e] = h[(e' + nr/2 - nr + 1) mod 2nr],  h[i] = sinc((i - nr/2) dr / res)
//  kind 1 (convolvePhasorChirp with genChirp):
//     g[e] = h[(e' + 1 - nr) mod nr],  h[i] = sinc((r-rst)/res) + sinc((r-rst-nr dr)/res)
// out[j] = sum_m x[m] g[j + nr - m] for j < nr: no wrap-around for P >= 2nr.
__global__ void k_gen_kernel(float2* __restrict__ g, int P, int nr, double dr, double res,
                             double rst, int kind) {
    const int e0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (e0 >= P) return;
    float v = 0.f;
    const int e = e0 - 1;
    if (kind == 0 || kind == 2) {
        // output bin j = input bin i + d sits at y[ws + j]: g[ws + d] for
        // |d| < nr, the 2nr-circular, kc-cropped kernel of the reference;
        // ws = nr (kind 0) or P / 2 (kind 2, the cropped per-pair chain)
        const int ws = kind == 0 ? nr : P / 2;
        const int d = e0 - ws;
        if (d > -nr && d < nr) {
            const int kc = nr / 2;
            const int m = (int)(((long long)d + kc + 2LL * nr) % (2LL * nr));
            const double x = ((double)(m - kc) * dr) / res;
            v = (fabs(x) < 1e-6) ? 1.f : (float)(sin(M_PI * x) / (M_PI * x));
        }
    } else if (e >= 0 && e < 2 * nr - 1) {
        // kind 1: the nr-circular kernel of one receive window start
        const int m = (int)(((long long)e + 1 - nr + 2LL * nr) % (long long)nr);
        const double r = rst + (double)m * dr;
        const double x1 = (r - rst) / res;
        const double x2 = (r - (rst + (double)nr * dr)) / res;
        const double a = (fabs(x1) < 1e-6) ? 1.0 : sin(M_PI * x1) / (M_PI * x1);
        const double b = (fabs(x2) < 1e-6) ? 1.0 : sin(M_PI * x2) / (M_PI * x2);
        v = (float)(a + b);
    }
    g[e0] = make_float2(v, 0.f);
}

__global__ void k_scale(float2* __restrict__ p, int n, float s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        p[i].x *= s;
        p[i].y *= s;
    }
}

// Bluestein tables for DFT_L: w[n] = exp(-i pi n^2 / L); b[m] = conj(w[|m|])
// circularly in M; crop mask of cropSpectrum (elementary_kernels.cu:258-278).
__global__ void k_blue_tables(float2* __restrict__ w, float2* __restrict__ b,
                              unsigned char* __restrict__ mask, int L, int M, double smpl,
                              double Bw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < L) {
        const long long q = ((long long)i * i) % (2LL * L);
        double sn, cs;
        sincospi(-(double)q / (double)L, &sn, &cs);
        w[i] = make_float2((float)cs, (float)sn);
        const double freq = (i <= L / 2) ? ((double)i * smpl) / (double)L
                                         : ((double)(i - L) * smpl) / (double)L;
        mask[i] = (fabs(freq) > Bw) ? 0 : 1;
    }
    if (i < M) {
        int m = -1;
        if (i < L) m = i;
        else if (i > M - L) m = M - i;
        float2 v = make_float2(0.f, 0.f);
        if (m >= 0) {
            const long long q = ((long long)m * m) % (2LL * L);
            double sn, cs;
            sincospi((double)q / (double)L, &sn, &cs);
            v = make_float2((float)cs, (float)sn);
        }
        b[i] = v;
    }
}

// ---------------------------------------------------------------------------
// 6. target selection (host aperture test of sim.cu:808-826, in FP64)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_target_select(
    const float* __restrict__ tg, int64_t Tpad, int T, const unsigned char* __restrict__ ttype,
    const erp_shot* __restrict__ shots, const int* __restrict__ Vs, double cos_ap,
    int* __restrict__ list1, int* __restrict__ list0, int* __restrict__ tcnt, int Tcap,
    const TrackK* __restrict__ trk, const RowK* __restrict__ rows) {
#pragma clang fp contract(off)
    const int s = blockIdx.x;
    if (trk) {
        // multi-track: the row's track's targets (Tcap >= every track's T)
        const TrackK& tk = trk[rows[s].trk];
        tg = tk.tg;
        Tpad = tk.Tpad;
        T = tk.T;
        ttype = tk.ttype;
        cos_ap = tk.kp.cos_ap;
    }
    const ShotK sh = load_shot(shots, s);
    // each chunk of 1024 targets: a wave's rank from its ballots, the waves'
    // counts through LDS (two buffers by chunk parity, so one barrier per
    // chunk), the running totals in registers
    __shared__ int wn[2][2][16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    int run1 = 0, run0 = 0;
    const bool live = Vs[s] > 0;
    for (int base = 0, it = 0; base < T; base += 1024, ++it) {
        const int t = base + threadIdx.x;
        bool in1 = false, in0 = false;
        if (live && t < T) {
            const double vx = (double)tg[t] - sh.sx, vy = (double)tg[Tpad + t] - sh.sy,
                         vz = (double)tg[2 * Tpad + t] - sh.sz;
            const double mag = sqrt(vx * vx + vy * vy + vz * vz);
            const double ux = vx / mag, uy = vy / mag, uz = vz / mag;
            const double dp = (-(double)sh.snx) * ux + (-(double)sh.sny) * uy +
                              (-(double)sh.snz) * uz;
            if (!(dp < cos_ap)) {
                if (ttype[t] != 0) in1 = true; else in0 = true;
            }
        }
        const unsigned long long b1 = __ballot(in1), b0 = __ballot(in0);
        int* n1 = wn[it & 1][0];
        int* n0 = wn[it & 1][1];
        if (lane == 0) {
            n1[wv] = __popcll(b1);
            n0[wv] = __popcll(b0);
        }
        __syncthreads();
        int o1 = run1 + __popcll(b1 & below), o0 = run0 + __popcll(b0 & below);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int a = n1[k], b = n0[k];
            if (k < wv) {
                o1 += a;
                o0 += b;
            }
            run1 += a;
            run0 += b;
        }
        if (in1) list1[(size_t)s * Tcap + o1] = t;
        if (in0) list0[(size_t)s * Tcap + o0] = t;
    }
    if (threadIdx.x == 0) {
        tcnt[2 * s] = run1;
        tcnt[2 * s + 1] = run0;
    }
}

// ---------------------------------------------------------------------------
// 7. facet x target accumulation
// ---------------------------------------------------------------------------
struct PairCtx {
    const int* pair_shot;   // per pair in chunk: local shot index
    const int* pair_tgt;    // per pair: target index
};

// geometry_kernels.cu:207-269
__device__ __forceinline__ float prism_len(float ox, float oy, float oz, float dx, float dy,
                                           float dz, const erp_prism& p, float t) {
    const float INF = 1e30f;
    float tx1, tx2, ty1, ty2, tz1, tz2;
    if (fabsf(dx) < 1e-12f) {
        if (ox < p.xmin || ox > p.xmax) return 0.f;
        tx1 = -INF; tx2 = INF;
    } else {
        tx1 = (p.xmin - ox) / dx; tx2 = (p.xmax - ox) / dx;
        if (tx1 > tx2) { float tt = tx1; tx1 = tx2; tx2 = tt; }
    }
    if (fabsf(dy) < 1e-12f) {
        if (oy < p.ymin || oy > p.ymax) return 0.f;
        ty1 = -INF; ty2 = INF;
    } else {
        ty1 = (p.ymin - oy) / dy; ty2 = (p.ymax - oy) / dy;
        if (ty1 > ty2) { float tt = ty1; ty1 = ty2; ty2 = tt; }
    }
    if (fabsf(dz) < 1e-12f) {
        if (oz < p.zmin || oz > p.zmax) return 0.f;
        tz1 = -INF; tz2 = INF;
    } else {
        tz1 = (p.zmin - oz) / dz; tz2 = (p.zmax - oz) / dz;
        if (tz1 > tz2) { float tt = tz1; tz1 = tz2; tz2 = tt; }
    }
    float te = fmaxf(fmaxf(tx1, ty1), tz1), tx = fminf(fminf(tx2, ty2), tz2);
    if (te < 0.f) te = 0.f;
    if (tx > t) tx = t;
    if (te >= tx) return 0.f;
    return tx - te;
}

__device__ __forceinline__ float rerad_f(int fn, float th) {
    const float pi = ERP_PI_F;
    switch (fn) {
    case 0: return 1.f;
    case 1: return (th * (180.f / pi) < 1.f) ? 1.f : 0.f;
    case 2: return (th * (180.f / pi) < 0.5f) ? 1.f : 0.f;
    case 3: return (th * (pi / 180.f) < 3.f) ? 1.f : 0.f;
    case 4: {
        const float sg = pi / 180.f;
        return expf(-0.5f * (th * th) / (sg * sg));
    }
    default: return 1.f;
    }
}

// MODE 0: type != 0, non-specular: one trace A (the inward trace PTtarg);
//         the outward trace is -A/sqrt(P) exactly.
// MODE 1: type != 0, specular: inward (A) and outward (B) traces separately.
// MODE 2: type == 0 (genRefrPhasor): single bin, summed into one per-shot
//         trace (out row = per shot, global atomics after the LDS window).
// DET (the deterministic mode; MODE 1 and 2): every term is rounded once to
// 64-bit fixed point and summed with integer atomics, so no sum depends on
// the order of arrival.  MODE 1: a first sweep over the facets takes the
// pair's largest term component (the scale), then each window [wb, wb +
// WIN) of [base, top) is summed from a sweep of its own; a term outside
// [base, top) breaks a bound and is counted in *oow.  MODE 2 (all type-0
// targets of a shot add to the shot's row): pass 1 only takes the largest
// component of each pair into tmax[shot]; pass 2 adds the terms, scaled by
// the shot's bound, into the shot's row as fixed-point (int64) pairs.
template <int MODE, bool DET = false>
__global__ void __launch_bounds__(NTT) k_target_accum(
    const float* __restrict__ tg, int64_t Tpad, const int* __restrict__ pshot,
    const int* __restrict__ ptgt, const erp_shot* __restrict__ shots,
    const int* __restrict__ Vs, const int64_t* __restrict__ rec_base, Rec rec,
    const Bounds* __restrict__ bounds, const erp_prism* __restrict__ prisms, int np, KPar kp,
    float2* __restrict__ outA, float2* __restrict__ outB, int* __restrict__ pbase,
    unsigned* __restrict__ tmax, const int* __restrict__ tcnt, int pass, int* __restrict__ oow) {
    extern __shared__ double2 lds[];
    constexpr int WIN = (MODE == 1) ? TWIN / 2 : TWIN;  // MODE 1 holds two windows
    double2* winA = lds;
    double2* winB = lds + WIN;  // MODE 1 only
    const int pi = blockIdx.x;
    const int s = pshot[pi], t = ptgt[pi];
    const ShotK sh = load_shot(shots, s);
    const int V = Vs[s];
    const int64_t r0 = rec_base[s];
    const float txf = tg[t], tyf = tg[Tpad + t], tzf = tg[2 * Tpad + t];
    const float tnx = tg[3 * Tpad + t], tny = tg[4 * Tpad + t], tnz = tg[5 * Tpad + t];
    const double tx = txf, ty = tyf, tz = tzf;
    const double rst = (double)sh.rst;
    const int nr = kp.nr, P = kp.P_fft;
    const bool lossless = kp.mode & ERP_MODE_LOSSLESS;

    // window base from the shot bounds: rngt >= Itd_min + n * dist(target, AABB)
    const Bounds bd = bounds[s];
    const float bx = fmaxf(fmaxf(o2f(bd.xmin) - txf, txf - o2f(bd.xmax)), 0.f);
    const float by = fmaxf(fmaxf(o2f(bd.ymin) - tyf, tyf - o2f(bd.ymax)), 0.f);
    const float bz = fmaxf(fmaxf(o2f(bd.zmin) - tzf, tzf - o2f(bd.zmax)), 0.f);
    const double tmin = sqrt((double)bx * bx + (double)by * by + (double)bz * bz);
    const double itmin = __longlong_as_double((long long)bd.itd_min);
    double rlo;
    if (MODE == 2) rlo = (itmin + tmin) - tmin + tmin * kp.c1c2;
    else rlo = itmin + kp.n_d * tmin;
    int base = (int)floor((rlo - rst) / kp.drd) - 2;
    base = max(base, 0);
    // upper bound: Itd_max + n * (distance to the farthest AABB corner)
    bool fits = false;
    int top = nr;
    if (DET && MODE == 1) {
        const float fx = fmaxf(fabsf(o2f(bd.xmin) - txf), fabsf(o2f(bd.xmax) - txf));
        const float fy = fmaxf(fabsf(o2f(bd.ymin) - tyf), fabsf(o2f(bd.ymax) - tyf));
        const float fz = fmaxf(fabsf(o2f(bd.zmin) - tzf), fabsf(o2f(bd.zmax) - tzf));
        const double tmax_d = sqrt((double)fx * fx + (double)fy * fy + (double)fz * fz);
        const double ithmax = __longlong_as_double((long long)bd.itd_max);
        const double rhi = ithmax + kp.n_d * tmax_d * (1.0 + 1e-6) + 1.0;
        top = min((int)ceil((rhi - rst) / kp.drd) + 2, nr);
    }
    if (MODE == 0 && pbase) {
        const float fx = fmaxf(fabsf(o2f(bd.xmin) - txf), fabsf(o2f(bd.xmax) - txf));
        const float fy = fmaxf(fabsf(o2f(bd.ymin) - tyf), fabsf(o2f(bd.ymax) - tyf));
        const float fz = fmaxf(fabsf(o2f(bd.zmin) - tzf), fabsf(o2f(bd.zmax) - tzf));
        const double tmax = sqrt((double)fx * fx + (double)fy * fy + (double)fz * fz);
        const double ithmax = __longlong_as_double((long long)bd.itd_max);
        const double rhi = ithmax + kp.n_d * tmax * (1.0 + 1e-6) + 1.0;
        const int top = (int)ceil((rhi - rst) / kp.drd) + 2;
        fits = (top < base + WIN - 1) || (base + WIN >= nr);
        if (threadIdx.x == 0) pbase[pi] = fits ? base : -1;
    }

    float2* rowA = (MODE == 2) ? outA + (size_t)s * P : outA + (size_t)pi * P;
    float2* rowB = (MODE == 1) ? outB + (size_t)pi * P : nullptr;

    // one facet x target term of facet k: add(b, w, c, d) with the bin, the
    // split weight, the inward term and (MODE 1) the outward term, unless
    // its range falls outside the bins
    auto eval = [&](int k, auto&& add) {
        const int64_t r = r0 + k;
        const float fx = rec.geo[r], fy = rec.geo[rec.cap + r], fz = rec.geo[2 * rec.cap + r];
        const float nx = rec.geo[3 * rec.cap + r], ny = rec.geo[4 * rec.cap + r],
                    nz = rec.geo[5 * rec.cap + r];
        const float ux = rec.geo[6 * rec.cap + r], uy = rec.geo[7 * rec.cap + r],
                    uz = rec.geo[8 * rec.cap + r];
        const float vx = rec.geo[9 * rec.cap + r], vy = rec.geo[10 * rec.cap + r],
                    vz = rec.geo[11 * rec.cap + r];
        const double Itd = rec.itd[r];
        // compTargetRays: forced ray facet -> target
        const double dx = tx - (double)fx, dy = ty - (double)fy, dz = tz - (double)fz;
        const double Ttd = dist3(dx, dy, dz);
        double rngt;
        if (MODE == 2) {
            // genRefrPhasor: sltrng = Itd + Ttd; rngt = (sltrng - Ttd) + Ttd*(c1/c2)
            const double sl = Itd + Ttd;
            rngt = madd_nc(sl - Ttd, Ttd, kp.c1c2);
        } else {
            rngt = madd_nc(Itd, Ttd, kp.n_d);
        }
        int b;
        float w;
        if (!bin_of(rngt, rst, kp.drd, nr, &b, &w)) return;
        const float Ttdf = (float)Ttd;
        const double invd = 1.0 / Ttd;
        const double ix = dx * invd, iy = dy * invd, iz = dz * invd;
        const float Ix = (float)ix, Iy = (float)iy, Iz = (float)iz;
        const float2 csT = cos_sin_ref(ix, iy, iz, nx, ny, nz);
        const float2 csTT = cos_sin_ref(ix, iy, iz, tnx, tny, tnz);
        const float cT = csT.x, sT = csT.y, cTT = csTT.x, sTT = csTT.y;
        float Iu = Ix * ux + Iy * uy + Iz * uz;
        float Iv = Ix * vx + Iy * vy + Iz * vz;
        if (!kp.ortho) {
            const float2 e = uv_dots_exact(ix, iy, iz, ux, uy, uz, vx, vy, vz);
            Iu = e.x;
            Iv = e.y;
        }
        // attenuation along the forced ray (prisms + background alpha2)
        float expo = 0.f;
        {
            float tot = 0.f;
            for (int p = 0; p < np; ++p) {
                const float d = prism_len(fx, fy, fz, Ix, Iy, Iz, prisms[p], Ttdf);
                tot += d;
                expo += prisms[p].alpha * d;
            }
            expo += kp.alpha2 * (Ttdf - tot);
        }
        const float kc = kp.ks * cT;
        if (MODE == 2) {
            // EI = atten*fRfrC*rough (*rerad_spec), EO likewise; rerad(TargetTh)
            const float fRfrC = rec.fRfrC[r];
            float EI = expf(-4.f * expo) * fRfrC * expf(-4.f * kc * kc);
            float EO = EI;
            if (kp.mode & ERP_MODE_SPECULAR) {
                const float Ith = rec.ang[r], Iph = rec.ang[rec.cap + r],
                            Rth = rec.ang[2 * rec.cap + r];
                const float Tth = atan2f(sT, cT);
                const float Tph = atan2f(Iu, Iv);
                const float dth = Rth - Tth, dph = Iph - Tph;
                const float si = sinf(dth);
                const float a1 = sinc_f(kp.fs_lam * si * cosf(dph)),
                            a2 = sinc_f(kp.fs_lam * si * sinf(dph));
                const float vi = kp.fs2_lam * a1 * a2;
                EI *= vi * vi;
                float sR = kp.sqrt_e2e1 * sinf(ERP_PI_F - Tth);
                const float RrTh = (sR > 1.f) ? 1e4f : asinf(sR);
                const float so = sinf(RrTh), dpo = Tph - Iph;
                const float b1 = sinc_f(kp.fs_lam * so * cosf(dpo)),
                            b2 = sinc_f(kp.fs_lam * so * sinf(dpo));
                const float vo = kp.fs2_lam * b1 * b2;
                EO *= vo * vo;
                (void)Ith;
            }
            const float TTh = atan2f(sTT, cTT);
            const double sl = Itd + Ttd;
            const double sl2 = sl * sl;
            const float req = (float)((double)V * ((double)(kp.P * sh.grefr * sh.grefr * kp.fs *
                                                           kp.fs * kp.fs * kp.fs) /
                                                   (ERP_FOURPI5_SQ * sl2 * sl2)));
            const float rc = rerad_f(kp.rerad, TTh) * EI * EO * req;
            const float2 ph = phasor_cycles(-2.0 * rngt * kp.inv_cyc);
            const float2 c = make_float2(ph.x * rc, ph.y * rc);
            add(b, 0.f, c, make_float2(0.f, 0.f));
            return;
        }
        // accumulateTarget / radiateTarget amplitude:
        // sqrt(Pray n) = F_f * sqrt(G(Tth,Tph) G(TargetTh,Tph)) / Ttd * sqrt(EI-part)
        // (cos Tph, sin Tph) = (I.v, I.u) / |(I.u, I.v)|
        const float rp = sqrtf(Iu * Iu + Iv * Iv);
        const float cph = (rp > 0.f) ? Iv / rp : 1.f;
        const float sph = (rp > 0.f) ? Iu / rp : 0.f;
        const float2 scT = make_float2(sT * cph, sT * sph);
        const float s1 = sinc_f(kp.piLlam * scT.x), s2 = sinc_f(kp.piLlam * scT.y);
        const float s3 = sinc_f(kp.piLlam * sTT * cph), s4 = sinc_f(kp.piLlam * sTT * sph);
        const float gg = fabsf(s1 * s2) * fabsf(s3 * s4) * kp.Dmax;
        float amp = rec.F[r] * gg / Ttdf;
        float ampB = 0.f;
        if (!lossless) {
            const float e = expf(-2.f * (expo + kc * kc));
            amp *= e;
            if (MODE == 1) {
                // specular EI/EO facet reradiation factors (geometry_kernels.cu:530-539,565-572)
                const float Iph = rec.ang[rec.cap + r], Rth = rec.ang[2 * rec.cap + r];
                const float Tth = atan2f(sT, cT);
                const float Tph = atan2f(Iu, Iv);
                const float si = sinf(Rth - Tth), dph = Iph - Tph;
                const float a1 = sinc_f(kp.fs_lam * si * cosf(dph)),
                            a2 = sinc_f(kp.fs_lam * si * sinf(dph));
                const float vi = kp.fs2_lam * a1 * a2;
                const float sR = kp.sqrt_e2e1 * sinf(ERP_PI_F - Tth);
                const float RrTh = (sR > 1.f) ? 1e4f : asinf(sR);
                const float so = sinf(RrTh), dpo = Tph - Iph;
                const float b1 = sinc_f(kp.fs_lam * so * cosf(dpo)),
                            b2 = sinc_f(kp.fs_lam * so * sinf(dpo));
                const float vo = kp.fs2_lam * b1 * b2;
                ampB = amp * fabsf(vo) / kp.sqrtP;
                amp *= fabsf(vi);
            }
        } else if (MODE == 1) {
            ampB = amp / kp.sqrtP;
        }
        const float2 ph = phasor_cycles(rngt * kp.inv_cyc);
        const float2 c = make_float2(ph.x * amp, ph.y * amp);
        float2 d = make_float2(0.f, 0.f);
        if (MODE == 1) {
            // radiateTarget: phasor(rngt - lam/2), P = 1 W
            const float2 pb = phasor_cycles(rngt * kp.inv_cyc - 0.5 * (double)kp.lam * kp.inv_cyc);
            d = make_float2(pb.x * ampB, pb.y * ampB);
        }
        add(b, w, c, d);
    };

    if constexpr (!DET) {
        // zero the output rows outside the window (MODE 2 rows are pre-zeroed);
        // unnecessary when every contribution provably lands inside the window
        if (MODE != 2 && !fits) {
            for (int k = threadIdx.x; k < P; k += NTT) {
                if (k < base || k >= base + WIN || k >= nr) {
                    rowA[k] = make_float2(0.f, 0.f);
                    if (MODE == 1) rowB[k] = make_float2(0.f, 0.f);
                }
            }
        }
        for (int k = threadIdx.x; k < WIN; k += NTT) {
            winA[k] = make_double2(0.0, 0.0);
            if (MODE == 1) winB[k] = make_double2(0.0, 0.0);
        }
        // only the row zeros above must land before other threads' atomics
        if (MODE != 2 && !fits) __threadfence();
        __syncthreads();

        for (int k = threadIdx.x; k < V; k += NTT) {
            eval(k, [&](int b, float w, float2 c, float2 d) {
                const int k0 = b - base;
                if (MODE == 2) {
                    if (k0 >= 0 && k0 < WIN) lds_add2(winA, k0, c); else glb_add2(rowA, b, c);
                    return;
                }
                const float2 c0 = make_float2(c.x * (1.f - w), c.y * (1.f - w));
                const float2 c1 = make_float2(c.x * w, c.y * w);
                if (k0 >= 0 && k0 < WIN) lds_add2(winA, k0, c0); else glb_add2(rowA, b, c0);
                if (b + 1 < nr) {
                    if (k0 + 1 >= 0 && k0 + 1 < WIN) lds_add2(winA, k0 + 1, c1);
                    else glb_add2(rowA, b + 1, c1);
                }
                if (MODE == 1) {
                    const float2 d0 = make_float2(d.x * (1.f - w), d.y * (1.f - w));
                    const float2 d1 = make_float2(d.x * w, d.y * w);
                    if (k0 >= 0 && k0 < WIN) lds_add2(winB, k0, d0); else glb_add2(rowB, b, d0);
                    if (b + 1 < nr) {
                        if (k0 + 1 >= 0 && k0 + 1 < WIN) lds_add2(winB, k0 + 1, d1);
                        else glb_add2(rowB, b + 1, d1);
                    }
                }
            });
        }
        __syncthreads();
        for (int k = threadIdx.x; k < WIN; k += NTT) {
            const int b = base + k;
            if (b >= nr) break;
            if (MODE == 2) {
                const double2 v = winA[k];
                if (v.x != 0.0 || v.y != 0.0) glb_add2(rowA, b, make_float2((float)v.x, (float)v.y));
            } else {
                rowA[b] = make_float2((float)winA[k].x, (float)winA[k].y);
                if (MODE == 1) rowB[b] = make_float2((float)winB[k].x, (float)winB[k].y);
            }
        }
    } else {
        static_assert(MODE != 0, "the deterministic mode of MODE 0 is k_target_accum_fast");
        // the largest component of this pair's terms (not needed by the
        // second pass of MODE 2, which reads its shot's bound)
        __shared__ float red[NTT / 64];
        float m = 0.f;
        if (MODE == 1 || pass == 1) {
            for (int k = threadIdx.x; k < V; k += NTT)
                eval(k, [&](int, float, float2 c, float2 d) {
                    m = fmaxf(m, fmaxf(fmaxf(fabsf(c.x), fabsf(c.y)),
                                       fmaxf(fabsf(d.x), fabsf(d.y))));
                });
            for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
            __syncthreads();
            m = red[0];
            for (int wv = 1; wv < NTT / 64; ++wv) m = fmaxf(m, red[wv]);
        }
        if constexpr (MODE == 2) {
            if (pass == 1) {
                if (threadIdx.x == 0 && m > 0.f) atomicMax(&tmax[s], __float_as_uint(m));
                return;
            }
            // pass 2: the shot's scale (V facets x its type-0 targets terms)
            const double scale = t0_det_scale(V, tcnt[2 * s + 1], tmax[s]);
            unsigned long long* wx = reinterpret_cast<unsigned long long*>(lds);
            unsigned long long* wy = wx + WIN;
            unsigned long long* ru = reinterpret_cast<unsigned long long*>(rowA);
            for (int k = threadIdx.x; k < WIN; k += NTT) {
                wx[k] = 0ull;
                wy[k] = 0ull;
            }
            __syncthreads();
            for (int k = threadIdx.x; k < V; k += NTT)
                eval(k, [&](int b, float, float2 c, float2) {
                    const int k0 = b - base;
                    if (k0 >= 0 && k0 < WIN) lds_add2_fix(wx, wy, k0, c, scale);
                    else glb_add2_fix(ru, b, c, scale);
                });
            __syncthreads();
            for (int k = threadIdx.x; k < WIN; k += NTT) {
                const int b = base + k;
                if (b >= nr) break;
                const unsigned long long x = wx[k], y = wy[k];
                if (x | y) {
                    atomicAdd(&ru[2 * b], x);
                    atomicAdd(&ru[2 * b + 1], y);
                }
            }
        } else {
            // MODE 1: rows outside [base, top) stay zero (no term reaches them)
            const int hi = max(top, base);
            for (int k = threadIdx.x; k < P; k += NTT) {
                if (k < base || k >= hi || k >= nr) {
                    rowA[k] = make_float2(0.f, 0.f);
                    rowB[k] = make_float2(0.f, 0.f);
                }
            }
            // 2 V terms per row, each at most m (1e-4 margin for the split)
            const double scale = det_scale(2.0 * (double)V, 1.0001 * (double)m);
            const double inv = scale > 0.0 ? 1.0 / scale : 0.0;
            unsigned long long* wax = reinterpret_cast<unsigned long long*>(lds);
            unsigned long long* way = wax + WIN;
            unsigned long long* wbx = way + WIN;
            unsigned long long* wby = wbx + WIN;
            for (int wb = base; wb < top || wb == base; wb += WIN) {
                const int wl = min(WIN, max(0, top - wb));
                for (int k = threadIdx.x; k < wl; k += NTT) {
                    wax[k] = 0ull; way[k] = 0ull; wbx[k] = 0ull; wby[k] = 0ull;
                }
                __syncthreads();
                const bool first = wb == base, last = wb + WIN >= top;
                for (int k = threadIdx.x; k < V; k += NTT)
                    eval(k, [&](int b, float w, float2 c, float2 d) {
                        auto put = [&](int bb, float2 v, float2 u) {
                            const int kk = bb - wb;
                            if (kk >= 0 && kk < wl) {
                                lds_add2_fix(wax, way, kk, v, scale);
                                lds_add2_fix(wbx, wby, kk, u, scale);
                            } else if ((bb < base && first) || (bb >= top && last)) {
                                atomicAdd(oow, 1);
                            }
                        };
                        put(b, make_float2(c.x * (1.f - w), c.y * (1.f - w)),
                            make_float2(d.x * (1.f - w), d.y * (1.f - w)));
                        if (b + 1 < nr)
                            put(b + 1, make_float2(c.x * w, c.y * w), make_float2(d.x * w, d.y * w));
                    });
                __syncthreads();
                for (int k = threadIdx.x; k < wl; k += NTT) {
                    const int b = wb + k;
                    if (b >= nr) break;
                    rowA[b] = make_float2((float)det_unfix(wax[k], inv), (float)det_unfix(way[k], inv));
                    rowB[b] = make_float2((float)det_unfix(wbx[k], inv), (float)det_unfix(wby[k], inv));
                }
                __syncthreads();
            }
        }
    }
}

// One facet x target term of the plain-mode target kernel: range bin, split
// weight and complex amplitude (compTargetRays + compRefrEnergyIn +
// accumulateTarget, geometry_kernels.cu:272-350,520-554, radar_kernels.cu:
// 767-830) for facet record r.
struct Tri {
    int b;
    double w;  // kept in FP64: the FP64 two-bin split takes it without conversions
    float2 c;
    bool valid;
};

// one facet record in registers
struct RecV {
    float4 qa, qb, qc;
    float qd;
    double itd;
};
__device__ __forceinline__ RecV load_rec(const Rec& rec, int64_t r) {
    RecV v;
    v.qa = rec.qa[r];
    v.qb = rec.qb[r];
    v.qc = rec.qc[r];
    v.qd = rec.qd[r];
    v.itd = rec.itd[r];
    return v;
}
// The packed records of one shot, based at the shot's first record (a value
// uniform over the workgroup, so the five bases sit in scalar registers).
// Record k of the shot is then one 32-bit byte offset per array off a scalar
// base (k < 2^27: 16 k fits 32 bits), which the loads take as is, instead of
// five 64-bit address sums per record.
struct RecShot {
    const char* qa;
    const char* qb;
    const char* qc;
    const char* qd;
    const char* itd;
};
__device__ __forceinline__ RecShot rec_shot(const Rec& rec, int64_t r0) {
    return RecShot{reinterpret_cast<const char*>(rec.qa + r0),
                   reinterpret_cast<const char*>(rec.qb + r0),
                   reinterpret_cast<const char*>(rec.qc + r0),
                   reinterpret_cast<const char*>(rec.qd + r0),
                   reinterpret_cast<const char*>(rec.itd + r0)};
}
__device__ __forceinline__ RecV load_rec(const RecShot& rs, unsigned k) {
    RecV v;
    const unsigned o16 = k << 4;
    v.qa = *reinterpret_cast<const float4*>(rs.qa + o16);
    v.qb = *reinterpret_cast<const float4*>(rs.qb + o16);
    v.qc = *reinterpret_cast<const float4*>(rs.qc + o16);
    v.qd = *reinterpret_cast<const float*>(rs.qd + (k << 2));
    v.itd = *reinterpret_cast<const double*>(rs.itd + (k << 3));
    return v;
}

// The second subsurface medium for one pair: deep = the target lies below
// the interface z = zi, so part of each facet's forced ray runs in eps_3
// (a uniform flag per pair).
struct Strat {
    bool deep;
    double zi, tz, dn;
    float tsq;
};

// Fraction of the straight facet (fz) -> target (tz) segment below z = zi
// for a target below it: (zi - tz) / (fz - tz), or 1 when the facet is not
// above the interface.  Exact: the oracle's FP64 division.
__device__ __forceinline__ double below_frac_exact(float fz, const Strat& st) {
#pragma clang fp contract(off)
    if (!((double)fz > st.zi)) return 1.0;
    return (st.zi - st.tz) / ((double)fz - st.tz);
}
// The same from an FP32 reciprocal and one Newton correction in FP64
// (relative error about 1e-14)
__device__ __forceinline__ double below_frac(float fz, const Strat& st) {
    if (!((double)fz > st.zi)) return 1.0;
    const double num = st.zi - st.tz, den = (double)fz - st.tz;
    const double r0 = (double)__builtin_amdgcn_rcpf((float)den);
    const double q = num * r0;
    return fma(fma(-q, den, num), r0, q);
}

// Bin window of a (shot, target) pair from the shot's bounds (the box and
// the Itd range of its illuminated facets): every term's range lies between
// Itd_min + n_lo Ttd_min and Itd_max + n_hi Ttd_max, so its bins b and b + 1
// lie in [base, top).  fits: a window of cap bins at base holds [base, top);
// len is then its live part, top - base rounded up to 64 bins (the spectral
// chain reads whole groups of 64), else cap.  need_top false: base only.
struct PairWin {
    int base, top, len;
    bool fits;
};
__device__ __forceinline__ PairWin pair_window(const Bounds& bd, float txf, float tyf, float tzf,
                                               double n_lo, double n_hi, double rst,
                                               const KPar& kp, int nr, int cap, bool need_top) {
    PairWin pw;
    const float bx = fmaxf(fmaxf(o2f(bd.xmin) - txf, txf - o2f(bd.xmax)), 0.f);
    const float by = fmaxf(fmaxf(o2f(bd.ymin) - tyf, tyf - o2f(bd.ymax)), 0.f);
    const float bz = fmaxf(fmaxf(o2f(bd.zmin) - tzf, tzf - o2f(bd.zmax)), 0.f);
    const double tmin = sqrt((double)bx * bx + (double)by * by + (double)bz * bz);
    const double itmin = __longlong_as_double((long long)bd.itd_min);
    int base = (int)floor((itmin + n_lo * tmin - rst) / kp.drd) - 2;
    // win_shift > 0 only to test the out-of-window check: a window that
    // starts above the lowest bin the bounds allow
    pw.base = max(base, 0) + kp.win_shift;
    pw.fits = false;
    pw.len = cap;
    pw.top = nr;
    if (need_top) {
        const float fx = fmaxf(fabsf(o2f(bd.xmin) - txf), fabsf(o2f(bd.xmax) - txf));
        const float fy = fmaxf(fabsf(o2f(bd.ymin) - tyf), fabsf(o2f(bd.ymax) - tyf));
        const float fz = fmaxf(fabsf(o2f(bd.zmin) - tzf), fabsf(o2f(bd.zmax) - tzf));
        double tmax = sqrt((double)fx * fx + (double)fy * fy + (double)fz * fz);
        if (bd.rad >= 0.f) {
            // |f - t| <= |c - t| + rad for every lit facet f
            const double dx = (double)txf - bd.cx, dy = (double)tyf - bd.cy,
                         dz = (double)tzf - bd.cz;
            tmax = fmin(tmax, sqrt(dx * dx + dy * dy + dz * dz) + (double)bd.rad);
        }
        const double ithmax = __longlong_as_double((long long)bd.itd_max);
        const double rhi = ithmax + n_hi * tmax * (1.0 + 1e-6) + 1.0;
        pw.top = min((int)ceil((rhi - rst) / kp.drd) + 2, nr);
        pw.fits = pw.top - pw.base <= cap;
        if (pw.fits) pw.len = max(0, min(cap, (pw.top - pw.base + 63) & ~63));
    }
    return pw;
}

template <int PH = PH_F64, bool S3 = false>
__device__ __forceinline__ Tri tri_eval(const RecV& rv, bool frame_ok, double sx,
                                        double sy, double sz, double tx, double ty, double tz,
                                        float tnx, float tny, float tnz, bool tn_up, double rst,
                                        double inv_dr, double n_d, double two_cyc, int nr,
                                        bool lossless, const erp_prism* __restrict__ prisms,
                                        int np, const KPar& kp, const Strat& st) {
    Tri o;
    const float4 qa = rv.qa, qb = rv.qb, qc = rv.qc;
    const float qd = rv.qd;
    const float fx = qa.x, fy = qa.y, fz = qa.z;
    // Itd = |facet - source|: the surface kernel's exact FP64 value
    const double Itd = rv.itd;
    const double dx = tx - (double)fx, dy = ty - (double)fy, dz = tz - (double)fz;
    const double t2 = fma(dx, dx, fma(dy, dy, dz * dz));
    // hardware reciprocal square root: t2 is a squared facet-target
    // distance, far from the denormal range rsqrtf guards against
    const float rs = __builtin_amdgcn_rsqf((float)t2);
    const double rd = (double)rs, hr = 0.5 * rd;
    const double y0 = t2 * rd;
    const double Ttd = fma(fma(-y0, y0, t2), hr, y0);
    double rngt = 0.0;
    float2 ph_st = make_float2(1.f, 0.f);
    if constexpr (PH == PH_F64) {
        rngt = fma(Ttd, n_d, Itd);
        // second medium: + (n3 - n2) times the ray's length below the interface
        if (S3 && st.deep) rngt = fma(Ttd * below_frac(fz, st), st.dn, rngt);
        // (rngt - rst) / dr to within 1e-11 bin; the 1e-7 margin below
        // covers it
        const double xq = fma(rngt, inv_dr, -rst * inv_dr);
        o.b = (int)xq;
        o.w = xq - (double)o.b;
        if (o.w < 1e-7 || o.w > 1.0 - 1e-7) {
            // within about 1e-7 bin of an edge (every negative coordinate
            // included): the oracle's exact operation sequence
            const double Te = dist3(dx, dy, dz);
            rngt = madd_nc(Itd, Te, n_d);
            if (S3 && st.deep) {
#pragma clang fp contract(off)
                const double L3 = Te * below_frac_exact(fz, st);
                rngt = madd_nc(rngt, L3, st.dn);
            }
            o.valid = bin_of(rngt, rst, kp.drd, nr, &o.b, &o.w);
        } else {
            o.valid = (xq > -1.0) && (xq < (double)nr);
        }
    } else if constexpr (PH == PH_F32) {
        // Itd, Ttd, rngt, the bin coordinate and the phase in FP32 as in
        // compIncidentRays / compTargetRays / accumulateTarget / phasor
        const float ax = fx - (float)sx, ay = fy - (float)sy, az = fz - (float)sz;
        const float bx = (float)tx - fx, by = (float)ty - fy, bz = (float)tz - fz;
        const float I32 = sqrtf(ax * ax + ay * ay + az * az);
        const float T32 = sqrtf(bx * bx + by * by + bz * bz);
        float r32 = I32 + T32 * kp.nf;
        if (S3 && st.deep) {
            const float fr = fz > (float)st.zi ? ((float)st.zi - (float)st.tz) / (fz - (float)st.tz)
                                                : 1.f;
            r32 = r32 + (T32 * fr) * (float)st.dn;
        }
        const float x32 = (r32 - (float)rst) / kp.dr;
        o.valid = (x32 > -1.f) && (x32 < (float)nr);
        o.b = (int)x32;
        o.w = x32 - (float)o.b;
        float sn, cs;
        sincosf(-(2.0f * 3.14159265f / kp.lam) * r32, &sn, &cs);
        ph_st = make_float2(cs, sn);
    } else {
        // double-float ranges and phase from the FP32 facet and target
        // coordinates and the shot position
        const DF Id = df_dist(fx, fy, fz, df_of(sx), df_of(sy), df_of(sz));
        const DF Td = df_dist(fx, fy, fz, df_of(tx), df_of(ty), df_of(tz));
        DF rd2 = df_add(Id, df_mul(Td, df_of(n_d)));
        if (S3 && st.deep)
            rd2 = df_add(rd2, df_mul(df_mul(Td, df_of(below_frac_exact(fz, st))), df_of(st.dn)));
        const DF xd = df_mul(df_add(rd2, df_neg(df_of(rst))), df_of(inv_dr));
        const double xq = (double)xd.hi + (double)xd.lo;
        o.valid = (xq > -1.0) && (xq < (double)nr);
        o.b = (int)xq;
        o.w = xq - (double)o.b;
        // exp(-i pi t), t = rngt * two_cyc half cycles, reduced in double-float
        const DF t = df_mul(rd2, df_of(0.5 * two_cyc));
        const float rr = rintf(t.hi);
        const float fr = (t.hi - rr) + t.lo;
        ph_st = make_float2(__builtin_amdgcn_cosf(fr), -__builtin_amdgcn_sinf(fr));
    }
    const float nx = qb.x, ny = qb.y, nz = qb.z;
    const float ux = qb.w, uy = qc.x, uz = qc.y;
    const float vx = qc.z, vy = qc.w, vz = qd;
    // unit vector facet -> target in FP64: 1/Ttd by one Newton step from rs,
    // rd (1.5 - t2 rd^2 / 2), with t2 rd^2 / 2 = y0 hr exactly
    const double invd = rd * fma(-y0, hr, 1.5);
    const double ix = dx * invd, iy = dy * invd, iz = dz * invd;
    const float Ix = (float)ix, Iy = (float)iy, Iz = (float)iz;
    const float2 csT = cos_sin_hw(ix, iy, iz, nx, ny, nz);
    const float cT = csT.x, sT = csT.y;
    float Iu = Ix * ux + Iy * uy + Iz * uz;
    float Iv = Ix * vx + Iy * vy + Iz * vz;
    if (!frame_ok) {
        const float2 e = uv_dots_exact(ix, iy, iz, ux, uy, uz, vx, vy, vz);
        Iu = e.x;
        Iv = e.y;
    }
    // sin(TargetTh); tn = (0, 0, 1) is the common case
    const float sTT = tn_up ? __builtin_amdgcn_sqrtf(fmaxf(0.f, (float)((1.0 + iz) * (1.0 - iz))))
                            : cos_sin_hw(ix, iy, iz, tnx, tny, tnz).y;
    // azimuth (cos Tph, sin Tph) = (Iv, Iu) / |(Iu, Iv)|; below 1e-30 the
    // sinc arguments s sin(Tph), s cos(Tph) are < 1e-6 whatever the azimuth
    // (s <= |(Iu, Iv)| for a unit I), so (1, 0) changes nothing and keeps
    // the hardware rsq away from denormal inputs
    const float q2 = Iu * Iu + Iv * Iv;
    const bool az = q2 > 1e-30f;
    const float rinv = az ? __builtin_amdgcn_rsqf(q2) : 0.f;
    const float cph = az ? Iv * rinv : 1.f;
    const float sph = Iu * rinv;
    const float a = kp.piLlam;
    float n1, d1, n2, d2, n3, d3, n4, d4;
    // sin(Tth) (cos Tph, sin Tph)
    const float scx = sT * cph, scy = sT * sph;
    pisinc_nd_hw(a * scx, &n1, &d1);
    pisinc_nd_hw(a * scy, &n2, &d2);
    pisinc_nd_hw(a * sTT * cph, &n3, &d3);
    pisinc_nd_hw(a * sTT * sph, &n4, &d4);
    const float den = fabsf((d1 * d2) * (d3 * d4));
    // qa.w = F Dmax / pi^4 (the 1/pi^4 takes the four factors pi of the
    // pi-sincs back out)
    float amp = qa.w * fabsf((n1 * n2) * (n3 * n4)) * __builtin_amdgcn_rcpf(den) * rs;
    // interface transmission of a ray that crosses into the second medium
    // (a loss factor like fRfrC: not applied to lossless runs)
    if (S3 && !lossless && st.deep && (double)fz > st.zi) amp *= st.tsq;
    if (!lossless) {
        float expo = 0.f, tot = 0.f;
        for (int k = 0; k < np; ++k) {
            const float d = prism_len(fx, fy, fz, Ix, Iy, Iz, prisms[k], (float)Ttd);
            tot += d;
            expo += prisms[k].alpha * d;
        }
        expo += kp.alpha2 * ((float)Ttd - tot);
        const float kc = kp.ks * cT;
        amp *= __expf(-2.f * (expo + kc * kc));
    }
    const float2 ph = (PH == PH_F64) ? phasor_rc_hw(rngt, kp.inv_cyc) : ph_st;
    o.c = make_float2(ph.x * amp, ph.y * amp);
    return o;
}

// two-bin split of one term into the LDS window [base, base + len) or, past
// it, the global row.  flag set: the pair's window holds every bin its
// bounds allow, so a term past it breaks a bound; it is counted in *flag
// (the run then fails) instead of being added to row bins nothing reads.
template <int ACC = ACC_F64>
__device__ __forceinline__ void tri_add(const Tri& t, int base, int len, int nr,
                                        WinS winA, float2* rowA, int* flag) {
    const int k0 = t.b - base;
    if constexpr (ACC == ACC_F64) {
        // split in FP64, the window's precision: c1 = c w, c0 = c - c1
        const double cx = t.c.x, cy = t.c.y, w = t.w;
        const double c1x = cx * w, c1y = cy * w;
        const double c0x = cx - c1x, c0y = cy - c1y;
        // 0 <= k < len as one unsigned compare (len >= 0)
        if ((unsigned)k0 < (unsigned)len) {
            atomicAdd(&winA.x[k0], c0x);
            atomicAdd(&winA.y[k0], c0y);
        } else if (flag) {
            atomicAdd(flag, 1);
        } else {
            glb_add2(rowA, t.b, make_float2((float)c0x, (float)c0y));
        }
        if (t.b < nr - 1) {
            if ((unsigned)(k0 + 1) < (unsigned)len) {
                atomicAdd(&winA.x[k0 + 1], c1x);
                atomicAdd(&winA.y[k0 + 1], c1y);
            } else if (flag) {
                atomicAdd(flag, 1);
            } else {
                glb_add2(rowA, t.b + 1, make_float2((float)c1x, (float)c1y));
            }
        }
        return;
    }
    const float wf = (float)t.w;
    const float2 c0 = make_float2(t.c.x * (1.f - wf), t.c.y * (1.f - wf));
    const float2 c1 = make_float2(t.c.x * wf, t.c.y * wf);
    if (k0 >= 0 && k0 < len) win_add<ACC>(winA, k0, c0);
    else if (flag) atomicAdd(flag, 1);
    else glb_add2(rowA, t.b, c0);
    if (t.b + 1 < nr) {
        if (k0 + 1 >= 0 && k0 + 1 < len) win_add<ACC>(winA, k0 + 1, c1);
        else if (flag) atomicAdd(flag, 1);
        else glb_add2(rowA, t.b + 1, c1);
    }
}

// Plain-mode (type != 0, non-specular) facet x target kernel with
// instruction-lean arithmetic: Newton-refined FP64 range from an FP32 rsqrt
// seed, one reciprocal for the product of the four sinc factors of
// facet_G(Tth,Tph) * facet_G(TargetTh,Tph), hardware sin/cos on exactly
// reduced arguments.  Bins are
// decided on the approximate range unless it lies within 1e-7 bin of a bin
// edge, where the exact FP64 sqrt and division of the oracle are repeated, so
// bin indices stay identical to the FP64 reference.
// S3: the second subsurface medium.  Its instantiations are held to 128
// VGPRs (4 waves per SIMD, the two 512-thread workgroups per CU that the LDS
// window allows; unbounded they take 132 and one workgroup per CU); the
// others keep the default bound.
// MT: a multi-track batch: each pair's row names its track, whose targets,
// parameters and facet frames the pair uses; the pair count is read from
// device memory (the launch is replayed from a graph).
template <bool DET, int PH = PH_F64, int ACC = ACC_F64, bool S3 = false, bool MT = false>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, NTF),
                               amdgpu_waves_per_eu(S3 ? 4 : 1)))
k_target_accum_fast(const float* __restrict__ tg0, int64_t Tpad0, const int* __restrict__ pshot,
                    const int* __restrict__ ptgt, const erp_shot* __restrict__ shots,
                    const int* __restrict__ Vs, const int64_t* __restrict__ rec_base, Rec rec,
                    const Bounds* __restrict__ bounds, const erp_prism* __restrict__ prisms,
                    int np, KPar kp0, float2* __restrict__ outA, int* __restrict__ pbase,
                    int* __restrict__ pwin, int frame_ok0, int npairs0,
                    int* __restrict__ next_pair, const TrackK* __restrict__ trk,
                    const RowK* __restrict__ rows, const int* __restrict__ npairs_dev,
                    int* __restrict__ oow) {
    __shared__ double winS[2 * TWF];
    __shared__ int s_pi;
    const WinS winA{winS, winS + TWF};
    const int npairs = MT ? *npairs_dev : npairs0;
    // persistent: each workgroup takes the next pair from a counter until
    // none is left (the counter starts at 0: k_pair_targets clears it), so
    // short pairs pay no workgroup launch and uneven pairs balance
    for (;;) {
        if (threadIdx.x == 0) s_pi = atomicAdd(next_pair, 1);
        __syncthreads();
        const int pi = s_pi;
        if (pi >= npairs) break;
        const int s = pshot[pi], t = ptgt[pi];
        const TrackK* tk = MT ? &trk[rows[s].trk] : nullptr;
        const KPar& kp = MT ? tk->kp : kp0;
        const float* __restrict__ tg = MT ? tk->tg : tg0;
        const int64_t Tpad = MT ? tk->Tpad : Tpad0;
        const int frame_ok = MT ? tk->frame_ok : frame_ok0;
        const ShotK sh = load_shot(shots, s);
        const int V = Vs[s];
        const int64_t r0 = rec_base[s];
        const RecShot rs = rec_shot(rec, r0);
        const float txf = tg[t], tyf = tg[Tpad + t], tzf = tg[2 * Tpad + t];
        const float tnx = tg[3 * Tpad + t], tny = tg[4 * Tpad + t], tnz = tg[5 * Tpad + t];
        const double tx = txf, ty = tyf, tz = tzf;
        const double rst = (double)sh.rst;
        const int nr = kp.nr, P = kp.P_fft;
        const bool lossless = kp.mode & ERP_MODE_LOSSLESS;

        // second medium: a uniform flag per pair
        const Strat st{S3 && (double)tzf < kp.zi, kp.zi, (double)tzf, kp.dn, kp.tsq};
        // window from the shot bounds (see k_target_accum); rngt lies between
        // Itd + n_lo Ttd and Itd + n_hi Ttd
        const double n_lo = st.deep ? kp.n_lo : kp.n_d, n_hi = st.deep ? kp.n_hi : kp.n_d;
        // live window: [base, base + len) (only it is zeroed, flushed and
        // read by the spectral chain)
        const PairWin pw = pair_window(bounds[s], txf, tyf, tzf, n_lo, n_hi, rst, kp, nr, TWF,
                                       pbase || DET);
        const int base = pw.base, top = pw.top, len = pw.len;
        const bool fits = pw.fits;
        if (pbase && threadIdx.x == 0) {
            pbase[pi] = fits ? base : -1;
            pwin[pi] = len;
        }
        float2* rowA = outA + (size_t)pi * P;
        // out-of-window terms: counted as broken bounds where the window
        // covers every bin the bounds allow (deterministic: always, so no
        // float atomic can reach the row)
        int* const flag = (DET || fits) ? oow : nullptr;
        if (!fits || (DET && !pbase)) {
            // bins outside the window(s) [base, base + len) or, deterministic,
            // [base, top): zero, then reached only by out-of-window atomics
            const int hi = fits ? base + len : (DET ? max(top, base) : base + TWF);
            for (int k = threadIdx.x; k < kp.row_live; k += NTF)
                if (k < base || k >= hi || k >= nr) rowA[k] = make_float2(0.f, 0.f);
        }

        const double inv_dr = kp.inv_dr, n_d = kp.n_d, two_cyc = 2.0 * kp.inv_cyc;
        const bool tn_up = tnx == 0.f && tny == 0.f && tnz == 1.f;

        if constexpr (DET) {
            // bound of every term of this pair: |component| <= F Dmax / Ttd (the
            // four sinc factors, the loss factor and the two-bin weights are <= 1)
            __shared__ float red[NTF / 64];
            float m = 0.f;
            for (int k = threadIdx.x; k < V; k += NTF) {
                const float4 qa = rec.qa[r0 + k];
                const double dx = tx - (double)qa.x, dy = ty - (double)qa.y, dz = tz - (double)qa.z;
                const double t2 = fma(dx, dx, fma(dy, dy, dz * dz));
                m = fmaxf(m, qa.w * __builtin_amdgcn_rsqf((float)t2));
            }
            for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
            __syncthreads();
            m = red[0];
            for (int w = 1; w < NTF / 64; ++w) m = fmaxf(m, red[w]);
            // 2x margin for the rounding of the sinc quotients
            // qa.w = F Dmax / pi^4 and each pi-sinc factor is <= pi
            const double scale = det_scale((double)V, 2.0 * (double)m * 97.40909103400244);
            const double inv = scale > 0.0 ? 1.0 / scale : 0.0;
            unsigned long long* wux = reinterpret_cast<unsigned long long*>(winS);
            unsigned long long* wuy = wux + TWF;
            // one LDS window of TWF bins at a time over [base, top)
            for (int wb = base; wb < top || wb == base; wb += TWF) {
                const int wl = fits ? len : min(TWF, max(0, top - wb));
                for (int k = threadIdx.x; k < wl; k += NTF) {
                    winS[k] = 0.0;
                    winS[TWF + k] = 0.0;
                }
                __syncthreads();
                const bool first = wb == base, last = wb + TWF >= top;
                auto add_det = [&](const Tri& ta) {
                    const float wf = (float)ta.w;
                    const float2 c0 = make_float2(ta.c.x * (1.f - wf), ta.c.y * (1.f - wf));
                    const float2 c1 = make_float2(ta.c.x * wf, ta.c.y * wf);
                    const int k0 = ta.b - wb;
                    if (k0 >= 0 && k0 < wl) lds_add2_fix(wux, wuy, k0, c0, scale);
                    else if ((ta.b < base && first) || (ta.b >= top && last)) atomicAdd(flag, 1);
                    if (ta.b + 1 < nr) {
                        if (k0 + 1 >= 0 && k0 + 1 < wl) lds_add2_fix(wux, wuy, k0 + 1, c1, scale);
                        else if ((ta.b + 1 < base && first) || (ta.b + 1 >= top && last))
                            atomicAdd(flag, 1);
                    }
                };
                // one facet per thread and step, the next step's record loaded
                // ahead into the other register set (two chains, as the plain
                // path runs, would cost this variant a wave per SIMD)
                if (V > 0) {
                    const int vl = V - 1;
                    auto rec_at = [&](int k) { return load_rec(rs, (unsigned)min(k, vl)); };
                    auto step = [&](const RecV& ca) {
                        const Tri ta = tri_eval<PH_F64, S3>(ca, frame_ok, sh.sx, sh.sy, sh.sz, tx, ty, tz, tnx,
                                                tny, tnz, tn_up, rst, inv_dr, n_d, two_cyc, nr,
                                                lossless, prisms, np, kp, st);
                        if (ta.valid) add_det(ta);
                    };
                    RecV a0 = rec_at(threadIdx.x);
                    for (int k = threadIdx.x; k < V; k += 2 * NTF) {
                        const RecV a1 = rec_at(k + NTF);
                        step(a0);
                        if (k + NTF >= V) break;
                        a0 = rec_at(k + 2 * NTF);
                        step(a1);
                    }
                }
                __syncthreads();
                for (int k = threadIdx.x; k < wl; k += NTF) {
                    const int b = wb + k;
                    if (b >= nr) break;
                    rowA[b] = make_float2((float)det_unfix(wux[k], inv),
                                          (float)det_unfix(wuy[k], inv));
                }
                __syncthreads();
                if (fits) break;
            }
            continue;  // the window loop ended on a barrier
        } else {
            for (int k = threadIdx.x; k < len; k += NTF) win_zero<ACC>(winA, k);
            // the zeros written to the row outside the window must land
            // before any thread's out-of-window atomics; a fitting pair
            // wrote none (and the LDS zeros need only the barrier)
            if (!fits) __threadfence();
            __syncthreads();

            // two facets per thread and step (two independent dependency
            // chains); the records of the next step are loaded before this
            // step's arithmetic so that their latency overlaps it.  Two register
            // sets (a0 b0, a1 b1) alternate, and every load reads a clamped,
            // in-bounds index, so no set is copied or kept conditionally.
            if (V > 0) {
                const int vl = V - 1;
                auto rec_at = [&](int k) { return load_rec(rs, (unsigned)min(k, vl)); };
                auto step = [&](const RecV& ca, const RecV& cb, int k) {
                    const Tri ta = tri_eval<PH, S3>(ca, frame_ok, sh.sx, sh.sy, sh.sz, tx, ty, tz, tnx,
                                                tny, tnz, tn_up, rst, inv_dr, n_d, two_cyc, nr,
                                                lossless, prisms, np, kp, st);
                    Tri tb;
                    tb.valid = false;
                    if (k + NTF < V)
                        tb = tri_eval<PH, S3>(cb, frame_ok, sh.sx, sh.sy, sh.sz, tx, ty, tz, tnx, tny,
                                          tnz, tn_up, rst, inv_dr, n_d, two_cyc, nr, lossless,
                                          prisms, np, kp, st);
                    if (ta.valid) tri_add<ACC>(ta, base, len, nr, winA, rowA, flag);
                    if (tb.valid) tri_add<ACC>(tb, base, len, nr, winA, rowA, flag);
                };
                RecV a0 = rec_at(threadIdx.x), b0 = rec_at(threadIdx.x + NTF);
                for (int k = threadIdx.x; k < V; k += 4 * NTF) {
                    const RecV a1 = rec_at(k + 2 * NTF), b1 = rec_at(k + 3 * NTF);
                    step(a0, b0, k);
                    if (k + 2 * NTF >= V) break;
                    a0 = rec_at(k + 4 * NTF);
                    b0 = rec_at(k + 5 * NTF);
                    step(a1, b1, k + 2 * NTF);
                }
            }
            __syncthreads();
            for (int k = threadIdx.x; k < len; k += NTF) {
                const int b = base + k;
                if (b >= nr) break;
                rowA[b] = win_get<ACC>(winA, k);
            }
        }
        // the next pair zeroes the window this one has just read
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// 7b. shot-pair target kernel
// ---------------------------------------------------------------------------
// Of a facet x target term only the source leg depends on the shot: Itd and
// the refracted gain F Dmax / pi^4 of the facet record.  The target leg (Ttd,
// the direction facet -> target in the facet frame, the four sinc factors of
// facet_G(Tth, Tph) facet_G(TargetTh, Tph), the loss factor, the path below
// a second interface) depends on the facet and the target alone.
// Consecutive shots of a track light nearly the same facets, so
// k_target_accum_pair takes one target with two consecutive shots, evaluates
// the target leg once per facet that either lights, and adds one term into
// each shot's bin window: the arithmetic of k_target_accum_fast's PH_F64 /
// ACC_F64 path, with the product of the gain and the four sinc factors
// formed as gain x (sincs x loss) instead of (gain x sincs) x loss.

// the target leg of one facet: Ttd, its optical path RT = n2 Ttd (+ (n3 -
// n2) times the part below a second interface) and its amplitude times its
// phasor exp(-i 2 pi RT / lam'), without the shot's gain
struct TLeg {
    double Ttd, RT;
    float2 Gc;
};

// one block record in registers
struct BRecV {
    float4 qa, qb, qc;
    float qd;
    double2 itd;
    float4 src;
};
// the records of one block, based at its first (uniform over the workgroup);
// record u at 32-bit byte offsets (u < 2^27)
struct BRecBlk {
    const char* qa;
    const char* qb;
    const char* qc;
    const char* qd;
    const char* itd;
    const char* src;
};
__device__ __forceinline__ BRecV load_brec(const BRecBlk& rb, unsigned u) {
    BRecV v;
    const unsigned o16 = u << 4;
    v.qa = *reinterpret_cast<const float4*>(rb.qa + o16);
    v.qb = *reinterpret_cast<const float4*>(rb.qb + o16);
    v.qc = *reinterpret_cast<const float4*>(rb.qc + o16);
    v.qd = *reinterpret_cast<const float*>(rb.qd + (u << 2));
    v.itd = *reinterpret_cast<const double2*>(rb.itd + o16);
    v.src = *reinterpret_cast<const float4*>(rb.src + o16);
    return v;
}

template <bool S3>
__device__ __forceinline__ TLeg tri_leg(const BRecV& rv, bool frame_ok, double tx, double ty,
                                        double tz, float tnx, float tny, float tnz, bool tn_up,
                                        bool lossless, const erp_prism* __restrict__ prisms,
                                        int np, const KPar& kp, const Strat& st) {
    TLeg o;
    const float fx = rv.qa.x, fy = rv.qa.y, fz = rv.qa.z;
    const double dx = tx - (double)fx, dy = ty - (double)fy, dz = tz - (double)fz;
    const double t2 = fma(dx, dx, fma(dy, dy, dz * dz));
    // Ttd = sqrt(t2) by one Newton step from the hardware reciprocal root
    // (t2 lies far from the denormal range rsqrtf guards against)
    const float rs = __builtin_amdgcn_rsqf((float)t2);
    const double rd = (double)rs, hr = 0.5 * rd;
    const double y0 = t2 * rd;
    o.Ttd = fma(fma(-y0, y0, t2), hr, y0);
    o.RT = o.Ttd * kp.n_d;
    if (S3 && st.deep) o.RT = fma(o.Ttd * below_frac(fz, st), st.dn, o.RT);
    const float nx = rv.qb.x, ny = rv.qb.y, nz = rv.qb.z;
    const float ux = rv.qb.w, uy = rv.qc.x, uz = rv.qc.y;
    const float vx = rv.qc.z, vy = rv.qc.w, vz = rv.qd;
    // unit vector facet -> target in FP64 (1/Ttd = rd (1.5 - t2 rd^2 / 2))
    const double invd = rd * fma(-y0, hr, 1.5);
    const double ix = dx * invd, iy = dy * invd, iz = dz * invd;
    const float Ix = (float)ix, Iy = (float)iy, Iz = (float)iz;
    const float2 csT = cos_sin_hw(ix, iy, iz, nx, ny, nz);
    const float cT = csT.x, sT = csT.y;
    float Iu = Ix * ux + Iy * uy + Iz * uz;
    float Iv = Ix * vx + Iy * vy + Iz * vz;
    if (!frame_ok) {
        const float2 e = uv_dots_exact(ix, iy, iz, ux, uy, uz, vx, vy, vz);
        Iu = e.x;
        Iv = e.y;
    }
    const float sTT = tn_up ? __builtin_amdgcn_sqrtf(fmaxf(0.f, (float)((1.0 + iz) * (1.0 - iz))))
                            : cos_sin_hw(ix, iy, iz, tnx, tny, tnz).y;
    // azimuth (cos Tph, sin Tph) = (Iv, Iu) / |(Iu, Iv)|, (1, 0) below 1e-30
    // (where every sinc argument is < 1e-6 whatever the azimuth)
    const float q2 = Iu * Iu + Iv * Iv;
    const bool az = q2 > 1e-30f;
    const float rinv = az ? __builtin_amdgcn_rsqf(q2) : 0.f;
    const float cph = az ? Iv * rinv : 1.f;
    const float sph = Iu * rinv;
    // pi-sincs from half their arguments h = x / 2 in revolutions, where the
    // hardware sine's own reduction is as accurate as an explicit one
    // (1.2e-7 absolute up to |h| = 260, checked against double on gfx950;
    // the host takes this kernel only for pi fs / lam <= 512, so |h| <= 256):
    // sin(pi |x|) / |x| = sin(2 pi |h|) / (2 |h|), the four factors 2 in
    // the final 1/16
    const float a = 0.5f * kp.piLlam;
    float n1, d1, n2, d2, n3, d3, n4, d4;
    auto pisinc_h = [](float h, float* num, float* den) {
        const float ha = fmaxf(fabsf(h), 5e-7f);
        *num = __builtin_amdgcn_sinf(ha);
        *den = ha;
    };
    pisinc_h(a * (sT * cph), &n1, &d1);
    pisinc_h(a * (sT * sph), &n2, &d2);
    pisinc_h(a * sTT * cph, &n3, &d3);
    pisinc_h(a * sTT * sph, &n4, &d4);
    const float den = fabsf((d1 * d2) * (d3 * d4));
    float g = fabsf((n1 * n2) * (n3 * n4)) * __builtin_amdgcn_rcpf(den) * (0.0625f * rs);
    if (S3 && !lossless && st.deep && (double)fz > st.zi) g *= st.tsq;
    if (!lossless) {
        // exp(-2 (alpha path + (ks cT)^2)) as exp2 with -2 log2(e) folded into
        // the coefficients
        constexpr float m2l2e = -2.f * 1.44269504088896341f;
        float expo = 0.f, tot = 0.f;
        for (int k = 0; k < np; ++k) {
            const float d = prism_len(fx, fy, fz, Ix, Iy, Iz, prisms[k], (float)o.Ttd);
            tot += d;
            expo += (m2l2e * prisms[k].alpha) * d;
        }
        expo = fmaf(m2l2e * kp.alpha2, (float)o.Ttd - tot, expo);
        const float kc2 = (m2l2e * kp.ks * kp.ks) * cT;
        g *= __builtin_amdgcn_exp2f(fmaf(kc2, cT, expo));
    }
    const float2 ph = phasor_rc_hw(o.RT, kp.inv_cyc);
    o.Gc = make_float2(ph.x * g, ph.y * g);
    return o;
}

// the source leg of one shot: range Itd + RT, bin and weight as in tri_eval
// (the exact FP64 sum within 1e-7 bin of a bin edge), the term the shot's
// gain x source phasor times the target leg's
template <bool S3>
__device__ __forceinline__ Tri tri_shot(const TLeg& g, double itd, float2 src, double rst,
                                        float fx, float fy, float fz, double tx, double ty,
                                        double tz, const KPar& kp, const Strat& st, int nr) {
    Tri o;
    double rngt = itd + g.RT;
    const double xq = fma(rngt, kp.inv_dr, -rst * kp.inv_dr);
    o.b = (int)xq;
    o.w = xq - (double)o.b;
    if (o.w < 1e-7 || o.w > 1.0 - 1e-7) {
        // the facet's coordinates pass through an empty asm statement, so
        // that the exact root below stays on this rare path (the two shots'
        // copies of it would otherwise be merged and hoisted out of it)
        float ex = fx, ey = fy, ez = fz;
        asm volatile("" : "+v"(ex), "+v"(ey), "+v"(ez));
        const double dx = tx - (double)ex, dy = ty - (double)ey, dz = tz - (double)ez;
        const double Te = dist3(dx, dy, dz);
        rngt = madd_nc(itd, Te, kp.n_d);
        if (S3 && st.deep) {
#pragma clang fp contract(off)
            const double L3 = Te * below_frac_exact(ez, st);
            rngt = madd_nc(rngt, L3, st.dn);
        }
        o.valid = bin_of(rngt, rst, kp.drd, nr, &o.b, &o.w);
    } else {
        o.valid = (xq > -1.0) && (xq < (double)nr);
    }
    o.c = cmul(src, g.Gc);
    return o;
}

// (target, row of the first shot, row of the second, block) per work item;
// a row of -1: that shot does not take the target
constexpr int NTB = 1024;  // threads of k_target_accum_pair (one workgroup per CU)
// bins of each of its two FP64 windows: 2 x 79 KB of the CU's 160 KB of LDS.
// On the 2048^2 fBm workload every pair's bounds (the nearer of the shot's
// box and bounding-sphere bounds) fit 5056 bins.
constexpr int TWB = 5056;

// MT: a multi-track batch: an item's block is two rows of one track, whose
// targets, parameters and facet frames it uses; the item count is read from
// device memory (the launch is replayed from a graph).
// DET: the deterministic mode.  Each shot's terms are summed as 64-bit fixed
// point, as in the single-shot kernel's deterministic path.  The scale comes
// from a bound of the shot's own terms (its lit facets only), so a row's
// bits do not depend on the shot it is paired with, or on whether it is
// paired at all; a pair whose window cannot hold its bounds is summed in
// several passes of one window each.
template <bool S3, bool MT = false, bool DET = false>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, NTB)))
k_target_accum_pair(const float* __restrict__ tg0, int64_t Tpad0, const int4* __restrict__ items,
                    int nitems0, int* __restrict__ next_item, const erp_shot* __restrict__ shots,
                    const int64_t* __restrict__ ubase, const int* __restrict__ Ucnt, BRec rec,
                    const Bounds* __restrict__ bounds, const erp_prism* __restrict__ prisms,
                    int np, KPar kp0, float2* __restrict__ outA, int* __restrict__ pbase,
                    int* __restrict__ pwin, int frame_ok0, int* __restrict__ oow,
                    const TrackK* __restrict__ trk, const RowK* __restrict__ rows,
                    const int* __restrict__ nitems_dev, const int* __restrict__ Vs) {
    __shared__ double winS[4 * TWB];
    __shared__ int s_it;
    __shared__ float s_m[NTB / 64][2];
    const int nitems = MT ? *nitems_dev : nitems0;
    // the windows are zero between items (and passes): each flush clears
    // what it read, and terms only reach the bins below a window's length
    for (int i = threadIdx.x; i < 4 * TWB; i += NTB) winS[i] = 0.0;
    // persistent: each workgroup takes items from a counter (cleared by the
    // host before the launch) until none is left; the next item's index is
    // fetched while the current one runs
    if (threadIdx.x == 0) s_it = atomicAdd(next_item, 1);
    __syncthreads();
    // the item index is uniform over the workgroup: what derives from it
    // stays in scalar registers
    int it = __builtin_amdgcn_readfirstlane(s_it);
    while (it < nitems) {
        int nxt = 0;
        if (threadIdx.x == 0) nxt = atomicAdd(next_item, 1);
        const int4 item = items[it];
        const int t = item.x, blk = item.w;
        const int row[2] = {item.y, item.z};
        const TrackK* tkp = MT ? &trk[rows[2 * blk].trk] : nullptr;
        const KPar& kp = MT ? tkp->kp : kp0;
        const float* __restrict__ tg = MT ? tkp->tg : tg0;
        const int64_t Tpad = MT ? tkp->Tpad : Tpad0;
        const int frame_ok = MT ? tkp->frame_ok : frame_ok0;
        const int nr = kp.nr;
        const bool lossless = kp.mode & ERP_MODE_LOSSLESS;
        const float txf = tg[t], tyf = tg[Tpad + t], tzf = tg[2 * Tpad + t];
        const float tnx = tg[3 * Tpad + t], tny = tg[4 * Tpad + t], tnz = tg[5 * Tpad + t];
        const double tx = txf, ty = tyf, tz = tzf;
        const bool tn_up = tnx == 0.f && tny == 0.f && tnz == 1.f;
        const Strat st{S3 && (double)tzf < kp.zi, kp.zi, (double)tzf, kp.dn, kp.tsq};
        const double n_lo = st.deep ? kp.n_lo : kp.n_d, n_hi = st.deep ? kp.n_hi : kp.n_d;

        // each shot's window (as k_target_accum_fast's); a shot without the
        // target has an empty one
        int base[2], len[2], top[2];
        bool fits[2];
        double rst[2];
        float2* rowp[2];
        int* flag[2];
        bool spill = false;
        WinS win[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            win[k] = WinS{winS + 2 * k * TWB, winS + (2 * k + 1) * TWB};
            base[k] = 0;
            len[k] = 0;
            top[k] = 0;
            fits[k] = true;
            rst[k] = 0.0;
            rowp[k] = nullptr;
            flag[k] = nullptr;
            if (row[k] < 0) continue;
            const int s = 2 * blk + k;
            rst[k] = (double)shots[s].rst;
            const PairWin pw = pair_window(bounds[s], txf, tyf, tzf, n_lo, n_hi, rst[k], kp, nr,
                                           TWB, true);
            base[k] = pw.base;
            len[k] = pw.len;
            top[k] = pw.top;
            fits[k] = pw.fits;
            rowp[k] = outA + (size_t)row[k] * kp.P_fft;
            // out-of-window terms of a pair whose window covers its bounds
            // (deterministic: of every pair) are broken bounds (counted);
            // the others go to the row
            flag[k] = (DET || pw.fits) ? oow : nullptr;
            if (threadIdx.x == 0) {
                pbase[row[k]] = pw.fits ? pw.base : -1;
                pwin[row[k]] = pw.len;
            }
            if (!pw.fits) {
                spill = true;
                // bins outside what the window passes write: zero
                const int hi = DET ? max(pw.top, pw.base) : pw.base + TWB;
                for (int i = threadIdx.x; i < kp.row_live; i += NTB)
                    if (i < pw.base || i >= hi || i >= nr) rowp[k][i] = make_float2(0.f, 0.f);
            }
        }
        // the row's zeros land before any thread's out-of-window atomics
        // (none in the deterministic mode)
        if (!DET && spill) {
            __threadfence();
            __syncthreads();
        }

        const int U = Ucnt[blk];
        const int64_t r0 = ubase[blk];
        const BRecBlk rb{reinterpret_cast<const char*>(rec.qa + r0),
                         reinterpret_cast<const char*>(rec.qb + r0),
                         reinterpret_cast<const char*>(rec.qc + r0),
                         reinterpret_cast<const char*>(rec.qd + r0),
                         reinterpret_cast<const char*>(rec.itd + r0),
                         reinterpret_cast<const char*>(rec.src + r0)};
        const unsigned ul = (unsigned)max(U - 1, 0);
        // every union facet through add(k, term): one facet per thread and
        // step, the next step's record loaded before this step's arithmetic
        // into the other of two register sets (clamped, in-bounds indices:
        // no set is copied)
        auto for_terms = [&](auto&& add) {
            if (U <= 0) return;
            auto step = [&](const BRecV& a) {
                const TLeg g = tri_leg<S3>(a, frame_ok, tx, ty, tz, tnx, tny, tnz, tn_up,
                                           lossless, prisms, np, kp, st);
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const double itd = k ? a.itd.y : a.itd.x;
                    if (row[k] < 0 || !(itd >= 0.0)) continue;
                    const float2 src = k ? make_float2(a.src.z, a.src.w)
                                         : make_float2(a.src.x, a.src.y);
                    const Tri o = tri_shot<S3>(g, itd, src, rst[k], a.qa.x, a.qa.y, a.qa.z, tx,
                                               ty, tz, kp, st, nr);
                    if (o.valid) add(k, o);
                }
            };
            BRecV a0 = load_brec(rb, min((unsigned)threadIdx.x, ul));
            for (int u = threadIdx.x; u < U; u += 2 * NTB) {
                const BRecV a1 = load_brec(rb, min((unsigned)(u + NTB), ul));
                step(a0);
                if (u + NTB >= U) break;
                a0 = load_brec(rb, min((unsigned)(u + 2 * NTB), ul));
                step(a1);
            }
        };

        if constexpr (!DET) {
            for_terms([&](int k, const Tri& o) {
                tri_add<ACC_F64>(o, base[k], len[k], nr, win[k], rowp[k], flag[k]);
            });
            __syncthreads();
            // flush each window into its row and clear what was read
#pragma unroll
            for (int k = 0; k < 2; ++k)
                for (int i = threadIdx.x; i < len[k]; i += NTB) {
                    const int b = base[k] + i;
                    if (b < nr) rowp[k][b] = make_float2((float)win[k].x[i], (float)win[k].y[i]);
                    win[k].x[i] = 0.0;
                    win[k].y[i] = 0.0;
                }
        } else {
            // bound of each shot's terms: |component| <= |gain source
            // phasor| x (four pi-sincs, each <= pi, times the loss factor and
            // the split weights, each <= 1) / Ttd; max(|re|, |im|) x 1.4143
            // bounds the phasor's modulus
            float m[2] = {0.f, 0.f};
            for (int u = threadIdx.x; u < U; u += NTB) {
                const unsigned o16 = (unsigned)u << 4;
                const float4 qa = *reinterpret_cast<const float4*>(rb.qa + o16);
                const double2 itd = *reinterpret_cast<const double2*>(rb.itd + o16);
                const float4 sr = *reinterpret_cast<const float4*>(rb.src + o16);
                const double dx = tx - (double)qa.x, dy = ty - (double)qa.y, dz = tz - (double)qa.z;
                const float rs = __builtin_amdgcn_rsqf((float)fma(dx, dx, fma(dy, dy, dz * dz)));
                if (row[0] >= 0 && itd.x >= 0.0)
                    m[0] = fmaxf(m[0], fmaxf(fabsf(sr.x), fabsf(sr.y)) * rs);
                if (row[1] >= 0 && itd.y >= 0.0)
                    m[1] = fmaxf(m[1], fmaxf(fabsf(sr.z), fabsf(sr.w)) * rs);
            }
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                for (int o = 32; o > 0; o >>= 1) m[k] = fmaxf(m[k], __shfl_xor(m[k], o));
                if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6][k] = m[k];
            }
            __syncthreads();
            double scale[2], inv[2];
            // a shot of at least 1024 lit facets bounds every term below
            // 2^61 / 2048 = 2^50 in fixed point: det_fix51 applies
            bool fast[2];
            int npass = 1;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                float mk = 0.f;
                for (int w = 0; w < NTB / 64; ++w) mk = fmaxf(mk, s_m[w][k]);
                // 2x margin for the rounding of the sinc quotients; pi^4 =
                // 97.409...
                const int V = row[k] >= 0 ? Vs[2 * blk + k] : 0;
                scale[k] = det_scale((double)V, 2.0 * 1.4143 * (double)mk * 97.40909103400244);
                inv[k] = scale[k] > 0.0 ? 1.0 / scale[k] : 0.0;
                fast[k] = V >= 1024;
                if (row[k] >= 0 && !fits[k])
                    npass = max(npass, (top[k] - base[k] + TWB - 1) / TWB);
            }
            // shot k's fixed-point window: real parts at 2k TWB, imaginary
            // parts at (2k + 1) TWB of the LDS array
            unsigned long long* const wu0 = reinterpret_cast<unsigned long long*>(winS);
            auto wux = [&](int k) { return wu0 + 2 * k * TWB; };
            auto wuy = [&](int k) { return wu0 + (2 * k + 1) * TWB; };
            // one window per shot and pass: [base + p TWB, ...) below top
            for (int p = 0; p < npass; ++p) {
                int wb[2], wl[2];
                bool first[2], last[2];
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    wb[k] = base[k] + p * TWB;
                    wl[k] = row[k] < 0 ? 0 : fits[k] ? (p == 0 ? len[k] : 0)
                                                     : min(TWB, max(0, top[k] - wb[k]));
                    first[k] = p == 0;
                    last[k] = fits[k] || wb[k] + TWB >= top[k];
                }
                for_terms([&](int k, const Tri& o) {
                    const float wf = (float)o.w;
                    const float2 c0 = make_float2(o.c.x * (1.f - wf), o.c.y * (1.f - wf));
                    const float2 c1 = make_float2(o.c.x * wf, o.c.y * wf);
                    const int k0 = o.b - wb[k];
                    auto put = [&](int i, float2 c) {
                        if (fast[k]) lds_add2_fix51(wux(k), wuy(k), i, c, scale[k]);
                        else lds_add2_fix(wux(k), wuy(k), i, c, scale[k]);
                    };
                    if (k0 >= 0 && k0 < wl[k]) put(k0, c0);
                    else if ((o.b < base[k] && first[k]) || (o.b >= top[k] && last[k]))
                        atomicAdd(flag[k], 1);
                    if (o.b + 1 < nr) {
                        if (k0 + 1 >= 0 && k0 + 1 < wl[k])
                            put(k0 + 1, c1);
                        else if ((o.b + 1 < base[k] && first[k]) || (o.b + 1 >= top[k] && last[k]))
                            atomicAdd(flag[k], 1);
                    }
                });
                __syncthreads();
#pragma unroll
                for (int k = 0; k < 2; ++k)
                    for (int i = threadIdx.x; i < wl[k]; i += NTB) {
                        const int b = wb[k] + i;
                        if (b < nr)
                            rowp[k][b] = make_float2((float)det_unfix(wux(k)[i], inv[k]),
                                                     (float)det_unfix(wuy(k)[i], inv[k]));
                        wux(k)[i] = 0ull;
                        wuy(k)[i] = 0ull;
                    }
                if (p + 1 < npass) __syncthreads();
            }
        }
        if (threadIdx.x == 0) s_it = nxt;
        __syncthreads();
        it = __builtin_amdgcn_readfirstlane(s_it);
    }
}

// Union of the aperture masks of shots 2j and 2j + 1 of a batch (the last
// block of an odd batch: shot 2j alone) and its count per facet block; one
// workgroup of MASKW threads per (facet block, block of shots).  Rows of a
// multi-track batch (trk, rows): a padding row, or a block past its track's
// mesh, adds nothing (k_illum_mask leaves their masks unwritten).
__global__ void k_union_mask(const unsigned long long* __restrict__ mask, int nblk, int S,
                             unsigned long long* __restrict__ umask, int* __restrict__ ucnt,
                             const TrackK* __restrict__ trk, const RowK* __restrict__ rows) {
    const int blk = blockIdx.x, j = blockIdx.y, w = threadIdx.x;
    const int s0 = 2 * j;
    auto lit = [&](int s) {
        if (s >= S) return false;
        return !trk || (rows[s].shot >= 0 && blk < trk[rows[s].trk].nblk);
    };
    unsigned long long m = 0;
    if (lit(s0)) m = mask[((size_t)s0 * nblk + blk) * MASKW + w];
    if (lit(s0 + 1)) m |= mask[((size_t)(s0 + 1) * nblk + blk) * MASKW + w];
    umask[((size_t)j * nblk + blk) * MASKW + w] = m;
    __shared__ int c_s;
    if (w == 0) c_s = 0;
    __syncthreads();
    atomicAdd(&c_s, __popcll(m));
    __syncthreads();
    if (w == 0) ucnt[(size_t)j * nblk + blk] = c_s;
}

// block records not written by a shot read as "not lit": Itd -1
__global__ void k_fill_itd(double2* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = make_double2(-1.0, -1.0);
}


// S[s][k] += scale * sum over the shot's pairs of X[p][k] * Z[p][k]
__global__ void k_accum_spectra(float2* __restrict__ S, const float2* __restrict__ X,
                                const float2* __restrict__ Z, const int* __restrict__ seg_shot,
                                const int* __restrict__ seg_beg, const int* __restrict__ seg_end,
                                int P, float scale) {
    const int sg = blockIdx.y;
    const int s = seg_shot[sg], p0 = seg_beg[sg], p1 = seg_end[sg];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < P; k += gridDim.x * blockDim.x) {
        float2 acc = make_float2(0.f, 0.f);
        for (int p = p0; p < p1; ++p) {
            const float2 x = X[(size_t)p * P + k], z = Z[(size_t)p * P + k];
            acc.x += x.x * z.x - x.y * z.y;
            acc.y += x.x * z.y + x.y * z.x;
        }
        float2* d = S + (size_t)s * P + k;
        d->x += scale * acc.x;
        d->y += scale * acc.y;
    }
}

// Bluestein step 1: a[n] = c[n] * cscale * w[n] (n < L), 0 for L <= n < M
__global__ void k_blue_in(float2* __restrict__ a, int M, const float2* __restrict__ c, int P,
                          const float2* __restrict__ w, int L, float cscale, int rows) {
    const size_t n = (size_t)M * rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / M), k = (int)(i - (size_t)r * M);
        float2 v = make_float2(0.f, 0.f);
        if (k < L) {
            float2 x = c[(size_t)r * P + k];
            x.x *= cscale;
            x.y *= cscale;
            v = cmul(x, w[k]);
        }
        a[i] = v;
    }
}

// Bluestein step 2 (between DFT_L and IDFT_L): a[n] = mask[n] conj(conv[n])
__global__ void k_blue_mid(float2* __restrict__ a, int M, const unsigned char* __restrict__ mask,
                           int L, int rows) {
    const size_t n = (size_t)M * rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(i % M);
        float2 v = make_float2(0.f, 0.f);
        if (k < L && mask[k]) {
            v = a[i];
            v.y = -v.y;
        }
        a[i] = v;
    }
}

// Bluestein output + decimation: refr[i] = conj(w[2i] conv2[2i]) / L
__global__ void k_blue_out(float2* __restrict__ refr, int nr, const float2* __restrict__ a,
                           int M, const float2* __restrict__ w, float invL, int rows,
                           const int* __restrict__ live) {
    const size_t n = (size_t)nr * rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / nr), j = (int)(i - (size_t)r * nr);
        float2 v = make_float2(0.f, 0.f);
        if (live[r]) {
            v = cmul(w[2 * j], a[(size_t)r * M + 2 * j]);
            v = make_float2(v.x * invL, -v.y * invL);
        }
        refr[i] = v;
    }
}

// out = refl + refr (+ type-0 conv), zeros for shots with V == 0
__global__ void k_finalize(float2* __restrict__ out, const float2* __restrict__ refl,
                           const float2* __restrict__ refr, const float2* __restrict__ refr0,
                           const int* __restrict__ Vs, int nr, int rows,
                           int32_t* __restrict__ nvalid) {
    const size_t n = (size_t)nr * rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / nr);
        float2 v = make_float2(0.f, 0.f);
        if (Vs[r] > 0) {
            const float2 a = refl ? refl[i] : make_float2(0.f, 0.f);
            const float2 b = refr ? refr[i] : make_float2(0.f, 0.f);
            const float2 c = refr0 ? refr0[i] : make_float2(0.f, 0.f);
            v = make_float2(a.x + b.x + c.x, a.y + b.y + c.y);
        }
        out[i] = v;
        if (nvalid && (i % nr) == 0) nvalid[r] = Vs[r];
    }
}

// ---------------------------------------------------------------------------
// multi-track batches: the row-wise forms of the per-shot kernels above,
// each row taking its track's tables from the track descriptors
// ---------------------------------------------------------------------------
// dst[r][k] = src[r][k] * table[k]: which = 0 the track's surface range
// kernel spectrum (none: a track without a surface echo gets zeros),
// which = 1 its Bluestein spectrum
__global__ void k_mul_spec_rows(float2* __restrict__ dst, const float2* __restrict__ src, int P,
                                int rows, const TrackK* __restrict__ trk,
                                const RowK* __restrict__ rk, int which) {
    const size_t n = (size_t)P * rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const TrackK& t = trk[rk[i / P].trk];
        const float2* spec = which == 0 ? t.g_surf : t.blue_Bw;
        dst[i] = spec ? cmul(src[i], spec[i & (size_t)(P - 1)]) : make_float2(0.f, 0.f);
    }
}

__global__ void k_blue_in_rows(float2* __restrict__ a, int M, const float2* __restrict__ c, int P,
                               int L, float cscale, int rows, const TrackK* __restrict__ trk,
                               const RowK* __restrict__ rk) {
    const size_t n = (size_t)M * rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / M), k = (int)(i - (size_t)r * M);
        float2 v = make_float2(0.f, 0.f);
        if (k < L) {
            float2 x = c[(size_t)r * P + k];
            x.x *= cscale;
            x.y *= cscale;
            v = cmul(x, trk[rk[r].trk].blue_w[k]);
        }
        a[i] = v;
    }
}

__global__ void k_blue_mid_rows(float2* __restrict__ a, int M, int L, int rows,
                                const TrackK* __restrict__ trk, const RowK* __restrict__ rk) {
    const size_t n = (size_t)M * rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / M), k = (int)(i - (size_t)r * M);
        float2 v = make_float2(0.f, 0.f);
        if (k < L && trk[rk[r].trk].blue_mask[k]) {
            v = a[i];
            v.y = -v.y;
        }
        a[i] = v;
    }
}

__global__ void k_blue_out_rows(float2* __restrict__ refr, int nr, const float2* __restrict__ a,
                                int M, float invL, int rows, const int* __restrict__ live,
                                const TrackK* __restrict__ trk, const RowK* __restrict__ rk) {
    const size_t n = (size_t)nr * rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / nr), j = (int)(i - (size_t)r * nr);
        float2 v = make_float2(0.f, 0.f);
        if (live[r]) {
            v = cmul(trk[rk[r].trk].blue_w[2 * j], a[(size_t)r * M + 2 * j]);
            v = make_float2(v.x * invL, -v.y * invL);
        }
        refr[i] = v;
    }
}

// out = refl + refr (refr null: no target in the batch) into each row's
// place in its track's radargram (rows of shot -1 pad a batch and write
// nothing)
__global__ void k_finalize_rows(const float2* __restrict__ refl, const float2* __restrict__ refr,
                                const int* __restrict__ Vs, int nr, int rows,
                                const TrackK* __restrict__ trk, const RowK* __restrict__ rk) {
    const size_t n = (size_t)nr * rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / nr), j = (int)(i - (size_t)r * nr);
        const RowK row = rk[r];
        if (row.shot < 0) continue;
        const TrackK& t = trk[row.trk];
        float2 v = make_float2(0.f, 0.f);
        if (Vs[r] > 0) {
            const float2 b = refr ? refr[i] : make_float2(0.f, 0.f);
            v = make_float2(refl[i].x + b.x, refl[i].y + b.y);
        }
        t.out[(size_t)row.shot * nr + j] = v;
        if (t.nvalid && j == 0) t.nvalid[row.shot] = Vs[r];
    }
}

// pairs of one chunk of a multi-track batch: the target id from the row's
// list, the track's range kernel spectrum per pair, and the accumulation
// kernel's pair counter cleared; the chunk's pair count is on the device
__global__ void k_pair_targets_rows(int* __restrict__ ptgt, const int* __restrict__ pshot,
                                    const int* __restrict__ list1, int Tcap,
                                    const int* __restrict__ n_dev, int* __restrict__ next_pair,
                                    const float4** __restrict__ Gpair,
                                    const TrackK* __restrict__ trk,
                                    const RowK* __restrict__ rk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < *n_dev) {
        const int r = pshot[i];
        ptgt[i] = list1[(size_t)r * Tcap + ptgt[i]];
        Gpair[i] = reinterpret_cast<const float4*>(trk[rk[r].trk].g_targ_nat);
    }
    if (i == 0) *next_pair = 0;
}

// ---------------------------------------------------------------------------
// NoConv direct sums (reflRadarSignal / refrRadarSignal, radar_kernels.cu:
// 58-159, 201-299): one thread per (shot, bin), facets streamed through LDS.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NT) k_refl_direct(
    float2* __restrict__ refl, const erp_shot* __restrict__ shots, const int* __restrict__ Vs,
    const int64_t* __restrict__ rec_base, Rec rec, KPar kp) {
    const int s = blockIdx.y;
    const int ir = blockIdx.x * NT + threadIdx.x;
    const ShotK sh = load_shot(shots, s);
    const int V = Vs[s];
    const int64_t r0 = rec_base[s];
    __shared__ double it_s[NT];
    __shared__ float fre_s[NT];
    const double r = (double)sh.rst + (double)ir * kp.drd;
    const bool lossless = kp.mode & ERP_MODE_LOSSLESS;
    (void)lossless;
    // FP64 sums: V terms of either sign per bin (sinc sidelobes of facets at
    // other ranges) cancel to a bin far smaller than its largest terms
    double ax = 0.0, ay = 0.0;
    for (int k0 = 0; k0 < V; k0 += NT) {
        const int k = k0 + threadIdx.x;
        if (k < V) {
            it_s[threadIdx.x] = rec.itd[r0 + k];
            fre_s[threadIdx.x] = rec.F[r0 + k];  // fRe stored in F for this mode
        }
        __syncthreads();
        const int m = min(NT, V - k0);
        for (int j = 0; j < m; ++j) {
            const double itd = it_s[j];
            const float dl = (float)((r - itd) / (double)kp.rng_res);
            const float2 ph = phasor_cycles(-2.0 * itd * kp.inv_cyc);
            const float sc = sinc_f(dl) * fre_s[j];
            ax += (double)(ph.x * sc);
            ay += (double)(ph.y * sc);
        }
        __syncthreads();
    }
    if (ir < kp.nr) refl[(size_t)s * kp.nr + ir] = make_float2((float)ax, (float)ay);
}

__global__ void __launch_bounds__(NT) k_refr0_direct(
    float2* __restrict__ refr, const float* __restrict__ tg, int64_t Tpad,
    const int* __restrict__ list0, const int* __restrict__ tcnt, int Tcap,
    const erp_shot* __restrict__ shots, const int* __restrict__ Vs,
    const int64_t* __restrict__ rec_base, Rec rec, const erp_prism* __restrict__ prisms, int np,
    KPar kp) {
    const int s = blockIdx.y;
    const int ir = blockIdx.x * NT + threadIdx.x;
    const ShotK sh = load_shot(shots, s);
    const int V = Vs[s];
    const int n0 = tcnt[2 * s + 1];
    const int64_t r0 = rec_base[s];
    __shared__ double rng_s[NT];
    __shared__ float rc_s[NT];
    const double r = (double)sh.rst + (double)ir * kp.drd;
    double ax = 0.0, ay = 0.0;  // FP64 sums, as in k_refl_direct
    for (int ti = 0; ti < n0; ++ti) {
        const int t = list0[(size_t)s * Tcap + ti];
        const double tx = tg[t], ty = tg[Tpad + t], tz = tg[2 * Tpad + t];
        for (int k0 = 0; k0 < V; k0 += NT) {
            const int k = k0 + threadIdx.x;
            if (k < V) {
                const int64_t rr = r0 + k;
                const float fx = rec.geo[rr], fy = rec.geo[rec.cap + rr], fz = rec.geo[2 * rec.cap + rr];
                const float nx = rec.geo[3 * rec.cap + rr], ny = rec.geo[4 * rec.cap + rr],
                            nz = rec.geo[5 * rec.cap + rr];
                const double Itd = rec.itd[rr];
                const double dx = tx - fx, dy = ty - fy, dz = tz - fz;
                const double Ttd = dist3(dx, dy, dz);
                const double invd = 1.0 / Ttd;
                const double ix = dx * invd, iy = dy * invd, iz = dz * invd;
                const float Ix = (float)ix, Iy = (float)iy, Iz = (float)iz;
                const float2 csT = cos_sin_ref(ix, iy, iz, nx, ny, nz);
                const float cT = csT.x;
                float expo = 0.f, tot = 0.f;
                for (int p = 0; p < np; ++p) {
                    const float d = prism_len(fx, fy, fz, Ix, Iy, Iz, prisms[p], (float)Ttd);
                    tot += d;
                    expo += prisms[p].alpha * d;
                }
                expo += kp.alpha2 * ((float)Ttd - tot);
                const float kc = kp.ks * cT;
                const float fRfrC = rec.fRfrC[rr];
                float EI = expf(-4.f * expo) * fRfrC * expf(-4.f * kc * kc);
                float EO = EI;
                {
                    float Iu = Ix * rec.geo[6 * rec.cap + rr] + Iy * rec.geo[7 * rec.cap + rr] +
                               Iz * rec.geo[8 * rec.cap + rr];
                    float Iv = Ix * rec.geo[9 * rec.cap + rr] + Iy * rec.geo[10 * rec.cap + rr] +
                               Iz * rec.geo[11 * rec.cap + rr];
                    const float sT = csT.y;
                    if (!kp.ortho) {
                        const float2 e = uv_dots_exact(
                            ix, iy, iz, rec.geo[6 * rec.cap + rr], rec.geo[7 * rec.cap + rr],
                            rec.geo[8 * rec.cap + rr], rec.geo[9 * rec.cap + rr],
                            rec.geo[10 * rec.cap + rr], rec.geo[11 * rec.cap + rr]);
                        Iu = e.x;
                        Iv = e.y;
                    }
                    const float Iph = rec.ang[rec.cap + rr], Rth = rec.ang[2 * rec.cap + rr];
                    const float Tth = atan2f(sT, cT);
                    const float Tph = atan2f(Iu, Iv);
                    const float si = sinf(Rth - Tth), dph = Iph - Tph;
                    const float a1 = sinc_f(kp.fs_lam * si * cosf(dph)),
                                a2 = sinc_f(kp.fs_lam * si * sinf(dph));
                    const float vi = kp.fs2_lam * a1 * a2;
                    EI *= vi * vi;
                    const float sR = kp.sqrt_e2e1 * sinf(ERP_PI_F - Tth);
                    const float RrTh = (sR > 1.f) ? 1e4f : asinf(sR);
                    const float so = sinf(RrTh), dpo = Tph - Iph;
                    const float b1 = sinc_f(kp.fs_lam * so * cosf(dpo)),
                                b2 = sinc_f(kp.fs_lam * so * sinf(dpo));
                    const float vo = kp.fs2_lam * b1 * b2;
                    EO *= vo * vo;
                }
                const double sl = Itd + Ttd;
                const double sl2 = sl * sl;
                const float req = (float)((double)V * ((double)(kp.P * sh.grefr * sh.grefr * kp.fs *
                                                               kp.fs * kp.fs * kp.fs) /
                                                       (ERP_FOURPI5_SQ * sl2 * sl2)));
                // refrRadarSignal passes d_Tth to rerad_funct (sim.cu:930-934)
                rc_s[threadIdx.x] =
                    rerad_f(kp.rerad, atan2f(csT.y, cT)) * EI * EO * req;
                rng_s[threadIdx.x] = madd_nc(sl - Ttd, Ttd, kp.cc2);
            }
            __syncthreads();
            const int m = min(NT, V - k0);
            for (int j = 0; j < m; ++j) {
                const double rngt = rng_s[j];
                const float dl = (float)((r - rngt) / (double)kp.rng_res);
                const float2 ph = phasor_cycles(-2.0 * rngt * kp.inv_cyc);
                const float sc = sinc_f(dl) * rc_s[j];
                ax += (double)(ph.x * sc);
                ay += (double)(ph.y * sc);
            }
            __syncthreads();
        }
    }
    if (ir < kp.nr) refr[(size_t)s * kp.nr + ir] = make_float2((float)ax, (float)ay);
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
inline int grid_for(size_t n, int per = 256, int cap = 1 << 16) {
    size_t g = (n + per - 1) / per;
    if (g < 1) g = 1;
    if (g > (size_t)cap) g = cap;
    return (int)g;
}

KPar make_kpar(const erp_params& p, int P_fft) {
    KPar k;
    k.lam = p.lam; k.fs = p.fs; k.P = p.P; k.ks = p.ks; k.eps1 = p.eps_1; k.eps2 = p.eps_2;
    k.nu1 = p.nu_1; k.nu2 = p.nu_2; k.alpha1 = p.alpha1; k.alpha2 = p.alpha2;
    k.rng_res = p.rng_res; k.dr = p.dr; k.c = p.c; k.c1 = p.c_1; k.c2 = p.c_2;
    k.piLlam = (ERP_PI_F * p.fs) / p.lam;
    k.ortho = 0;
    k.Dmax = (4.f * ERP_PI_F * p.fs * p.fs) / (p.lam * p.lam);
    k.fs_lam = p.fs / p.lam;
    k.fs2_lam = p.fs * p.fs / p.lam;
    k.nf = sqrtf(p.eps_2);
    k.sqrtP = sqrtf(p.P);
    k.invK = (float)(1.0 / ERP_FOURPI5_SQ);
    k.sqrt_e1e2 = sqrtf(p.eps_1 / p.eps_2);
    k.sqrt_e2e1 = sqrtf(p.eps_2 / p.eps_1);
    k.drd = (double)p.dr;
    k.inv_dr = 1.0 / (double)p.dr;
    k.cos_ap = p.cos_ap;
    k.n_d = std::sqrt((double)p.eps_2);
    k.c1c2 = (double)p.c_1 / (double)p.c_2;
    k.cc2 = (double)p.c / (double)p.c_2;
    k.has3 = p.eps_3 > 0.f ? 1 : 0;
    k.zi = p.z_int;
    k.dn = 0.0;
    k.tsq = 1.f;
    k.n_lo = k.n_hi = k.n_d;
    if (k.has3) {
        const double n3 = std::sqrt((double)p.eps_3);
        const double rho = (k.n_d - n3) / (k.n_d + n3);
        k.dn = n3 - k.n_d;
        k.tsq = (float)std::sqrt(1.0 - rho * rho);
        k.n_lo = std::min(k.n_d, n3);
        k.n_hi = std::max(k.n_d, n3);
    }
    // phasor(r) = exp(-i (2 PI_F / lam) r) = exp(-i 2 pi r * inv_cyc)
    k.inv_cyc = (double)ERP_PI_F / (M_PI * (double)p.lam);
    k.lam_d = p.lam;
    k.B = p.B;
    k.smpl = p.smpl;
    k.nr = p.nr;
    k.P_fft = P_fft;
    k.row_live = P_fft;
    k.rerad = p.rerad_funct;
    k.mode = p.mode;
    k.seed = p.seed;
    k.win_shift = 0;
    return k;
}

// Range-compression kernel spectrum of (nr, dr, res, kind, rst) into dst:
// P complex64 in the order fft_forward produces, scaled by 1/P.
void build_chirp(erp_ctx* ctx, float2* dst, int nr, float dr, float res, int kind, float rst,
                 int P, const FftPlan& plan, hipStream_t st) {
    hipLaunchKernelGGL(k_gen_kernel, dim3((P + 255) / 256), dim3(256), 0, st, dst, P, nr,
                       (double)dr, (double)res, (double)rst, kind);
    ERP_CHECK_LAUNCH("k_gen_kernel");
    fft_forward(ctx, plan, dst, 1, st);
    hipLaunchKernelGGL(k_scale, dim3((P + 255) / 256), dim3(256), 0, st, dst, P, 1.0f / (float)P);
    ERP_CHECK_LAUNCH("k_scale");
}

// Range-compression kernel spectrum, cached per context.  natural = true
// returns it in natural frequency order (the per-pair LDS chain), otherwise
// in the order fft_forward produces.  Only kind-1 entries (one per receive
// window start) are evicted, so kind-0 pointers held by a run stay valid.
// layout: 0 the forward FFT's order, 1 natural, 2 natural in the layout of
// k_tpair_wl (wl_perm_table)
const float2* get_chirp(erp_ctx* ctx, int nr, float dr, float res, int kind, float rst, int P,
                        const FftPlan& plan, hipStream_t st, int layout = 0) {
    ChirpKey key{nr, dr, res, kind, kind == 1 ? rst : 0.f};
    std::lock_guard<std::recursive_mutex> lk(ctx->cache_mu);
    ChirpSpec* c = nullptr;
    for (ChirpSpec* e : ctx->chirps)
        if (e->key == key) c = e;
    if (!c) {
        c = new ChirpSpec();
        c->key = key;
        c->spec.ensure(sizeof(float2) * (size_t)P);
        build_chirp(ctx, c->spec.as<float2>(), nr, dr, res, kind, rst, P, plan, st);
        // other streams may use the new spectrum as soon as the lock is released
        ERP_HIP_TRY(hipStreamSynchronize(st));
        int n1 = 0;
        for (ChirpSpec* e : ctx->chirps) n1 += e->key.kind == 1;
        // evict only while no asynchronous job can hold a pointer into the cache
        if (n1 > 16 && ctx->active.load() == 0) {
            for (size_t i = 0; i < ctx->chirps.size(); ++i)
                if (ctx->chirps[i]->key.kind == 1) {
                    ERP_HIP_TRY(hipStreamSynchronize(st));
                    delete ctx->chirps[i];
                    ctx->chirps.erase(ctx->chirps.begin() + (long)i);
                    break;
                }
        }
        ctx->chirps.push_back(c);
    }
    if (layout == 0) return c->spec.as<float2>();
    if (!c->nat.p) {
        c->nat.ensure(sizeof(float2) * (size_t)P);
        fft_natural_order(plan, c->spec.as<float2>(), c->nat.as<float2>(), st);
        ERP_HIP_TRY(hipStreamSynchronize(st));
    }
    if (layout == 1) return c->nat.as<float2>();
    if (!c->wl.p) {
        c->wl.ensure(sizeof(float2) * (size_t)P);
        wl_perm_table(c->nat.as<float2>(), c->wl.as<float2>(), P, st);
        ERP_HIP_TRY(hipStreamSynchronize(st));
    }
    return c->wl.as<float2>();
}

// Bluestein tables of b (L, M, smpl, B set; w, Bw and mask sized) built on st
void build_blue(erp_ctx* ctx, BlueTables* b, hipStream_t st) {
    hipLaunchKernelGGL(k_blue_tables, dim3((b->M + 255) / 256), dim3(256), 0, st,
                       b->w.as<float2>(), b->Bw.as<float2>(), b->mask.as<unsigned char>(), b->L,
                       b->M, (double)b->smpl, (double)b->B);
    ERP_CHECK_LAUNCH("k_blue_tables");
    FftPlan pm = fft_plan(b->M);
    fft_forward(ctx, pm, b->Bw.as<float2>(), 1, st);
    hipLaunchKernelGGL(k_scale, dim3((b->M + 255) / 256), dim3(256), 0, st, b->Bw.as<float2>(),
                       b->M, 1.0f / (float)b->M);
    ERP_CHECK_LAUNCH("k_scale");
}

BlueTables* get_blue(erp_ctx* ctx, int L, float smpl, float B, hipStream_t st) {
    std::lock_guard<std::recursive_mutex> lk(ctx->cache_mu);
    for (BlueTables* b : ctx->blues)
        if (b->L == L && b->smpl == smpl && b->B == B) return b;
    BlueTables* b = new BlueTables();
    b->L = L;
    b->M = next_pow2(2 * (int64_t)L - 1);
    b->smpl = smpl;
    b->B = B;
    b->w.ensure(sizeof(float2) * (size_t)L);
    b->Bw.ensure(sizeof(float2) * (size_t)b->M);
    b->mask.ensure((size_t)L);
    build_blue(ctx, b, st);
    if (ctx->blues.size() > 8 && ctx->active.load() == 0) {
        ERP_HIP_TRY(hipStreamSynchronize(st));
        delete ctx->blues.front();
        ctx->blues.erase(ctx->blues.begin());
    }
    ctx->blues.push_back(b);
    ERP_HIP_TRY(hipStreamSynchronize(st));
    return b;
}

void zero(float2* p, size_t n, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_zero, dim3(grid_for(n)), dim3(256), 0, st, p, n);
    ERP_CHECK_LAUNCH("k_zero");
}

// windowed convolution of `rows` padded rows (P each) with spectrum g:
// data := FFT, *g, IFFT in place; then out[r][j] = data[r][j + nr - 1]
void window_conv(erp_ctx* ctx, const FftPlan& plan, float2* data, int rows, const float2* g,
                 float2* out, int nr, hipStream_t st) {
    if (rows <= 0) return;
    const int P = plan.N;
    fft_forward(ctx, plan, data, rows, st);
    hipLaunchKernelGGL(k_mul_spec, dim3(grid_for((size_t)P * rows)), dim3(256), 0, st, data, data,
                       g, P, rows);
    ERP_CHECK_LAUNCH("k_mul_spec");
    fft_inverse(ctx, plan, data, rows, st);
    hipLaunchKernelGGL(k_window, dim3(grid_for((size_t)nr * rows)), dim3(256), 0, st, out, nr,
                       data, P, nr, rows, 1.0f, 0);
    ERP_CHECK_LAUNCH("k_window");
}

// After a run: the count of target terms that fell outside a bin window
// their pair's bounds should have covered (k_target_accum_fast); any is a
// broken bound, whose terms were dropped, so the run fails.
void check_windows(const int* oow, hipStream_t st) {
    int n = 0;
    ERP_HIP_TRY(hipMemcpyAsync(&n, oow, sizeof(int), hipMemcpyDeviceToHost, st));
    ERP_HIP_TRY(hipStreamSynchronize(st));
    if (n != 0)
        throw ErpError{ERP_E_INTERNAL,
                       std::to_string(n) + " facet x target terms fell outside the bin window "
                                           "of their pair (a window bound failed)"};
}

// The (row, target) pairs of one target launch, taken in order from row
// cs, list index cj: at most C pairs, row r holding tc[r].  Segments are the
// runs of one row's pairs, slices runs of at most GS pairs of one segment.
// canon: a row whose pairs go on into the next launch is cut at a multiple
// of GS of its own pair index, so every row's slices are the same however
// the launches fall (deterministic mode, multi-track batches).
struct ChunkFill {
    std::vector<int> pshot, ptgt;       // per pair: row, index in the row's list
    std::vector<int> sg_s, sg_b, sg_e;  // per segment: row, first pair, end
    std::vector<int> gb, ge;            // per slice: first pair, end
    std::vector<int> gsb, gse;          // per segment: its slices
    int n() const { return (int)pshot.size(); }
};

void fill_chunk(const int* tc, int rows, int& cs, int& cj, int C, int GS, bool canon,
                ChunkFill& f) {
    f.pshot.clear(); f.ptgt.clear();
    f.sg_s.clear(); f.sg_b.clear(); f.sg_e.clear();
    f.gb.clear(); f.ge.clear(); f.gsb.clear(); f.gse.clear();
    canon = canon && C >= GS;
    while (cs < rows && f.n() < C) {
        const int rem = tc[cs] - cj;
        if (rem <= 0) { ++cs; cj = 0; continue; }
        int take = std::min(rem, C - f.n());
        if (canon && take < rem) take -= take % GS;
        if (take == 0) break;
        const int b = f.n();
        for (int j = 0; j < take; ++j) {
            f.pshot.push_back(cs);
            f.ptgt.push_back(cj + j);  // list index; resolved on the device
        }
        f.sg_s.push_back(cs);
        f.sg_b.push_back(b);
        f.sg_e.push_back(b + take);
        f.gsb.push_back((int)f.gb.size());
        for (int i = b; i < b + take; i += GS) {
            f.gb.push_back(i);
            f.ge.push_back(std::min(i + GS, b + take));
        }
        f.gse.push_back((int)f.gb.size());
        cj += take;
    }
}

// Work of the shot-pair target kernel for the blocks of two shots of a
// batch: per block, the union of its shots' target lists (each ascending)
// as (target, index in the first shot's list or -1, index in the second's
// or -1), ascending.  lists: S rows of pitch ints; tc: per shot its count
// at stride 2 (the target counts as k_target_select leaves them).
struct PairIdx {
    int t, i0, i1;
};
void merge_pair_lists(const int* lists, size_t pitch, const int* tc, int S,
                      std::vector<std::vector<PairIdx>>& out) {
    const int nbk = (S + 1) / 2;
    out.resize(nbk);
    for (int j = 0; j < nbk; ++j) {
        std::vector<PairIdx>& o = out[j];
        o.clear();
        const int* a = lists + (size_t)(2 * j) * pitch;
        const int na = tc[2 * (2 * j)];
        const int* b = 2 * j + 1 < S ? lists + (size_t)(2 * j + 1) * pitch : nullptr;
        const int nb = b ? tc[2 * (2 * j + 1)] : 0;
        int i = 0, k = 0;
        while (i < na || k < nb) {
            if (k >= nb || (i < na && a[i] < b[k])) {
                o.push_back({a[i], i, -1});
                ++i;
            } else if (i >= na || b[k] < a[i]) {
                o.push_back({b[k], -1, k});
                ++k;
            } else {
                o.push_back({a[i], i, k});
                ++i;
                ++k;
            }
        }
    }
}

// One launch of the shot-pair kernel: whole items in block order while
// their rows fit C.  Rows are laid out per (block, shot) segment: a block's
// first-shot rows of this launch in item order, then its second-shot rows.
struct PairChunk {
    std::vector<int4> items;          // (target, row of shot 0, row of shot 1, block)
    std::vector<int> rshot, ridx;     // per row: its shot, the target's index in its list
    std::vector<int> sg_s, gsb, gse;  // per segment: shot, its first and end slice
    std::vector<int> gb, ge;          // per slice of GS rows: first row, end
    int rows = 0;
    double triples = 0;               // facet x target terms (for the stage timer)
};
void fill_pair_chunk(const std::vector<std::vector<PairIdx>>& bl, int& jb, size_t& ib, int C,
                     int GS, const int* V, PairChunk& f) {
    f.items.clear();
    f.rshot.clear(); f.ridx.clear();
    f.sg_s.clear(); f.gsb.clear(); f.gse.clear(); f.gb.clear(); f.ge.clear();
    f.rows = 0;
    f.triples = 0;
    auto segment = [&](int shot, int b, int e) {
        f.sg_s.push_back(shot);
        f.gsb.push_back((int)f.gb.size());
        for (int i = b; i < e; i += GS) {
            f.gb.push_back(i);
            f.ge.push_back(std::min(i + GS, e));
        }
        f.gse.push_back((int)f.gb.size());
        f.triples += (double)(e - b) * V[shot];
    };
    while (jb < (int)bl.size()) {
        const std::vector<PairIdx>& its = bl[jb];
        if (ib >= its.size()) {
            ++jb;
            ib = 0;
            continue;
        }
        size_t e = ib;
        int n0 = 0, n1 = 0;
        while (e < its.size()) {
            const int a = its[e].i0 >= 0, b = its[e].i1 >= 0;
            if (f.rows + n0 + n1 + a + b > C) break;
            n0 += a;
            n1 += b;
            ++e;
        }
        if (e == ib) break;  // this launch is full
        const int r = f.rows;
        f.rshot.resize((size_t)(r + n0 + n1));
        f.ridx.resize((size_t)(r + n0 + n1));
        int c0 = 0, c1 = 0;
        for (size_t q = ib; q < e; ++q) {
            const PairIdx& p = its[q];
            int r0 = -1, r1 = -1;
            if (p.i0 >= 0) {
                r0 = r + c0++;
                f.rshot[r0] = 2 * jb;
                f.ridx[r0] = p.i0;
            }
            if (p.i1 >= 0) {
                r1 = r + n0 + c1++;
                f.rshot[r1] = 2 * jb + 1;
                f.ridx[r1] = p.i1;
            }
            f.items.push_back(make_int4(p.t, r0, r1, jb));
        }
        if (n0 > 0) segment(2 * jb, r, r + n0);
        if (n1 > 0) segment(2 * jb + 1, r + n0, r + n0 + n1);
        f.rows += n0 + n1;
        ib = e;
    }
}

// Deterministic launches of the shot-pair kernel: every shot's rows are cut
// only at multiples of GS rows of its own target index, so that its slices
// (and their sums, added to the shot in order) do not depend on the launches.
// Each shot of a block keeps its own place in the block's items.  A launch
// takes merged items as far as its room allows, then pads each shot's rows
// with items of that shot alone up to its next multiple of GS; the next
// launch resumes each shot where it stopped (the padded items then carry the
// other shot alone).  A block that cannot fit one merged item and its pads
// into a whole launch (launches of under 2 GS rows) is taken one shot at a
// time.  A row's values are the same either way.
struct PairCursor {
    int jb = 0;          // block
    size_t p0 = 0, p1 = 0;  // per shot of the block: its first item not yet taken
};
void fill_pair_chunk_canon(const std::vector<std::vector<PairIdx>>& bl, PairCursor& cur, int C,
                           int GS, const int* V, PairChunk& f) {
    f.items.clear();
    f.rshot.clear(); f.ridx.clear();
    f.sg_s.clear(); f.gsb.clear(); f.gse.clear(); f.gb.clear(); f.ge.clear();
    f.rows = 0;
    f.triples = 0;
    auto segment = [&](int shot, int b, int e) {
        f.sg_s.push_back(shot);
        f.gsb.push_back((int)f.gb.size());
        for (int i = b; i < e; i += GS) {
            f.gb.push_back(i);
            f.ge.push_back(std::min(i + GS, e));
        }
        f.gse.push_back((int)f.gb.size());
        f.triples += (double)(e - b) * V[shot];
    };
    const bool aligned = C >= GS;
    std::vector<int> c0, c1;  // rows of each shot before item q
    while (cur.jb < (int)bl.size()) {
        const std::vector<PairIdx>& its = bl[cur.jb];
        const size_t ni = its.size();
        c0.assign(ni + 1, 0);
        c1.assign(ni + 1, 0);
        for (size_t q = 0; q < ni; ++q) {
            c0[q + 1] = c0[q] + (its[q].i0 >= 0);
            c1[q + 1] = c1[q] + (its[q].i1 >= 0);
        }
        const int m0 = c0[ni], m1 = c1[ni];
        const int a0 = c0[std::min(cur.p0, ni)], a1 = c1[std::min(cur.p1, ni)];
        if (a0 >= m0 && a1 >= m1) {
            ++cur.jb;
            cur.p0 = cur.p1 = 0;
            continue;
        }
        const int room = C - f.rows;
        // rows of each shot when the merged items end at item e, with its pad
        auto pad = [&](int a, int m) {
            return (!aligned || a % GS == 0 || a >= m) ? 0 : std::min(GS - a % GS, m - a);
        };
        const size_t qs = std::min(cur.p0, cur.p1);
        size_t e = qs;
        int n0 = 0, n1 = 0, d0 = 0, d1 = 0;
        for (size_t q = ni; q > qs; --q) {
            const int x0 = c0[std::max(q, cur.p0)] - a0, x1 = c1[std::max(q, cur.p1)] - a1;
            const int y0 = pad(a0 + x0, m0), y1 = pad(a1 + x1, m1);
            if (x0 + x1 + y0 + y1 <= room && x0 + x1 > 0) {
                e = q;
                n0 = x0; n1 = x1; d0 = y0; d1 = y1;
                break;
            }
        }
        if (e == qs) {
            if (f.rows > 0) break;  // the next launch starts here
            // one shot at a time: the first shot with rows left, cut at a
            // multiple of GS
            const int k = a0 < m0 ? 0 : 1;
            const int left = k ? m1 - a1 : m0 - a0;
            int take = std::min(left, C);
            if (aligned && take < left) take -= take % GS;
            size_t& pk = k ? cur.p1 : cur.p0;
            f.rshot.resize((size_t)take);
            f.ridx.resize((size_t)take);
            int c = 0;
            for (; pk < ni && c < take; ++pk) {
                const PairIdx& p = its[pk];
                const int i = k ? p.i1 : p.i0;
                if (i < 0) continue;
                f.rshot[c] = 2 * cur.jb + k;
                f.ridx[c] = i;
                f.items.push_back(make_int4(p.t, k ? -1 : c, k ? c : -1, cur.jb));
                ++c;
            }
            segment(2 * cur.jb + k, 0, take);
            f.rows = take;
            break;
        }
        // shot 0's rows [r, r + n0 + d0), then shot 1's
        const int r = f.rows, w0 = n0 + d0, w1 = n1 + d1;
        f.rshot.resize((size_t)(r + w0 + w1));
        f.ridx.resize((size_t)(r + w0 + w1));
        int k0 = 0, k1 = 0;
        for (size_t q = qs; q < e; ++q) {
            const PairIdx& p = its[q];
            int r0 = -1, r1 = -1;
            if (q >= cur.p0 && p.i0 >= 0) {
                r0 = r + k0++;
                f.rshot[r0] = 2 * cur.jb;
                f.ridx[r0] = p.i0;
            }
            if (q >= cur.p1 && p.i1 >= 0) {
                r1 = r + w0 + k1++;
                f.rshot[r1] = 2 * cur.jb + 1;
                f.ridx[r1] = p.i1;
            }
            if (r0 >= 0 || r1 >= 0) f.items.push_back(make_int4(p.t, r0, r1, cur.jb));
        }
        cur.p0 = std::max(cur.p0, e);
        cur.p1 = std::max(cur.p1, e);
        for (; k0 < w0; ++cur.p0) {  // the pads
            const PairIdx& p = its[cur.p0];
            if (p.i0 < 0) continue;
            f.rshot[r + k0] = 2 * cur.jb;
            f.ridx[r + k0] = p.i0;
            f.items.push_back(make_int4(p.t, r + k0, -1, cur.jb));
            ++k0;
        }
        for (; k1 < w1; ++cur.p1) {
            const PairIdx& p = its[cur.p1];
            if (p.i1 < 0) continue;
            f.rshot[r + w0 + k1] = 2 * cur.jb + 1;
            f.ridx[r + w0 + k1] = p.i1;
            f.items.push_back(make_int4(p.t, -1, r + w0 + k1, cur.jb));
            ++k1;
        }
        if (w0 > 0) segment(2 * cur.jb, r, r + w0);
        if (w1 > 0) segment(2 * cur.jb + 1, r + w0, r + w0 + w1);
        f.rows += w0 + w1;
        if (e < ni) break;  // the launch is full
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// mesh from a height grid: positions, central-difference normals and
// cross-product tangents in FP64 with the operation order of gen_normals
// (simple_surfaces.py:3-39: numpy's elementwise arithmetic, np.cross,
// np.linalg.norm over three components), rounded once to float
// ---------------------------------------------------------------------------
__global__ void k_mesh_from_heights(const double* __restrict__ z, int64_t nx, int64_t ny,
                                    double fs, double ox, double oy, float* __restrict__ soa,
                                    int64_t Fpad) {
#pragma clang fp contract(off)
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nx * ny) return;
    const int64_t j = f / nx, i = f - j * nx;
    const double two_fs = 2.0 * fs;
    // d/dx at column c (interior columns only)
    auto ddx = [&](int64_t c) { return (z[j * nx + c + 1] - z[j * nx + c - 1]) / two_fs; };
    auto ddy = [&](int64_t r) { return (z[(r + 1) * nx + i] - z[(r - 1) * nx + i]) / two_fs; };
    double dx = 0.0, dy = 0.0;
    if (nx > 2) dx = ddx(i == 0 ? 1 : (i == nx - 1 ? nx - 2 : i));
    if (ny > 2) dy = ddy(j == 0 ? 1 : (j == ny - 1 ? ny - 2 : j));
    const double norms = sqrt(dx * dx + dy * dy + 1.0);
    const double n0 = -dx / norms, n1 = -dy / norms, n2 = 1.0 / norms;
    double r0 = 1.0, r1 = 0.0, r2 = 0.0;
    if (fabs(n0) > 0.9) {
        r0 = 0.0;
        r1 = 1.0;
    }
    double u0 = r1 * n2 - r2 * n1, u1 = r2 * n0 - r0 * n2, u2 = r0 * n1 - r1 * n0;
    const double un = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
    u0 = u0 / un;
    u1 = u1 / un;
    u2 = u2 / un;
    const double v0 = n1 * u2 - n2 * u1, v1 = n2 * u0 - n0 * u2, v2 = n0 * u1 - n1 * u0;
    const double vals[12] = {(double)i * fs + ox, (double)j * fs + oy, z[f], n0, n1, n2,
                             u0, u1, u2, v0, v1, v2};
#pragma unroll
    for (int a = 0; a < 12; ++a) soa[(size_t)a * Fpad + f] = (float)vals[a];
}

void mesh_from_heights(const double* heights, int64_t nx, int64_t ny, double fs, double ox,
                       double oy, float* soa, int64_t Fpad, hipStream_t st) {
    const int64_t F = nx * ny;
    if (F == 0) return;
    hipLaunchKernelGGL(k_mesh_from_heights, dim3((unsigned)((F + 255) / 256)), dim3(256), 0, st,
                       heights, nx, ny, fs, ox, oy, soa, Fpad);
    ERP_CHECK_LAUNCH("k_mesh_from_heights");
}

using TargetAccumFn = decltype(&k_target_accum_fast<false>);

TargetAccumFn pick_target_accum(bool det, int ph, int acc, bool s3) {
    if (s3) {
        // second subsurface medium: its own instantiations, so that runs
        // without it keep the kernel as it was tuned
        if (det) return k_target_accum_fast<true, PH_F64, ACC_F64, true>;
        switch (ph * 3 + acc) {
        case PH_F64 * 3 + ACC_F32: return k_target_accum_fast<false, PH_F64, ACC_F32, true>;
        case PH_F64 * 3 + ACC_COMP: return k_target_accum_fast<false, PH_F64, ACC_COMP, true>;
        case PH_F32 * 3 + ACC_F64: return k_target_accum_fast<false, PH_F32, ACC_F64, true>;
        case PH_F32 * 3 + ACC_F32: return k_target_accum_fast<false, PH_F32, ACC_F32, true>;
        case PH_F32 * 3 + ACC_COMP: return k_target_accum_fast<false, PH_F32, ACC_COMP, true>;
        case PH_DF * 3 + ACC_F64: return k_target_accum_fast<false, PH_DF, ACC_F64, true>;
        case PH_DF * 3 + ACC_F32: return k_target_accum_fast<false, PH_DF, ACC_F32, true>;
        case PH_DF * 3 + ACC_COMP: return k_target_accum_fast<false, PH_DF, ACC_COMP, true>;
        default: return k_target_accum_fast<false, PH_F64, ACC_F64, true>;
        }
    }
    if (det) return k_target_accum_fast<true>;
    switch (ph * 3 + acc) {
    case PH_F64 * 3 + ACC_F32: return k_target_accum_fast<false, PH_F64, ACC_F32>;
    case PH_F64 * 3 + ACC_COMP: return k_target_accum_fast<false, PH_F64, ACC_COMP>;
    case PH_F32 * 3 + ACC_F64: return k_target_accum_fast<false, PH_F32, ACC_F64>;
    case PH_F32 * 3 + ACC_F32: return k_target_accum_fast<false, PH_F32, ACC_F32>;
    case PH_F32 * 3 + ACC_COMP: return k_target_accum_fast<false, PH_F32, ACC_COMP>;
    case PH_DF * 3 + ACC_F64: return k_target_accum_fast<false, PH_DF, ACC_F64>;
    case PH_DF * 3 + ACC_F32: return k_target_accum_fast<false, PH_DF, ACC_F32>;
    case PH_DF * 3 + ACC_COMP: return k_target_accum_fast<false, PH_DF, ACC_COMP>;
    default: return k_target_accum_fast<false>;
    }
}

// Device bytes of one run, following the allocations of erp_run_impl: the
// work buffers of one batch of up to kShotsPerBatch shots (facet records counted for
// every facet illuminated by every shot of the batch, an upper bound) plus
// the ns x nr output and counts.
int64_t erp_footprint_impl(const erp_params* par, int64_t F, int64_t T, int64_t ns) {
    const int64_t nr = par->nr;
    const int64_t P = next_pow2(2 * nr - 1);
    const int64_t M = next_pow2(2 * (2 * nr - 1) - 1);
    const int64_t S = std::max<int64_t>(1, std::min<int64_t>(ns, kShotsPerBatch));
    const int64_t nblk = (F + FB - 1) / FB;
    const bool spec = par->mode & ERP_MODE_SPECULAR;
    int64_t b = 0;
    b += S * nblk * (MASKW * 8 + 8);          // mask + counts + offsets
    b += S * F * 136 + 2560;                  // facet records (136 B per facet-shot)
    b += S * P * 8 * 3 + S * M * 8;           // surface, spectral sum, type-0 rows; Bluestein
    b += S * nr * (16 + 8 + 16) + S * 4;      // FP64 surface bins, reflected, refracted
    if (T > 0) {
        const int64_t nbuf = spec ? 4 : 3;
        int64_t C = std::max<int64_t>(1, kPairRowBytes / (P * 8 * nbuf));
        C = std::min<int64_t>(C, kMaxPairsPerLaunch);
        const int64_t gcap = C / 16 + S + 2;
        b += S * T * 8 + S * 8;               // target lists and counts
        b += S * nblk * 32;                   // facet block boxes (bounding spheres)
        b += P * 8 * (C * nbuf + gcap);       // per-target chunk rows and partial sums
        b += 4 * (C * 4 + 2 * gcap + 3 * S + 16);
    }
    b += ns * nr * 8 + ns * 4;                // the radargram and the counts
    return b;
}

void erp_run_impl(erp_ctx* ctx, const erp_params* par, const erp_mesh* mesh,
                  const erp_targets* targets, const erp_shot* shots, int64_t ns,
                  const erp_prism* prisms, int64_t nprisms, float2* out_dev,
                  int32_t* nvalid_dev, uint32_t flags, hipStream_t st, WorkBufs& W,
                  StageTimer* timer, Arena* arena) {
    const erp_params& p = *par;
    const uint32_t mode = p.mode;
    const bool conv = mode & ERP_MODE_CONVOLUTION;
    const bool lin = mode & ERP_MODE_CONV_LINEAR;
    const bool spec = mode & ERP_MODE_SPECULAR;
    const int nr = p.nr;
    if (nr <= 0 || nr > 32767) throw ErpError{ERP_E_INVALID, "nr must be in [1, 32767]"};
    const uint32_t known = ERP_RUN_RAW_SURFACE | ERP_RUN_DETERMINISTIC | ERP_RUN_PHASE_F32 |
                           ERP_RUN_PHASE_DF | ERP_RUN_ACC_F32 | ERP_RUN_ACC_COMP;
    if (flags & ~known) throw ErpError{ERP_E_INVALID, "unknown run flags"};
    const bool det = flags & ERP_RUN_DETERMINISTIC;
    // precision-study arithmetic of the target accumulation
    const int ph = (flags & ERP_RUN_PHASE_F32) ? PH_F32 : (flags & ERP_RUN_PHASE_DF) ? PH_DF : PH_F64;
    const int acc = (flags & ERP_RUN_ACC_F32) ? ACC_F32 : (flags & ERP_RUN_ACC_COMP) ? ACC_COMP
                                                                                 : ACC_F64;
    if ((flags & ERP_RUN_PHASE_F32) && (flags & ERP_RUN_PHASE_DF))
        throw ErpError{ERP_E_INVALID, "PHASE_F32 and PHASE_DF are exclusive"};
    if ((flags & ERP_RUN_ACC_F32) && (flags & ERP_RUN_ACC_COMP))
        throw ErpError{ERP_E_INVALID, "ACC_F32 and ACC_COMP are exclusive"};
    if ((ph != PH_F64 || acc != ACC_F64) && (det || spec))
        throw ErpError{ERP_E_UNSUPPORTED,
                       "the precision-study flags apply to the non-specular target "
                       "accumulation without ERP_RUN_DETERMINISTIC"};
    if (!spec && !(conv && lin))
        throw ErpError{ERP_E_UNSUPPORTED,
                       "non-specular runs need convolution and convolution_linear "
                       "(sim.cu:431-439,778 otherwise write past the chirp buffer)"};
    const int64_t F = mesh->F;
    const int64_t T = targets ? targets->T : 0;
    bool any_typed = false, any_t0 = false;
    std::vector<unsigned char> htype;
    if (T > 0) {
        htype.resize(T);
        ERP_HIP_TRY(hipMemcpyAsync(htype.data(), targets->type.p, T, hipMemcpyDeviceToHost, st));
        ERP_HIP_TRY(hipStreamSynchronize(st));
        for (auto t : htype) (t ? any_typed : any_t0) = true;
    }
    if (any_typed && !(conv && lin))
        throw ErpError{ERP_E_UNSUPPORTED,
                       "targets of type != 0 need convolution and convolution_linear "
                       "(sim.cu:410,872,912 otherwise use unallocated buffers)"};
    if (p.eps_3 > 0.f && (spec || any_t0))
        throw ErpError{ERP_E_UNSUPPORTED,
                       "the second subsurface medium (eps_3) is defined for the plain "
                       "method's targets of type != 0"};
    if (F == 0 || ns == 0) {
        // no facets: no surface echo and no refracted path, so every trace is
        // zero and no facet is illuminated
        zero(out_dev, (size_t)ns * nr, st);
        if (nvalid_dev && ns > 0) {
            const std::vector<int32_t> z((size_t)ns, 0);
            ERP_HIP_TRY(hipMemcpyAsync(nvalid_dev, z.data(), sizeof(int32_t) * ns,
                                       hipMemcpyHostToDevice, st));
            ERP_HIP_TRY(hipStreamSynchronize(st));  // z is a host temporary
        }
        return;
    }

    const int P = next_pow2(2 * (int64_t)nr - 1);
    const FftPlan planP = fft_plan(P);
    KPar kp = make_kpar(p, P);
    kp.ortho = mesh->frame_ok ? 1 : 0;
    kp.win_shift = (int)ctx->tune_shift;
    const int L = 2 * nr - 1;
    const int nblk = (int)((F + FB - 1) / FB);
    // facet order of the surface kernels: tiles where a wide aperture lights
    // most of the mesh (each block's ranges then stay within its window);
    // rows where a narrow one lights a small disk (rows give more, smaller
    // lit blocks and a better balanced launch).  Measured on the 2048^2 fBm
    // mesh: 60 deg, tiles 4.2x faster; 10 deg, rows 1.7x faster.
    const int tile_nx = p.cos_ap < 0.94 ? mesh->tile_nx : 0;
    const bool need_rec = (T > 0) || (spec && !conv);
    const bool do_trace = !spec || (conv && lin);
    const float res_t = spec ? p.rng_res : 2.f * p.rng_res;  // sim.cu:853-858

    // chirp spectra: cached per context, or (arena) built into the region
    auto chirp = [&](float res, int layout, int kind = 0) -> const float2* {
        if (!arena) return get_chirp(ctx, nr, p.dr, res, kind, 0.f, P, planP, st, layout);
        float2* spec_ = arena->take_n<float2>((size_t)P, "chirp spectrum");
        build_chirp(ctx, spec_, nr, p.dr, res, kind, 0.f, P, planP, st);
        if (layout == 0) return spec_;
        float2* nat = arena->take_n<float2>((size_t)P, "chirp spectrum");
        fft_natural_order(planP, spec_, nat, st);
        if (layout == 1) return nat;
        float2* wl = arena->take_n<float2>((size_t)P, "chirp spectrum");
        wl_perm_table(nat, wl, P, st);
        return wl;
    };
    const float2* g_surf = nullptr;
    if (!spec || (conv && lin)) g_surf = chirp(p.rng_res, 0);
    const float2* g_targ = nullptr;
    if (T > 0 && conv && lin) g_targ = chirp(res_t, 0);
    // plain mode: per-pair spectral chain in one workgroup when P / 2 fits
    // LDS (nr below P / 2: its cropped variant, whose range kernel has the
    // kept window at P / 2 and whose rows are read up to P / 2)
    const bool half = !spec && T > 0 && conv && lin && target_half_supported(planP, nr);
    kp.row_live = half ? P / 2 : P;
    // the per-pair chain's spectrum: natural, or in the wave-local layout
    const bool wl = half && target_wl(planP);
    const float2* g_targ_nat = half ? chirp(res_t, wl ? 2 : 1, nr < P / 2 ? 2 : 0) : nullptr;
    // circular convolution of type-0 targets: one genChirp table per rx window
    // start; with an arena one table is rebuilt for each shot
    float2* g_circ = (arena && any_t0 && conv && !lin)
                         ? arena->take_n<float2>((size_t)P, "circular chirp spectrum")
                         : nullptr;
    BlueTables* blue = nullptr;
    BlueTables blue_local;
    FftPlan planM;
    if (any_typed) {
        if (arena) {
            blue = &blue_local;
            blue->L = L;
            blue->M = next_pow2(2 * (int64_t)L - 1);
            blue->smpl = p.smpl;
            blue->B = p.B;
            arena->lend(blue->w, sizeof(float2) * (size_t)L, "Bluestein chirp");
            arena->lend(blue->Bw, sizeof(float2) * (size_t)blue->M, "Bluestein spectrum");
            arena->lend(blue->mask, (size_t)L, "band mask");
            build_blue(ctx, blue, st);
        } else {
            blue = get_blue(ctx, L, p.smpl, p.B, st);
        }
        planM = fft_plan(blue->M);
    }

    // batch size: bounded by memory of per-shot buffers
    int Smax = kShotsPerBatch;
    if (ctx->tune_shots > 0) Smax = (int)std::min<int64_t>(Smax, ctx->tune_shots);
    if ((int64_t)P * 8 * 4 * Smax > (int64_t)1 << 30) Smax = std::max<int>(1, (int)(((int64_t)1 << 30) / ((int64_t)P * 32)));
    const int64_t Tpad = targets ? targets->Tpad : 0;

    // (shot, target) pairs per target launch and the per-pair buffer layout
    const bool fused = !spec && !half && (!planP.two_pass || nr % planP.N1 == 0);
    // pairs per multiply-accumulate group (fused) or per summed slice (half)
    const int GS = half ? target_half_slice() : 16;
    const int nbuf = half ? 2 : (fused ? 3 : (spec ? 4 : 3));
    int C = (int)std::max<int64_t>(1, kPairRowBytes / ((int64_t)P * 8 * nbuf));
    C = std::min<int>(C, kMaxPairsPerLaunch);
    if (ctx->tune_pairs > 0) C = (int)std::min<int64_t>(C, ctx->tune_pairs);
    // deterministic: launches of at least one slice, so that every shot's
    // rows are cut at multiples of GS of its own targets whatever the tuning
    const int Cmin_det = det && half ? GS : 1;
    C = std::max(C, Cmin_det);
    // shot-pair target kernel (k_target_accum_pair): plain runs of the
    // product's arithmetic whose targets all have a type != 0, unless
    // ERP_TUNE_SINGLE_SHOT asks for the single-shot kernel
    const bool pairk = half && ph == PH_F64 && acc == ACC_F64 && !any_t0 && C >= 2 &&
                       ctx->tune_single == 0 && kp.piLlam <= 512.f;

    // bytes of the facet records of cap illuminated facets (only the layouts
    // the run reads; each array 256-byte aligned, as the batch loop carves them)
    auto rec_need = [&](size_t cap) -> size_t {
        size_t b = Arena::round(cap * 8);
        if (spec || any_t0)
            b += Arena::round(cap * 48) + 2 * Arena::round(cap * 4) + Arena::round(cap * 12);
        if (!spec && any_typed) b += 3 * Arena::round(cap * 16) + Arena::round(cap * 4);
        return b;
    };
    // Allocation-free run: every batch buffer is lent from the arena before
    // the loop (the loop's ensure calls then find it large enough).  Shots per
    // batch and pairs per launch shrink until the fixed buffers take at most
    // three quarters of what is left; the facet records get the rest.
    if (arena) {
        auto layout = [&](int S_, int C_, bool lend) -> size_t {
            size_t tot = 0;
            auto buf = [&](DevBuf& b, size_t n, const char* what) {
                tot += Arena::round(std::max<size_t>(n, 1));
                if (lend) arena->lend(b, n, what);
            };
            const size_t s = (size_t)S_;
            buf(W.shots, sizeof(erp_shot) * s, "shot table");
            if (nprisms > 0) buf(W.prisms, sizeof(erp_prism) * (size_t)nprisms, "prism table");
            buf(W.cnt, sizeof(int) * s * nblk, "block counts");
            buf(W.off, sizeof(int) * s * nblk, "block offsets");
            buf(W.V, sizeof(int) * s, "illuminated counts");
            buf(W.Vd, sizeof(unsigned long long) * s * nblk * MASKW, "aperture mask");
            buf(W.bounds, sizeof(Bounds) * s, "shot bounds");
            buf(W.tcount, sizeof(int64_t) * s, "record bases");
            buf(W.psurf, sizeof(float2) * s * P, "surface traces");
            buf(W.srow, sizeof(float2) * (s * nr * 2 + (s + 1) / 2), "surface bins");
            buf(W.refl, sizeof(float2) * s * nr, "reflected traces");
            buf(W.out, sizeof(float2) * s * nr * 2, "refracted traces");
            if (any_typed) {
                buf(W.shotS, sizeof(float2) * s * P, "spectral sums");
                buf(W.blu, sizeof(float2) * s * planM.N, "Bluestein rows");
            }
            if (any_t0) buf(W.t0trace, sizeof(float2) * s * P, "type-0 traces");
            if (T > 0) {
                buf(W.tlist, sizeof(int) * s * (size_t)T * 2, "target lists");
                buf(W.tcnt, sizeof(int) * s * 2, "target counts");
                buf(W.bbox, sizeof(float4) * 2 * s * nblk, "facet block boxes");
            }
            size_t pb = 0;
            if (any_typed) {
                const size_t gc = (size_t)C_ / GS + s + 2;
                buf(W.tA, sizeof(float2) * (size_t)P * ((size_t)C_ * (nbuf - 1) + (half ? gc : 0)),
                    "pair rows");
                buf(W.tC, sizeof(float2) * (size_t)C_ * P, "pair products");
                pb = sizeof(int) * ((size_t)C_ * 4 + 2 * gc + 3 * s + 16);
            }
            if (any_t0) pb = std::max(pb, sizeof(int) * 2 * s * (size_t)T + 64);
            if (pb) buf(W.pairs, pb, "pair lists");
            if (any_typed) buf(W.oow, 256, "window check");
            if (pairk) {
                const size_t nb = (s + 1) / 2;
                buf(W.umask, sizeof(unsigned long long) * nb * nblk * MASKW, "union masks");
                buf(W.ucnt, sizeof(int) * nb * nblk, "union counts");
                buf(W.uoff, sizeof(int) * nb * nblk, "union offsets");
                buf(W.U, sizeof(int) * nb, "union sizes");
                buf(W.ubase, sizeof(int64_t) * nb, "block record bases");
                buf(W.items, sizeof(int4) * (size_t)C_, "pair items");
            }
            return tot;
        };
        const size_t avail = arena->left();
        while (layout(Smax, C, false) > avail / 4 * 3) {
            const int Cmin = std::max(Cmin_det, pairk ? 2 : 1);
            if (any_typed && C > 256) C /= 2;
            else if (Smax > 1) Smax = (Smax + 1) / 2;
            else if (any_typed && C > Cmin) C = std::max(C / 2, Cmin);
            else
                throw ErpError{ERP_E_NOMEM,
                               "job too large for the allocation-free entry point: one shot's "
                               "work buffers need " + std::to_string(layout(1, 1, false)) +
                                   " bytes, the scratch has " + std::to_string(avail) +
                                   " left after the tables"};
        }
        layout(Smax, C, true);
        if (need_rec) arena->lend(pairk ? W.brec : W.rec, arena->left(), "facet records");
    }

    W.shots.ensure(sizeof(erp_shot) * (size_t)Smax);
    if (nprisms > 0) {
        W.prisms.ensure(sizeof(erp_prism) * (size_t)nprisms);
        ERP_HIP_TRY(hipMemcpyAsync(W.prisms.p, prisms, sizeof(erp_prism) * nprisms,
                                   hipMemcpyHostToDevice, st));
    }
    W.cnt.ensure(sizeof(int) * (size_t)Smax * nblk);
    W.off.ensure(sizeof(int) * (size_t)Smax * nblk);
    W.V.ensure(sizeof(int) * (size_t)Smax);
    W.Vd.ensure(sizeof(unsigned long long) * (size_t)Smax * nblk * MASKW);  // mask
    W.bounds.ensure(sizeof(Bounds) * (size_t)Smax);
    W.tcount.ensure(sizeof(int64_t) * (size_t)Smax);  // rec_base
    W.psurf.ensure(sizeof(float2) * (size_t)Smax * P);
    const size_t srow_f2 = (size_t)Smax * nr * 2 + ((size_t)Smax + 1) / 2;  // in float2 units
    W.srow.ensure(sizeof(float2) * srow_f2);
    double2* srow = W.srow.as<double2>();
    unsigned* smax = reinterpret_cast<unsigned*>(srow + (size_t)Smax * nr);
    W.refl.ensure(sizeof(float2) * (size_t)Smax * nr);
    W.out.ensure(sizeof(float2) * (size_t)Smax * nr * 2);  // refr + refr0
    if (any_typed) {
        W.shotS.ensure(sizeof(float2) * (size_t)Smax * P);
        W.blu.ensure(sizeof(float2) * (size_t)Smax * planM.N);
    }
    if (any_t0) W.t0trace.ensure(sizeof(float2) * (size_t)Smax * P);
    if (T > 0) {
        W.tlist.ensure(sizeof(int) * (size_t)Smax * T * 2);
        W.tcnt.ensure(sizeof(int) * (size_t)Smax * 2);
        W.bbox.ensure(sizeof(float4) * 2 * (size_t)Smax * nblk);
    }

    // terms of the plain target kernel that fall outside their pair's bin
    // window where the window holds every bin the pair's bounds allow
    const bool check_win = any_typed && (!spec || det);
    if (check_win) {
        W.oow.ensure(256);
        ERP_HIP_TRY(hipMemsetAsync(W.oow.p, 0, sizeof(int), st));
    }

    // the shot-pair kernel's buffers: union masks and records per block of
    // two shots
    const int nbk_max = (Smax + 1) / 2;
    if (pairk) {
        W.umask.ensure(sizeof(unsigned long long) * (size_t)nbk_max * nblk * MASKW);
        W.ucnt.ensure(sizeof(int) * (size_t)nbk_max * nblk);
        W.uoff.ensure(sizeof(int) * (size_t)nbk_max * nblk);
        W.U.ensure(sizeof(int) * (size_t)nbk_max);
        W.ubase.ensure(sizeof(int64_t) * (size_t)nbk_max);
    }
    std::vector<int> hU(nbk_max);
    std::vector<int64_t> hubase(nbk_max + 1);

    std::vector<int> hV(Smax), htc(2 * Smax);
    std::vector<int64_t> hbase(Smax + 1);
    int* dV = W.V.as<int>();
    int64_t* dbase = W.tcount.as<int64_t>();
    unsigned long long* dmask = W.Vd.as<unsigned long long>();
    Bounds* dbounds = W.bounds.as<Bounds>();
    float2* refr = W.out.as<float2>();
    float2* refr0 = refr + (size_t)Smax * nr;

    // S: shots of the current batch (an allocation-free run may take fewer
    // than Smax when their facet records do not fit the region)
    int S = 0;
    for (int64_t s0 = 0; s0 < ns; s0 += S) {
        S = (int)std::min<int64_t>(Smax, ns - s0);
        ERP_HIP_TRY(hipMemcpyAsync(W.shots.p, shots + s0, sizeof(erp_shot) * S,
                                   hipMemcpyHostToDevice, st));
        const erp_shot* dshots = W.shots.as<erp_shot>();

        StageTimer none_;
        StageTimer& PT_ = timer ? *timer : none_;
        // 1-2: aperture mask, counts, offsets
        PT_.begin(ERP_STAGE_MASK, st, (double)F * S);
        hipLaunchKernelGGL(k_illum_mask, dim3(nblk, (S + MASK_SB - 1) / MASK_SB), dim3(NT), 0, st,
                           mesh->soa.as<float>(), mesh->Fpad, F, dshots, p.cos_ap, dmask,
                           W.cnt.as<int>(), nblk, tile_nx, (const TrackK*)nullptr,
                           (const RowK*)nullptr, S, MASK_SB);
        ERP_CHECK_LAUNCH("k_illum_mask");
        hipLaunchKernelGGL(k_scan_blocks, dim3(S), dim3(1024), 0, st, W.cnt.as<int>(),
                           W.off.as<int>(), dV, nblk, (const TrackK*)nullptr,
                           (const RowK*)nullptr);
        ERP_CHECK_LAUNCH("k_scan_blocks");
        // blocks of two shots: the union of their masks, its offsets and size
        int nbk = (S + 1) / 2;
        if (pairk) {
            hipLaunchKernelGGL(k_union_mask, dim3(nblk, nbk), dim3(MASKW), 0, st, dmask, nblk, S,
                               W.umask.as<unsigned long long>(), W.ucnt.as<int>(),
                               (const TrackK*)nullptr, (const RowK*)nullptr);
            ERP_CHECK_LAUNCH("k_union_mask");
            hipLaunchKernelGGL(k_scan_blocks, dim3(nbk), dim3(1024), 0, st, W.ucnt.as<int>(),
                               W.uoff.as<int>(), W.U.as<int>(), nblk, (const TrackK*)nullptr,
                               (const RowK*)nullptr);
            ERP_CHECK_LAUNCH("k_scan_blocks");
            ERP_HIP_TRY(hipMemcpyAsync(hU.data(), W.U.p, sizeof(int) * nbk,
                                       hipMemcpyDeviceToHost, st));
        }
        PT_.end(st);
        ERP_HIP_TRY(hipMemcpyAsync(hV.data(), dV, sizeof(int) * S, hipMemcpyDeviceToHost, st));
        ERP_HIP_TRY(hipStreamSynchronize(st));
        auto brec_need = [](size_t cap) {
            return 5 * Arena::round(cap * 16) + Arena::round(cap * 4);
        };
        if (arena && need_rec && pairk) {
            // keep the leading blocks whose union records fit the region (or,
            // failing that, the first shot alone); the other shots are masked
            // again as the next batch
            int fitb = 0;
            size_t acc = 0;
            while (fitb < nbk && brec_need(acc + (size_t)hU[fitb]) <= W.brec.bytes)
                acc += hU[fitb++];
            if (fitb == 0 && S > 1) {
                S = 1;
                nbk = 1;
                hipLaunchKernelGGL(k_union_mask, dim3(nblk, 1), dim3(MASKW), 0, st, dmask, nblk, 1,
                                   W.umask.as<unsigned long long>(), W.ucnt.as<int>(),
                                   (const TrackK*)nullptr, (const RowK*)nullptr);
                ERP_CHECK_LAUNCH("k_union_mask");
                hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, W.ucnt.as<int>(),
                                   W.uoff.as<int>(), W.U.as<int>(), nblk, (const TrackK*)nullptr,
                                   (const RowK*)nullptr);
                ERP_CHECK_LAUNCH("k_scan_blocks");
                ERP_HIP_TRY(hipMemcpyAsync(hU.data(), W.U.p, sizeof(int), hipMemcpyDeviceToHost,
                                           st));
                ERP_HIP_TRY(hipStreamSynchronize(st));
                if (brec_need((size_t)hU[0]) <= W.brec.bytes) fitb = 1;
            }
            if (fitb == 0)
                throw ErpError{ERP_E_NOMEM,
                               "shot " + std::to_string(s0) + " illuminates " +
                                   std::to_string(hV[0]) + " facets: their records need " +
                                   std::to_string(brec_need((size_t)hU[0])) +
                                   " bytes, the allocation-free scratch has " +
                                   std::to_string(W.brec.bytes) + " for records"};
            S = std::min(S, 2 * fitb);
            nbk = (S + 1) / 2;
        } else if (arena && need_rec) {
            // keep the leading shots whose records fit the region; the others
            // are masked again as the next batch
            int fit = 0;
            size_t acc = 0;
            while (fit < S && rec_need(acc + (size_t)hV[fit]) <= W.rec.bytes) acc += hV[fit++];
            if (fit == 0)
                throw ErpError{ERP_E_NOMEM,
                               "shot " + std::to_string(s0) + " illuminates " +
                                   std::to_string(hV[0]) + " facets: their records need " +
                                   std::to_string(rec_need((size_t)hV[0])) +
                                   " bytes, the allocation-free scratch has " +
                                   std::to_string(W.rec.bytes) + " for records"};
            S = fit;
        }
        hbase[0] = 0;
        for (int s = 0; s < S; ++s) hbase[s + 1] = hbase[s] + (need_rec ? hV[s] : 0);
        // the target kernel reads a shot's records at 32-bit byte offsets
        for (int s = 0; s < S; ++s)
            if (need_rec && hV[s] >= (1 << 28))
                throw ErpError{ERP_E_INVALID, "a shot illuminates 2^28 facets or more"};
        const int64_t Vtot = hbase[S];
        ERP_HIP_TRY(hipMemcpyAsync(dbase, hbase.data(), sizeof(int64_t) * S,
                                   hipMemcpyHostToDevice, st));
        Rec rec{};
        BlkRec brk{};
        if (pairk && Vtot > 0) {
            // block records: 84 bytes per facet of a block's union
            hubase[0] = 0;
            for (int j = 0; j < nbk; ++j) {
                if (hU[j] >= (1 << 28))
                    throw ErpError{ERP_E_INVALID, "two shots illuminate 2^28 facets or more"};
                hubase[j + 1] = hubase[j] + hU[j];
            }
            const size_t cap = (size_t)hubase[nbk];
            auto rnd = [](size_t n) { return (n + 255) & ~(size_t)255; };
            const size_t bytes = brec_need(cap);
            if (arena) W.brec.ensure(bytes);  // fits: the batch was cut to fit
            else W.brec.ensure_grow(bytes, (size_t)nbk * (size_t)F * 84 + 256 * 6);
            char* b = W.brec.as<char>();
            auto take = [&](size_t n) { char* r = b; b += rnd(n); return r; };
            BRec& br = brk.rec;
            br.qa = (float4*)take(cap * 16);
            br.qb = (float4*)take(cap * 16);
            br.qc = (float4*)take(cap * 16);
            br.qd = (float*)take(cap * 4);
            br.itd = (double2*)take(cap * 16);
            br.src = (float4*)take(cap * 16);
            brk.umask = W.umask.as<unsigned long long>();
            brk.uoff = W.uoff.as<int>();
            brk.U = W.U.as<int>();
            brk.ubase = W.ubase.as<int64_t>();
            ERP_HIP_TRY(hipMemcpyAsync(W.ubase.p, hubase.data(), sizeof(int64_t) * nbk,
                                       hipMemcpyHostToDevice, st));
            if (cap > 0) {
                hipLaunchKernelGGL(k_fill_itd, dim3(grid_for(cap)), dim3(256), 0, st, br.itd,
                                   (int64_t)cap);
                ERP_CHECK_LAUNCH("k_fill_itd");
            }
        } else if (need_rec && Vtot > 0) {
            const size_t cap = (size_t)Vtot;
            // at most every facet of every shot of the batch (the bound that
            // erp_run_footprint counts)
            const size_t rec_unit = 12 * 4 + 8 + 4 + 4 + 3 * 4 + 52;
            const size_t bytes = rec_need(cap);
            if (arena) W.rec.ensure(bytes);  // fits: the batch was cut to fit
            else W.rec.ensure_grow(bytes, (size_t)S * (size_t)F * rec_unit + 256 * 10);
            char* b = W.rec.as<char>();
            auto take = [&](size_t n) { char* r = b; b += (n + 255) & ~(size_t)255; return r; };
            rec.itd = (double*)take(cap * 8);
            // the per-field arrays feed the specular and type-0 kernels, the
            // packed records the plain target kernel: write only what runs
            if (spec || any_t0) {
                rec.geo = (float*)take(cap * 12 * 4);
                rec.F = (float*)take(cap * 4);
                rec.fRfrC = (float*)take(cap * 4);
                rec.ang = (float*)take(cap * 3 * 4);
            }
            if (!spec && any_typed) {
                rec.qa = (float4*)take(cap * 16);
                rec.qb = (float4*)take(cap * 16);
                rec.qc = (float4*)take(cap * 16);
                rec.qd = (float*)take(cap * 4);
            }
            rec.cap = (int64_t)cap;
        }

        // 3: fused surface kernel
        zero(W.psurf.as<float2>(), (size_t)S * P, st);
        hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, st, dbounds, S);
        ERP_CHECK_LAUNCH("k_init_bounds");
        double Vsum = 0;
        for (int s = 0; s < S; ++s) Vsum += hV[s];
        PT_.begin(ERP_STAGE_SURFACE, st, Vsum);
        if (do_trace) zero(W.srow.as<float2>(), srow_f2, st);
        const int write = (need_rec && Vtot > 0) ? 1 : 0;
        hipLaunchKernelGGL(k_surface, dim3(nblk, S), dim3(NT), 0, st, mesh->soa.as<float>(),
                           mesh->Fpad, F, dshots, dmask, W.off.as<int>(), dV, dbase, nblk, kp, rec,
                           write, do_trace ? 1 : 0, srow, s0, dbounds, det ? 1 : 0, smax,
                           tile_nx, ph, acc, (const TrackK*)nullptr, (const RowK*)nullptr, brk,
                           T > 0 ? W.bbox.as<float4>() : nullptr);
        ERP_CHECK_LAUNCH("k_surface");
        if (T > 0) {
            hipLaunchKernelGGL(k_shot_sphere, dim3(S), dim3(256), 0, st, dbounds,
                               (const float4*)W.bbox.as<float4>(), (const int*)W.cnt.as<int>(),
                               nblk, (const TrackK*)nullptr, (const RowK*)nullptr);
            ERP_CHECK_LAUNCH("k_shot_sphere");
        }
        if (det && do_trace) {
            // second pass: the same terms as fixed-point values scaled by the
            // per-shot maxima of the first
            hipLaunchKernelGGL(k_surface, dim3(nblk, S), dim3(NT), 0, st, mesh->soa.as<float>(),
                               mesh->Fpad, F, dshots, dmask, W.off.as<int>(), dV, dbase, nblk, kp,
                               rec, 0, 1, srow, s0, (Bounds*)nullptr, 2, smax, tile_nx, 0,
                               0, (const TrackK*)nullptr, (const RowK*)nullptr, BlkRec{},
                               (float4*)nullptr);
            ERP_CHECK_LAUNCH("k_surface");
        }
        if (do_trace) {
            hipLaunchKernelGGL(k_surf_rows, dim3((nr + 255) / 256, S), dim3(256), 0, st,
                               W.psurf.as<float2>(), P, srow, nr, S, det ? 1 : 0, smax, dV);
            ERP_CHECK_LAUNCH("k_surf_rows");
        }
        PT_.end(st);

        if (flags & ERP_RUN_RAW_SURFACE) {
            // debug_surface: the raw surface phasor trace, no range compression
            hipLaunchKernelGGL(k_copy_rows, dim3(grid_for((size_t)nr * S)), dim3(256), 0, st,
                               out_dev + (size_t)s0 * nr, nr, W.psurf.as<float2>(), P, nr, S);
            ERP_CHECK_LAUNCH("k_copy_rows");
            if (nvalid_dev)
                ERP_HIP_TRY(hipMemcpyAsync(nvalid_dev + s0, dV, sizeof(int) * S,
                                           hipMemcpyDeviceToDevice, st));
            continue;
        }

        // 4: reflected signal
        float2* refl = W.refl.as<float2>();
        bool have_refl = false;
        if (!spec || (conv && lin)) {
            PT_.begin(ERP_STAGE_SURFACE_CONV, st, (double)S);
            if (mode & ERP_MODE_DISABLE_SURFACE && !spec) {
                zero(refl, (size_t)S * nr, st);
            } else {
                window_conv(ctx, planP, W.psurf.as<float2>(), S, g_surf, refl, nr, st);
            }
            PT_.end(st);
            have_refl = true;
        } else if (spec && !conv) {
            PT_.begin(ERP_STAGE_OTHER, st, 0.0);
            hipLaunchKernelGGL(k_refl_direct, dim3((nr + NT - 1) / NT, S), dim3(NT), 0, st, refl,
                               dshots, dV, dbase, rec, kp);
            ERP_CHECK_LAUNCH("k_refl_direct");
            PT_.end(st);
            have_refl = true;
        }

        // 5-7: targets
        bool have_refr = false, have_refr0 = false;
        if (T > 0) {
            const int Tcap = (int)T;
            int* list1 = W.tlist.as<int>();
            int* list0 = list1 + (size_t)S * Tcap;
            PT_.begin(ERP_STAGE_TARGET_SELECT, st, (double)T * S);
            hipLaunchKernelGGL(k_target_select, dim3(S), dim3(1024), 0, st,
                               targets->xyzn.as<float>(), Tpad, (int)T,
                               targets->type.as<unsigned char>(), dshots, dV, p.cos_ap, list1,
                               list0, W.tcnt.as<int>(), Tcap, (const TrackK*)nullptr,
                               (const RowK*)nullptr);
            ERP_CHECK_LAUNCH("k_target_select");
            PT_.end(st);
            ERP_HIP_TRY(hipMemcpyAsync(htc.data(), W.tcnt.p, sizeof(int) * 2 * S,
                                       hipMemcpyDeviceToHost, st));
            ERP_HIP_TRY(hipStreamSynchronize(st));
            const erp_prism* dpr = nprisms > 0 ? W.prisms.as<erp_prism>() : nullptr;

            // type != 0 targets: per-pair traces + spectral sum
            int64_t npairs = 0;
            for (int s = 0; s < S; ++s) npairs += htc[2 * s];
            if (npairs > 0) {
                float2* Sacc = W.shotS.as<float2>();
                zero(Sacc, (size_t)S * P, st);
                const size_t gcap = (size_t)C / GS + S + 2;
                // half: A rows, then one partial row per slice
                W.tA.ensure(sizeof(float2) * (size_t)P *
                            ((size_t)C * (nbuf - 1) + (half ? gcap : 0)));
                W.tC.ensure(sizeof(float2) * (size_t)C * P);
                float2* A = W.tA.as<float2>();
                float2* Bq = A + (size_t)C * P;                 // y / Y' (and MODE 1: B)
                float2* Bs = spec ? Bq + (size_t)C * P : nullptr;
                float2* Z = W.tC.as<float2>();                  // Z or fused workspace
                // each shot's list holds Tcap slots of which htc[2s] are used:
                // copy only the leading columns that some shot uses
                W.pairs.ensure(sizeof(int) * ((size_t)C * 4 + 2 * gcap + 3 * (size_t)S + 16));
                int* dps = W.pairs.as<int>();
                int* dpt = dps + C;
                int* dpb = dpt + C;                  // per-pair window base (fused)
                int* dgb = dpb + C;                  // group begin / end
                int* dge = dgb + gcap;
                int* dseg = dge + gcap;              // segment shot / begin / end
                int* dpw = dseg + 3 * S;             // per-pair window length (fused)
                if (pairk) {
                    // the shot-pair kernel: items of a target and two
                    // consecutive shots, from the shots' target lists
                    int maxc = 0;
                    for (int s = 0; s < S; ++s) maxc = std::max(maxc, htc[2 * s]);
                    std::vector<int> hl((size_t)S * maxc);
                    if (maxc > 0) {
                        ERP_HIP_TRY(hipMemcpy2DAsync(hl.data(), sizeof(int) * maxc, list1,
                                                     sizeof(int) * Tcap, sizeof(int) * maxc, S,
                                                     hipMemcpyDeviceToHost, st));
                        ERP_HIP_TRY(hipStreamSynchronize(st));
                    }
                    std::vector<std::vector<PairIdx>> bl;
                    merge_pair_lists(hl.data(), (size_t)maxc, htc.data(), S, bl);
                    W.items.ensure(sizeof(int4) * (size_t)C);
                    int4* ditems = W.items.as<int4>();
                    int* dnext = dpw + C;  // item counter, in the spare tail of the list
                    PairChunk pc;
                    int jb = 0;
                    size_t ib = 0;
                    PairCursor cur;
                    for (;;) {
                        // deterministic: every shot's slices independent of the launches
                        if (det) fill_pair_chunk_canon(bl, cur, C, GS, hV.data(), pc);
                        else fill_pair_chunk(bl, jb, ib, C, GS, hV.data(), pc);
                        const int n = pc.rows;
                        if (n == 0) break;
                        const int nseg = (int)pc.sg_s.size();
                        const int ngroups = (int)pc.gb.size();
                        const int nit = (int)pc.items.size();
                        // slices and segments at the offsets the chain reads them
                        std::vector<int> up(2 * gcap + 3 * (size_t)S, 0);
                        std::copy(pc.gb.begin(), pc.gb.end(), up.begin());
                        std::copy(pc.ge.begin(), pc.ge.end(), up.begin() + gcap);
                        std::copy(pc.sg_s.begin(), pc.sg_s.end(), up.begin() + 2 * gcap);
                        std::copy(pc.gsb.begin(), pc.gsb.end(), up.begin() + 2 * gcap + S);
                        std::copy(pc.gse.begin(), pc.gse.end(), up.begin() + 2 * gcap + 2 * S);
                        ERP_HIP_TRY(hipMemcpyAsync(dgb, up.data(), sizeof(int) * up.size(),
                                                   hipMemcpyHostToDevice, st));
                        ERP_HIP_TRY(hipMemcpyAsync(ditems, pc.items.data(), sizeof(int4) * nit,
                                                   hipMemcpyHostToDevice, st));
                        ERP_HIP_TRY(hipMemsetAsync(dnext, 0, sizeof(int), st));
                        PT_.begin(ERP_STAGE_TARGET_ACCUM, st, pc.triples);
                        // persistent grid: one workgroup per CU (its two windows
                        // take 128 KB of LDS), each taking items until none is left
                        const int grid = std::max(1, std::min(nit, ctx->num_cu));
                        auto kern = det ? (kp.has3 ? k_target_accum_pair<true, false, true>
                                                   : k_target_accum_pair<false, false, true>)
                                        : (kp.has3 ? k_target_accum_pair<true>
                                                   : k_target_accum_pair<false>);
                        hipLaunchKernelGGL(kern, dim3(grid), dim3(NTB), 0, st,
                                           targets->xyzn.as<float>(), Tpad,
                                           (const int4*)ditems, nit, dnext, dshots,
                                           brk.ubase, brk.U, brk.rec, dbounds, dpr, (int)nprisms,
                                           kp, A, dpb, dpw, mesh->frame_ok ? 1 : 0,
                                           W.oow.as<int>(), (const TrackK*)nullptr,
                                           (const RowK*)nullptr, (const int*)nullptr,
                                           (const int*)dV);
                        ERP_CHECK_LAUNCH("k_target_accum_pair");
                        PT_.end(st);
                        PT_.begin(ERP_STAGE_TARGET_FFT, st, (double)n);
                        target_pairs_half(ctx, planP, A, dpb, dpw, nr, n, g_targ_nat, Z, st);
                        PT_.end(st);
                        PT_.begin(ERP_STAGE_TARGET_SUM, st, (double)n);
                        target_sum_half(planP, Z, A + (size_t)C * P, dgb, dge, ngroups, dseg,
                                        dseg + S, dseg + 2 * S, nseg, Sacc, -1.0f / kp.sqrtP, st,
                                        nullptr, det, wl);
                        PT_.end(st);
                    }
                }
                // iterate pairs in (shot, target) order
                std::vector<int> tc1((size_t)S);
                for (int s = 0; s < S; ++s) tc1[s] = htc[2 * s];
                ChunkFill cf;
                int cs = 0, cj = 0;
                while (!pairk && cs < S) {
                    // deterministic: every shot's slices independent of the launches
                    fill_chunk(tc1.data(), S, cs, cj, C, GS, det && half, cf);
                    const int n = cf.n();
                    if (n == 0) break;
                    const std::vector<int>& pshot = cf.pshot;
                    const std::vector<int>& ptgt = cf.ptgt;
                    const std::vector<int>&sg_s = cf.sg_s, &sg_b = cf.sg_b, &sg_e = cf.sg_e;
                    const std::vector<int>&gb = cf.gb, &ge = cf.ge, &gsb = cf.gsb, &gse = cf.gse;
                    const int nseg = (int)sg_s.size();
                    const int ngroups = (int)gb.size();
                    std::vector<int> up((size_t)C * 3 + 2 * gcap + 3 * (size_t)S, 0);
                    std::copy(pshot.begin(), pshot.end(), up.begin());
                    std::copy(ptgt.begin(), ptgt.end(), up.begin() + C);
                    std::copy(gb.begin(), gb.end(), up.begin() + 3 * C);
                    std::copy(ge.begin(), ge.end(), up.begin() + 3 * C + gcap);
                    const size_t so = 3 * (size_t)C + 2 * gcap;
                    std::copy(sg_s.begin(), sg_s.end(), up.begin() + so);
                    const bool two = planP.two_pass;
                    const bool by_group = (fused && two) || half;
                    const std::vector<int>& segb = by_group ? gsb : sg_b;
                    const std::vector<int>& sege = by_group ? gse : sg_e;
                    std::copy(segb.begin(), segb.end(), up.begin() + so + S);
                    std::copy(sege.begin(), sege.end(), up.begin() + so + 2 * S);
                    ERP_HIP_TRY(hipMemcpyAsync(dps, up.data(), sizeof(int) * up.size(),
                                               hipMemcpyHostToDevice, st));
                    int* dnext = dpw + C;  // pair counter, in the spare tail of the list
                    hipLaunchKernelGGL(k_pair_targets, dim3((n + 255) / 256), dim3(256), 0, st,
                                       dpt, dps, list1, Tcap, n, dnext);
                    ERP_CHECK_LAUNCH("k_pair_targets");
                    double triples = 0;
                    for (int i = 0; i < n; ++i) triples += hV[pshot[i]];
                    PT_.begin(ERP_STAGE_TARGET_ACCUM, st, triples);
                    if (!spec) {
                        auto kern = pick_target_accum(det, ph, acc, kp.has3 != 0);
                        // persistent grid: two workgroups per CU (the LDS window
                        // allows two), each taking pairs until none is left
                        const int grid = std::max(1, std::min(n, 2 * ctx->num_cu));
                        hipLaunchKernelGGL(kern, dim3(grid), dim3(NTF), 0, st,
                                           targets->xyzn.as<float>(), Tpad, dps, dpt, dshots, dV,
                                           dbase, rec, dbounds, dpr, (int)nprisms, kp, A,
                                           (fused || half) ? dpb : nullptr, dpw,
                                           mesh->frame_ok ? 1 : 0, n, dnext,
                                           (const TrackK*)nullptr, (const RowK*)nullptr,
                                           (const int*)nullptr, W.oow.as<int>());
                        ERP_CHECK_LAUNCH("k_target_accum_fast");
                    } else {
                        auto kern1 = det ? k_target_accum<1, true> : k_target_accum<1, false>;
                        hipLaunchKernelGGL(kern1, dim3(n), dim3(NTT), sizeof(double2) * TWIN, st,
                                           targets->xyzn.as<float>(), Tpad, dps, dpt, dshots, dV,
                                           dbase, rec, dbounds, dpr, (int)nprisms, kp, A, Bs,
                                           nullptr, (unsigned*)nullptr, (const int*)nullptr, 0,
                                           det ? W.oow.as<int>() : nullptr);
                        ERP_CHECK_LAUNCH("k_target_accum<1>");
                    }
                    PT_.end(st);
                    PT_.begin(ERP_STAGE_TARGET_FFT, st, (double)n);
                    if (half) {
                        // outward trace = -A/sqrt(P): scale -1/sqrt(P)
                        target_pairs_half(ctx, planP, A, dpb, dpw, nr, n, g_targ_nat, Z, st);
                        PT_.end(st);
                        PT_.begin(ERP_STAGE_TARGET_SUM, st, (double)n);
                        target_sum_half(planP, Z, A + (size_t)C * P, dgb, dge, ngroups, dseg,
                                        dseg + S, dseg + 2 * S, nseg, Sacc, -1.0f / kp.sqrtP, st,
                                        nullptr, det, wl);
                        PT_.end(st);
                        continue;
                    }
                    if (fused) {
                        // outward trace = -A/sqrt(P): scale -1/sqrt(P)
                        target_spectra_fused(ctx, planP, A, Bq, dpb, dpw, nr, n, g_targ, Z,
                                             ngroups, dgb, dge, dseg, dseg + S, dseg + 2 * S,
                                             nseg, Sacc,
                                             -1.0f / kp.sqrtP, st);
                        PT_.end(st);
                        continue;
                    }
                    // X = FFT(A); y = IFFT(X * g_targ); Ptarg = window(y); Z = FFT(Ptarg)
                    fft_forward(ctx, planP, A, n, st);
                    hipLaunchKernelGGL(k_mul_spec, dim3(grid_for((size_t)P * n)), dim3(256), 0,
                                       st, Bq, A, g_targ, P, n);
                    ERP_CHECK_LAUNCH("k_mul_spec");
                    fft_inverse(ctx, planP, Bq, n, st);
                    hipLaunchKernelGGL(k_window, dim3(grid_for((size_t)P * n)), dim3(256), 0, st,
                                       Z, P, Bq, P, nr, n, 1.0f, 1);
                    ERP_CHECK_LAUNCH("k_window");
                    fft_forward(ctx, planP, Z, n, st);
                    const float2* Xs = A;
                    float sc = -1.0f / kp.sqrtP;  // outward trace = -A/sqrt(P)
                    if (spec) {
                        fft_forward(ctx, planP, Bs, n, st);
                        Xs = Bs;
                        sc = 1.0f;
                    }
                    hipLaunchKernelGGL(k_accum_spectra, dim3(std::max(1, P / 256), nseg), dim3(256),
                                       0, st, Sacc, Xs, Z, dseg, dseg + S, dseg + 2 * S, P, sc);
                    ERP_CHECK_LAUNCH("k_accum_spectra");
                    PT_.end(st);
                }
                // per shot: c = IFFT(S)/P; Bluestein DFT_L, crop, IDFT_L, decimate
                PT_.begin(ERP_STAGE_SHOT_FINAL, st, (double)S);
                fft_inverse(ctx, planP, Sacc, S, st);
                float2* bu = W.blu.as<float2>();
                const int M = planM.N;
                hipLaunchKernelGGL(k_blue_in, dim3(grid_for((size_t)M * S)), dim3(256), 0, st, bu,
                                   M, Sacc, P, blue->w.as<float2>(), L, 1.0f / (float)P, S);
                ERP_CHECK_LAUNCH("k_blue_in");
                fft_forward(ctx, planM, bu, S, st);
                hipLaunchKernelGGL(k_mul_spec, dim3(grid_for((size_t)M * S)), dim3(256), 0, st, bu,
                                   bu, blue->Bw.as<float2>(), M, S);
                ERP_CHECK_LAUNCH("k_mul_spec");
                fft_inverse(ctx, planM, bu, S, st);
                hipLaunchKernelGGL(k_blue_mid, dim3(grid_for((size_t)M * S)), dim3(256), 0, st, bu,
                                   M, blue->mask.as<unsigned char>(), L, S);
                ERP_CHECK_LAUNCH("k_blue_mid");
                fft_forward(ctx, planM, bu, S, st);
                hipLaunchKernelGGL(k_mul_spec, dim3(grid_for((size_t)M * S)), dim3(256), 0, st, bu,
                                   bu, blue->Bw.as<float2>(), M, S);
                ERP_CHECK_LAUNCH("k_mul_spec");
                fft_inverse(ctx, planM, bu, S, st);
                hipLaunchKernelGGL(k_blue_out, dim3(grid_for((size_t)nr * S)), dim3(256), 0, st,
                                   refr, nr, bu, M, blue->w.as<float2>(), 1.0f / (float)L, S, dV);
                ERP_CHECK_LAUNCH("k_blue_out");
                PT_.end(st);
                have_refr = true;
            }

            // type 0 targets
            int64_t n0 = 0;
            for (int s = 0; s < S; ++s) n0 += htc[2 * s + 1];
            if (n0 > 0) {
                PT_.begin(ERP_STAGE_OTHER, st, (double)n0);
                if (!conv) {
                    hipLaunchKernelGGL(k_refr0_direct, dim3((nr + NT - 1) / NT, S), dim3(NT), 0, st,
                                       refr0, targets->xyzn.as<float>(), Tpad, list0,
                                       W.tcnt.as<int>(), Tcap, dshots, dV, dbase, rec, dpr,
                                       (int)nprisms, kp);
                    ERP_CHECK_LAUNCH("k_refr0_direct");
                } else {
                    float2* t0 = W.t0trace.as<float2>();
                    zero(t0, (size_t)S * P, st);
                    std::vector<int> hl0((size_t)S * Tcap);
                    ERP_HIP_TRY(hipMemcpyAsync(hl0.data(), list0, sizeof(int) * (size_t)S * Tcap,
                                               hipMemcpyDeviceToHost, st));
                    ERP_HIP_TRY(hipStreamSynchronize(st));
                    std::vector<int> ps, pt;
                    for (int s = 0; s < S; ++s)
                        for (int j = 0; j < htc[2 * s + 1]; ++j) {
                            ps.push_back(s);
                            pt.push_back(hl0[(size_t)s * Tcap + j]);
                        }
                    const int n = (int)ps.size();
                    W.pairs.ensure_grow(sizeof(int) * 2 * (size_t)n + 64,
                                        sizeof(int) * 2 * (size_t)S * Tcap + 64);
                    int* dps = W.pairs.as<int>();
                    std::vector<int> up(ps);
                    up.insert(up.end(), pt.begin(), pt.end());
                    ERP_HIP_TRY(hipMemcpyAsync(dps, up.data(), sizeof(int) * up.size(),
                                               hipMemcpyHostToDevice, st));
                    if (!det) {
                        hipLaunchKernelGGL(k_target_accum<2>, dim3(n), dim3(NTT),
                                           sizeof(double2) * TWIN, st, targets->xyzn.as<float>(),
                                           Tpad, dps, dps + n, dshots, dV, dbase, rec, dbounds,
                                           dpr, (int)nprisms, kp, t0, nullptr, nullptr,
                                           (unsigned*)nullptr, (const int*)nullptr, 0,
                                           (int*)nullptr);
                        ERP_CHECK_LAUNCH("k_target_accum<2>");
                    } else {
                        // per-shot bounds (the surface's maxima are consumed), then
                        // the fixed-point rows, then FP32 rows into the free
                        // surface buffer
                        zero(reinterpret_cast<float2*>(smax), ((size_t)S + 1) / 2, st);
                        auto kern2 = k_target_accum<2, true>;
                        for (int pass = 1; pass <= 2; ++pass) {
                            hipLaunchKernelGGL(kern2, dim3(n), dim3(NTT),
                                               sizeof(double2) * TWIN, st,
                                               targets->xyzn.as<float>(), Tpad, dps, dps + n,
                                               dshots, dV, dbase, rec, dbounds, dpr, (int)nprisms,
                                               kp, t0, nullptr, nullptr, smax,
                                               (const int*)W.tcnt.as<int>(), pass, (int*)nullptr);
                            ERP_CHECK_LAUNCH("k_target_accum<2>");
                        }
                        hipLaunchKernelGGL(k_fix_rows, dim3(grid_for((size_t)P * S)), dim3(256), 0,
                                           st, W.psurf.as<float2>(),
                                           reinterpret_cast<const unsigned long long*>(t0), P, nr,
                                           S, dV, (const int*)W.tcnt.as<int>(), smax);
                        ERP_CHECK_LAUNCH("k_fix_rows");
                        t0 = W.psurf.as<float2>();
                    }
                    if (lin) {
                        window_conv(ctx, planP, t0, S, g_targ, refr0, nr, st);
                    } else {
                        // circular: genChirp table depends on rst (per shot)
                        std::vector<erp_shot> hs(shots + s0, shots + s0 + S);
                        for (int s = 0; s < S; ++s) {
                            const float2* gc = g_circ;
                            if (g_circ)  // rebuilt in place: stream order keeps the last use
                                build_chirp(ctx, g_circ, nr, p.dr, p.rng_res, 1, hs[s].rst, P,
                                            planP, st);
                            else
                                gc = get_chirp(ctx, nr, p.dr, p.rng_res, 1, hs[s].rst, P, planP,
                                               st);
                            window_conv(ctx, planP, t0 + (size_t)s * P, 1, gc,
                                        refr0 + (size_t)s * nr, nr, st);
                        }
                    }
                }
                PT_.end(st);
                have_refr0 = true;
            }
        }

        // 8: combine
        hipLaunchKernelGGL(k_finalize, dim3(grid_for((size_t)nr * S)), dim3(256), 0, st,
                           out_dev + (size_t)s0 * nr, have_refl ? W.refl.as<float2>() : nullptr,
                           have_refr ? refr : nullptr, have_refr0 ? refr0 : nullptr, dV, nr, S,
                           nvalid_dev ? nvalid_dev + s0 : nullptr);
        ERP_CHECK_LAUNCH("k_finalize");
    }
    if (check_win) check_windows(W.oow.as<int>(), st);
}

// ---------------------------------------------------------------------------
// multi-track batches (erp_run_tracks)
//
// The (track, shot) rows of several tracks of one range length run in
// batches of R rows: every launch of a batch covers all its rows, and each
// row takes its mesh, targets, parameters, range kernel spectra and
// Bluestein tables from its track's descriptor (TrackK).  Per batch:
//   A1  aperture masks and counts of the R rows
//       -> illuminated counts to the host: facet record bases
//   A2  surface kernel, surface range compression, target selection
//       -> pair counts to the host: pair lists, cut into launches of C
//   chunk (per launch of C pairs) pair resolution, facet x target
//       accumulation, per-pair LDS chain, slice and row sums
//   tail  Bluestein crop and decimation of every row, then each row's trace
//       into its track's radargram
// Each of the four sequences has fixed launch shapes (a short last batch is
// padded with rows of shot -1, a short launch exits past its pair count on
// the device), so it is captured once as a HIP graph and replayed; the
// tables the host writes between them live at fixed device addresses.
// ---------------------------------------------------------------------------
namespace {

struct MtKey {
    int R = 0, nblk = 0, nr = 0, Tcap = 0, C = 0, ntrk = 0;
    size_t reccap = 0;
    bool det = false, s3 = false, targ = false, pair = false;
    bool operator==(const MtKey& o) const {
        return R == o.R && nblk == o.nblk && nr == o.nr && Tcap == o.Tcap && C == o.C &&
               ntrk == o.ntrk && reccap == o.reccap && det == o.det && s3 == o.s3 &&
               targ == o.targ && pair == o.pair;
    }
};

struct MtGraph {
    hipGraphExec_t exec = nullptr;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }
};

constexpr int kMtRowsDefault = 64;
constexpr int kMtRowsMax = 256;

}  // namespace

struct MtState {
    MtKey key;
    DevBuf trk, rows, shots, cnt, off, V, mask, bounds, base, psurf, srow, refl, refr, Sacc, blu,
        list, tcnt, A, Z, tabs, dtabs, rec;
    // shot-pair kernel: union masks, counts, offsets, sizes, first records
    // and first rows per block of two rows; work items and their count
    DevBuf umask, ucnt, uoff, U, ubase, brows, items;
    DevBuf bbox;  // facet block boxes of the rows (bounding spheres)
    MtGraph gA1, gA2, gChunk, gTail;
    void reset_graphs() {
        gA1.reset();
        gA2.reset();
        gChunk.reset();
        gTail.reset();
    }
    void release() {
        reset_graphs();
        DevBuf* b[] = {&trk, &rows, &shots, &cnt, &off, &V, &mask, &bounds, &base, &psurf, &srow,
                       &refl, &refr, &Sacc, &blu, &list, &tcnt, &A, &Z, &tabs, &dtabs, &rec,
                       &umask, &ucnt, &uoff, &U, &ubase, &brows, &items, &bbox};
        for (DevBuf* d : b) d->release();
        key = MtKey{};
    }
};

void mt_state_free(MtState* m) {
    if (!m) return;
    m->release();
    delete m;
}

namespace {

// body() either launched directly or, the first time, captured into g and
// then replayed from it
template <class Fn>
void mt_launch(MtGraph& g, bool direct, hipStream_t st, Fn&& body) {
    if (direct) {
        body();
        return;
    }
    if (!g.exec) {
        ERP_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            hipGraph_t t = nullptr;
            (void)hipStreamEndCapture(st, &t);
            if (t) (void)hipGraphDestroy(t);
            throw;
        }
        hipGraph_t graph = nullptr;
        ERP_HIP_TRY(hipStreamEndCapture(st, &graph));
        const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g.exec = nullptr;
            ERP_HIP_TRY(e);
        }
    }
    ERP_HIP_TRY(hipGraphLaunch(g.exec, st));
}

// the tracks a multi-track batch can take: the plain method with linear
// convolution, a range length whose per-pair LDS chain exists, no type-0
// target (their direct histograms and circular chirps are per shot)
bool mt_eligible(const erp_track& t, uint32_t flags) {
    const erp_params& p = *t.par;
    if (flags & ~ERP_RUN_DETERMINISTIC) return false;
    const uint32_t m = p.mode;
    if ((m & ERP_MODE_SPECULAR) || !(m & ERP_MODE_CONVOLUTION) || !(m & ERP_MODE_CONV_LINEAR))
        return false;
    if (p.nr <= 512 || p.nr > 16384) return false;
    if (t.targets && t.targets->n0 > 0) return false;
    if (t.ns <= 0 || t.mesh->F == 0) return false;
    return true;
}

// one group: tracks trk_idx of range length nr, all with (s3) or all
// without the second subsurface medium
void mt_run_group(erp_ctx* ctx, const erp_track* tracks, const std::vector<int64_t>& trk_idx,
                  int nr, bool s3, bool det, hipStream_t st) {
    const int P = next_pow2(2 * (int64_t)nr - 1);
    const FftPlan planP = fft_plan(P);
    if (!target_half_supported(planP, nr))
        throw ErpError{ERP_E_INVALID, "multi-track batch: range length without a pair chain"};
    const int tkind = nr < P / 2 ? 2 : 0;  // the cropped pair chain's kernel placement
    const int L = 2 * nr - 1;
    const FftPlan planM = fft_plan(next_pow2(2 * (int64_t)L - 1));
    const int M = planM.N;
    const int ntr = (int)trk_idx.size();

    // track descriptors
    std::vector<TrackK> tk((size_t)ntr);
    int nblk_max = 1, Tcap = 0;
    bool targ = false;
    int64_t nrows = 0;
    for (int k = 0; k < ntr; ++k) {
        const erp_track& t = tracks[trk_idx[k]];
        const erp_params& p = *t.par;
        const erp_mesh* mesh = t.mesh;
        TrackK& d = tk[k];
        d = TrackK{};
        d.kp = make_kpar(p, P);
        d.kp.ortho = mesh->frame_ok ? 1 : 0;
        d.kp.win_shift = (int)ctx->tune_shift;
        d.kp.row_live = P / 2;
        d.mesh = mesh->soa.as<float>();
        d.Fpad = mesh->Fpad;
        d.F = mesh->F;
        d.nblk = (int)((mesh->F + FB - 1) / FB);
        d.tile_nx = p.cos_ap < 0.94 ? mesh->tile_nx : 0;
        d.frame_ok = mesh->frame_ok ? 1 : 0;
        const int64_t T = t.targets ? t.targets->T : 0;
        d.T = (int)T;
        if (T > 0) {
            d.tg = t.targets->xyzn.as<float>();
            d.Tpad = t.targets->Tpad;
            d.ttype = t.targets->type.as<unsigned char>();
            // sim.cu:853-858: the target range kernel has twice the surface's
            d.g_targ_nat = get_chirp(ctx, nr, p.dr, 2.f * p.rng_res, tkind, 0.f, P, planP, st,
                                     target_wl(planP) ? 2 : 1);
            targ = true;
        }
        // a disabled surface echo: no spectrum, so the row's echo is zero
        d.g_surf = (p.mode & ERP_MODE_DISABLE_SURFACE)
                       ? nullptr
                       : get_chirp(ctx, nr, p.dr, p.rng_res, 0, 0.f, P, planP, st);
        d.out = static_cast<float2*>(t.out_dev);
        d.nvalid = static_cast<int32_t*>(t.nvalid_dev);
        d.tscale = -1.0f / d.kp.sqrtP;  // outward trace = -A / sqrt(P)
        nblk_max = std::max(nblk_max, d.nblk);
        Tcap = std::max<int>(Tcap, (int)T);
        nrows += t.ns;
    }

    // Bluestein tables of every track once any has targets: the tail runs
    // over all rows of a batch (a row without targets crops a zero sum)
    if (targ)
        for (int k = 0; k < ntr; ++k) {
            const erp_params& p = *tracks[trk_idx[k]].par;
            const BlueTables* b = get_blue(ctx, L, p.smpl, p.B, st);
            tk[k].blue_w = b->w.as<float2>();
            tk[k].blue_Bw = b->Bw.as<float2>();
            tk[k].blue_mask = b->mask.as<unsigned char>();
        }

    // shapes
    int R = kMtRowsDefault;
    if (ctx->tune_rows > 0) R = (int)std::min<int64_t>(ctx->tune_rows, kMtRowsMax);
    const int GS = target_half_slice();
    int C = (int)std::max<int64_t>(1, kPairRowBytes / ((int64_t)P * 8 * 2));
    C = std::min<int>(C, kMaxPairsPerLaunch);
    if (ctx->tune_pairs > 0) C = (int)std::min<int64_t>(C, ctx->tune_pairs);
    // deterministic: launches of at least one slice (as in a single track)
    if (det) C = std::max(C, GS);
    // the shot-pair kernel (default arithmetic, unless ERP_TUNE_SINGLE_SHOT):
    // blocks of two rows of one track, so R is even and a track with an odd
    // shot count ends on a padding row
    bool pairk = targ && C >= 2 && ctx->tune_single == 0;
    for (int k = 0; k < ntr; ++k) pairk = pairk && tk[k].kp.piLlam <= 512.f;
    if (pairk && (R & 1)) ++R;
    const int nbk = R / 2;
    const int gcap = C / GS + R + 2;
    const int segcap = R;
    const bool direct = ctx->tune_direct > 0;

    if (!ctx->mt) ctx->mt = new MtState();
    MtState& m = *ctx->mt;
    MtKey key;
    key.R = R; key.nblk = nblk_max; key.nr = nr; key.Tcap = Tcap; key.C = C; key.ntrk = ntr;
    key.reccap = m.key.reccap; key.det = det; key.s3 = s3; key.targ = targ; key.pair = pairk;
    if (!(key == m.key)) {
        m.reset_graphs();
        m.key = key;
    }
    const size_t r = (size_t)R;
    m.trk.ensure(sizeof(TrackK) * (size_t)ntr);
    m.rows.ensure(sizeof(RowK) * r);
    m.shots.ensure(sizeof(erp_shot) * r);
    m.cnt.ensure(sizeof(int) * r * nblk_max);
    m.off.ensure(sizeof(int) * r * nblk_max);
    m.V.ensure(sizeof(int) * r);
    m.mask.ensure(sizeof(unsigned long long) * r * nblk_max * MASKW);
    m.bounds.ensure(sizeof(Bounds) * r);
    m.base.ensure(sizeof(int64_t) * r);
    m.psurf.ensure(sizeof(float2) * r * P);
    const size_t srow_f2 = r * nr * 2 + (r + 1) / 2;  // FP64 bins, then per-row maxima
    m.srow.ensure(sizeof(float2) * srow_f2);
    m.refl.ensure(sizeof(float2) * r * nr);
    if (targ) {
        m.refr.ensure(sizeof(float2) * r * nr);
        m.Sacc.ensure(sizeof(float2) * r * P);
        m.blu.ensure(sizeof(float2) * r * M);
        m.list.ensure(sizeof(int) * r * (size_t)Tcap * 2);
        m.tcnt.ensure(sizeof(int) * r * 2);
        m.bbox.ensure(sizeof(float4) * 2 * r * nblk_max);
        m.A.ensure(sizeof(float2) * (size_t)P * ((size_t)C + gcap));  // pair rows, slice sums
        m.Z.ensure(sizeof(float2) * (size_t)P * C);                   // per-pair products
    }
    // host-written tables of a launch: pair rows and list indices, slices,
    // segments, the pair count, segment scales
    const size_t tab_ints = 2 * (size_t)C + 2 * (size_t)gcap + 3 * (size_t)segcap + 1 + segcap;
    m.tabs.ensure(sizeof(int) * tab_ints);
    // device-written: pair window bases and lengths, the pair counter, the
    // per-pair range kernel spectra
    m.dtabs.ensure(sizeof(int) * (2 * (size_t)C + 64) + sizeof(void*) * (size_t)C);
    if (pairk) {
        m.umask.ensure(sizeof(unsigned long long) * (size_t)nbk * nblk_max * MASKW);
        m.ucnt.ensure(sizeof(int) * (size_t)nbk * nblk_max);
        m.uoff.ensure(sizeof(int) * (size_t)nbk * nblk_max);
        m.U.ensure(sizeof(int) * (size_t)nbk);
        m.ubase.ensure(sizeof(int64_t) * (size_t)nbk);
        m.brows.ensure(sizeof(RowK) * (size_t)nbk);
        m.items.ensure(sizeof(int4) * (size_t)C + 64);  // items, then their count
    }
    int4* d_items = m.items.as<int4>();
    int* d_nitems = reinterpret_cast<int*>(d_items + C);
    int* t_pshot = m.tabs.as<int>();
    int* t_ptgt = t_pshot + C;
    int* t_slb = t_ptgt + C;
    int* t_sle = t_slb + gcap;
    int* t_sgs = t_sle + gcap;
    int* t_sgb = t_sgs + segcap;
    int* t_sge = t_sgb + segcap;
    int* t_n = t_sge + segcap;
    float* t_scale = reinterpret_cast<float*>(t_n + 1);
    int* d_pbase = m.dtabs.as<int>();
    int* d_pwin = d_pbase + C;
    int* d_next = d_pwin + C;  // 64 ints: the pair counter, then padding
    const float4** d_G = reinterpret_cast<const float4**>(d_next + 64);

    ERP_HIP_TRY(hipMemcpyAsync(m.trk.p, tk.data(), sizeof(TrackK) * (size_t)ntr,
                               hipMemcpyHostToDevice, st));
    const TrackK* dtrk = m.trk.as<TrackK>();
    const RowK* drows = m.rows.as<RowK>();
    const erp_shot* dshots = m.shots.as<erp_shot>();
    int* dV = m.V.as<int>();
    Bounds* dbounds = m.bounds.as<Bounds>();
    int64_t* dbase = m.base.as<int64_t>();
    double2* srow = m.srow.as<double2>();
    unsigned* smax = reinterpret_cast<unsigned*>(srow + r * nr);
    int* list1 = targ ? m.list.as<int>() : nullptr;
    int* list0 = targ ? list1 + r * (size_t)Tcap : nullptr;
    const KPar kp0 = tk[0].kp;  // launch argument only: every row reads its own

    // all rows, track-major
    std::vector<RowK> all;
    all.reserve((size_t)nrows);
    for (int k = 0; k < ntr; ++k) {
        const int64_t ns_k = tracks[trk_idx[k]].ns;
        for (int64_t s = 0; s < ns_k; ++s) all.push_back(RowK{k, (int)s});
        if (pairk && (ns_k & 1)) all.push_back(RowK{k, -1});
    }

    if (targ) ERP_HIP_TRY(hipMemsetAsync(d_next + 1, 0, sizeof(int), st));  // window check
    std::vector<RowK> hrows(r);
    std::vector<erp_shot> hshots(r);
    std::vector<int> hV(r), htc(2 * r), tc1(r), hU(nbk);
    std::vector<int64_t> hbase(r), hubase(nbk);
    std::vector<RowK> hbrows(nbk);
    std::vector<std::vector<PairIdx>> bl;
    PairChunk pc;
    std::vector<int4> hitems;
    std::vector<int> htab(tab_ints);
    StageTimer none_;
    StageTimer& PT_ = ctx->prof.on ? ctx->prof : none_;
    ChunkFill cf;

    for (size_t b0 = 0; b0 < all.size(); b0 += r) {
        const size_t nb = std::min(r, all.size() - b0);
        for (size_t i = 0; i < r; ++i) {
            const RowK row = i < nb ? all[b0 + i] : RowK{0, -1};
            hrows[i] = row;
            const erp_track& t = tracks[trk_idx[row.trk]];
            hshots[i] = t.shots[row.shot < 0 ? 0 : row.shot];
        }
        ERP_HIP_TRY(hipMemcpyAsync(m.rows.p, hrows.data(), sizeof(RowK) * r,
                                   hipMemcpyHostToDevice, st));
        ERP_HIP_TRY(hipMemcpyAsync(m.shots.p, hshots.data(), sizeof(erp_shot) * r,
                                   hipMemcpyHostToDevice, st));
        if (pairk) {
            // a block's rows share a track: its first row names it
            for (int j = 0; j < nbk; ++j) hbrows[j] = hrows[2 * j];
            ERP_HIP_TRY(hipMemcpyAsync(m.brows.p, hbrows.data(), sizeof(RowK) * nbk,
                                       hipMemcpyHostToDevice, st));
        }

        // A1: aperture masks, counts
        PT_.begin(ERP_STAGE_MASK, st, 0.0);
        mt_launch(m.gA1, direct, st, [&] {
            hipLaunchKernelGGL(k_illum_mask, dim3(nblk_max, R), dim3(NT), 0, st,
                               (const float*)nullptr, (int64_t)0, (int64_t)0, dshots, 0.0,
                               m.mask.as<unsigned long long>(), m.cnt.as<int>(), nblk_max, 0,
                               dtrk, drows, R, 1);
            ERP_CHECK_LAUNCH("k_illum_mask");
            hipLaunchKernelGGL(k_scan_blocks, dim3(R), dim3(1024), 0, st, m.cnt.as<int>(),
                               m.off.as<int>(), dV, nblk_max, dtrk, drows);
            ERP_CHECK_LAUNCH("k_scan_blocks");
            if (pairk) {
                hipLaunchKernelGGL(k_union_mask, dim3(nblk_max, nbk), dim3(MASKW), 0, st,
                                   m.mask.as<unsigned long long>(), nblk_max, R,
                                   m.umask.as<unsigned long long>(), m.ucnt.as<int>(), dtrk,
                                   drows);
                ERP_CHECK_LAUNCH("k_union_mask");
                hipLaunchKernelGGL(k_scan_blocks, dim3(nbk), dim3(1024), 0, st, m.ucnt.as<int>(),
                                   m.uoff.as<int>(), m.U.as<int>(), nblk_max, dtrk,
                                   m.brows.as<RowK>());
                ERP_CHECK_LAUNCH("k_scan_blocks");
            }
        });
        PT_.end(st);
        ERP_HIP_TRY(hipMemcpyAsync(hV.data(), dV, sizeof(int) * r, hipMemcpyDeviceToHost, st));
        if (pairk)
            ERP_HIP_TRY(hipMemcpyAsync(hU.data(), m.U.p, sizeof(int) * nbk, hipMemcpyDeviceToHost,
                                       st));
        ERP_HIP_TRY(hipStreamSynchronize(st));

        // facet records of the rows whose track has targets
        int64_t Vtot = 0;
        for (size_t i = 0; i < r; ++i) {
            hbase[i] = Vtot;
            const bool keep = hrows[i].shot >= 0 && tk[hrows[i].trk].T > 0;
            if (keep && hV[i] >= (1 << 28))
                throw ErpError{ERP_E_INVALID, "a shot illuminates 2^28 facets or more"};
            if (keep && !pairk) Vtot += hV[i];
        }
        if (pairk) {
            // block records: the union of each block's two rows
            for (int j = 0; j < nbk; ++j) {
                hubase[j] = Vtot;
                const bool keep = hrows[2 * j].shot >= 0 && tk[hrows[2 * j].trk].T > 0;
                if (keep && hU[j] >= (1 << 28))
                    throw ErpError{ERP_E_INVALID, "two shots illuminate 2^28 facets or more"};
                if (keep) Vtot += hU[j];
            }
            ERP_HIP_TRY(hipMemcpyAsync(m.ubase.p, hubase.data(), sizeof(int64_t) * nbk,
                                       hipMemcpyHostToDevice, st));
        }
        auto rnd = [](size_t n) { return (n + 255) & ~(size_t)255; };
        // bytes of cap records in this group's layout: block records (84 B)
        // or a shot's packed records (60 B); one buffer serves both, so a
        // capacity reached in one layout does not hold in the other
        auto rec_bytes = [&](size_t cap) {
            return pairk ? 5 * rnd(cap * 16) + rnd(cap * 4)
                         : cap * 8 + 3 * cap * 16 + cap * 4 + 5 * 256;
        };
        if (targ && ((size_t)Vtot > m.key.reccap || rec_bytes(m.key.reccap) > m.rec.bytes)) {
            // more records than the buffers hold: grow by a quarter more, and
            // capture again the graphs that read the record arrays
            const size_t cap = std::max((size_t)Vtot + (size_t)Vtot / 4 + 1024, m.key.reccap);
            m.rec.ensure(rec_bytes(cap));
            m.key.reccap = cap;
            m.gA2.reset();
            m.gChunk.reset();
        }
        Rec rec{};
        BlkRec brk{};
        if (targ && pairk) {
            const size_t cap = m.key.reccap;
            char* pb = m.rec.as<char>();
            auto take = [&](size_t n) { char* q = pb; pb += rnd(n); return q; };
            brk.rec.qa = (float4*)take(cap * 16);
            brk.rec.qb = (float4*)take(cap * 16);
            brk.rec.qc = (float4*)take(cap * 16);
            brk.rec.qd = (float*)take(cap * 4);
            brk.rec.itd = (double2*)take(cap * 16);
            brk.rec.src = (float4*)take(cap * 16);
            brk.umask = m.umask.as<unsigned long long>();
            brk.uoff = m.uoff.as<int>();
            brk.U = m.U.as<int>();
            brk.ubase = m.ubase.as<int64_t>();
        } else if (targ) {
            const size_t cap = m.key.reccap;
            char* pb = m.rec.as<char>();
            auto take = [&](size_t n) { char* q = pb; pb += (n + 255) & ~(size_t)255; return q; };
            rec.itd = (double*)take(cap * 8);
            rec.qa = (float4*)take(cap * 16);
            rec.qb = (float4*)take(cap * 16);
            rec.qc = (float4*)take(cap * 16);
            rec.qd = (float*)take(cap * 4);
            rec.cap = (int64_t)cap;
        }
        ERP_HIP_TRY(hipMemcpyAsync(dbase, hbase.data(), sizeof(int64_t) * r,
                                   hipMemcpyHostToDevice, st));

        // A2: surface kernel, surface range compression, target selection
        double Vsum = 0;
        for (size_t i = 0; i < r; ++i) Vsum += hV[i];
        PT_.begin(ERP_STAGE_SURFACE, st, Vsum);
        mt_launch(m.gA2, direct, st, [&] {
            if (pairk) {
                hipLaunchKernelGGL(k_fill_itd, dim3(grid_for(m.key.reccap)), dim3(256), 0, st,
                                   brk.rec.itd, (int64_t)m.key.reccap);
                ERP_CHECK_LAUNCH("k_fill_itd");
            }
            zero(m.psurf.as<float2>(), r * P, st);
            hipLaunchKernelGGL(k_init_bounds, dim3((R + 63) / 64), dim3(64), 0, st, dbounds, R);
            ERP_CHECK_LAUNCH("k_init_bounds");
            zero(m.srow.as<float2>(), srow_f2, st);
            hipLaunchKernelGGL(k_surface, dim3(nblk_max, R), dim3(NT), 0, st,
                               (const float*)nullptr, (int64_t)0, (int64_t)0, dshots,
                               m.mask.as<unsigned long long>(), m.off.as<int>(), dV, dbase,
                               nblk_max, kp0, rec, targ ? 1 : 0, 1, srow, (int64_t)0, dbounds,
                               det ? 1 : 0, smax, 0, 0, 0, dtrk, drows, brk,
                               targ ? m.bbox.as<float4>() : nullptr);
            ERP_CHECK_LAUNCH("k_surface");
            if (targ) {
                hipLaunchKernelGGL(k_shot_sphere, dim3(R), dim3(256), 0, st, dbounds,
                                   (const float4*)m.bbox.as<float4>(),
                                   (const int*)m.cnt.as<int>(), nblk_max, dtrk, drows);
                ERP_CHECK_LAUNCH("k_shot_sphere");
            }
            if (det) {
                hipLaunchKernelGGL(k_surface, dim3(nblk_max, R), dim3(NT), 0, st,
                                   (const float*)nullptr, (int64_t)0, (int64_t)0, dshots,
                                   m.mask.as<unsigned long long>(), m.off.as<int>(), dV, dbase,
                                   nblk_max, kp0, rec, 0, 1, srow, (int64_t)0, (Bounds*)nullptr,
                                   2, smax, 0, 0, 0, dtrk, drows, BlkRec{}, (float4*)nullptr);
                ERP_CHECK_LAUNCH("k_surface");
            }
            hipLaunchKernelGGL(k_surf_rows, dim3((nr + 255) / 256, R), dim3(256), 0, st,
                               m.psurf.as<float2>(), P, srow, nr, R, det ? 1 : 0, smax, dV);
            ERP_CHECK_LAUNCH("k_surf_rows");
            // reflected signal: each row with its track's range kernel
            float2* ps = m.psurf.as<float2>();
            fft_forward(ctx, planP, ps, R, st);
            hipLaunchKernelGGL(k_mul_spec_rows, dim3(grid_for(r * P)), dim3(256), 0, st, ps, ps,
                               P, R, dtrk, drows, 0);
            ERP_CHECK_LAUNCH("k_mul_spec_rows");
            fft_inverse(ctx, planP, ps, R, st);
            hipLaunchKernelGGL(k_window, dim3(grid_for(r * nr)), dim3(256), 0, st,
                               m.refl.as<float2>(), nr, ps, P, nr, R, 1.0f, 0);
            ERP_CHECK_LAUNCH("k_window");
            if (targ) {
                hipLaunchKernelGGL(k_target_select, dim3(R), dim3(1024), 0, st,
                                   (const float*)nullptr, (int64_t)0, 0,
                                   (const unsigned char*)nullptr, dshots, dV, 0.0, list1, list0,
                                   m.tcnt.as<int>(), Tcap, dtrk, drows);
                ERP_CHECK_LAUNCH("k_target_select");
                zero(m.Sacc.as<float2>(), r * P, st);
            } else {
                hipLaunchKernelGGL(k_finalize_rows, dim3(grid_for(r * nr)), dim3(256), 0, st,
                                   m.refl.as<float2>(), (const float2*)nullptr, dV, nr, R, dtrk,
                                   drows);
                ERP_CHECK_LAUNCH("k_finalize_rows");
            }
        });
        PT_.end(st);
        if (!targ) continue;

        ERP_HIP_TRY(hipMemcpyAsync(htc.data(), m.tcnt.p, sizeof(int) * 2 * r,
                                   hipMemcpyDeviceToHost, st));
        ERP_HIP_TRY(hipStreamSynchronize(st));
        for (size_t i = 0; i < r; ++i) tc1[i] = htc[2 * i];

        float2* A = m.A.as<float2>();
        float2* Z = m.Z.as<float2>();
        if (pairk) {
            // shot-pair kernel: items of a target and a block's two rows, from
            // the rows' target lists; launches of whole items of at most C rows
            int maxc = 0;
            for (size_t i = 0; i < r; ++i) maxc = std::max(maxc, htc[2 * i]);
            std::vector<int> hl(r * (size_t)maxc);
            if (maxc > 0) {
                ERP_HIP_TRY(hipMemcpy2DAsync(hl.data(), sizeof(int) * maxc, list1,
                                             sizeof(int) * Tcap, sizeof(int) * maxc, r,
                                             hipMemcpyDeviceToHost, st));
                ERP_HIP_TRY(hipStreamSynchronize(st));
            }
            merge_pair_lists(hl.data(), (size_t)maxc, htc.data(), R, bl);
            int jb = 0;
            size_t ib = 0;
            PairCursor cur;
            for (;;) {
                // deterministic: every row's slices independent of the launches
                if (det) fill_pair_chunk_canon(bl, cur, C, GS, hV.data(), pc);
                else fill_pair_chunk(bl, jb, ib, C, GS, hV.data(), pc);
                const int n = pc.rows;
                if (n == 0) break;
                const int nsl = (int)pc.gb.size(), nseg = (int)pc.sg_s.size();
                const int nit = (int)pc.items.size();
                if (nsl > gcap || nseg > segcap)
                    throw ErpError{ERP_E_INVALID, "multi-track launch tables overflow"};
                std::fill(htab.begin(), htab.end(), 0);
                std::copy(pc.rshot.begin(), pc.rshot.end(), htab.begin());
                std::copy(pc.ridx.begin(), pc.ridx.end(), htab.begin() + C);
                std::copy(pc.gb.begin(), pc.gb.end(), htab.begin() + 2 * C);
                std::copy(pc.ge.begin(), pc.ge.end(), htab.begin() + 2 * C + gcap);
                const size_t so = 2 * (size_t)C + 2 * gcap;
                for (int k = 0; k < nseg; ++k) {
                    htab[so + k] = pc.sg_s[k];
                    htab[so + segcap + k] = pc.gsb[k];
                    htab[so + 2 * segcap + k] = pc.gse[k];
                    const float sc = tk[hrows[pc.sg_s[k]].trk].tscale;
                    std::memcpy(&htab[so + 3 * segcap + 1 + k], &sc, sizeof(float));
                }
                htab[so + 3 * segcap] = n;
                ERP_HIP_TRY(hipMemcpyAsync(m.tabs.p, htab.data(), sizeof(int) * tab_ints,
                                           hipMemcpyHostToDevice, st));
                hitems.resize((size_t)C + 1);
                std::copy(pc.items.begin(), pc.items.end(), hitems.begin());
                hitems[C] = make_int4(nit, 0, 0, 0);
                ERP_HIP_TRY(hipMemcpyAsync(d_items, hitems.data(), sizeof(int4) * nit,
                                           hipMemcpyHostToDevice, st));
                ERP_HIP_TRY(hipMemcpyAsync(d_items + C, hitems.data() + C, sizeof(int4),
                                           hipMemcpyHostToDevice, st));
                PT_.begin(ERP_STAGE_TARGET_ACCUM, st, pc.triples);
                mt_launch(m.gChunk, direct, st, [&] {
                    // range kernel spectra per row, the item counter cleared
                    hipLaunchKernelGGL(k_pair_targets_rows, dim3((C + 255) / 256), dim3(256), 0,
                                       st, t_ptgt, t_pshot, list1, Tcap, t_n, d_next, d_G, dtrk,
                                       drows);
                    ERP_CHECK_LAUNCH("k_pair_targets_rows");
                    auto kern = det ? (s3 ? k_target_accum_pair<true, true, true>
                                          : k_target_accum_pair<false, true, true>)
                                    : (s3 ? k_target_accum_pair<true, true>
                                          : k_target_accum_pair<false, true>);
                    // persistent: one workgroup per CU takes items until *d_nitems
                    hipLaunchKernelGGL(kern, dim3(std::max(1, std::min(C, ctx->num_cu))),
                                       dim3(NTB), 0, st, (const float*)nullptr, (int64_t)0,
                                       (const int4*)d_items, 0, d_next, dshots,
                                       (const int64_t*)m.ubase.as<int64_t>(),
                                       (const int*)m.U.as<int>(), brk.rec, dbounds,
                                       (const erp_prism*)nullptr, 0, kp0, A, d_pbase, d_pwin, 0,
                                       d_next + 1, dtrk, drows, (const int*)d_nitems,
                                       (const int*)dV);
                    ERP_CHECK_LAUNCH("k_target_accum_pair");
                    target_pairs_half(ctx, planP, A, d_pbase, d_pwin, nr, C, nullptr, Z, st, d_G,
                                      t_n);
                    target_sum_half(planP, Z, A + (size_t)C * P, t_slb, t_sle, gcap, t_sgs, t_sgb,
                                    t_sge, segcap, m.Sacc.as<float2>(), 1.0f, st, t_scale, det,
                                    target_wl(planP));
                });
                PT_.end(st);
            }
        }

        // chunks: launches of at most C pairs in (row, target) order
        int cs = 0, cj = 0;
        for (; !pairk;) {
            fill_chunk(tc1.data(), R, cs, cj, C, GS, det, cf);
            const int n = cf.n();
            if (n == 0) break;
            const int nsl = (int)cf.gb.size(), nseg = (int)cf.sg_s.size();
            if (nsl > gcap || nseg > segcap)
                throw ErpError{ERP_E_INVALID, "multi-track launch tables overflow"};
            std::fill(htab.begin(), htab.end(), 0);
            std::copy(cf.pshot.begin(), cf.pshot.end(), htab.begin());
            std::copy(cf.ptgt.begin(), cf.ptgt.end(), htab.begin() + C);
            std::copy(cf.gb.begin(), cf.gb.end(), htab.begin() + 2 * C);
            std::copy(cf.ge.begin(), cf.ge.end(), htab.begin() + 2 * C + gcap);
            const size_t so = 2 * (size_t)C + 2 * gcap;
            for (int k = 0; k < nseg; ++k) {
                htab[so + k] = cf.sg_s[k];
                htab[so + segcap + k] = cf.gsb[k];
                htab[so + 2 * segcap + k] = cf.gse[k];
                const float sc = tk[hrows[cf.sg_s[k]].trk].tscale;
                std::memcpy(&htab[so + 3 * segcap + 1 + k], &sc, sizeof(float));
            }
            htab[so + 3 * segcap] = n;
            ERP_HIP_TRY(hipMemcpyAsync(m.tabs.p, htab.data(), sizeof(int) * tab_ints,
                                       hipMemcpyHostToDevice, st));
            double triples = 0;
            for (int i = 0; i < n; ++i) triples += hV[cf.pshot[i]];
            PT_.begin(ERP_STAGE_TARGET_ACCUM, st, triples);
            mt_launch(m.gChunk, direct, st, [&] {
                hipLaunchKernelGGL(k_pair_targets_rows, dim3((C + 255) / 256), dim3(256), 0, st,
                                   t_ptgt, t_pshot, list1, Tcap, t_n, d_next, d_G, dtrk, drows);
                ERP_CHECK_LAUNCH("k_pair_targets_rows");
                auto kern = det ? (s3 ? k_target_accum_fast<true, PH_F64, ACC_F64, true, true>
                                      : k_target_accum_fast<true, PH_F64, ACC_F64, false, true>)
                                : (s3 ? k_target_accum_fast<false, PH_F64, ACC_F64, true, true>
                                      : k_target_accum_fast<false, PH_F64, ACC_F64, false, true>);
                // persistent: two workgroups per CU take pairs until *t_n
                hipLaunchKernelGGL(kern, dim3(std::max(1, std::min(C, 2 * ctx->num_cu))),
                                   dim3(NTF), 0, st, (const float*)nullptr, (int64_t)0, t_pshot,
                                   t_ptgt, dshots, dV, dbase, rec, dbounds,
                                   (const erp_prism*)nullptr, 0, kp0, A, d_pbase, d_pwin, 0, 0,
                                   d_next, dtrk, drows, (const int*)t_n, d_next + 1);
                ERP_CHECK_LAUNCH("k_target_accum_fast");
                target_pairs_half(ctx, planP, A, d_pbase, d_pwin, nr, C, nullptr, Z, st, d_G,
                                  t_n);
                target_sum_half(planP, Z, A + (size_t)C * P, t_slb, t_sle, gcap, t_sgs, t_sgb,
                                t_sge, segcap, m.Sacc.as<float2>(), 1.0f, st, t_scale, det,
                                target_wl(planP));
            });
            PT_.end(st);
        }

        // tail: per row c = IFFT(S)/P; Bluestein DFT_L, crop, IDFT_L, decimate;
        // then refl + refr into the track's radargram
        PT_.begin(ERP_STAGE_SHOT_FINAL, st, (double)nb);
        mt_launch(m.gTail, direct, st, [&] {
            float2* Sacc = m.Sacc.as<float2>();
            float2* bu = m.blu.as<float2>();
            fft_inverse(ctx, planP, Sacc, R, st);
            hipLaunchKernelGGL(k_blue_in_rows, dim3(grid_for(r * M)), dim3(256), 0, st, bu, M,
                               Sacc, P, L, 1.0f / (float)P, R, dtrk, drows);
            ERP_CHECK_LAUNCH("k_blue_in_rows");
            fft_forward(ctx, planM, bu, R, st);
            hipLaunchKernelGGL(k_mul_spec_rows, dim3(grid_for(r * M)), dim3(256), 0, st, bu, bu,
                               M, R, dtrk, drows, 1);
            ERP_CHECK_LAUNCH("k_mul_spec_rows");
            fft_inverse(ctx, planM, bu, R, st);
            hipLaunchKernelGGL(k_blue_mid_rows, dim3(grid_for(r * M)), dim3(256), 0, st, bu, M,
                               L, R, dtrk, drows);
            ERP_CHECK_LAUNCH("k_blue_mid_rows");
            fft_forward(ctx, planM, bu, R, st);
            hipLaunchKernelGGL(k_mul_spec_rows, dim3(grid_for(r * M)), dim3(256), 0, st, bu, bu,
                               M, R, dtrk, drows, 1);
            ERP_CHECK_LAUNCH("k_mul_spec_rows");
            fft_inverse(ctx, planM, bu, R, st);
            hipLaunchKernelGGL(k_blue_out_rows, dim3(grid_for(r * nr)), dim3(256), 0, st,
                               m.refr.as<float2>(), nr, bu, M, 1.0f / (float)L, R, dV, dtrk,
                               drows);
            ERP_CHECK_LAUNCH("k_blue_out_rows");
            hipLaunchKernelGGL(k_finalize_rows, dim3(grid_for(r * nr)), dim3(256), 0, st,
                               m.refl.as<float2>(), m.refr.as<float2>(), dV, nr, R, dtrk, drows);
            ERP_CHECK_LAUNCH("k_finalize_rows");
        });
        PT_.end(st);
    }
    if (targ) check_windows(d_next + 1, st);
}

}  // namespace

void erp_run_tracks_impl(erp_ctx* ctx, const erp_track* tracks, int64_t ntracks, uint32_t flags,
                         hipStream_t st) {
    const uint32_t known = ERP_RUN_RAW_SURFACE | ERP_RUN_DETERMINISTIC | ERP_RUN_PHASE_F32 |
                           ERP_RUN_PHASE_DF | ERP_RUN_ACC_F32 | ERP_RUN_ACC_COMP;
    if (flags & ~known) throw ErpError{ERP_E_INVALID, "unknown run flags"};
    const bool det = flags & ERP_RUN_DETERMINISTIC;
    // groups of one range length and second-medium setting, in track order
    std::vector<std::pair<std::pair<int, bool>, std::vector<int64_t>>> groups;
    std::vector<int64_t> solo;
    for (int64_t i = 0; i < ntracks; ++i) {
        const erp_track& t = tracks[i];
        if (!mt_eligible(t, flags)) {
            solo.push_back(i);
            continue;
        }
        const std::pair<int, bool> gk{t.par->nr, t.par->eps_3 > 0.f};
        auto it = std::find_if(groups.begin(), groups.end(),
                               [&](const auto& g) { return g.first == gk; });
        if (it == groups.end()) groups.push_back({gk, {i}});
        else it->second.push_back(i);
    }
    for (int64_t i : solo) {
        const erp_track& t = tracks[i];
        erp_run_impl(ctx, t.par, t.mesh, t.targets, t.shots, t.ns, nullptr, 0,
                     static_cast<float2*>(t.out_dev), static_cast<int32_t*>(t.nvalid_dev), flags,
                     st, ctx->work, &ctx->prof);
    }
    for (const auto& g : groups) mt_run_group(ctx, tracks, g.second, g.first.first, g.first.second,
                                              det, st);
}
