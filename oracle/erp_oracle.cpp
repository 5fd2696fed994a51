/*
 * erp_oracle.cpp — CPU restatement of the EuropaRaypaths per-shot path.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle the tests compare with.  The product
 * (europaraypaths_amd/) never links, loads or calls it; only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg do.
 *
 * The reference (src/CUDA, CUDA + cuFFT + Thrust) cannot be built in this
 * image (no nvcc / cuFFT / Thrust / nlohmann::json), so this file restates
 * its arithmetic op for op, templated on the working precision R:
 *   R = double : the FP64 parity oracle (every step in IEEE double).
 *   R = float  : the reference's FP32 arithmetic (the "reference CPU path").
 * Every function cites the reference file:line it follows.  Constants are
 * the reference's literals (3.14159265f, 3.14159, M_PI, 299792458, 1e-6f),
 * kept as written so that oracle and GPU agree on the same formulas.
 *
 * Reference quirks (numbered as in the reference review) and how they are
 * handled here — identically in the HIP product:
 *   Q2  range bin = trunc((R - rst)/dr), valid iff -1 < x < nr; the second
 *       bin of the split is dropped when it equals nr (the reference writes it
 *       one past the end, radar_kernels.cu:757-758).
 *   Q8  incoherent phase from a counter-based hash of (seed, shot, facet)
 *       instead of clock64() (radar_kernels.cu:689-711).
 *   Q9  a shot with no illuminated facet yields an all-zero trace
 *       (sim.cu:693-697 writes no file for it).
 *   Q12 FP64 aperture predicate: acos(c) < ap  <=>  c > cos(ap) && c <= 1
 *       (the reference's unclamped acos drops c > 1 as NaN).
 *   Q13 specular + circular convolution: no reflected signal (sim.cu:765-767).
 * Mode combinations the reference cannot execute (they read or write
 * unallocated buffers) return ERP_E_UNSUPPORTED:
 *   non-specular without linear convolution (sim.cu:431-439 allocates nr
 *   floats, sim.cu:778 writes 2nr); convolution=false without specular
 *   (sim.cu:778,796 use the never-allocated d_chirp); type!=0 targets without
 *   convolution (convolveComplex writes the NULL d_refr_temp, sim.cu:410,912)
 *   or with circular convolution (sim.cu:872 reads 2nr chirp floats from an
 *   nr buffer).
 */
#include <cmath>
#include <complex>
#include <cstring>
#include <cstdint>
#include <vector>
#include <type_traits>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "erp_oracle.h"

namespace {

template <class R> using cx = std::complex<R>;
template <class R> constexpr bool is_f32 = std::is_same<R, float>::value;

/* geometry_kernels.cu:45  const float pi = 3.14159265f; also the literal
 * 3.14159265f of radar_kernels.cu:105,250,324,374,682 and file_io.cu:45. */
template <class R> constexpr R PI_F() { return (R)3.14159265f; }
/* friis/radarEq/alpha use the double literal 3.14159 (geometry_kernels.cu:388,
 * 411; file_io.cu:196-205). */
constexpr double PI_5 = 3.14159;

/* Planted mistakes for the pin tests (tests/test_oracle_closed_forms.py):
 * 0 = none, the oracle as it is.  Each other value swaps one detail of the
 * restatement for a plausible slip, so that a test can show that its pin
 * catches that slip.  Set only through erp_oracle_set_mutation. */
enum {
    MUT_NONE = 0,
    MUT_FACET_G_PI_ONCE = 1,    /* facet_G's sinc argument without the 2nd pi */
    MUT_RADAREQ_NO_V = 2,       /* radarEq without the coherent-sum factor V */
    MUT_RADAREQ_4PI_CUBED = 3,  /* radarEq over (4 pi)^3 as its comment reads */
    MUT_TPH_SWAPPED = 4,        /* Tph = atan2(I.v, I.u): u and v exchanged */
    MUT_TARGETTH_FACET_N = 5,   /* TargetTh against the facet's normal */
    MUT_INCOHERENT_HALF = 6,    /* incoherent phases drawn from [0, pi) */
    MUT_FACETRERAD_PI_TWICE = 7,/* facetReradiation's sinc argument times pi */
    MUT_REFR_INDEX_DROPPED = 8, /* target range Itd + Ttd instead of Itd + n Ttd */
    MUT_NO_TRANSMISSION = 9,    /* target power without the (1 - rho^2) factor */
};
static int g_mut = MUT_NONE;

/* ------------------------------------------------------------------ math */

/* elementary_kernels.cu:50-53: sinc with the double M_PI. */
template <class R> inline R sinc_ref(R x) {
    if (std::fabs(x) < (R)1e-6f) return (R)1;
    if constexpr (is_f32<R>) {
        double a = M_PI * (double)x;
        return (float)((double)sinf((float)a) / a);
    } else {
        double a = M_PI * x;
        return std::sin(a) / a;
    }
}

/* radar_kernels.cu:658-677 (pi appears twice: here and inside sinc). */
template <class R> inline R facet_G(R th, R ph, R lam, R fs) {
    R piLlam = (PI_F<R>() * fs) / lam;
    if (g_mut == MUT_FACET_G_PI_ONCE) piLlam = fs / lam;
    R s1 = sinc_ref<R>(piLlam * std::sin(th) * std::cos(ph));
    R s2 = sinc_ref<R>(piLlam * std::sin(th) * std::sin(ph));
    R d = s1 * s2;
    R Dn = d * d;
    R Dmax = ((R)4 * PI_F<R>() * fs * fs) / (lam * lam);
    return Dn * Dmax;
}

/* geometry_kernels.cu:410-412 (the denominator is evaluated in double). */
template <class R> inline R friis(R P, R Gt, R Gr, R lam, R dist) {
    R num = P * lam * lam * Gt * Gr;
    double den = std::pow(4 * PI_5, 2) * (double)dist * (double)dist;
    return (R)((double)num / den);
}

/* geometry_kernels.cu:384-393: (4*pi)^2 with the coherent-sum factor nf. */
template <class R> inline R radarEq(R P, R G, R fs, R lam, R dist, int64_t nf) {
    (void)lam;
    R num = P * G * G * fs * fs * fs * fs;
    double den = std::pow(4 * PI_5, 2) * (double)dist * (double)dist *
                 (double)dist * (double)dist;
    if (g_mut == MUT_RADAREQ_NO_V) nf = 1;
    if (g_mut == MUT_RADAREQ_4PI_CUBED) den *= 4 * PI_5;
    return (R)((double)nf * ((double)num / den));
}

/* geometry_kernels.cu:353-381 (the complex prefactor is computed but unused
 * by the reference; only the squared sinc product is returned). */
template <class R> inline R facetReradiation(R dist, R th, R ph, R lam, R fs) {
    (void)dist;
    const R a = (g_mut == MUT_FACETRERAD_PI_TWICE) ? PI_F<R>() * fs / lam : fs / lam;
    R s1 = sinc_ref<R>(a * std::sin(th) * std::cos(ph));
    R s2 = sinc_ref<R>(a * std::sin(th) * std::sin(ph));
    R v = (fs * fs / lam) * s1 * s2;
    return v * v;
}

/* geometry_kernels.cu:490-501 */
template <class R> inline R snellsLaw(R th, R eps_1, R eps_2) {
    R sin_th = std::sin(th);
    R sin_Rth = std::sqrt(eps_1 / eps_2) * sin_th;
    if (sin_Rth > (R)1) return (R)1e4;
    return std::asin(sin_Rth);
}

/* radar_kernels.cu:167-189 */
template <class R> inline R rerad_funct(int fn, R th1) {
    const R pi = PI_F<R>();
    switch (fn) {
    case 0: return 1;
    case 1: return (th1 * ((R)180 / pi) < (R)1) ? 1 : 0;
    case 2: return (th1 * ((R)180 / pi) < (R)0.5) ? 1 : 0;
    case 3: return (th1 * (pi / (R)180) < (R)3) ? 1 : 0;
    case 4: {
        R sigma = pi / (R)180;
        return (R)std::exp((R)(-0.5 * (double)((th1 - 0) * (th1 - 0)) /
                               (double)(sigma * sigma)));
    }
    default: return 1;
    }
}

/* geometry_kernels.cu:207-269 (direction must be normalised). */
template <class R>
inline R prismIntersectionDistance(R ox, R oy, R oz, R dx, R dy, R dz, R xmin,
                                   R xmax, R ymin, R ymax, R zmin, R zmax, R t) {
    const R INF = (R)1e30f;
    R tx1, tx2, ty1, ty2, tz1, tz2;
    if (std::fabs(dx) < (R)1e-12f) {
        if (ox < xmin || ox > xmax) return 0;
        tx1 = -INF; tx2 = INF;
    } else {
        tx1 = (xmin - ox) / dx; tx2 = (xmax - ox) / dx;
        if (tx1 > tx2) std::swap(tx1, tx2);
    }
    if (std::fabs(dy) < (R)1e-12f) {
        if (oy < ymin || oy > ymax) return 0;
        ty1 = -INF; ty2 = INF;
    } else {
        ty1 = (ymin - oy) / dy; ty2 = (ymax - oy) / dy;
        if (ty1 > ty2) std::swap(ty1, ty2);
    }
    if (std::fabs(dz) < (R)1e-12f) {
        if (oz < zmin || oz > zmax) return 0;
        tz1 = -INF; tz2 = INF;
    } else {
        tz1 = (zmin - oz) / dz; tz2 = (zmax - oz) / dz;
        if (tz1 > tz2) std::swap(tz1, tz2);
    }
    R t_entry = std::max(std::max(tx1, ty1), tz1);
    R t_exit = std::min(std::min(tx2, ty2), tz2);
    if (t_entry < 0) t_entry = 0;
    if (t_exit > t) t_exit = t;
    if (t_entry >= t_exit) return 0;
    return t_exit - t_entry;
}

/* Counter-based replacement for randomPhasor's clock64() seed
 * (radar_kernels.cu:689-711): splitmix64 finaliser of (seed, shot, facet);
 * the top 24 bits give u in [0,1) and the phase is 2*pi*u.  The HIP kernel
 * evaluates the same integer hash, so incoherent runs are reproducible and
 * comparable value by value. */
inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double incoherent_u(uint64_t seed, int64_t shot, int64_t facet) {
    uint64_t h = mix64(seed ^ mix64((uint64_t)shot * 0x100000001B3ull ^
                                    mix64((uint64_t)facet)));
    return (double)(h >> 40) * (1.0 / 16777216.0);
}

/* ------------------------------------------------------------------- FFT */

template <class R> struct Pow2Plan {
    int64_t n = 0;
    std::vector<int64_t> rev;
    std::vector<cx<R>> tw; /* exp(-2 pi i k / n), k < n/2, from double */
    void init(int64_t nn) {
        n = nn;
        int lg = 0;
        while (((int64_t)1 << lg) < n) ++lg;
        rev.assign(n, 0);
        for (int64_t i = 0; i < n; ++i) {
            int64_t r = 0;
            for (int b = 0; b < lg; ++b)
                if (i & ((int64_t)1 << b)) r |= (int64_t)1 << (lg - 1 - b);
            rev[i] = r;
        }
        tw.resize(std::max<int64_t>(1, n / 2));
        for (int64_t k = 0; k < n / 2; ++k) {
            double a = -2.0 * M_PI * (double)k / (double)n;
            tw[k] = cx<R>((R)std::cos(a), (R)std::sin(a));
        }
    }
};

template <class R> void fft_pow2(cx<R>* a, const Pow2Plan<R>& p, bool inverse) {
    const int64_t n = p.n;
    for (int64_t i = 0; i < n; ++i)
        if (i < p.rev[i]) std::swap(a[i], a[p.rev[i]]);
    for (int64_t len = 2; len <= n; len <<= 1) {
        const int64_t half = len >> 1, step = n / len;
        for (int64_t i = 0; i < n; i += len) {
            for (int64_t j = 0; j < half; ++j) {
                cx<R> w = p.tw[j * step];
                if (inverse) w = std::conj(w);
                cx<R> u = a[i + j], v = a[i + j + half] * w;
                a[i + j] = u + v;
                a[i + j + half] = u - v;
            }
        }
    }
}

inline bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

/* Generic-length DFT (radix-2 or Bluestein), unnormalised, cuFFT sign
 * convention: forward exp(-2 pi i jk/n), inverse exp(+2 pi i jk/n). */
template <class R> struct DFT {
    int64_t n = 0, m = 0;
    Pow2Plan<R> p2;
    std::vector<cx<R>> w;  /* exp(-pi i k^2 / n) */
    std::vector<cx<R>> Bf; /* FFT_m of the conjugate chirp */
    std::vector<cx<R>> buf;
    void init(int64_t nn) {
        n = nn;
        if (is_pow2(n)) {
            p2.init(n);
            return;
        }
        m = 1;
        while (m < 2 * n - 1) m <<= 1;
        p2.init(m);
        w.resize(n);
        for (int64_t k = 0; k < n; ++k) {
            int64_t q = (int64_t)(((__int128)k * k) % (2 * n));
            double a = -M_PI * (double)q / (double)n;
            w[k] = cx<R>((R)std::cos(a), (R)std::sin(a));
        }
        Bf.assign(m, cx<R>(0, 0));
        Bf[0] = std::conj(w[0]);
        for (int64_t k = 1; k < n; ++k) {
            Bf[k] = std::conj(w[k]);
            Bf[m - k] = std::conj(w[k]);
        }
        fft_pow2<R>(Bf.data(), p2, false);
        buf.resize(m);
    }
    void run(cx<R>* a, bool inverse) {
        if (m == 0) {
            fft_pow2<R>(a, p2, inverse);
            return;
        }
        if (inverse)
            for (int64_t i = 0; i < n; ++i) a[i] = std::conj(a[i]);
        std::fill(buf.begin(), buf.end(), cx<R>(0, 0));
        for (int64_t i = 0; i < n; ++i) buf[i] = a[i] * w[i];
        fft_pow2<R>(buf.data(), p2, false);
        for (int64_t i = 0; i < m; ++i) buf[i] *= Bf[i];
        fft_pow2<R>(buf.data(), p2, true);
        const R sc = (R)1 / (R)m;
        for (int64_t i = 0; i < n; ++i) a[i] = w[i] * buf[i] * sc;
        if (inverse)
            for (int64_t i = 0; i < n; ++i) a[i] = std::conj(a[i]);
    }
};

/* ------------------------------------------------------------ parameters */

template <class R> struct Par {
    R lam, rng_res, P, fs, ks, eps_1, eps_2, c, c_1, c_2, nu_1, nu_2, alpha1,
        alpha2, aperture, smpl, B, dr;
    int nr, rerad, pol;
    uint32_t mode;
    double cos_ap;
    uint64_t seed;
    bool conv, linear, specular, lossless, incoherent, disable_surface;
    R eps_3, z_int;  /* second subsurface medium (eps_3 <= 0: none) */
    explicit Par(const erp_params& p) {
        lam = p.lam; rng_res = p.rng_res; P = p.P; fs = p.fs; ks = p.ks;
        eps_1 = p.eps_1; eps_2 = p.eps_2; c = p.c; c_1 = p.c_1; c_2 = p.c_2;
        nu_1 = p.nu_1; nu_2 = p.nu_2; alpha1 = p.alpha1; alpha2 = p.alpha2;
        aperture = p.aperture; smpl = p.smpl; B = p.B; dr = p.dr;
        nr = p.nr; rerad = p.rerad_funct; pol = p.pol; mode = p.mode;
        cos_ap = p.cos_ap; seed = p.seed;
        conv = mode & ERP_MODE_CONVOLUTION;
        linear = mode & ERP_MODE_CONV_LINEAR;
        specular = mode & ERP_MODE_SPECULAR;
        lossless = mode & ERP_MODE_LOSSLESS;
        incoherent = mode & ERP_MODE_INCOHERENT;
        disable_surface = mode & ERP_MODE_DISABLE_SURFACE;
        eps_3 = p.eps_3; z_int = p.z_int;
    }
};

/* Second subsurface medium (an extension of the reference, which has one
 * medium eps_2, file_io.cu:81): the part of the straight facet -> target
 * segment that lies below the horizontal interface z = z_int, as a fraction
 * of its length.  0 when the target is not below the interface, 1 when the
 * facet is not above it. */
template <class R> inline R below_fraction(R fz, R tz, R zi) {
    if (!(tz < zi)) return 0;
    if (!(fz > zi)) return 1;
    return (zi - tz) / (fz - tz);
}
/* Optical range of a target term: rngt = Itd + n Ttd (accumulateTarget,
 * radar_kernels.cu:767-830) plus (n_3 - n) times the length below the
 * interface; *tr receives the two-way-per-leg power transmission 1 - rho^2
 * of the interface (1 when the ray does not cross it). */
template <class R>
inline R target_range(const Par<R>& p, R Itd, R Ttd, R n, R fz, R tz, R* tr) {
    R rngt = Itd + Ttd * n;
    *tr = 1;
    if (p.eps_3 > 0) {
        const R frac = below_fraction<R>(fz, tz, p.z_int);
        if (frac > 0) {
            const R n3 = std::sqrt(p.eps_3);
            rngt = rngt + (Ttd * frac) * (n3 - n);
            if (frac < 1) {
                const R rho = (n - n3) / (n + n3);
                *tr = 1 - rho * rho;
            }
        }
    }
    return rngt;
}

struct MeshView {
    const float *x, *y, *z, *nx, *ny, *nz, *ux, *uy, *uz, *vx, *vy, *vz;
    int64_t F;
    MeshView(const float* soa, int64_t n) : F(n) {
        const float* a[12];
        for (int i = 0; i < 12; ++i) a[i] = soa + (size_t)i * (size_t)n;
        x = a[0]; y = a[1]; z = a[2]; nx = a[3]; ny = a[4]; nz = a[5];
        ux = a[6]; uy = a[7]; uz = a[8]; vx = a[9]; vy = a[10]; vz = a[11];
    }
};

/* Per illuminated facet quantities of one shot. */
template <class R> struct FacetRec {
    int64_t idx;
    R Itd, Ith, Iph, Rth, fRe, fRfrC;
};

/* Aperture predicate of compSourceInclination + cropByAperture
 * (geometry_kernels.cu:75-100, 143-204).  FP32: the reference's unclamped
 * acosf compared against (pi/180)*aperture.  FP64: the same comparison as
 * c > cos(ap) && c <= 1 on the double dot product. */
template <class R>
inline bool illuminated(const Par<R>& p, const MeshView& m, int64_t i,
                        const erp_shot& s) {
    if constexpr (is_f32<R>) {
        float sx = (float)s.sx, sy = (float)s.sy, sz = (float)s.sz;
        float dx = m.x[i] - sx, dy = m.y[i] - sy, dz = m.z[i] - sz;
        float Itd = sqrtf(dx * dx + dy * dy + dz * dz);
        float Ix = (m.x[i] - sx) / Itd, Iy = (m.y[i] - sy) / Itd,
              Iz = (m.z[i] - sz) / Itd;
        float c = (-1 * Ix) * s.snx + (-1 * Iy) * s.sny + (-1 * Iz) * s.snz;
        float th = acosf(c);
        return th < (PI_F<float>() / 180) * p.aperture;
    } else {
        double dx = (double)m.x[i] - s.sx, dy = (double)m.y[i] - s.sy,
               dz = (double)m.z[i] - s.sz;
        double Itd = std::sqrt(dx * dx + dy * dy + dz * dz);
        double Ix = dx / Itd, Iy = dy / Itd, Iz = dz / Itd;
        double c = (-Ix) * (double)s.snx + (-Iy) * (double)s.sny +
                   (-Iz) * (double)s.snz;
        return c > p.cos_ap && c <= 1.0;
    }
}

/* compIncidentRays + compRefractedRays + compReflectedEnergy for facet i
 * (geometry_kernels.cu:103-140, 490-517, 426-487).  V = illuminated count
 * (the nfacets argument of compReflectedEnergy, sim.cu:728-730). */
template <class R>
inline FacetRec<R> facet_record(const Par<R>& p, const MeshView& m, int64_t i,
                                const erp_shot& s, int64_t V) {
    FacetRec<R> r;
    r.idx = i;
    R Ix, Iy, Iz;
    if constexpr (is_f32<R>) {
        float dx = m.x[i] - (float)s.sx, dy = m.y[i] - (float)s.sy,
              dz = m.z[i] - (float)s.sz;
        r.Itd = sqrtf(dx * dx + dy * dy + dz * dz);
        Ix = dx / r.Itd; Iy = dy / r.Itd; Iz = dz / r.Itd;
    } else {
        double dx = (double)m.x[i] - s.sx, dy = (double)m.y[i] - s.sy,
               dz = (double)m.z[i] - s.sz;
        r.Itd = std::sqrt(dx * dx + dy * dy + dz * dz);
        Ix = dx / r.Itd; Iy = dy / r.Itd; Iz = dz / r.Itd;
    }
    R cI = (-Ix) * (R)m.nx[i] + (-Iy) * (R)m.ny[i] + (-Iz) * (R)m.nz[i];
    r.Ith = std::acos(std::max((R)-1, std::min((R)1, cI)));
    r.Iph = std::atan2(Ix * (R)m.ux[i] + Iy * (R)m.uy[i] + Iz * (R)m.uz[i],
                       Ix * (R)m.vx[i] + Iy * (R)m.vy[i] + Iz * (R)m.vz[i]);
    r.Rth = snellsLaw<R>(r.Ith, p.eps_1, p.eps_2);
    /* compReflectedEnergy */
    R fRe;
    if (p.specular) {
        fRe = facetReradiation<R>(r.Itd, 2 * r.Ith, -1 * r.Iph, p.lam, p.fs);
        fRe = fRe * radarEq<R>(p.P, (R)s.grefl_lin, p.fs, p.lam, r.Itd, V);
    } else {
        fRe = 1;
    }
    R rho = 0, fRfrC = 0;
    if (p.pol == 0) {
        rho = (p.nu_2 * std::cos(r.Ith) - p.nu_1 * std::cos(r.Rth)) /
              (p.nu_2 * std::cos(r.Ith) + p.nu_1 * std::cos(r.Rth));
        fRfrC = 1 - (rho * rho);
    } else if (p.pol == 1) {
        rho = (p.nu_2 * std::cos(r.Rth) - p.nu_1 * std::cos(r.Ith)) /
              (p.nu_2 * std::cos(r.Rth) + p.nu_1 * std::cos(r.Ith));
        fRfrC = 1 - (rho * rho);
    }
    fRe = fRe * rho * rho;
    if (r.Rth > (R)1e3) fRfrC = 0;
    R atm = std::exp((R)-4 * p.alpha1 * r.Itd);
    fRe = fRe * atm;
    fRfrC = fRfrC * atm;
    R ci = p.ks * std::cos(r.Ith);
    R rough = std::exp((R)-4 * (ci * ci));
    fRe = fRe * rough;
    r.fRe = fRe;
    r.fRfrC = fRfrC;
    return r;
}

/* Range-bin rule shared by every scatter: (short)((R - rst)/dr) with the
 * CUDA saturating conversion; valid iff -1 < x < nr (radar_kernels.cu:
 * 335-341, 736-750, 803-817, 872-878). */
template <class R> inline bool bin_of(R rng, R rst, R dr, int nr, int* b, R* w) {
    R x = (rng - rst) / dr;
    if (!(x > (R)-1) || !(x < (R)nr)) return false;
    int bi = (int)x;
    *b = bi;
    *w = x - (R)bi;
    return true;
}

/* phasor(r, lam) = exp(-i 2 pi/lam r)  (radar_kernels.cu:681-686). */
template <class R> inline cx<R> phasor_neg(R r, R lam) {
    if constexpr (is_f32<R>) {
        float ph = -(2.0f * 3.14159265f / lam) * r;
        return cx<R>(cosf(ph), sinf(ph));
    } else {
        double ph = -(2.0 * PI_F<double>() / lam) * r;
        return cx<R>(std::cos(ph), std::sin(ph));
    }
}
/* exp(+i 4 pi/lam r)  (radar_kernels.cu:105,250,324,374). */
template <class R> inline cx<R> phasor_pos4(R r, R lam) {
    R ph = ((R)4 * PI_F<R>() / lam) * r;
    return cx<R>(std::cos(ph), std::sin(ph));
}

template <class R> inline void split_add(cx<R>* tr, int nr, int b, R w, cx<R> c) {
    tr[b] += c * ((R)1 - w);
    if (b + 1 < nr) tr[b + 1] += c * w; /* Q2: drop the write past the end */
}

/* Chirp tables (radar_kernels.cu:454-466, 489-503). */
template <class R> void gen_chirp_padded(std::vector<R>& h, int nr, R dr, R res) {
    const int n2 = 2 * nr, kc = nr / 2;
    h.resize(n2);
    for (int i = 0; i < n2; ++i) {
        R off = (R)(i - kc) * dr;
        h[i] = sinc_ref<R>(off / res);
    }
}
template <class R> void gen_chirp_circ(std::vector<R>& h, int nr, R rst, R dr, R res) {
    h.resize(nr);
    for (int ir = 0; ir < nr; ++ir) {
        R r = rst + (R)ir * dr;
        h[ir] = sinc_ref<R>((r - rst) / res) + sinc_ref<R>((r - (rst + (R)nr * dr)) / res);
    }
}

/* Per-thread working state: DFT plans and chirp spectra cache. */
template <class R> struct Work {
    int nr = -1;
    DFT<R> d2, dL, d1; /* lengths 2nr, 2nr-1, nr */
    R cached_res[2] = {-1, -1};
    std::vector<cx<R>> Hlin[2];
    R circ_key[2] = {-1, -1};
    std::vector<cx<R>> Hcirc;
    std::vector<cx<R>> pad, padk, ptr, tmp;
    void ensure(int n) {
        if (nr == n) return;
        nr = n;
        d2.init(2 * (int64_t)n);
        dL.init(2 * (int64_t)n - 1);
        d1.init(n);
        cached_res[0] = cached_res[1] = -1;
        circ_key[0] = circ_key[1] = -1;
    }
    /* FFT of the padded chirp for resolution res (cached for 2 values). */
    const std::vector<cx<R>>& chirp_lin(R dr, R res) {
        int slot = -1;
        for (int s = 0; s < 2; ++s)
            if (cached_res[s] == res) slot = s;
        if (slot < 0) {
            slot = (cached_res[0] < 0) ? 0 : 1;
            if (cached_res[0] >= 0 && cached_res[1] >= 0) slot = 0;
            std::vector<R> h;
            gen_chirp_padded<R>(h, nr, dr, res);
            Hlin[slot].assign(2 * nr, cx<R>(0, 0));
            for (int i = 0; i < 2 * nr; ++i) Hlin[slot][i] = cx<R>(h[i], 0);
            d2.run(Hlin[slot].data(), false);
            cached_res[slot] = res;
        }
        return Hlin[slot];
    }
    const std::vector<cx<R>>& chirp_circ(R rst, R dr, R res) {
        if (!(circ_key[0] == rst && circ_key[1] == res)) {
            std::vector<R> h;
            gen_chirp_circ<R>(h, nr, rst, dr, res);
            Hcirc.assign(nr, cx<R>(0, 0));
            for (int i = 0; i < nr; ++i) Hcirc[i] = cx<R>(h[i], 0);
            d1.run(Hcirc.data(), false);
            circ_key[0] = rst;
            circ_key[1] = res;
        }
        return Hcirc;
    }
};

/* convolvePhasorChirpLinear (radar_kernels.cu:554-649): zero-pad to 2nr,
 * FFT, multiply by the chirp spectrum, IFFT, scale 1/(2nr), keep
 * [nr/2, nr/2 + nr). */
template <class R>
void conv_linear(Work<R>& w, const cx<R>* x, R dr, R res, cx<R>* y) {
    const int nr = w.nr, n2 = 2 * nr, kc = nr / 2;
    const auto& H = w.chirp_lin(dr, res);
    w.pad.assign(n2, cx<R>(0, 0));
    std::copy(x, x + nr, w.pad.begin());
    w.d2.run(w.pad.data(), false);
    for (int i = 0; i < n2; ++i) w.pad[i] *= H[i];
    w.d2.run(w.pad.data(), true);
    const R sc = (R)1.0f / (R)n2;
    for (int j = 0; j < nr; ++j) y[j] = w.pad[kc + j] * sc;
}

/* convolvePhasorChirp (radar_kernels.cu:516-552): circular, length nr. */
template <class R>
void conv_circ(Work<R>& w, const cx<R>* x, R rst, R dr, R res, cx<R>* y) {
    const int nr = w.nr;
    const auto& H = w.chirp_circ(rst, dr, res);
    w.pad.assign(x, x + nr);
    w.d1.run(w.pad.data(), false);
    for (int i = 0; i < nr; ++i) w.pad[i] *= H[i];
    w.d1.run(w.pad.data(), true);
    const R sc = (R)1.0f / (R)nr;
    for (int j = 0; j < nr; ++j) y[j] = w.pad[j] * sc;
}

/* convolveComplex + cropSpectrum + takeEveryOtherComplex
 * (elementary_kernels.cu:258-353, 187-194): L = 2nr-1. */
template <class R>
void conv_complex(Work<R>& w, const cx<R>* sig, const cx<R>* ker, R Fs, R Bw,
                  cx<R>* out) {
    const int nr = w.nr, L = 2 * nr - 1;
    w.pad.assign(L, cx<R>(0, 0));
    w.padk.assign(L, cx<R>(0, 0));
    std::copy(sig, sig + nr, w.pad.begin());
    std::copy(ker, ker + nr, w.padk.begin());
    w.dL.run(w.pad.data(), false);
    w.dL.run(w.padk.data(), false);
    for (int k = 0; k < L; ++k) {
        R freq;
        if (k <= L / 2)
            freq = (R)k * Fs / (R)L;
        else
            freq = (R)(k - L) * Fs / (R)L;
        if (std::fabs(freq) > Bw) w.padk[k] = cx<R>(0, 0);
    }
    for (int k = 0; k < L; ++k) w.pad[k] *= w.padk[k];
    w.dL.run(w.pad.data(), true);
    const R sc = (R)1.0f / (R)L;
    for (int i = 0; i < nr; ++i) out[i] = w.pad[2 * i] * sc;
}

/* Target forced-ray quantities (geometry_kernels.cu:272-350, 520-584). */
template <class R> struct PairRec {
    R Ttd, Tth, TargetTh, Tph, EI, EO;
};

template <class R>
inline PairRec<R> pair_record(const Par<R>& p, const MeshView& m,
                              const FacetRec<R>& f, R tx, R ty, R tz, R tnx,
                              R tny, R tnz, const erp_prism* pr, int64_t np) {
    PairRec<R> q;
    const int64_t i = f.idx;
    R Ix, Iy, Iz;
    if constexpr (is_f32<R>) {
        float dx = tx - m.x[i], dy = ty - m.y[i], dz = tz - m.z[i];
        q.Ttd = sqrtf(dx * dx + dy * dy + dz * dz);
        Ix = dx / q.Ttd; Iy = dy / q.Ttd; Iz = dz / q.Ttd;
    } else {
        double dx = tx - (double)m.x[i], dy = ty - (double)m.y[i],
               dz = tz - (double)m.z[i];
        q.Ttd = std::sqrt(dx * dx + dy * dy + dz * dz);
        Ix = dx / q.Ttd; Iy = dy / q.Ttd; Iz = dz / q.Ttd;
    }
    R c1 = (-Ix) * (R)m.nx[i] + (-Iy) * (R)m.ny[i] + (-Iz) * (R)m.nz[i];
    q.Tth = std::acos(std::max((R)-1, std::min((R)1, c1)));
    if (g_mut == MUT_TARGETTH_FACET_N) {
        tnx = (R)m.nx[i];
        tny = (R)m.ny[i];
        tnz = (R)m.nz[i];
    }
    R c2 = (-Ix) * tnx + (-Iy) * tny + (-Iz) * tnz;
    q.TargetTh = std::acos(std::max((R)-1, std::min((R)1, c2)));
    q.Tph = std::atan2(Ix * (R)m.ux[i] + Iy * (R)m.uy[i] + Iz * (R)m.uz[i],
                       Ix * (R)m.vx[i] + Iy * (R)m.vy[i] + Iz * (R)m.vz[i]);
    if (g_mut == MUT_TPH_SWAPPED)
        q.Tph = std::atan2(Ix * (R)m.vx[i] + Iy * (R)m.vy[i] + Iz * (R)m.vz[i],
                           Ix * (R)m.ux[i] + Iy * (R)m.uy[i] + Iz * (R)m.uz[i]);
    R expo = 0, tot = 0;
    for (int64_t k = 0; k < np; ++k) {
        R d = prismIntersectionDistance<R>((R)m.x[i], (R)m.y[i], (R)m.z[i], Ix,
                                           Iy, Iz, pr[k].xmin, pr[k].xmax,
                                           pr[k].ymin, pr[k].ymax, pr[k].zmin,
                                           pr[k].zmax, q.Ttd);
        tot += d;
        expo += (R)pr[k].alpha * d;
    }
    expo += p.alpha2 * (q.Ttd - tot);
    R atten = std::exp((R)-4 * expo);
    q.EI = atten;
    q.EO = atten;
    /* compRefrEnergyIn (geometry_kernels.cu:520-554) */
    if (p.specular) {
        R dth = f.Rth - q.Tth, dph = f.Iph - q.Tph;
        q.EI *= facetReradiation<R>(q.Ttd, dth, dph, p.lam, p.fs);
    }
    if (g_mut != MUT_NO_TRANSMISSION) q.EI = q.EI * f.fRfrC;
    R ct = p.ks * std::cos(q.Tth);
    R rough = std::exp((R)-4 * (ct * ct));
    q.EI = q.EI * rough;
    /* compRefrEnergyOut (geometry_kernels.cu:557-584) */
    if (p.specular) {
        R RrTh = snellsLaw<R>(PI_F<R>() - q.Tth, p.eps_2, p.eps_1);
        q.EO *= facetReradiation<R>(f.Itd, RrTh, q.Tph - f.Iph, p.lam, p.fs);
    }
    if (g_mut != MUT_NO_TRANSMISSION) q.EO = q.EO * f.fRfrC;
    q.EO = q.EO * rough;
    return q;
}

/* Host FP32 target aperture test (sim.cu:808-826); FP64: include iff
 * !(c < cos_ap), the same comparison on the double dot product. */
template <class R>
inline bool target_in_aperture(const Par<R>& p, const erp_shot& s, float tx,
                               float ty, float tz) {
    if constexpr (is_f32<R>) {
        float vx = tx - (float)s.sx, vy = ty - (float)s.sy, vz = tz - (float)s.sz;
        float mag = sqrtf(vx * vx + vy * vy + vz * vz);
        vx /= mag; vy /= mag; vz /= mag;
        float dp = (-1 * s.snx) * vx + (-1 * s.sny) * vy + (-1 * s.snz) * vz;
        float th = acosf(dp);
        return !(th > p.aperture * (PI_F<float>() / 180.0f));
    } else {
        double vx = (double)tx - s.sx, vy = (double)ty - s.sy,
               vz = (double)tz - s.sz;
        double mag = std::sqrt(vx * vx + vy * vy + vz * vz);
        vx /= mag; vy /= mag; vz /= mag;
        double dp = (-(double)s.snx) * vx + (-(double)s.sny) * vy +
                    (-(double)s.snz) * vz;
        return !(dp < p.cos_ap);
    }
}

inline int check_modes(const erp_params& p, const uint8_t* ttype, int64_t T) {
    const bool conv = p.mode & ERP_MODE_CONVOLUTION;
    const bool lin = p.mode & ERP_MODE_CONV_LINEAR;
    const bool spec = p.mode & ERP_MODE_SPECULAR;
    if (p.nr <= 0 || p.nr > 32767) return ERP_E_INVALID;
    if (!spec && !(conv && lin)) return ERP_E_UNSUPPORTED;
    bool any_typed = false;
    for (int64_t t = 0; t < T; ++t)
        if (ttype && ttype[t] != 0) any_typed = true;
    if (any_typed && !(conv && lin)) return ERP_E_UNSUPPORTED;
    /* the second medium is defined for the plain method's typed targets */
    if (p.eps_3 > 0) {
        if (spec) return ERP_E_UNSUPPORTED;
        for (int64_t t = 0; t < T; ++t)
            if (ttype && ttype[t] == 0) return ERP_E_UNSUPPORTED;
    }
    return ERP_OK;
}

/* One in-aperture target of one shot (the body of the target loop,
 * sim.cu:801-986); the refracted contribution is added to refr. */
template <class R>
void target_one(const Par<R>& p, const MeshView& m, const std::vector<FacetRec<R>>& fr,
                const float* tgt, const uint8_t* ttype, int64_t T, int64_t it,
                const erp_shot& s, const erp_prism* pr, int64_t np, R n, R res_t, int64_t V,
                Work<R>& w, std::vector<PairRec<R>>& q, std::vector<cx<R>>& tmp,
                std::vector<cx<R>>& PT, std::vector<cx<R>>& Ptg, std::vector<cx<R>>& Ps,
                std::vector<cx<R>>& refr) {
    const int nr = p.nr;
    const R rst = (R)s.rst, dr = p.dr;
    const R Grefr = (R)s.grefr_lin;

    const float tx = tgt[it], ty = tgt[T + it], tz = tgt[2 * T + it];
    const float tnx = tgt[3 * T + it], tny = tgt[4 * T + it],
                tnz = tgt[5 * T + it];
    const uint8_t ty8 = ttype ? ttype[it] : 0;
    for (int64_t k = 0; k < V; ++k)
        q[k] = pair_record<R>(p, m, fr[k], (R)tx, (R)ty, (R)tz, (R)tnx,
                              (R)tny, (R)tnz, pr, np);
    if (ty8 != 0) {
        /* accumulateTarget (radar_kernels.cu:767-830) */
        std::fill(PT.begin(), PT.end(), cx<R>(0, 0));
        std::fill(Ps.begin(), Ps.end(), cx<R>(0, 0));
        for (int64_t k = 0; k < V; ++k) {
            const FacetRec<R>& f = fr[k];
            R Gd = facet_G<R>(q[k].TargetTh, q[k].Tph, p.lam, p.fs);
            R G1 = facet_G<R>(f.Ith, f.Iph, p.lam, p.fs);
            R G2 = facet_G<R>(q[k].Tth, q[k].Tph, p.lam, p.fs);
            R tr;
            R rngt = target_range<R>(p, f.Itd, q[k].Ttd,
                                     g_mut == MUT_REFR_INDEX_DROPPED ? (R)1 : n,
                                     (R)m.z[f.idx], (R)tz, &tr);
            int b;
            R wgt;
            if (!bin_of<R>(rngt, rst, dr, nr, &b, &wgt)) continue;
            R Pin = friis<R>(p.P, Grefr, G1, p.lam, f.Itd);
            Pin = friis<R>(Pin, G2, Gd, p.lam, q[k].Ttd);
            /* radiateTarget (radar_kernels.cu:834-892): P = 1 W */
            R Pout = friis<R>((R)1, Grefr, G1, p.lam, f.Itd);
            Pout = friis<R>(Pout, G2, Gd, p.lam, q[k].Ttd);
            if (!p.lossless) {
                Pin = Pin * q[k].EI * tr;
                Pout = Pout * q[k].EO * tr;
            }
            split_add<R>(PT.data(), nr, b, wgt,
                         phasor_neg<R>(rngt, p.lam) * std::sqrt(Pin * n));
            split_add<R>(Ps.data(), nr, b, wgt,
                         phasor_neg<R>(rngt - (p.lam / (R)2.0f), p.lam) *
                             std::sqrt(Pout * n));
        }
        conv_linear<R>(w, PT.data(), dr, res_t, Ptg.data());
        /* convolveComplex(Psour, Ptarg) (sim.cu:912) */
        conv_complex<R>(w, Ps.data(), Ptg.data(), p.smpl, p.B, tmp.data());
        for (int j = 0; j < nr; ++j) refr[j] += tmp[j];
    } else if (!p.conv) {
        /* refrRadarSignal (radar_kernels.cu:201-299), rerad on Tth */
        for (int ir = 0; ir < nr; ++ir) {
            cx<R> acc(0, 0);
            R r = rst + (R)ir * dr;
            for (int64_t k = 0; k < V; ++k) {
                R sl = fr[k].Itd + q[k].Ttd;
                R rc = rerad_funct<R>(p.rerad, q[k].Tth) * q[k].EI * q[k].EO;
                rc = rc * radarEq<R>(p.P, Grefr, p.fs, p.lam, sl, V);
                R rngt = (sl - q[k].Ttd) + q[k].Ttd * (p.c / p.c_2);
                R dl = (r - rngt) / p.rng_res;
                acc += phasor_pos4<R>(rngt, p.lam) * (sinc_ref<R>(dl) * rc);
            }
            refr[ir] += acc;
        }
    } else {
        /* genRefrPhasor + genPhasorTrace (radar_kernels.cu:351-398) */
        std::fill(PT.begin(), PT.end(), cx<R>(0, 0));
        for (int64_t k = 0; k < V; ++k) {
            R sl = fr[k].Itd + q[k].Ttd;
            R rc = rerad_funct<R>(p.rerad, q[k].TargetTh) * q[k].EI * q[k].EO;
            rc = rc * radarEq<R>(p.P, Grefr, p.fs, p.lam, sl, V);
            R rngt = (sl - q[k].Ttd) + q[k].Ttd * (p.c_1 / p.c_2);
            cx<R> ph = phasor_pos4<R>(rngt, p.lam) * rc;
            int b;
            R wgt;
            if (!bin_of<R>(rngt, rst, dr, nr, &b, &wgt)) {
                b = 0;
                ph = cx<R>(0, 0);
            }
            PT[b] += ph;
        }
        if (!p.linear)
            conv_circ<R>(w, PT.data(), rst, dr, p.rng_res, tmp.data());
        else
            conv_linear<R>(w, PT.data(), dr, res_t, tmp.data());
        for (int j = 0; j < nr; ++j) refr[j] += tmp[j];
    }
}

/* One shot of sim.cu:615-1008. */
template <class R>
int run_shot(const Par<R>& p, const MeshView& m, const float* tgt,
             const uint8_t* ttype, int64_t T, const erp_shot& s, int64_t is,
             const erp_prism* pr, int64_t np, Work<R>& w, double* out,
             double* parts, int32_t* nvalid, int inner_threads) {
    const int nr = p.nr;
    w.ensure(nr);
    std::vector<cx<R>> psurf(nr, cx<R>(0, 0)), refl(nr, cx<R>(0, 0)),
        refr(nr, cx<R>(0, 0));
    const R rst = (R)s.rst, dr = p.dr;
    const R Grefr = (R)s.grefr_lin;

    /* compSourceInclination + cropByAperture (sim.cu:667-689) */
    std::vector<int64_t> ill;
    ill.reserve(1024);
    for (int64_t i = 0; i < m.F; ++i)
        if (illuminated<R>(p, m, i, s)) ill.push_back(i);
    const int64_t V = (int64_t)ill.size();
    if (nvalid) *nvalid = (int32_t)V;
    auto emit = [&]() {
        for (int j = 0; j < nr; ++j) {
            cx<R> v = refl[j] + refr[j];
            out[2 * j] = (double)v.real();
            out[2 * j + 1] = (double)v.imag();
        }
        if (parts) {
            for (int j = 0; j < nr; ++j) {
                parts[2 * j] = (double)psurf[j].real();
                parts[2 * j + 1] = (double)psurf[j].imag();
                parts[2 * nr + 2 * j] = (double)refl[j].real();
                parts[2 * nr + 2 * j + 1] = (double)refl[j].imag();
                parts[4 * nr + 2 * j] = (double)refr[j].real();
                parts[4 * nr + 2 * j + 1] = (double)refr[j].imag();
            }
        }
    };
    if (V == 0) { /* Q9 */
        emit();
        return ERP_OK;
    }
    std::vector<FacetRec<R>> fr(V);
    for (int64_t k = 0; k < V; ++k) fr[k] = facet_record<R>(p, m, ill[k], s, V);

    /* ---- reflected (surface) signal, sim.cu:733-799 ---- */
    if (!p.specular) {
        /* surfacePT (radar_kernels.cu:714-764) */
        for (int64_t k = 0; k < V; ++k) {
            const FacetRec<R>& f = fr[k];
            R G = facet_G<R>(f.Ith, f.Iph, p.lam, p.fs);
            R Pf = friis<R>(p.P, Grefr, G, p.lam, f.Itd);
            R Psrc = friis<R>(Pf, G, Grefr, p.lam, f.Itd);
            if (!p.lossless) Psrc = Psrc * f.fRe;
            int b;
            R wgt;
            if (!bin_of<R>(f.Itd, rst, dr, nr, &b, &wgt)) continue;
            cx<R> ph;
            if (p.incoherent) {
                double u = incoherent_u(p.seed, is, f.idx);
                if (g_mut == MUT_INCOHERENT_HALF) u *= 0.5;
                R a = (R)(2.0 * M_PI * u);
                ph = cx<R>(std::cos(a), std::sin(a));
            } else {
                ph = phasor_neg<R>(2 * f.Itd, p.lam);
            }
            cx<R> c = ph * std::sqrt(Psrc);
            split_add<R>(psurf.data(), nr, b, wgt, c);
        }
        if (p.disable_surface) std::fill(psurf.begin(), psurf.end(), cx<R>(0, 0));
        conv_linear<R>(w, psurf.data(), dr, p.rng_res, refl.data());
    } else if (!p.conv) {
        /* reflRadarSignal (radar_kernels.cu:58-159): direct sinc sum */
        for (int ir = 0; ir < nr; ++ir) {
            cx<R> acc(0, 0);
            R r = rst + (R)ir * dr;
            for (int64_t k = 0; k < V; ++k) {
                R dl = (r - fr[k].Itd) / p.rng_res;
                cx<R> pe = phasor_pos4<R>(fr[k].Itd, p.lam);
                acc += pe * (sinc_ref<R>(dl) * fr[k].fRe);
            }
            refl[ir] = acc;
        }
    } else {
        /* genReflPhasor + genPhasorTrace (radar_kernels.cu:314-348, 400-448) */
        for (int64_t k = 0; k < V; ++k) {
            int b;
            R wgt;
            cx<R> ph = phasor_pos4<R>(fr[k].Itd, p.lam) * fr[k].fRe;
            if (!bin_of<R>(fr[k].Itd, rst, dr, nr, &b, &wgt)) {
                b = 0;
                ph = cx<R>(0, 0);
            }
            psurf[b] += ph;
        }
        if (p.linear) conv_linear<R>(w, psurf.data(), dr, p.rng_res, refl.data());
        /* circular: the reflected path is never produced (Q13) */
    }

    /* ---- refracted (target) signal, sim.cu:801-986 ---- */
    const R n = std::sqrt(p.eps_2);
    const R res_t = p.specular ? p.rng_res : (R)2 * p.rng_res; /* sim.cu:853-858 */
    std::vector<int64_t> tin;
    for (int64_t it = 0; it < T; ++it)
        if (target_in_aperture<R>(p, s, tgt[it], tgt[T + it], tgt[2 * T + it])) tin.push_back(it);
    const int64_t NTi = (int64_t)tin.size();
    int nth = 1;
#ifdef _OPENMP
    nth = (inner_threads > 1 && NTi > 1) ? inner_threads : 1;
#endif
    std::vector<std::vector<cx<R>>> acc(nth, std::vector<cx<R>>(nr, cx<R>(0, 0)));
#ifdef _OPENMP
#pragma omp parallel num_threads(nth) if (nth > 1)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        Work<R> wl;
        Work<R>& ww = (nth > 1) ? wl : w;
        ww.ensure(nr);
        std::vector<cx<R>> tmp(nr), PT(nr), Ptg(nr), Ps(nr);
        std::vector<PairRec<R>> q(V);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t ii = 0; ii < NTi; ++ii) {
            target_one<R>(p, m, fr, tgt, ttype, T, tin[ii], s, pr, np, n, res_t, V, ww,
                          q, tmp, PT, Ptg, Ps, acc[tid]);
        }
    }
    for (int t = 0; t < nth; ++t)
        for (int j = 0; j < nr; ++j) refr[j] += acc[t][j];
    emit();
    return ERP_OK;
}

template <class R>
int run_all(const erp_params* par, const float* soa, int64_t F, const float* tgt,
            const uint8_t* ttype, int64_t T, const erp_shot* shots, int64_t ns,
            const erp_prism* pr, int64_t np, int nthreads, double* out,
            double* parts, int32_t* nvalid) {
    int rc = check_modes(*par, ttype, T);
    if (rc) return rc;
    Par<R> p(*par);
    MeshView m(soa, F);
    const int nr = p.nr;
    int err = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    if (ns >= nthreads || T == 0) {
        /* shots are independent: one shot per thread */
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
        {
            Work<R> w;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
            for (int64_t is = 0; is < ns; ++is) {
                int r = run_shot<R>(p, m, tgt, ttype, T, shots[is], is, pr, np, w,
                                    out + (size_t)is * 2 * nr,
                                    parts ? parts + (size_t)is * 6 * nr : nullptr,
                                    nvalid ? nvalid + is : nullptr, 1);
                if (r) err = r;
            }
        }
    } else {
        /* few shots: parallel over each shot's targets */
        Work<R> w;
        for (int64_t is = 0; is < ns; ++is) {
            int r = run_shot<R>(p, m, tgt, ttype, T, shots[is], is, pr, np, w,
                                out + (size_t)is * 2 * nr,
                                parts ? parts + (size_t)is * 6 * nr : nullptr,
                                nvalid ? nvalid + is : nullptr, nthreads);
            if (r) err = r;
        }
    }
    return err;
}

} // namespace

extern "C" int erp_oracle_run(const erp_params* par, const float* soa, int64_t F,
                   const float* tgt, const uint8_t* ttype, int64_t T,
                   const erp_shot* shots, int64_t ns, const erp_prism* prisms,
                   int64_t nprisms, int precision, int nthreads, double* out,
                   double* parts, int32_t* nvalid) {
    if (!par || !soa || F < 0 || !shots || ns < 0 || !out) return ERP_E_INVALID;
    if (T > 0 && !tgt) return ERP_E_INVALID;
    if (precision == 32)
        return run_all<float>(par, soa, F, tgt, ttype, T, shots, ns, prisms,
                              nprisms, nthreads, out, parts, nvalid);
    return run_all<double>(par, soa, F, tgt, ttype, T, shots, ns, prisms,
                           nprisms, nthreads, out, parts, nvalid);
}

template <class R>
static int facets_impl(const erp_params* par, const float* soa, int64_t F,
                       const erp_shot* s, double* out) {
    Par<R> p(*par);
    MeshView m(soa, F);
    int64_t V = 0;
    std::vector<char> ill(F);
    for (int64_t i = 0; i < F; ++i) {
        ill[i] = illuminated<R>(p, m, i, *s);
        V += ill[i];
    }
    for (int64_t i = 0; i < F; ++i) {
        double* o = out + (size_t)i * 12;
        std::fill(o, o + 12, 0.0);
        o[0] = ill[i];
        FacetRec<R> f = facet_record<R>(p, m, i, *s, V);
        R G = facet_G<R>(f.Ith, f.Iph, p.lam, p.fs);
        R Pf = friis<R>(p.P, (R)s->grefr_lin, G, p.lam, f.Itd);
        R Ps = friis<R>(Pf, G, (R)s->grefr_lin, p.lam, f.Itd);
        if (!p.lossless) Ps = Ps * f.fRe;
        int b = -1;
        R w = 0;
        if (!bin_of<R>(f.Itd, (R)s->rst, p.dr, p.nr, &b, &w)) b = -1;
        cx<R> ph = phasor_neg<R>(2 * f.Itd, p.lam);
        o[1] = f.Itd; o[2] = f.Ith; o[3] = f.Iph; o[4] = f.Rth; o[5] = f.fRe;
        o[6] = f.fRfrC; o[7] = G; o[8] = Ps; o[9] = b; o[10] = w;
        o[11] = std::arg(ph);
    }
    return ERP_OK;
}

extern "C" int erp_oracle_facets(const erp_params* par, const float* soa, int64_t F,
                      const erp_shot* shot, int precision, double* out) {
    if (!par || !soa || !shot || !out) return ERP_E_INVALID;
    if (precision == 32) return facets_impl<float>(par, soa, F, shot, out);
    return facets_impl<double>(par, soa, F, shot, out);
}

template <class R>
static int pairs_impl(const erp_params* par, const float* soa, int64_t F,
                      const float* tgt, const uint8_t* ttype, int64_t T,
                      int64_t it, const erp_shot* s, const erp_prism* pr,
                      int64_t np, double* out) {
    (void)ttype;
    Par<R> p(*par);
    MeshView m(soa, F);
    int64_t V = 0;
    std::vector<char> ill(F);
    for (int64_t i = 0; i < F; ++i) {
        ill[i] = illuminated<R>(p, m, i, *s);
        V += ill[i];
    }
    const R n = std::sqrt(p.eps_2);
    for (int64_t i = 0; i < F; ++i) {
        double* o = out + (size_t)i * 10;
        FacetRec<R> f = facet_record<R>(p, m, i, *s, V);
        PairRec<R> q = pair_record<R>(p, m, f, (R)tgt[it], (R)tgt[T + it],
                                      (R)tgt[2 * T + it], (R)tgt[3 * T + it],
                                      (R)tgt[4 * T + it], (R)tgt[5 * T + it], pr, np);
        R Gd = facet_G<R>(q.TargetTh, q.Tph, p.lam, p.fs);
        R G1 = facet_G<R>(f.Ith, f.Iph, p.lam, p.fs);
        R G2 = facet_G<R>(q.Tth, q.Tph, p.lam, p.fs);
        R Pin = friis<R>(p.P, (R)s->grefr_lin, G1, p.lam, f.Itd);
        Pin = friis<R>(Pin, G2, Gd, p.lam, q.Ttd);
        R Pout = friis<R>((R)1, (R)s->grefr_lin, G1, p.lam, f.Itd);
        Pout = friis<R>(Pout, G2, Gd, p.lam, q.Ttd);
        if (!p.lossless) {
            Pin = Pin * q.EI;
            Pout = Pout * q.EO;
        }
        o[0] = ill[i]; o[1] = q.Ttd; o[2] = q.Tth; o[3] = q.TargetTh;
        R tr;
        o[4] = q.Tph; o[5] = q.EI; o[6] = q.EO;
        o[7] = target_range<R>(p, f.Itd, q.Ttd, n, (R)m.z[i], (R)tgt[2 * T + it], &tr);
        o[8] = Pin; o[9] = Pout;
    }
    return ERP_OK;
}

extern "C" int erp_oracle_target_pairs(const erp_params* par, const float* soa, int64_t F,
                            const float* tgt, const uint8_t* ttype, int64_t T,
                            int64_t it, const erp_shot* shot,
                            const erp_prism* prisms, int64_t nprisms,
                            int precision, double* out) {
    if (!par || !soa || !tgt || !shot || !out || it < 0 || it >= T)
        return ERP_E_INVALID;
    if (precision == 32)
        return pairs_impl<float>(par, soa, F, tgt, ttype, T, it, shot, prisms,
                                 nprisms, out);
    return pairs_impl<double>(par, soa, F, tgt, ttype, T, it, shot, prisms,
                              nprisms, out);
}

extern "C" int erp_oracle_set_mutation(int m) {
    const int prev = g_mut;
    g_mut = m;
    return prev;
}

extern "C" int erp_oracle_dft(const double* in, int64_t n, int inverse, double* out) {
    if (!in || !out || n <= 0) return ERP_E_INVALID;
    DFT<double> d;
    d.init(n);
    std::vector<cx<double>> a(n);
    for (int64_t i = 0; i < n; ++i) a[i] = cx<double>(in[2 * i], in[2 * i + 1]);
    d.run(a.data(), inverse != 0);
    for (int64_t i = 0; i < n; ++i) {
        out[2 * i] = a[i].real();
        out[2 * i + 1] = a[i].imag();
    }
    return ERP_OK;
}


/* Convolution kernels of the reference, exposed for pinning against numpy:
 * kind 0: convolvePhasorChirpLinear with the padded chirp of resolution res
 * kind 1: convolvePhasorChirp (circular, genChirp table with rst)
 * kind 2: convolveComplex(sig=x, ker=y) with crop at B for sample rate smpl
 * x, y, out: nr interleaved complex (double).  precision 64 or 32. */
template <class R>
static int conv_impl(int kind, const double* x, const double* y, int nr, double dr,
                     double res, double rst, double smpl, double Bw, double* out) {
    Work<R> w;
    w.ensure(nr);
    std::vector<cx<R>> a(nr), b(nr), o(nr);
    for (int i = 0; i < nr; ++i) {
        a[i] = cx<R>((R)x[2 * i], (R)x[2 * i + 1]);
        if (y) b[i] = cx<R>((R)y[2 * i], (R)y[2 * i + 1]);
    }
    if (kind == 0) conv_linear<R>(w, a.data(), (R)dr, (R)res, o.data());
    else if (kind == 1) conv_circ<R>(w, a.data(), (R)rst, (R)dr, (R)res, o.data());
    else if (kind == 2) conv_complex<R>(w, a.data(), b.data(), (R)smpl, (R)Bw, o.data());
    else return ERP_E_INVALID;
    for (int i = 0; i < nr; ++i) {
        out[2 * i] = (double)o[i].real();
        out[2 * i + 1] = (double)o[i].imag();
    }
    return ERP_OK;
}

extern "C" int erp_oracle_conv(int kind, const double* x, const double* y, int nr, double dr,
                               double res, double rst, double smpl, double Bw, int precision,
                               double* out) {
    if (!x || !out || nr <= 0 || (kind == 2 && !y)) return ERP_E_INVALID;
    if (precision == 32) return conv_impl<float>(kind, x, y, nr, dr, res, rst, smpl, Bw, out);
    return conv_impl<double>(kind, x, y, nr, dr, res, rst, smpl, Bw, out);
}

extern "C" const char* erp_oracle_version(void) { return "erp_oracle 1 (fp64 + fp32 mirror)"; }

