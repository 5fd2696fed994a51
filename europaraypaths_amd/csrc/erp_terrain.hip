// Synthetic terrains on the device: the height grid of a Gaussian random
// field or a fractional Brownian surface by spectral synthesis, then its
// facet mesh (positions, central-difference normals, tangents), without a
// round trip through host memory.  The fields follow the seeded host
// generators of the benchmark meshes (europaraypaths_amd/surfaces.py
// gaussian_heights / fbm_heights): the same spectra on the same fftfreq
// grids, mean removed and rescaled to the requested RMS.  The reference
// builds its terrains on the host (src/PYTHON/simple_surfaces.py:3-39 for
// the normals; examples/CHAOS/code/setup.py draws its fields with gstools).
//
// Transforms of any grid size: a Bluestein DFT per dimension (chirp
// multiply, power-of-two FFT convolution, chirp multiply) over rows, with a
// tiled transpose between the two dimensions.  Arithmetic in FP32 except
// the chirp angles (FP64, exact integer reduction of k^2 mod 2L) and the
// mean and RMS sums (FP64).
#include "erp_internal.h"

#include <cmath>

namespace {

constexpr double kPi = 3.14159265358979323846;

// w[k] = exp(-i pi k^2 / L) for k < L; b = conj(w) placed circularly on M
// points (b[j] = b[M - j] = conj(w[j])), to be transformed and scaled by 1/M
__global__ void k_bdft_tables(float2* __restrict__ w, float2* __restrict__ b, int L, int M) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    auto chirp = [&](int k) {
        const long long m = ((long long)k * k) % (2LL * L);
        double s, c;
        sincos(kPi * (double)m / (double)L, &s, &c);
        return make_float2((float)c, (float)-s);
    };
    if (j < L) w[j] = chirp(j);
    float2 v = make_float2(0.f, 0.f);
    if (j < L) {
        const float2 c = chirp(j);
        v = make_float2(c.x, -c.y);
    } else if (M - j < L) {
        const float2 c = chirp(M - j);
        v = make_float2(c.x, -c.y);
    }
    b[j] = v;
}

__global__ void k_scale_c(float2* __restrict__ a, int n, float s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = make_float2(a[i].x * s, a[i].y * s);
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// a[r][j] = x[r][j] w[j] (x conjugated when conj_in) for j < L, else 0
__global__ void k_bdft_in(float2* __restrict__ a, int M, const float2* __restrict__ x, int L,
                          const float2* __restrict__ w, int rows, int conj_in) {
    const size_t n = (size_t)M * rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / M), j = (int)(i - (size_t)r * M);
        float2 v = make_float2(0.f, 0.f);
        if (j < L) {
            float2 xv = x[(size_t)r * L + j];
            if (conj_in) xv.y = -xv.y;
            v = cmulf(xv, w[j]);
        }
        a[i] = v;
    }
}

// a *= B (both in the order fft_forward produces)
__global__ void k_bdft_mul(float2* __restrict__ a, const float2* __restrict__ B, int M,
                           int rows) {
    const size_t n = (size_t)M * rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        a[i] = cmulf(a[i], B[i & (size_t)(M - 1)]);
}

// X[r][k] = w[k] a[r][k] for k < L (conjugated when conj_out)
__global__ void k_bdft_out(float2* __restrict__ X, int L, const float2* __restrict__ a, int M,
                           const float2* __restrict__ w, int rows, int conj_out) {
    const size_t n = (size_t)L * rows;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / L), k = (int)(i - (size_t)r * L);
        float2 v = cmulf(w[k], a[(size_t)r * M + k]);
        if (conj_out) v.y = -v.y;
        X[i] = v;
    }
}

// dst (cols x rows) = transpose of src (rows x cols), 32 x 32 tiles in LDS
__global__ void k_transpose(float2* __restrict__ dst, const float2* __restrict__ src,
                            int rows, int cols) {
    __shared__ float2 tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int j = threadIdx.y; j < 32; j += blockDim.y) {
        const int r = by + j, c = bx + threadIdx.x;
        if (r < rows && c < cols) tile[j][threadIdx.x] = src[(size_t)r * cols + c];
    }
    __syncthreads();
    for (int j = threadIdx.y; j < 32; j += blockDim.y) {
        const int c = bx + j, r = by + threadIdx.x;
        if (r < rows && c < cols) dst[(size_t)c * rows + r] = tile[threadIdx.x][j];
    }
}

// counter-based uniform in (0, 1]: SplitMix64 of (seed, stream, index)
__device__ __forceinline__ double unif(uint64_t seed, uint64_t stream, uint64_t idx) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (idx * 4 + stream + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((double)(z >> 11) + 1.0) * 0x1.0p-53;
}
__device__ __forceinline__ double normal_at(uint64_t seed, uint64_t idx) {
    // Box-Muller on two streams of the counter generator
    const double u1 = unif(seed, 1, idx), u2 = unif(seed, 2, idx);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
}

// fftfreq(n, d)[i] = (i < (n + 1) / 2 ? i : i - n) / (n d)
__device__ __forceinline__ double fftfreq(int64_t i, int64_t n, double d) {
    const int64_t k = i < (n + 1) / 2 ? i : i - n;
    return (double)k / ((double)n * d);
}

// the noise of the field, as the host generators draw it (noise given) or
// from the counter generator: Gaussian, ny x nx real white noise; fBm,
// amplitude K^-(H+1) times a normal times exp(i * uniform phase)
__global__ void k_terrain_spec(float2* __restrict__ spec, int64_t nx, int64_t ny, double fs,
                               int kind, double hurst, uint64_t seed,
                               const double* __restrict__ noise) {
    const int64_t F = nx * ny;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < F;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (kind == ERP_TERRAIN_GAUSSIAN) {
            const double w = noise ? noise[i] : normal_at(seed, (uint64_t)i);
            spec[i] = make_float2((float)w, 0.f);
            continue;
        }
        const int64_t iy = i / nx, ix = i - iy * nx;
        const double kx = fftfreq(ix, nx, fs), ky = fftfreq(iy, ny, fs);
        const double K = sqrt(kx * kx + ky * ky);
        const double amp = (ix == 0 && iy == 0) ? 0.0 : pow(K, -(hurst + 1.0));
        const double phase = noise ? noise[i] : 2.0 * kPi * (1.0 - unif(seed, 0, (uint64_t)i));
        const double nrm = noise ? noise[F + i] : normal_at(seed, (uint64_t)i);
        double s, c;
        sincos(phase, &s, &c);
        spec[i] = make_float2((float)(amp * nrm * c), (float)(amp * nrm * s));
    }
}

// Gaussian filter exp(-K^2 l^2 / 8), K the angular wavenumber on the fftfreq grid
__global__ void k_terrain_gauss_filter(float2* __restrict__ spec, int64_t nx, int64_t ny,
                                       double fs, double corr_len) {
    const int64_t F = nx * ny;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < F;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t iy = i / nx, ix = i - iy * nx;
        const double kx = 2.0 * kPi * fftfreq(ix, nx, fs), ky = 2.0 * kPi * fftfreq(iy, ny, fs);
        const float f = (float)exp(-(kx * kx + ky * ky) * corr_len * corr_len / 8.0);
        spec[i] = make_float2(spec[i].x * f, spec[i].y * f);
    }
}

// per-block FP64 partial sums of re(z) and re(z)^2 (fixed tree order)
__global__ void __launch_bounds__(256) k_terrain_moments(const float2* __restrict__ z, int64_t F,
                                                         double* __restrict__ part) {
    __shared__ double s1[256], s2[256];
    double a = 0.0, b = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < F;
         i += (int64_t)gridDim.x * 256) {
        const double v = z[i].x;
        a += v;
        b += v * v;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s1[threadIdx.x] += s1[threadIdx.x + o];
            s2[threadIdx.x] += s2[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1[0];
        part[2 * blockIdx.x + 1] = s2[0];
    }
}

// heights: (re(z) - mean) * rms / std, std the population deviation (FP64)
__global__ void k_terrain_heights(double* __restrict__ h, const float2* __restrict__ z, int64_t F,
                                  const double* __restrict__ part, int nblk, double rms) {
    __shared__ double mean_s, scale_s;
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < nblk; ++k) {
            a += part[2 * k];
            b += part[2 * k + 1];
        }
        const double mean = a / (double)F;
        const double var = fmax(b / (double)F - mean * mean, 0.0);
        mean_s = mean;
        scale_s = var > 0.0 ? rms / sqrt(var) : 1.0;
    }
    __syncthreads();
    const double mean = mean_s, sc = scale_s;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < F;
         i += (int64_t)gridDim.x * blockDim.x)
        h[i] = ((double)z[i].x - mean) * sc;
}

inline int grid_of(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 1 << 16); }

// DFT of `rows` rows of length L in place (inverse: unnormalised, the sign
// flipped by conjugating input and output), through a workspace of rows x M
struct Bdft {
    int L = 0, M = 0;
    FftPlan plan;
    DevBuf w, B;
    void init(erp_ctx* ctx, int L_, hipStream_t st) {
        L = L_;
        M = std::max(2, next_pow2(2 * (int64_t)L - 1));
        plan = fft_plan(M);
        w.ensure(sizeof(float2) * (size_t)L);
        B.ensure(sizeof(float2) * (size_t)M);
        hipLaunchKernelGGL(k_bdft_tables, dim3((M + 255) / 256), dim3(256), 0, st,
                           w.as<float2>(), B.as<float2>(), L, M);
        ERP_CHECK_LAUNCH("k_bdft_tables");
        fft_forward(ctx, plan, B.as<float2>(), 1, st);
        hipLaunchKernelGGL(k_scale_c, dim3((M + 255) / 256), dim3(256), 0, st, B.as<float2>(), M,
                           1.0f / (float)M);
        ERP_CHECK_LAUNCH("k_scale_c");
    }
    void run(erp_ctx* ctx, float2* x, int rows, float2* work, bool inverse, hipStream_t st) {
        const size_t nM = (size_t)M * rows, nL = (size_t)L * rows;
        hipLaunchKernelGGL(k_bdft_in, dim3(grid_of(nM)), dim3(256), 0, st, work, M, x, L,
                           w.as<float2>(), rows, inverse ? 1 : 0);
        ERP_CHECK_LAUNCH("k_bdft_in");
        fft_forward(ctx, plan, work, rows, st);
        hipLaunchKernelGGL(k_bdft_mul, dim3(grid_of(nM)), dim3(256), 0, st, work,
                           B.as<float2>(), M, rows);
        ERP_CHECK_LAUNCH("k_bdft_mul");
        fft_inverse(ctx, plan, work, rows, st);
        hipLaunchKernelGGL(k_bdft_out, dim3(grid_of(nL)), dim3(256), 0, st, x, L, work, M,
                           w.as<float2>(), rows, inverse ? 1 : 0);
        ERP_CHECK_LAUNCH("k_bdft_out");
    }
    void release() {
        w.release();
        B.release();
    }
};

void transpose(float2* dst, const float2* src, int rows, int cols, hipStream_t st) {
    hipLaunchKernelGGL(k_transpose, dim3((cols + 31) / 32, (rows + 31) / 32), dim3(32, 8), 0, st,
                       dst, src, rows, cols);
    ERP_CHECK_LAUNCH("k_transpose");
}

}  // namespace

// 2-D DFT of the ny x nx grid a (row-major, x fastest) in place; t and work
// are scratch of nx * ny and max(nx Mx, ny My) complex values
static void dft2(erp_ctx* ctx, Bdft& bx, Bdft& by, float2* a, float2* t, float2* work, int nx,
                 int ny, bool inverse, hipStream_t st) {
    bx.run(ctx, a, ny, work, inverse, st);  // along x: ny rows of nx
    transpose(t, a, ny, nx, st);            // t: nx rows of ny
    by.run(ctx, t, nx, work, inverse, st);  // along y
    transpose(a, t, nx, ny, st);
}

void terrain_heights(erp_ctx* ctx, const erp_terrain& tr, double* h, hipStream_t st) {
    const int nx = (int)tr.nx, ny = (int)tr.ny;
    const int64_t F = (int64_t)nx * ny;
    Bdft bx, by;
    DevBuf a, t, work, part, noise;
    try {
        bx.init(ctx, nx, st);
        by.init(ctx, ny, st);
        a.ensure(sizeof(float2) * (size_t)F);
        t.ensure(sizeof(float2) * (size_t)F);
        work.ensure(sizeof(float2) * std::max((size_t)bx.M * ny, (size_t)by.M * nx));
        const double* dnoise = nullptr;
        if (tr.noise) {
            const size_t n = (size_t)F * (tr.kind == ERP_TERRAIN_FBM ? 2 : 1);
            noise.ensure(sizeof(double) * n);
            ERP_HIP_TRY(hipMemcpyAsync(noise.p, tr.noise, sizeof(double) * n,
                                       hipMemcpyHostToDevice, st));
            dnoise = noise.as<double>();
        }
        float2* A = a.as<float2>();
        hipLaunchKernelGGL(k_terrain_spec, dim3(grid_of(F)), dim3(256), 0, st, A, (int64_t)nx,
                           (int64_t)ny, tr.fs, (int)tr.kind, tr.hurst, tr.seed, dnoise);
        ERP_CHECK_LAUNCH("k_terrain_spec");
        if (tr.kind == ERP_TERRAIN_GAUSSIAN) {
            // z = ifft2(fft2(w) * filt)
            dft2(ctx, bx, by, A, t.as<float2>(), work.as<float2>(), nx, ny, false, st);
            hipLaunchKernelGGL(k_terrain_gauss_filter, dim3(grid_of(F)), dim3(256), 0, st, A,
                               (int64_t)nx, (int64_t)ny, tr.fs, tr.corr_len);
            ERP_CHECK_LAUNCH("k_terrain_gauss_filter");
        }
        // z = ifft2(spec) (unnormalised: the RMS rescale takes the scale out)
        dft2(ctx, bx, by, A, t.as<float2>(), work.as<float2>(), nx, ny, true, st);
        const int nblk = (int)std::min<int64_t>((F + 255) / 256, 1024);
        part.ensure(sizeof(double) * 2 * (size_t)nblk);
        hipLaunchKernelGGL(k_terrain_moments, dim3(nblk), dim3(256), 0, st, A, F,
                           part.as<double>());
        ERP_CHECK_LAUNCH("k_terrain_moments");
        hipLaunchKernelGGL(k_terrain_heights, dim3(grid_of(F)), dim3(256), 0, st, h, A, F,
                           part.as<double>(), nblk, tr.rms);
        ERP_CHECK_LAUNCH("k_terrain_heights");
        ERP_HIP_TRY(hipStreamSynchronize(st));
    } catch (...) {
        (void)hipStreamSynchronize(st);
        bx.release(); by.release(); a.release(); t.release(); work.release(); part.release();
        noise.release();
        throw;
    }
    bx.release(); by.release(); a.release(); t.release(); work.release(); part.release();
    noise.release();
}
