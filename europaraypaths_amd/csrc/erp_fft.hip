// erp_fft.hip — hand-written power-of-two complex FFT for gfx950.
//
// Replaces the cuFFT calls of the reference (radar_kernels.cu:532-550,
// 589-645; elementary_kernels.cu:303-351).  Stockham autosort stages of
// radix 4 (plus one radix-2 stage for odd log2 N) run in LDS with 256-thread
// workgroups holding 4096 complex points (32 KB).
//
// N <= 4096: one pass, natural order in and out; several transforms share a
// workgroup when N < 4096.
// N  > 4096: N = N1 * N2 (N1 = row length, contiguous; N2 = number of rows).
//   forward : column FFTs (length N2, stride N1) * twiddle W_N^(n1 k2), then
//             row FFTs (length N1).  Output is "permuted": frequency
//             k = k2 + N2*k1 lives at address k2*N1 + k1.
//   inverse : row IFFTs * conj twiddle, then column IFFTs; takes the permuted
//             order and returns natural order.
// Pointwise spectral products never look at the order, so a forward/product/
// inverse chain needs no transpose at all.  Twiddles come from one table of
// exp(-2 pi i k / 2^18) computed in double.
#include "erp_internal.h"

namespace {

constexpr int NT = 256;         // threads per workgroup
constexpr int PPT = 16;         // points per thread
constexpr int NPTS = NT * PPT;  // points per workgroup (4096)
constexpr int LOG_TWN = 18;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int R> __device__ __forceinline__ void dft_r(float2* v, bool inv);

template <> __device__ __forceinline__ void dft_r<2>(float2* v, bool) {
    float2 a = v[0], b = v[1];
    v[0] = make_float2(a.x + b.x, a.y + b.y);
    v[1] = make_float2(a.x - b.x, a.y - b.y);
}

template <> __device__ __forceinline__ void dft_r<4>(float2* v, bool inv) {
    float2 a = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
    float2 b = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
    float2 c = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
    float2 d = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
    v[0] = make_float2(a.x + c.x, a.y + c.y);
    v[2] = make_float2(a.x - c.x, a.y - c.y);
    if (!inv) {  // -i*d
        v[1] = make_float2(b.x + d.y, b.y - d.x);
        v[3] = make_float2(b.x - d.y, b.y + d.x);
    } else {     // +i*d
        v[1] = make_float2(b.x - d.y, b.y + d.x);
        v[3] = make_float2(b.x + d.y, b.y - d.x);
    }
}

// One Stockham stage of radix R over nb transforms of length n stored in LDS
// at s[b*bstride + i].  Ls = product of the radices of the previous stages.
// All inputs are read into registers before the barrier, so the stage is
// in place.
template <int R>
__device__ __forceinline__ void stockham_stage(float2* s, int n, int Ls, int log2LsR,
                                               int nb, int bstride, bool inv,
                                               const float2* __restrict__ tw) {
    constexpr int Q = PPT / R;
    const int nR = n / R;
    const int nbf = nb * nR;
    float2 v[Q][R];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = threadIdx.x + q * NT;
        if (j < nbf) {
            const int b = j / nR, jj = j - b * nR;
            const float2* src = s + b * bstride + jj;
#pragma unroll
            for (int i = 0; i < R; ++i) v[q][i] = src[i * nR];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = threadIdx.x + q * NT;
        if (j < nbf) {
            const int b = j / nR, jj = j - b * nR;
            const int k = jj & (Ls - 1);
            if (Ls > 1) {
#pragma unroll
                for (int i = 1; i < R; ++i) {
                    float2 w = tw[(i * k) << (LOG_TWN - log2LsR)];
                    if (inv) w.y = -w.y;
                    v[q][i] = cmul(v[q][i], w);
                }
            }
            dft_r<R>(v[q], inv);
            float2* dst = s + b * bstride + (jj - k) * R + k;
#pragma unroll
            for (int i = 0; i < R; ++i) dst[i * Ls] = v[q][i];
        }
    }
    __syncthreads();
}

// Complete FFT of nb transforms of length n (power of two, nb*n <= NPTS).
__device__ void lds_fft(float2* s, int n, int log2n, int nb, int bstride, bool inv,
                        const float2* __restrict__ tw) {
    int Ls = 1, lg = 0, rem = log2n;
    if (rem & 1) {
        stockham_stage<2>(s, n, Ls, lg + 1, nb, bstride, inv, tw);
        Ls <<= 1; lg += 1; rem -= 1;
    }
    while (rem > 0) {
        stockham_stage<4>(s, n, Ls, lg + 2, nb, bstride, inv, tw);
        Ls <<= 2; lg += 2; rem -= 2;
    }
}

__global__ void k_twiddle_init(float2* tw) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < (1 << LOG_TWN)) {
        double s, c;
        sincospi(-2.0 * (double)k / (double)(1 << LOG_TWN), &s, &c);
        tw[k] = make_float2((float)c, (float)s);
    }
}

// Single-pass transform: each workgroup holds NPTS/n consecutive rows.
__global__ void __launch_bounds__(NT) k_fft_small(float2* __restrict__ data, int n,
                                                  int log2n, int batch, int inv,
                                                  const float2* __restrict__ tw) {
    __shared__ float2 s[NPTS];
    const int nb = NPTS / n;
    const int row0 = blockIdx.x * nb;
    const int rows = min(nb, batch - row0);
    if (rows <= 0) return;
    float2* g = data + (size_t)row0 * n;
    const int tot = rows * n;
    for (int i = threadIdx.x; i < NPTS; i += NT)
        s[i] = (i < tot) ? g[i] : make_float2(0.f, 0.f);
    __syncthreads();
    lds_fft(s, n, log2n, nb, n, inv != 0, tw);
    for (int i = threadIdx.x; i < tot; i += NT) g[i] = s[i];
}

// Column pass of the two-pass transform.  Workgroup = CT adjacent columns x
// N2 rows of one batch row; LDS column stride N2+1 avoids bank conflicts.
// fwd: FFT then multiply element (k2, n1) by W_N^(n1*k2).  inv: IFFT only.
__global__ void __launch_bounds__(NT) k_fft_cols(float2* __restrict__ data, int N1,
                                                 int N2, int log2N2, int log2N, int inv,
                                                 const float2* __restrict__ tw) {
    extern __shared__ float2 s[];
    const int CT = NPTS / N2;
    const int bstride = N2 + 1;
    const int c0 = blockIdx.x * CT;
    float2* g = data + (size_t)blockIdx.y * N1 * N2;
    for (int idx = threadIdx.x; idx < NPTS; idx += NT) {
        const int r = idx / CT, c = idx - r * CT;
        s[c * bstride + r] = g[(size_t)r * N1 + c0 + c];
    }
    __syncthreads();
    lds_fft(s, N2, log2N2, CT, bstride, inv != 0, tw);
    for (int idx = threadIdx.x; idx < NPTS; idx += NT) {
        const int r = idx / CT, c = idx - r * CT;
        float2 v = s[c * bstride + r];
        if (!inv) {
            const unsigned e = (unsigned)((c0 + c) * r) & ((1u << log2N) - 1u);
            v = cmul(v, tw[e << (LOG_TWN - log2N)]);
        }
        g[(size_t)r * N1 + c0 + c] = v;
    }
}

// Row pass of the two-pass transform.  Workgroup = RT rows of length N1.
// fwd: FFT only.  inv: IFFT then multiply element (k2, n1) by W_N^(-n1*k2).
__global__ void __launch_bounds__(NT) k_fft_rows(float2* __restrict__ data, int N1,
                                                 int log2N1, int N2, int log2N, int inv,
                                                 const float2* __restrict__ tw) {
    __shared__ float2 s[NPTS];
    const int RT = NPTS / N1;
    const int r0 = blockIdx.x * RT;
    float2* g = data + (size_t)blockIdx.y * N1 * N2 + (size_t)r0 * N1;
    for (int i = threadIdx.x; i < NPTS; i += NT) s[i] = g[i];
    __syncthreads();
    lds_fft(s, N1, log2N1, RT, N1, inv != 0, tw);
    for (int i = threadIdx.x; i < NPTS; i += NT) {
        float2 v = s[i];
        if (inv) {
            const int r = r0 + i / N1, n1 = i & (N1 - 1);
            const unsigned e = (unsigned)(n1 * r) & ((1u << log2N) - 1u);
            float2 w = tw[e << (LOG_TWN - log2N)];
            w.y = -w.y;
            v = cmul(v, w);
        }
        g[i] = v;
    }
}

}  // namespace

FftPlan fft_plan(int N) {
    FftPlan p;
    if (N < 2 || (N & (N - 1)) != 0 || N > (1 << LOG_TWN))
        throw ErpError{ERP_E_INVALID, "FFT size " + std::to_string(N) + " unsupported"};
    p.N = N;
    p.log2N = ilog2(N);
    if (N <= NPTS) {
        p.two_pass = false;
    } else {
        p.two_pass = true;
        const int l1 = (p.log2N + 1) / 2;
        p.N1 = 1 << l1;
        p.N2 = N >> l1;
    }
    return p;
}

void fft_init_twiddles(erp_ctx* ctx) {
    ctx->twiddle.ensure(sizeof(float2) * (size_t)(1 << LOG_TWN));
    hipLaunchKernelGGL(k_twiddle_init, dim3((1 << LOG_TWN) / 256), dim3(256), 0,
                       ctx->stream, ctx->twiddle.as<float2>());
    ERP_CHECK_LAUNCH("k_twiddle_init");
}

static void run_fft(erp_ctx* ctx, const FftPlan& p, float2* data, int batch, bool inv,
                    hipStream_t s) {
    if (batch <= 0) return;
    const float2* tw = ctx->twiddle.as<float2>();
    if (!p.two_pass) {
        const int nb = NPTS / p.N;
        dim3 grid((batch + nb - 1) / nb);
        hipLaunchKernelGGL(k_fft_small, grid, dim3(NT), 0, s, data, p.N, p.log2N, batch,
                           inv ? 1 : 0, tw);
        ERP_CHECK_LAUNCH("k_fft_small");
        return;
    }
    const int CT = NPTS / p.N2, RT = NPTS / p.N1;
    const size_t col_lds = sizeof(float2) * (size_t)CT * (p.N2 + 1);
    dim3 gcol(p.N1 / CT, batch), grow(p.N2 / RT, batch);
    if (!inv) {
        hipLaunchKernelGGL(k_fft_cols, gcol, dim3(NT), col_lds, s, data, p.N1, p.N2,
                           ilog2(p.N2), p.log2N, 0, tw);
        ERP_CHECK_LAUNCH("k_fft_cols");
        hipLaunchKernelGGL(k_fft_rows, grow, dim3(NT), 0, s, data, p.N1, ilog2(p.N1), p.N2,
                           p.log2N, 0, tw);
        ERP_CHECK_LAUNCH("k_fft_rows");
    } else {
        hipLaunchKernelGGL(k_fft_rows, grow, dim3(NT), 0, s, data, p.N1, ilog2(p.N1), p.N2,
                           p.log2N, 1, tw);
        ERP_CHECK_LAUNCH("k_fft_rows");
        hipLaunchKernelGGL(k_fft_cols, gcol, dim3(NT), col_lds, s, data, p.N1, p.N2,
                           ilog2(p.N2), p.log2N, 1, tw);
        ERP_CHECK_LAUNCH("k_fft_cols");
    }
}

void fft_forward(erp_ctx* ctx, const FftPlan& p, float2* data, int batch, hipStream_t s) {
    run_fft(ctx, p, data, batch, false, s);
}

void fft_inverse(erp_ctx* ctx, const FftPlan& p, float2* data, int batch, hipStream_t s) {
    run_fft(ctx, p, data, batch, true, s);
}
