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
// inverse chain needs no transpose at all.  Stage twiddles come from a
// "ladder" table lad[m + j] = exp(-2 pi i j / m) (m = 1, 2, 4, ..., 4096) so
// the lanes of a stage read consecutive entries; the inter-pass twiddles of
// each two-pass size live in a table laid out like the data (coalesced, L2
// resident).  All tables are computed in double and rounded once.
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
                    float2 w = tw[(1 << log2LsR) + i * k];
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

// Copy ladder entries [0, 2n) (the twiddles of every stage of a length-n
// transform) into LDS; callers sync before the first stage.
constexpr int TWL = 1024;  // LDS ladder capacity: transforms up to 512 points
__device__ __forceinline__ const float2* stage_ladder(float2* twl, const float2* __restrict__ tw,
                                                      int n) {
    if (2 * n > TWL) return tw;
    for (int i = threadIdx.x; i < 2 * n; i += NT) twl[i] = tw[i];
    return twl;
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

// ladder table: lad[m + j] = exp(-2 pi i j / m) for m = 2^0 .. 2^12, j < m
__global__ void k_twiddle_init(float2* lad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 2 * NPTS) {
        if (t == 0) {
            lad[0] = make_float2(1.f, 0.f);
            return;
        }
        const int m = 1 << (31 - __clz(t));
        const int j = t - m;
        double s, c;
        sincospi(-2.0 * (double)j / (double)m, &s, &c);
        lad[t] = make_float2((float)c, (float)s);
    }
}

// inter-pass twiddles of an N = N1 * N2 transform in data layout:
// tab[k2 * N1 + n1] = exp(-2 pi i n1 k2 / N)
__global__ void k_inter_init(float2* tab, int N1, int N) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N) {
        const int k2 = t / N1, n1 = t - k2 * N1;
        const long long e = ((long long)n1 * k2) % N;
        double s, c;
        sincospi(-2.0 * (double)e / (double)N, &s, &c);
        tab[t] = make_float2((float)c, (float)s);
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
                                                 const float2* __restrict__ tw,
                                                 const float2* __restrict__ it) {
    extern __shared__ float2 s[];
    const int CT = NPTS / N2;
    const int bstride = N2 + 1;
    const int c0 = blockIdx.x * CT;
    float2* g = data + (size_t)blockIdx.y * N1 * N2;
    __shared__ float2 twl[TWL];
    const float2* tws = stage_ladder(twl, tw, N2);
    for (int idx = threadIdx.x; idx < NPTS; idx += NT) {
        const int r = idx / CT, c = idx - r * CT;
        s[c * bstride + r] = g[(size_t)r * N1 + c0 + c];
    }
    __syncthreads();
    lds_fft(s, N2, log2N2, CT, bstride, inv != 0, tws);
    for (int idx = threadIdx.x; idx < NPTS; idx += NT) {
        const int r = idx / CT, c = idx - r * CT;
        float2 v = s[c * bstride + r];
        if (!inv) v = cmul(v, it[(size_t)r * N1 + c0 + c]);
        g[(size_t)r * N1 + c0 + c] = v;
    }
}

// Row pass of the two-pass transform.  Workgroup = RT rows of length N1.
// fwd: FFT only.  inv: IFFT then multiply element (k2, n1) by W_N^(-n1*k2).
__global__ void __launch_bounds__(NT) k_fft_rows(float2* __restrict__ data, int N1,
                                                 int log2N1, int N2, int log2N, int inv,
                                                 const float2* __restrict__ tw,
                                                 const float2* __restrict__ it) {
    __shared__ float2 s[NPTS];
    const int RT = NPTS / N1;
    const int r0 = blockIdx.x * RT;
    float2* g = data + (size_t)blockIdx.y * N1 * N2 + (size_t)r0 * N1;
    __shared__ float2 twl[TWL];
    const float2* tws = stage_ladder(twl, tw, N1);
    for (int i = threadIdx.x; i < NPTS; i += NT) s[i] = g[i];
    __syncthreads();
    lds_fft(s, N1, log2N1, RT, N1, inv != 0, tws);
    for (int i = threadIdx.x; i < NPTS; i += NT) {
        float2 v = s[i];
        if (inv) {
            float2 w = it[(size_t)r0 * N1 + i];
            w.y = -w.y;
            v = cmul(v, w);
        }
        g[i] = v;
    }
}


// ---------------------------------------------------------------------------
// Fused per-target spectral chain.  For every (shot, target) pair p with
// inward trace A_p (rows of length P, nonzero only in the bin window
// [base_p, base_p + win) unless base_p < 0) it produces
//     X_p = FFT(A_p),  y_p = IFFT(X_p G),  z_p[n] = y_p[n + nr] (n < nr),
//     Z_p = FFT(z_p)
// and sums X_p Z_p over the pairs of each shot.  Two-pass sizes use four
// kernels with every transform step held in LDS:
//   cols_fwd  : column FFT of the windowed rows (+ twiddle)      A  -> X'
//   rows_conv : row FFT, x G, row IFFT (+ conj twiddle)          X' -> Y'
//   cols_shift: column IFFT, shift by nr/N1 rows, column FFT (+ twiddle)
//                                                                Y' -> Z'
//   rows_mac  : row FFTs of X' and Z', multiply-accumulate over a group of
//               pairs into a partial spectrum
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NT) k_tfft_cols_fwd(float2* __restrict__ data, int N1, int N2,
                                                      int log2N2, int log2N,
                                                      const int* __restrict__ pbase, int win,
                                                      int nr, const float2* __restrict__ tw,
                                                      const float2* __restrict__ it) {
    extern __shared__ float2 s[];
    const int CT = NPTS / N2;
    const int bstride = N2 + 1;
    const int c0 = blockIdx.x * CT;
    const int base = pbase[blockIdx.y];
    float2* g = data + (size_t)blockIdx.y * N1 * N2;
    __shared__ float2 twl[TWL];
    const float2* tws = stage_ladder(twl, tw, N2);
    for (int idx = threadIdx.x; idx < NPTS; idx += NT) {
        const int r = idx / CT, c = idx - r * CT;
        const int n = r * N1 + c0 + c;
        const bool live = (n < nr) && (base < 0 || (n >= base && n < base + win));
        s[c * bstride + r] = live ? g[n] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    lds_fft(s, N2, log2N2, CT, bstride, false, tws);
    for (int idx = threadIdx.x; idx < NPTS; idx += NT) {
        const int r = idx / CT, c = idx - r * CT;
        g[(size_t)r * N1 + c0 + c] = cmul(s[c * bstride + r], it[(size_t)r * N1 + c0 + c]);
    }
}

__global__ void __launch_bounds__(NT) k_tfft_rows_conv(const float2* __restrict__ src,
                                                       float2* __restrict__ dst, int N1,
                                                       int log2N1, int N2, int log2N,
                                                       const float2* __restrict__ G,
                                                       const float2* __restrict__ tw,
                                                       const float2* __restrict__ it) {
    __shared__ float2 s[NPTS];
    const int RT = NPTS / N1;
    const int r0 = blockIdx.x * RT;
    const size_t off = (size_t)blockIdx.y * N1 * N2 + (size_t)r0 * N1;
    __shared__ float2 twl[TWL];
    const float2* tws = stage_ladder(twl, tw, N1);
    for (int i = threadIdx.x; i < NPTS; i += NT) s[i] = src[off + i];
    __syncthreads();
    lds_fft(s, N1, log2N1, RT, N1, false, tws);
    const float2* Gr = G + (size_t)r0 * N1;
    for (int i = threadIdx.x; i < NPTS; i += NT) s[i] = cmul(s[i], Gr[i]);
    __syncthreads();
    lds_fft(s, N1, log2N1, RT, N1, true, tws);
    for (int i = threadIdx.x; i < NPTS; i += NT) {
        float2 w = it[(size_t)r0 * N1 + i];
        w.y = -w.y;
        dst[off + i] = cmul(s[i], w);
    }
}

__global__ void __launch_bounds__(NT) k_tfft_cols_shift(float2* __restrict__ data, int N1, int N2,
                                                        int log2N2, int log2N, int shift,
                                                        const float2* __restrict__ tw,
                                                        const float2* __restrict__ it) {
    extern __shared__ float2 s[];
    const int CT = NPTS / N2;
    const int bstride = N2 + 1;
    const int c0 = blockIdx.x * CT;
    float2* g = data + (size_t)blockIdx.y * N1 * N2;
    __shared__ float2 twl[TWL];
    const float2* tws = stage_ladder(twl, tw, N2);
    for (int idx = threadIdx.x; idx < NPTS; idx += NT) {
        const int r = idx / CT, c = idx - r * CT;
        s[c * bstride + r] = g[(size_t)r * N1 + c0 + c];
    }
    __syncthreads();
    lds_fft(s, N2, log2N2, CT, bstride, true, tws);
    // z row r = y row r + shift for r < shift (shift = nr / N1), zero above
    float2 v[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + q * NT;
        const int c = idx / N2, r = idx - c * N2;
        v[q] = (r < shift) ? s[c * bstride + r + shift] : make_float2(0.f, 0.f);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + q * NT;
        const int c = idx / N2, r = idx - c * N2;
        s[c * bstride + r] = v[q];
    }
    __syncthreads();
    lds_fft(s, N2, log2N2, CT, bstride, false, tws);
    for (int idx = threadIdx.x; idx < NPTS; idx += NT) {
        const int r = idx / CT, c = idx - r * CT;
        g[(size_t)r * N1 + c0 + c] = cmul(s[c * bstride + r], it[(size_t)r * N1 + c0 + c]);
    }
}

__global__ void __launch_bounds__(NT) k_tfft_rows_mac(const float2* __restrict__ X,
                                                      const float2* __restrict__ Z, int N1,
                                                      int log2N1, int N2,
                                                      const int* __restrict__ gbeg,
                                                      const int* __restrict__ gend,
                                                      float2* __restrict__ partial,
                                                      const float2* __restrict__ tw) {
    __shared__ float2 s[NPTS];
    const int RT = NPTS / N1;
    const int r0 = blockIdx.x * RT;
    const int grp = blockIdx.y;
    const size_t P = (size_t)N1 * N2;
    float2 acc[PPT], xr[PPT];
    __shared__ float2 twl[TWL];
    const float2* tws = stage_ladder(twl, tw, N1);
#pragma unroll
    for (int q = 0; q < PPT; ++q) acc[q] = make_float2(0.f, 0.f);
    for (int p = gbeg[grp]; p < gend[grp]; ++p) {
        const size_t off = (size_t)p * P + (size_t)r0 * N1;
        for (int i = threadIdx.x; i < NPTS; i += NT) s[i] = X[off + i];
        __syncthreads();
        lds_fft(s, N1, log2N1, RT, N1, false, tws);
#pragma unroll
        for (int q = 0; q < PPT; ++q) xr[q] = s[threadIdx.x + q * NT];
        __syncthreads();
        for (int i = threadIdx.x; i < NPTS; i += NT) s[i] = Z[off + i];
        __syncthreads();
        lds_fft(s, N1, log2N1, RT, N1, false, tws);
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const float2 z = s[threadIdx.x + q * NT];
            acc[q].x += xr[q].x * z.x - xr[q].y * z.y;
            acc[q].y += xr[q].x * z.y + xr[q].y * z.x;
        }
        __syncthreads();
    }
    float2* o = partial + (size_t)grp * P + (size_t)r0 * N1;
#pragma unroll
    for (int q = 0; q < PPT; ++q) o[threadIdx.x + q * NT] = acc[q];
}

// Single-pass sizes (P <= 4096): the whole chain for one pair in one
// workgroup; writes the product X_p Z_p.
__global__ void __launch_bounds__(NT) k_tfft_small(const float2* __restrict__ A, int P, int log2P,
                                                   const int* __restrict__ pbase, int win, int nr,
                                                   const float2* __restrict__ G,
                                                   float2* __restrict__ prod,
                                                   const float2* __restrict__ tw) {
    __shared__ float2 s[NPTS];
    const int p = blockIdx.x;
    const int base = pbase[p];
    const float2* a = A + (size_t)p * P;
    for (int n = threadIdx.x; n < P; n += NT) {
        const bool live = (n < nr) && (base < 0 || (n >= base && n < base + win));
        s[n] = live ? a[n] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    lds_fft(s, P, log2P, 1, P, false, tw);
    float2 xr[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int n = threadIdx.x + q * NT;
        if (n < P) {
            xr[q] = s[n];
            s[n] = cmul(xr[q], G[n]);
        }
    }
    __syncthreads();
    lds_fft(s, P, log2P, 1, P, true, tw);
    float2 v[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int n = threadIdx.x + q * NT;
        if (n < P) v[q] = (n < nr) ? s[n + nr] : make_float2(0.f, 0.f);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int n = threadIdx.x + q * NT;
        if (n < P) s[n] = v[q];
    }
    __syncthreads();
    lds_fft(s, P, log2P, 1, P, false, tw);
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int n = threadIdx.x + q * NT;
        if (n < P) prod[(size_t)p * P + n] = cmul(xr[q], s[n]);
    }
}

// S[shot] += scale * sum of rows [beg, end) of src (one segment per shot)
__global__ void k_sum_rows(float2* __restrict__ S, const float2* __restrict__ src,
                           const int* __restrict__ seg_shot, const int* __restrict__ seg_beg,
                           const int* __restrict__ seg_end, int P, float scale) {
    const int sg = blockIdx.y;
    const int sh = seg_shot[sg], b = seg_beg[sg], e = seg_end[sg];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < P; k += gridDim.x * blockDim.x) {
        float2 acc = make_float2(0.f, 0.f);
        for (int r = b; r < e; ++r) {
            const float2 v = src[(size_t)r * P + k];
            acc.x += v.x;
            acc.y += v.y;
        }
        float2* d = S + (size_t)sh * P + k;
        d->x += scale * acc.x;
        d->y += scale * acc.y;
    }
}

}  // namespace

FftPlan fft_plan(int N) {
    FftPlan p;
    if (N < 2 || (N & (N - 1)) != 0 || N > NPTS * NPTS)
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
    ctx->twiddle.ensure(sizeof(float2) * (size_t)(2 * NPTS));
    hipLaunchKernelGGL(k_twiddle_init, dim3(2 * NPTS / 256), dim3(256), 0, ctx->stream,
                       ctx->twiddle.as<float2>());
    ERP_CHECK_LAUNCH("k_twiddle_init");
}

// inter-pass twiddle table of a two-pass plan (cached per size in the context)
static const float2* inter_table(erp_ctx* ctx, const FftPlan& p, hipStream_t s) {
    for (auto& t : ctx->inter)
        if (t.first == p.N) return t.second->as<float2>();
    DevBuf* b = new DevBuf();
    b->ensure(sizeof(float2) * (size_t)p.N);
    hipLaunchKernelGGL(k_inter_init, dim3((p.N + 255) / 256), dim3(256), 0, s, b->as<float2>(),
                       p.N1, p.N);
    ERP_CHECK_LAUNCH("k_inter_init");
    ctx->inter.push_back({p.N, b});
    return b->as<float2>();
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
    const float2* it = inter_table(ctx, p, s);
    dim3 gcol(p.N1 / CT, batch), grow(p.N2 / RT, batch);
    if (!inv) {
        hipLaunchKernelGGL(k_fft_cols, gcol, dim3(NT), col_lds, s, data, p.N1, p.N2,
                           ilog2(p.N2), p.log2N, 0, tw, it);
        ERP_CHECK_LAUNCH("k_fft_cols");
        hipLaunchKernelGGL(k_fft_rows, grow, dim3(NT), 0, s, data, p.N1, ilog2(p.N1), p.N2,
                           p.log2N, 0, tw, it);
        ERP_CHECK_LAUNCH("k_fft_rows");
    } else {
        hipLaunchKernelGGL(k_fft_rows, grow, dim3(NT), 0, s, data, p.N1, ilog2(p.N1), p.N2,
                           p.log2N, 1, tw, it);
        ERP_CHECK_LAUNCH("k_fft_rows");
        hipLaunchKernelGGL(k_fft_cols, gcol, dim3(NT), col_lds, s, data, p.N1, p.N2,
                           ilog2(p.N2), p.log2N, 1, tw, it);
        ERP_CHECK_LAUNCH("k_fft_cols");
    }
}

void fft_forward(erp_ctx* ctx, const FftPlan& p, float2* data, int batch, hipStream_t s) {
    run_fft(ctx, p, data, batch, false, s);
}

void fft_inverse(erp_ctx* ctx, const FftPlan& p, float2* data, int batch, hipStream_t s) {
    run_fft(ctx, p, data, batch, true, s);
}

// Host driver of the fused chain for n pairs (see the kernel comment).  A and
// Y hold n rows of P; partial/prod workspace holds max(ngroups, n) rows.
void target_spectra_fused(erp_ctx* ctx, const FftPlan& p, float2* A, float2* Y,
                          const int* pbase, int win, int nr, int n, const float2* G,
                          float2* work, int ngroups, const int* gbeg, const int* gend,
                          const int* gseg_shot, const int* gseg_beg, const int* gseg_end,
                          int nseg, float2* S, float scale, hipStream_t st) {
    const float2* tw = ctx->twiddle.as<float2>();
    const int P = p.N;
    if (!p.two_pass) {
        hipLaunchKernelGGL(k_tfft_small, dim3(n), dim3(NT), 0, st, A, P, p.log2N, pbase, win, nr,
                           G, work, tw);
        ERP_CHECK_LAUNCH("k_tfft_small");
        // groups are unused here: sum the per-pair products per shot segment
        hipLaunchKernelGGL(k_sum_rows, dim3(std::max(1, P / 256), nseg), dim3(256), 0, st, S,
                           work, gseg_shot, gseg_beg, gseg_end, P, scale);
        ERP_CHECK_LAUNCH("k_sum_rows");
        return;
    }
    const int CT = NPTS / p.N2, RT = NPTS / p.N1;
    const size_t col_lds = sizeof(float2) * (size_t)CT * (p.N2 + 1);
    const int l1 = ilog2(p.N1), l2 = ilog2(p.N2);
    if (nr % p.N1 != 0) throw ErpError{ERP_E_INVALID, "fused target chain needs nr % N1 == 0"};
    const float2* it = inter_table(ctx, p, st);
    hipLaunchKernelGGL(k_tfft_cols_fwd, dim3(p.N1 / CT, n), dim3(NT), col_lds, st, A, p.N1, p.N2,
                       l2, p.log2N, pbase, win, nr, tw, it);
    ERP_CHECK_LAUNCH("k_tfft_cols_fwd");
    hipLaunchKernelGGL(k_tfft_rows_conv, dim3(p.N2 / RT, n), dim3(NT), 0, st, A, Y, p.N1, l1,
                       p.N2, p.log2N, G, tw, it);
    ERP_CHECK_LAUNCH("k_tfft_rows_conv");
    hipLaunchKernelGGL(k_tfft_cols_shift, dim3(p.N1 / CT, n), dim3(NT), col_lds, st, Y, p.N1,
                       p.N2, l2, p.log2N, nr / p.N1, tw, it);
    ERP_CHECK_LAUNCH("k_tfft_cols_shift");
    hipLaunchKernelGGL(k_tfft_rows_mac, dim3(p.N2 / RT, ngroups), dim3(NT), 0, st, A, Y, p.N1, l1,
                       p.N2, gbeg, gend, work, tw);
    ERP_CHECK_LAUNCH("k_tfft_rows_mac");
    hipLaunchKernelGGL(k_sum_rows, dim3(std::max(1, P / 256), nseg), dim3(256), 0, st, S, work,
                       gseg_shot, gseg_beg, gseg_end, P, scale);
    ERP_CHECK_LAUNCH("k_sum_rows");
}
