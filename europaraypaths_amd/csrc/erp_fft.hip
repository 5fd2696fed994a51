// erp_fft.hip — hand-written power-of-two complex FFT for gfx950.
//
// Replaces the cuFFT calls of the reference (radar_kernels.cu:532-550,
// 589-645; elementary_kernels.cu:303-351).  Stockham autosort stages of
// radix 16 (plus one radix 2, 4 or 8 stage when log2 n is not a multiple of
// four) run on 256-thread workgroups holding 4096 complex points in LDS
// (32 KB); each thread keeps 16 points in registers, so a 256-point transform
// costs two LDS exchanges.  LDS indices go through an XOR swizzle that keeps
// the Stockham strided writes and the column transposes free of bank
// conflicts.
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
// "ladder" table lad[m + j] = exp(-2 pi i j / m) (m = 1, 2, 4, ..., 4096);
// the inter-pass twiddles of each two-pass size live in a table laid out
// like the data (coalesced, L2 resident).  All tables are computed in double
// and rounded once.
#include "erp_internal.h"

namespace {

constexpr int NT = 256;         // threads per workgroup
constexpr int PPT = 16;         // points per thread
constexpr int NPTS = NT * PPT;  // points per workgroup (4096)
constexpr int TWL = 1024;       // LDS ladder capacity: transforms up to 512 points
constexpr int LADN = 65536;     // ladder entries: m = 2^0 .. 2^15

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// LDS swizzle (a bijection on [0, 4096) that only moves bits 0-4): the low
// four index bits are XORed with bits 4-11, then bit 4 with bits 6-8.  Every
// group of 16 lanes of the Stockham reads/writes and of the column tile
// load/store then touches 16 distinct 8-byte bank pairs.
__device__ __forceinline__ int sw(int a) {
    a ^= ((a >> 4) ^ (a >> 8)) & 15;
    return a ^ ((((a >> 6) ^ (a >> 7) ^ (a >> 8)) & 1) << 4);
}

// Thread index hidden from the optimiser: addresses derived from it are
// recomputed where they are used instead of being hoisted and kept live
// across transforms (which would exhaust the register budget).
__device__ __forceinline__ int opaque_tid() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// v * W with W = (c, -s) forward, (c, +s) inverse
__device__ __forceinline__ float2 twc(float2 v, float c, float s, bool inv) {
    const float si = inv ? s : -s;
    return make_float2(v.x * c - v.y * si, v.x * si + v.y * c);
}

__device__ __forceinline__ void dft2(float2& a, float2& b) {
    const float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

__device__ __forceinline__ void dft4(float2& v0, float2& v1, float2& v2, float2& v3, bool inv) {
    const float2 a = cadd(v0, v2), b = csub(v0, v2), c = cadd(v1, v3), d = csub(v1, v3);
    v0 = cadd(a, c);
    v2 = csub(a, c);
    if (!inv) {  // -i d
        v1 = make_float2(b.x + d.y, b.y - d.x);
        v3 = make_float2(b.x - d.y, b.y + d.x);
    } else {     // +i d
        v1 = make_float2(b.x - d.y, b.y + d.x);
        v3 = make_float2(b.x + d.y, b.y - d.x);
    }
}

constexpr float K_C1 = 0.92387953251128674f;  // cos(pi/8)
constexpr float K_S1 = 0.38268343236508978f;  // sin(pi/8)
constexpr float K_R2 = 0.70710678118654752f;  // sqrt(1/2)

template <int R> __device__ __forceinline__ void dft_r(float2* v, bool inv);

template <> __device__ __forceinline__ void dft_r<2>(float2* v, bool) { dft2(v[0], v[1]); }

template <> __device__ __forceinline__ void dft_r<4>(float2* v, bool inv) {
    dft4(v[0], v[1], v[2], v[3], inv);
}

// 8-point DFT as 2 x 4: n = i2 + 2j (i2 < 2, j < 4), k = k1 + 4 k2
template <> __device__ __forceinline__ void dft_r<8>(float2* v, bool inv) {
    float2 a0 = v[0], a1 = v[2], a2 = v[4], a3 = v[6];
    float2 b0 = v[1], b1 = v[3], b2 = v[5], b3 = v[7];
    dft4(a0, a1, a2, a3, inv);
    dft4(b0, b1, b2, b3, inv);
    b1 = twc(b1, K_R2, K_R2, inv);   // W8^1
    b2 = twc(b2, 0.f, 1.f, inv);     // W8^2
    b3 = twc(b3, -K_R2, K_R2, inv);  // W8^3
    v[0] = cadd(a0, b0); v[4] = csub(a0, b0);
    v[1] = cadd(a1, b1); v[5] = csub(a1, b1);
    v[2] = cadd(a2, b2); v[6] = csub(a2, b2);
    v[3] = cadd(a3, b3); v[7] = csub(a3, b3);
}

// 16-point DFT as 4 x 4: n = i2 + 4j, k = k1 + 4 k2
template <> __device__ __forceinline__ void dft_r<16>(float2* v, bool inv) {
    float2 a[4][4];
#pragma unroll
    for (int i2 = 0; i2 < 4; ++i2) {
        a[i2][0] = v[i2]; a[i2][1] = v[i2 + 4]; a[i2][2] = v[i2 + 8]; a[i2][3] = v[i2 + 12];
        dft4(a[i2][0], a[i2][1], a[i2][2], a[i2][3], inv);
    }
    // W16^(i2 k1)
    a[1][1] = twc(a[1][1], K_C1, K_S1, inv);    // 1
    a[1][2] = twc(a[1][2], K_R2, K_R2, inv);    // 2
    a[1][3] = twc(a[1][3], K_S1, K_C1, inv);    // 3
    a[2][1] = twc(a[2][1], K_R2, K_R2, inv);    // 2
    a[2][2] = twc(a[2][2], 0.f, 1.f, inv);      // 4
    a[2][3] = twc(a[2][3], -K_R2, K_R2, inv);   // 6
    a[3][1] = twc(a[3][1], K_S1, K_C1, inv);    // 3
    a[3][2] = twc(a[3][2], -K_R2, K_R2, inv);   // 6
    a[3][3] = twc(a[3][3], -K_C1, -K_S1, inv);  // 9
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        float2 t0 = a[0][k1], t1 = a[1][k1], t2 = a[2][k1], t3 = a[3][k1];
        dft4(t0, t1, t2, t3, inv);
        v[k1] = t0; v[k1 + 4] = t1; v[k1 + 8] = t2; v[k1 + 12] = t3;
    }
}

// Multiply v[i] by the stage twiddle W_m^(i k) (conjugated for the inverse),
// i < R, m = Ls R.  Powers 1, 2, 4 and 8 come from the ladder table, the
// others are formed as products just before use, which keeps few twiddles
// live at once.
template <int R>
__device__ __forceinline__ void apply_twiddles(float2* v, const float2* __restrict__ lad, int m,
                                               int k, bool inv) {
    const float sg = inv ? -1.f : 1.f;
    auto ld = [&](int e) {
        const float2 w = lad[m + e * k];
        return make_float2(w.x, sg * w.y);
    };
    const float2 w1 = ld(1);
    v[1] = cmul(v[1], w1);
    if (R == 2) return;
    const float2 w2 = ld(2);
    v[2] = cmul(v[2], w2);
    if (R == 4) {
        v[3] = cmul(v[3], ld(3));
        return;
    }
    const float2 w3 = cmul(w1, w2);
    v[3] = cmul(v[3], w3);
    const float2 w4 = ld(4);
    v[4] = cmul(v[4], w4);
    v[5] = cmul(v[5], cmul(w4, w1));
    v[6] = cmul(v[6], cmul(w4, w2));
    v[7] = cmul(v[7], cmul(w4, w3));
    if (R == 8) return;
    const float2 w8 = ld(8);
    v[8] = cmul(v[8], w8);
    v[9] = cmul(v[9], cmul(w8, w1));
    v[10] = cmul(v[10], cmul(w8, w2));
    v[11] = cmul(v[11], cmul(w8, w3));
    const float2 w12 = cmul(w8, w4);
    v[12] = cmul(v[12], w12);
    v[13] = cmul(v[13], cmul(w12, w1));
    v[14] = cmul(v[14], cmul(w12, w2));
    v[15] = cmul(v[15], cmul(w12, w3));
}

// One Stockham stage of radix R over nb transforms of length n stored at
// logical LDS index b*n + i.  Ls = product of the previous radices.  All
// inputs are read into registers before the barrier, so the stage is in place.
template <int R, int NTB = NT>
__device__ __forceinline__ void stockham_stage(float2* s, int n, int Ls, int nb, bool inv,
                                               const float2* __restrict__ lad) {
    constexpr int Q = PPT / R;
    const int nR = n / R;
    const int nbf = nb * nR;
    float2 v[Q][R];
    const int tid = opaque_tid();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = tid + q * NTB;
        if (j < nbf) {
            const int b = j / nR, jj = j - b * nR;
            const int src = b * n + jj;
#pragma unroll
            for (int i = 0; i < R; ++i) v[q][i] = s[sw(src + i * nR)];
        }
    }
    __syncthreads();
    const int tid2 = opaque_tid();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = tid2 + q * NTB;
        if (j < nbf) {
            const int b = j / nR, jj = j - b * nR;
            const int k = jj & (Ls - 1);
            if (Ls > 1) apply_twiddles<R>(v[q], lad, Ls * R, k, inv);
            dft_r<R>(v[q], inv);
            const int dst = b * n + (jj - k) * R + k;
#pragma unroll
            for (int i = 0; i < R; ++i) s[sw(dst + i * Ls)] = v[q][i];
        }
    }
    __syncthreads();
}

// Complete FFT of nb transforms of length n (power of two, nb*n <= 16 NTB),
// logical layout b*n + i, by a workgroup of NTB threads.  Callers sync
// before the call.
template <int NTB = NT>
__device__ __forceinline__ void lds_fft(float2* s, int n, int log2n, int nb, bool inv,
                        const float2* __restrict__ lad) {
    int Ls = 1;
    const int r = log2n & 3;
    if (r == 1) { stockham_stage<2, NTB>(s, n, Ls, nb, inv, lad); Ls <<= 1; }
    if (r == 2) { stockham_stage<4, NTB>(s, n, Ls, nb, inv, lad); Ls <<= 2; }
    if (r == 3) { stockham_stage<8, NTB>(s, n, Ls, nb, inv, lad); Ls <<= 3; }
    for (int rem = log2n - r; rem > 0; rem -= 4) {
        stockham_stage<16, NTB>(s, n, Ls, nb, inv, lad);
        Ls <<= 4;
    }
}

// Copy ladder entries [0, 2n) (the twiddles of every stage of a length-n
// transform) into LDS; callers sync before the first stage.
__device__ __forceinline__ const float2* stage_ladder(float2* twl, const float2* __restrict__ tw,
                                                      int n) {
    if (2 * n > TWL) return tw;
    for (int i = threadIdx.x; i < 2 * n; i += NT) twl[i] = tw[i];
    return twl;
}

// ladder table: lad[m + j] = exp(-2 pi i j / m) for m = 2^0 .. 2^15, j < m
__global__ void k_twiddle_init(float2* lad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < LADN) {
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
        s[sw(i)] = (i < tot) ? g[i] : make_float2(0.f, 0.f);
    __syncthreads();
    lds_fft(s, n, log2n, nb, inv != 0, tw);
    for (int i = threadIdx.x; i < tot; i += NT) g[i] = s[sw(i)];
}

// Column tile load: CT = NPTS/N2 adjacent columns of N2 rows go to LDS
// logical index c*N2 + r.
__device__ __forceinline__ void load_cols(float2* s, const float2* __restrict__ g, int N1, int N2,
                                          int c0) {
    const int CT = NPTS / N2;
    for (int idx = threadIdx.x; idx < NPTS; idx += NT) {
        const int r = idx / CT, c = idx - r * CT;
        s[sw(c * N2 + r)] = g[(size_t)r * N1 + c0 + c];
    }
}

// Column pass of the two-pass transform.
// fwd: FFT then multiply element (k2, n1) by W_N^(n1*k2).  inv: IFFT only.
__global__ void __launch_bounds__(NT) k_fft_cols(float2* __restrict__ data, int N1,
                                                 int N2, int log2N2, int inv,
                                                 const float2* __restrict__ tw,
                                                 const float2* __restrict__ it) {
    __shared__ float2 s[NPTS];
    __shared__ float2 twl[TWL];
    const int CT = NPTS / N2;
    const int c0 = blockIdx.x * CT;
    float2* g = data + (size_t)blockIdx.y * N1 * N2;
    const float2* tws = stage_ladder(twl, tw, N2);
    load_cols(s, g, N1, N2, c0);
    __syncthreads();
    lds_fft(s, N2, log2N2, CT, inv != 0, tws);
    for (int idx = threadIdx.x; idx < NPTS; idx += NT) {
        const int r = idx / CT, c = idx - r * CT;
        float2 v = s[sw(c * N2 + r)];
        if (!inv) v = cmul(v, it[(size_t)r * N1 + c0 + c]);
        g[(size_t)r * N1 + c0 + c] = v;
    }
}

// Row pass of the two-pass transform.  Workgroup = RT rows of length N1.
// fwd: FFT only.  inv: IFFT then multiply element (k2, n1) by W_N^(-n1*k2).
__global__ void __launch_bounds__(NT) k_fft_rows(float2* __restrict__ data, int N1,
                                                 int log2N1, int N2, int inv,
                                                 const float2* __restrict__ tw,
                                                 const float2* __restrict__ it) {
    __shared__ float2 s[NPTS];
    __shared__ float2 twl[TWL];
    const int RT = NPTS / N1;
    const int r0 = blockIdx.x * RT;
    float2* g = data + (size_t)blockIdx.y * N1 * N2 + (size_t)r0 * N1;
    const float2* tws = stage_ladder(twl, tw, N1);
    for (int i = threadIdx.x; i < NPTS; i += NT) s[sw(i)] = g[i];
    __syncthreads();
    lds_fft(s, N1, log2N1, RT, inv != 0, tws);
    for (int i = threadIdx.x; i < NPTS; i += NT) {
        float2 v = s[sw(i)];
        if (inv) v = cmul(v, conjf2(it[(size_t)r0 * N1 + i]));
        g[i] = v;
    }
}

// ---------------------------------------------------------------------------
// Fused per-target spectral chain.  For every (shot, target) pair p with
// inward trace A_p (rows of length P, nonzero only in the bin window
// [base_p, base_p + win_p) unless base_p < 0) it produces
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
                                                      int log2N2, const int* __restrict__ pbase,
                                                      const int* __restrict__ pwin, int nr,
                                                      const float2* __restrict__ tw,
                                                      const float2* __restrict__ it) {
    __shared__ float2 s[NPTS];
    __shared__ float2 twl[TWL];
    const int CT = NPTS / N2;
    const int c0 = blockIdx.x * CT;
    const int base = pbase[blockIdx.y], win = pwin[blockIdx.y];
    float2* g = data + (size_t)blockIdx.y * N1 * N2;
    const float2* tws = stage_ladder(twl, tw, N2);
    for (int idx = threadIdx.x; idx < NPTS; idx += NT) {
        const int r = idx / CT, c = idx - r * CT;
        const int n = r * N1 + c0 + c;
        const bool live = (n < nr) && (base < 0 || (n >= base && n < base + win));
        s[sw(c * N2 + r)] = live ? g[n] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    lds_fft(s, N2, log2N2, CT, false, tws);
    for (int idx = threadIdx.x; idx < NPTS; idx += NT) {
        const int r = idx / CT, c = idx - r * CT;
        g[(size_t)r * N1 + c0 + c] = cmul(s[sw(c * N2 + r)], it[(size_t)r * N1 + c0 + c]);
    }
}

__global__ void __launch_bounds__(NT) k_tfft_rows_conv(const float2* __restrict__ src,
                                                       float2* __restrict__ dst, int N1,
                                                       int log2N1, int N2,
                                                       const float2* __restrict__ G,
                                                       const float2* __restrict__ tw,
                                                       const float2* __restrict__ it) {
    __shared__ float2 s[NPTS];
    __shared__ float2 twl[TWL];
    const int RT = NPTS / N1;
    const int r0 = blockIdx.x * RT;
    const size_t off = (size_t)blockIdx.y * N1 * N2 + (size_t)r0 * N1;
    const float2* tws = stage_ladder(twl, tw, N1);
    for (int i = threadIdx.x; i < NPTS; i += NT) s[sw(i)] = src[off + i];
    __syncthreads();
    lds_fft(s, N1, log2N1, RT, false, tws);
    const float2* Gr = G + (size_t)r0 * N1;
    for (int i = threadIdx.x; i < NPTS; i += NT) s[sw(i)] = cmul(s[sw(i)], Gr[i]);
    __syncthreads();
    lds_fft(s, N1, log2N1, RT, true, tws);
    for (int i = threadIdx.x; i < NPTS; i += NT)
        dst[off + i] = cmul(s[sw(i)], conjf2(it[(size_t)r0 * N1 + i]));
}

__global__ void __launch_bounds__(NT) k_tfft_cols_shift(float2* __restrict__ data, int N1, int N2,
                                                        int log2N2, int shift,
                                                        const float2* __restrict__ tw,
                                                        const float2* __restrict__ it) {
    __shared__ float2 s[NPTS];
    __shared__ float2 twl[TWL];
    const int CT = NPTS / N2;
    const int c0 = blockIdx.x * CT;
    float2* g = data + (size_t)blockIdx.y * N1 * N2;
    const float2* tws = stage_ladder(twl, tw, N2);
    load_cols(s, g, N1, N2, c0);
    __syncthreads();
    lds_fft(s, N2, log2N2, CT, true, tws);
    // z row r = y row r + shift for r < shift (shift = nr / N1 <= N2 / 2), zero above
    float2 v[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + q * NT;
        const int c = idx / N2, r = idx - c * N2;
        v[q] = (r < shift) ? s[sw(c * N2 + r + shift)] : make_float2(0.f, 0.f);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PPT; ++q) s[sw(threadIdx.x + q * NT)] = v[q];
    __syncthreads();
    lds_fft(s, N2, log2N2, CT, false, tws);
    for (int idx = threadIdx.x; idx < NPTS; idx += NT) {
        const int r = idx / CT, c = idx - r * CT;
        g[(size_t)r * N1 + c0 + c] = cmul(s[sw(c * N2 + r)], it[(size_t)r * N1 + c0 + c]);
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
    __shared__ float2 twl[TWL];
    const int RT = NPTS / N1;
    const int r0 = blockIdx.x * RT;
    const int grp = blockIdx.y;
    const size_t P = (size_t)N1 * N2;
    const float2* tws = stage_ladder(twl, tw, N1);
    float2 acc[PPT], xr[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) acc[q] = make_float2(0.f, 0.f);
    for (int p = gbeg[grp]; p < gend[grp]; ++p) {
        const size_t off = (size_t)p * P + (size_t)r0 * N1;
        for (int i = threadIdx.x; i < NPTS; i += NT) s[sw(i)] = X[off + i];
        __syncthreads();
        lds_fft(s, N1, log2N1, RT, false, tws);
#pragma unroll
        for (int q = 0; q < PPT; ++q) xr[q] = s[sw(threadIdx.x + q * NT)];
        __syncthreads();
        for (int i = threadIdx.x; i < NPTS; i += NT) s[sw(i)] = Z[off + i];
        __syncthreads();
        lds_fft(s, N1, log2N1, RT, false, tws);
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const float2 z = s[sw(threadIdx.x + q * NT)];
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
                                                   const int* __restrict__ pbase,
                                                   const int* __restrict__ pwin, int nr,
                                                   const float2* __restrict__ G,
                                                   float2* __restrict__ prod,
                                                   const float2* __restrict__ tw) {
    __shared__ float2 s[NPTS];
    const int p = blockIdx.x;
    const int base = pbase[p], win = pwin[p];
    const float2* a = A + (size_t)p * P;
    for (int n = threadIdx.x; n < P; n += NT) {
        const bool live = (n < nr) && (base < 0 || (n >= base && n < base + win));
        s[sw(n)] = live ? a[n] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    lds_fft(s, P, log2P, 1, false, tw);
    float2 xr[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int n = threadIdx.x + q * NT;
        if (n < P) {
            xr[q] = s[sw(n)];
            s[sw(n)] = cmul(xr[q], G[n]);
        }
    }
    __syncthreads();
    lds_fft(s, P, log2P, 1, true, tw);
    float2 v[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int n = threadIdx.x + q * NT;
        if (n < P) v[q] = (n < nr) ? s[sw(n + nr)] : make_float2(0.f, 0.f);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int n = threadIdx.x + q * NT;
        if (n < P) s[sw(n)] = v[q];
    }
    __syncthreads();
    lds_fft(s, P, log2P, 1, false, tw);
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int n = threadIdx.x + q * NT;
        if (n < P) prod[(size_t)p * P + n] = cmul(xr[q], s[sw(n)]);
    }
}

// S[shot] += scale * sum of rows [beg, end) of src (one segment per shot).
// pN1 > 0: src rows are in natural order and S in the permuted order of a
// two-pass N = pN1 * pN2 transform (frequency k at (k mod pN2) pN1 + k / pN2).
__global__ void k_sum_rows(float2* __restrict__ S, const float2* __restrict__ src,
                           const int* __restrict__ seg_shot, const int* __restrict__ seg_beg,
                           const int* __restrict__ seg_end, int P, float scale, int pN1,
                           int pN2) {
    const int sg = blockIdx.y;
    const int sh = seg_shot[sg], b = seg_beg[sg], e = seg_end[sg];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < P; k += gridDim.x * blockDim.x) {
        float2 acc = make_float2(0.f, 0.f);
        for (int r = b; r < e; ++r) {
            const float2 v = src[(size_t)r * P + k];
            acc.x += v.x;
            acc.y += v.y;
        }
        const int kd = pN1 > 0 ? (k % pN2) * pN1 + k / pN2 : k;
        float2* d = S + (size_t)sh * P + kd;
        d->x += scale * acc.x;
        d->y += scale * acc.y;
    }
}

// ---------------------------------------------------------------------------
// Whole per-pair chain in one workgroup for P = 2 nr, nr = 16 NTB (the nr
// points fill 16 registers per thread and nr * 8 bytes of LDS).  With
// m(j) = exp(-i pi j / nr), a P-point transform of x supported on [0, nr)
// splits into two nr-point ones:
//     X[2k] = DFT_nr(x)[k],   X[2k+1] = DFT_nr(x m)[k];
// and the upper half of an (unnormalised) inverse P-point transform is
//     y[j + nr] = IDFT_nr(Y_even)[j] - conj(m(j)) IDFT_nr(Y_odd)[j].
// Both the inward trace a and the cropped z = y[nr + j] live on [0, nr).
// With E = DFT(a), O = DFT(a m), u = IDFT(E G_e), v = IDFT(O G_o) and
// z = u - conj(m) v, and since DFT(IDFT(x)) = nr x and |m| = 1:
//     Ze = DFT(z)   = nr E G_e - T1,   T1 = DFT(conj(m) v)
//     Zo = DFT(z m) = T2 - nr O G_o,   T2 = DFT(m u)
//     out[2k] = E Ze,  out[2k+1] = O Zo.
// Six nr-point transforms per pair: E is parked in the even half of the
// pair's output row (which only the last step writes) instead of being
// transformed again, and at most one register array is held across a
// transform.
// Gn holds the chirp spectrum in natural order as (G[2k], G[2k+1]) pairs.
// Each output row (P complex) holds two planes: out[k] = product at
// frequency 2k, out[n + k] = product at frequency 2k + 1.
// ---------------------------------------------------------------------------

// LDS index with one float2 of padding per 16: for every access pattern of
// fft_regs the padded index is a per-thread base plus a compile-time
// constant, so the addresses cost no per-access arithmetic, and each group
// of 16 lanes still touches 16 distinct 8-byte bank pairs.
__device__ __forceinline__ int pad16(int a) { return a + (a >> 4); }

// Forward nr-point transform held in registers: x[q] = element tid + q NTB
// in and out (natural order); the inverse is conj(fft(conj(x))).  The first
// stage reads its inputs from registers and the last stage leaves its
// outputs there, so only the middle stages go through LDS.  Starts with a
// barrier (earlier readers of s are done).  s holds n + n/16 float2.
template <int NTB>
__device__ __forceinline__ void fft_regs(float2* s, float2 (&x)[PPT],
                                         const float2* __restrict__ lad) {
    constexpr int n = NTB * PPT;
    constexpr int L = NTB == 64 ? 10 : NTB == 128 ? 11 : NTB == 256 ? 12 : NTB == 512 ? 13 : 14;
    static_assert(NTB >= 64 && NTB <= 1024 && (NTB & (NTB - 1)) == 0, "NTB");
    constexpr int r0 = L & 3;
    constexpr int R0 = r0 == 0 ? 16 : (1 << r0);
    constexpr int Q0 = PPT / R0;
    constexpr int MID = (L - (r0 == 0 ? 4 : r0)) / 4 - 1;
    int tid = opaque_tid();
    __syncthreads();
    // first stage, Ls = 1 (no twiddles): input j + i n/R0 is x[q + Q0 i];
    // output j R0 + i
#pragma unroll
    for (int q = 0; q < Q0; ++q) {
        float2 v[R0];
#pragma unroll
        for (int i = 0; i < R0; ++i) v[i] = x[q + Q0 * i];
        dft_r<R0>(v, false);
        const int j = tid + q * NTB;
        float2* d = s + pad16(j * R0);
#pragma unroll
        for (int i = 0; i < R0; ++i) d[i] = v[i];  // j R0 + i stays in one 16-block
    }
    __syncthreads();
    // middle radix-16 stages: input tid + i n/16, output (tid - k) 16 + k + i Ls
    int Ls = R0;
#pragma unroll
    for (int st = 0; st < MID; ++st) {
        tid = opaque_tid();
        float2 v[16];
        const float2* src = s + pad16(tid);
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = src[i * (NTB + NTB / 16)];
        const int k = tid & (Ls - 1);
        apply_twiddles<16>(v, lad, Ls * 16, k, false);
        dft_r<16>(v, false);
        __syncthreads();
        const int dst = (tid - k) * 16 + k;
        if (Ls >= 16) {
            float2* d = s + pad16(dst);
#pragma unroll
            for (int i = 0; i < 16; ++i) d[i * (Ls + Ls / 16)] = v[i];
        } else {
#pragma unroll
            for (int i = 0; i < 16; ++i) s[pad16(dst + i * Ls)] = v[i];
        }
        __syncthreads();
        Ls *= 16;
    }
    // last stage, Ls = NTB: thread tid owns outputs tid + i NTB
    tid = opaque_tid();
    const float2* src = s + pad16(tid);
#pragma unroll
    for (int i = 0; i < PPT; ++i) x[i] = src[i * (NTB + NTB / 16)];
    apply_twiddles<16>(x, lad, n, tid, false);
    dft_r<16>(x, false);
}

// x[q] = a[i] (times mod[i]) for i = tid + q NTB inside the pair's window
// [base, base + win) (base < 0: the whole row), else 0; LIM: also 0 from
// i = lim on (the row's bins end there)
template <int NTB, bool LIM = false>
__device__ __forceinline__ void load_window(float2 (&x)[PPT], const float2* __restrict__ a,
                                            int base, int win, const float2* __restrict__ mod,
                                            int lim = 0) {
    const int tid = opaque_tid();
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int i = tid + q * NTB;
        bool live = base < 0 || (i >= base && i < base + win);
        if (LIM) live = live && i < lim;
        float2 v = live ? a[i] : make_float2(0.f, 0.f);
        if (mod) v = cmul(v, mod[i]);
        x[q] = v;
    }
}

// CROP: nr < n bins (P = 2n, n the next power of two): the kept window
// z[j] = y[n + j] holds only j < nr, so z is no longer u - conj(m) v over
// the whole half and the shortcut DFT(IDFT(x)) = n x does not apply; z is
// formed in the time domain and transformed twice (still six transforms):
//   0: a          -> E        -> even plane = E; x = conj(E G_e)
//   1: x          -> conj(u)  -> h = u
//   2: a m        -> O        -> odd plane = O; x = conj(O G_o)
//   3: x          -> conj(v)  -> z = u - conj(m) v for j < nr, 0 above;
//                                x = z, h = z m
//   4: x          -> Z_e      -> even plane = E Z_e (E read back)
//   5: h          -> Z_o      -> odd plane = O Z_o (O read back)
// The range kernel spectrum G is placed for the window at n (kind 2 of
// k_gen_kernel).  Every thread reads back only what it wrote.
template <int NTB, bool MT = false, bool CROP = false>
__global__ void __launch_bounds__(NTB) k_tpair_half(const float2* __restrict__ A, int P,
                                                    const int* __restrict__ pbase,
                                                    const int* __restrict__ pwin,
                                                    const float4* __restrict__ Gn,
                                                    float2* __restrict__ out,
                                                    const float2* __restrict__ lad,
                                                    const float4* const* __restrict__ Gpair,
                                                    const int* __restrict__ npairs, int nr) {
    constexpr int n = NTB * PPT;
    __shared__ float2 s[n + n / 16];
    const int p = blockIdx.x;
    // a multi-track launch: a grid of the pair bound, the count on the
    // device, and each pair's own range kernel spectrum
    if (MT) {
        if (p >= *npairs) return;
        Gn = Gpair[p];
    }
    const int base = pbase[p], win = pwin[p];
    const float2* a = A + (size_t)p * P;
    const float2* mod = lad + 2 * n;  // m(j) = lad[2n + j]
    float2* o2 = out + (size_t)p * P;  // even plane [0, n), odd plane [n, 2n)
    const float fn = (float)n;
    float2 x[PPT], h[PPT];
    if constexpr (CROP) {
#pragma unroll 1
        for (int ph = 0; ph < 6; ++ph) {
            if (ph == 0 || ph == 2)
                load_window<NTB, true>(x, a, base, win, ph == 2 ? mod : nullptr, nr);
            if (ph == 5) {
#pragma unroll
                for (int q = 0; q < PPT; ++q) x[q] = h[q];
            }
            fft_regs<NTB>(s, x, lad);
            const int tid = opaque_tid();
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                const int i = tid + q * NTB;
                if (ph == 0) {
                    const float4 g = Gn[i];
                    o2[i] = x[q];
                    x[q] = conjf2(cmul(x[q], make_float2(g.x, g.y)));
                } else if (ph == 1) {
                    h[q] = conjf2(x[q]);
                } else if (ph == 2) {
                    const float4 g = Gn[i];
                    o2[n + i] = x[q];
                    x[q] = conjf2(cmul(x[q], make_float2(g.z, g.w)));
                } else if (ph == 3) {
                    // v = conj(x); z = u - conj(m) v = u - conj(m x)
                    const float2 mx = cmul(mod[i], x[q]);
                    const float2 z = i < nr ? make_float2(h[q].x - mx.x, h[q].y + mx.y)
                                            : make_float2(0.f, 0.f);
                    x[q] = z;
                    h[q] = cmul(z, mod[i]);
                } else if (ph == 4) {
                    o2[i] = cmul(o2[i], x[q]);
                } else {
                    o2[n + i] = cmul(o2[n + i], x[q]);
                }
            }
        }
        return;
    }
    // Phase: input -> forward transform -> use; an inverse transform is
    // fft(conj(.)) with the conjugates folded into the neighbouring steps
    //   0: a          -> E        -> even plane = E (kept there); x = conj(E G_e)
    //   1: x          -> conj(u)  -> x = m u
    //   2: x          -> T2       -> h = T2
    //   3: a m        -> O        -> odd plane = O (T2 - n O G_o); x = conj(O G_o)
    //   4: x          -> conj(v)  -> x = conj(m) v
    //   5: x          -> T1       -> h = T1
    //   6: E read back from the even plane -> even plane = E (n E G_e - T1)
    // Each thread reads back only the elements it wrote, so phase 6 needs no
    // barrier and saves the seventh transform.
#pragma unroll 1
    for (int ph = 0; ph < 7; ++ph) {
        if (ph == 0 || ph == 3) load_window<NTB>(x, a, base, win, ph == 3 ? mod : nullptr);
        if (ph != 6) fft_regs<NTB>(s, x, lad);
        const int tid = opaque_tid();
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int i = tid + q * NTB;
            if (ph == 0) {
                const float4 g = Gn[i];
                o2[i] = x[q];
                x[q] = conjf2(cmul(x[q], make_float2(g.x, g.y)));
            } else if (ph == 1) {
                x[q] = cmul(conjf2(x[q]), mod[i]);
            } else if (ph == 2 || ph == 5) {
                h[q] = x[q];
            } else if (ph == 3) {
                const float4 g = Gn[i];
                const float2 og = cmul(x[q], make_float2(g.z, g.w));
                o2[n + i] = cmul(x[q], make_float2(h[q].x - fn * og.x, h[q].y - fn * og.y));
                x[q] = conjf2(og);
            } else if (ph == 4) {
                x[q] = conjf2(cmul(x[q], mod[i]));
            } else {
                const float4 g = Gn[i];
                const float2 e = o2[i];
                const float2 eg = cmul(e, make_float2(g.x, g.y));
                o2[i] = cmul(e, make_float2(fn * eg.x - h[q].x, fn * eg.y - h[q].y));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Wave-local transforms of the per-pair chain: n = 1024 W points held by W
// waves of 64 lanes, 16 registers each.  With t = 64 w + l the thread and
// its register r holding element t + 64 W r (the natural layout), a
// transform is 16 (registers) x W (waves) x 64 (lanes):
//   F, natural -> permuted: DFT16 over r, twiddle W_n^(t k1); one exchange
//     across waves (element (t, k1) at k1 64 W + t; wave w then holds the 16
//     / W values of k1 in [16 w / W, 16 (w + 1) / W), each over every wave
//     index) and DFT_W over the wave index, twiddle W_(64 W)^(l kw); then
//     two exchanges inside the wave's own 1024 slots: DFT16 over l_hi
//     (l = l_lo + 4 l_hi), twiddle W_64^(l_lo ka), and four DFT4 over l_lo.
//     Thread (w, l = g + 4 c) register h + 4 kb then holds element wl_perm.
//   G, permuted -> natural: the same steps transposed in reverse order.
// Both compute the forward DFT (the DFT matrix is symmetric).  Only F's
// writes cross waves (each wave then reads its own 1024 slots) and only G's
// reads do (each wave has written its own slots).  So G takes one
// workgroup barrier (between its writes and reads) and F one (between its
// writes and reads) plus, after another F, one before its writes (whose
// slots that F's readers may still hold), against six for fft_regs; every
// exchange inside a wave needs only a wave-level fence, as LDS operations
// of one wave complete in order.  G always follows an F in the chain.  A
// row position p = t + 64 W r holds element p in the natural layout and
// element wl_perm(p) in the permuted one, so tables read in the permuted
// layout are permuted copies (k_wl_perm_table).
// ---------------------------------------------------------------------------
constexpr int clog2(int v) { return v <= 1 ? 0 : 1 + clog2(v >> 1); }

__host__ __device__ __forceinline__ int wl_perm(int p, int lw) {
    // lw = log2 W
    const int rho = p >> (6 + lw), t = p & ((64 << lw) - 1), w = t >> 6, mu = t & 63;
    const int g = mu & 3, c = mu >> 2, h = rho & 3, kb = rho >> 2;
    const int k1 = (w << (4 - lw)) + (c >> lw), kw = c & ((1 << lw) - 1);
    return k1 + 16 * (kw + (1 << lw) * (4 * g + h + 16 * kb));
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// padded slot of element e: pad16(e); for the constant strides below
// pad16(a + 64 j) = pad16(a) + 68 j and pad16(a + 64 W k) = pad16(a) + 68 W k.
// after_f: the previous transform was an F, whose last reads (each wave in
// its own slots) the writes below may overtake.  After a G no barrier is
// needed: these writes go to the very slots this thread read last (G's
// final read, element k1 64 W + t), which no other thread reads or writes.
template <int W>
__device__ __forceinline__ void fft_wl_F(float2* s, float2 (&x)[16],
                                         const float2* __restrict__ lad, bool after_f) {
    constexpr int NTB = 64 * W, n = 1024 * W;
    int tid = opaque_tid();
    dft_r<16>(x, false);
    apply_twiddles<16>(x, lad, n, tid, false);
    if (after_f) __syncthreads();
    {
        float2* d = s + pad16(tid);
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) d[k1 * (NTB + NTB / 16)] = x[k1];
    }
    __syncthreads();
    tid = opaque_tid();
    const int w = tid >> 6, l = tid & 63;
    {
        // this wave's own slots: register j = c1 W + (wave index)
        const float2* src = s + pad16(w * 1024 + l);
#pragma unroll
        for (int j = 0; j < 16; ++j) x[j] = src[j * 68];
    }
    // (W = 16 on the whole array: the grouped form costs it registers)
    if constexpr (W == 16) {
        dft_r<16>(x, false);
        apply_twiddles<16>(x, lad, NTB, l, false);
    } else if constexpr (W > 1) {
#pragma unroll
        for (int c1 = 0; c1 < 16 / W; ++c1) {
            dft_r<W>(x + c1 * W, false);
            apply_twiddles<W>(x + c1 * W, lad, NTB, l, false);
        }
    }
    // the reads above were of this wave's own slots (elements 1024 w + l +
    // 64 j), which no other wave touches until the next transform's first
    // barrier: a wave-level fence orders them before the writes below
    wave_sync();
    float2* R = s + w * 1088;  // this wave's slots: elements [1024 w, 1024 w + 1024)
    {
        float2* d = R + pad16(l);
#pragma unroll
        for (int c = 0; c < 16; ++c) d[c * 68] = x[c];
    }
    wave_sync();
    const int lo = l & 3, kq = l >> 2;  // l_lo and the combination c of this lane's next steps
    {
        // element kq 64 + lo + 4 j at kq 68 + lo + 4 j + j / 4
        const float2* src = R + kq * 68 + lo;
#pragma unroll
        for (int j = 0; j < 16; ++j) x[j] = src[4 * j + (j >> 2)];
    }
    dft_r<16>(x, false);
    apply_twiddles<16>(x, lad, 64, lo, false);
    wave_sync();
    {
        float2* d = R + kq * 68 + lo;
#pragma unroll
        for (int ka = 0; ka < 16; ++ka) d[4 * ka + (ka >> 2)] = x[ka];
    }
    wave_sync();
    {
        // element kq 64 + 16 g + 4 h + m at kq 68 + 17 g + 4 h + m (g = lo)
        const float2* src = R + kq * 68 + 17 * lo;
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            float2 y0 = src[4 * h], y1 = src[4 * h + 1], y2 = src[4 * h + 2], y3 = src[4 * h + 3];
            dft4(y0, y1, y2, y3, false);
            x[h] = y0;
            x[h + 4] = y1;
            x[h + 8] = y2;
            x[h + 12] = y3;
        }
    }
}

template <int W>
__device__ __forceinline__ void fft_wl_G(float2* s, float2 (&x)[16],
                                         const float2* __restrict__ lad) {
    constexpr int NTB = 64 * W, n = 1024 * W;
    int tid = opaque_tid();
    const int w = tid >> 6, l = tid & 63;
    const int lo = l & 3, kq = l >> 2;
    float2* R = s + w * 1088;
    {
        float2* d = R + kq * 68 + 17 * lo;
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            float2 y0 = x[h], y1 = x[h + 4], y2 = x[h + 8], y3 = x[h + 12];
            dft4(y0, y1, y2, y3, false);
            d[4 * h] = y0;
            d[4 * h + 1] = y1;
            d[4 * h + 2] = y2;
            d[4 * h + 3] = y3;
        }
    }
    wave_sync();
    {
        const float2* src = R + kq * 68 + lo;
#pragma unroll
        for (int ka = 0; ka < 16; ++ka) x[ka] = src[4 * ka + (ka >> 2)];
    }
    apply_twiddles<16>(x, lad, 64, lo, false);
    dft_r<16>(x, false);
    wave_sync();
    {
        float2* d = R + kq * 68 + lo;
#pragma unroll
        for (int j = 0; j < 16; ++j) d[4 * j + (j >> 2)] = x[j];
    }
    wave_sync();
    {
        const float2* src = R + pad16(l);
#pragma unroll
        for (int c = 0; c < 16; ++c) x[c] = src[c * 68];
    }
    if constexpr (W == 16) {
        apply_twiddles<16>(x, lad, NTB, l, false);
        dft_r<16>(x, false);
    } else if constexpr (W > 1) {
#pragma unroll
        for (int c1 = 0; c1 < 16 / W; ++c1) {
            apply_twiddles<W>(x + c1 * W, lad, NTB, l, false);
            dft_r<W>(x + c1 * W, false);
        }
    }
    // these writes go to this wave's own slots too (elements 1024 w + l +
    // 64 j); no other wave has read them since the last F's barrier
    wave_sync();
    {
        float2* d = s + pad16(w * 1024 + l);
#pragma unroll
        for (int j = 0; j < 16; ++j) d[j * 68] = x[j];
    }
    __syncthreads();
    tid = opaque_tid();
    {
        const float2* src = s + pad16(tid);
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) x[k1] = src[k1 * (NTB + NTB / 16)];
    }
    apply_twiddles<16>(x, lad, n, tid, false);
    dft_r<16>(x, false);
}

// Gp[p] = Gn[wl_perm(p)], p < n = 1024 W (the (G[2k], G[2k+1]) pairs of a
// range kernel spectrum in the permuted layout)
__global__ void k_wl_perm_table(float4* __restrict__ Gp, const float4* __restrict__ Gn, int n,
                                int lw) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < n) Gp[p] = Gn[wl_perm(p, lw)];
}

// k_tpair_half with the wave-local transforms (n = 1024 W): the same
// phases, with F where a transform takes natural input and G where it takes
// permuted input (plain: F G F F G F; cropped: F G F G F F).  Gn is the
// permuted spectrum (k_wl_perm_table), the output planes are written in the
// permuted layout, and the inputs, m(j) and the cropped z stay natural.
template <int W, bool MT = false, bool CROP = false>
__global__ void __launch_bounds__(64 * W) k_tpair_wl(const float2* __restrict__ A, int P,
                                                   const int* __restrict__ pbase,
                                                   const int* __restrict__ pwin,
                                                   const float4* __restrict__ Gn,
                                                   float2* __restrict__ out,
                                                   const float2* __restrict__ lad,
                                                   const float4* const* __restrict__ Gpair,
                                                   const int* __restrict__ npairs, int nr) {
    constexpr int NTB = 64 * W, n = 1024 * W;
    __shared__ float2 s[n + n / 16];
    const int p = blockIdx.x;
    if (MT) {
        if (p >= *npairs) return;
        Gn = Gpair[p];
    }
    const int base = pbase[p], win = pwin[p];
    const float2* a = A + (size_t)p * P;
    const float2* mod = lad + 2 * n;
    float2* o2 = out + (size_t)p * P;
    const float fn = (float)n;
    float2 x[PPT], h[PPT];
    if constexpr (CROP) {
#pragma unroll 1
        for (int ph = 0; ph < 6; ++ph) {
            if (ph == 0 || ph == 2)
                load_window<NTB, true>(x, a, base, win, ph == 2 ? mod : nullptr, nr);
            if (ph == 5) {
#pragma unroll
                for (int q = 0; q < PPT; ++q) x[q] = h[q];
            }
            if (ph == 1 || ph == 3) fft_wl_G<W>(s, x, lad);
            else fft_wl_F<W>(s, x, lad, ph == 5);
            const int tid = opaque_tid();
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                const int i = tid + q * NTB;
                if (ph == 0) {
                    const float4 g = Gn[i];
                    o2[i] = x[q];
                    x[q] = conjf2(cmul(x[q], make_float2(g.x, g.y)));
                } else if (ph == 1) {
                    h[q] = conjf2(x[q]);
                } else if (ph == 2) {
                    const float4 g = Gn[i];
                    o2[n + i] = x[q];
                    x[q] = conjf2(cmul(x[q], make_float2(g.z, g.w)));
                } else if (ph == 3) {
                    const float2 mx = cmul(mod[i], x[q]);
                    const float2 z = i < nr ? make_float2(h[q].x - mx.x, h[q].y + mx.y)
                                            : make_float2(0.f, 0.f);
                    x[q] = z;
                    h[q] = cmul(z, mod[i]);
                } else if (ph == 4) {
                    o2[i] = cmul(o2[i], x[q]);
                } else {
                    o2[n + i] = cmul(o2[n + i], x[q]);
                }
            }
        }
        return;
    }
#pragma unroll 1
    for (int ph = 0; ph < 7; ++ph) {
        if (ph == 0 || ph == 3) load_window<NTB>(x, a, base, win, ph == 3 ? mod : nullptr);
        if (ph == 1 || ph == 4) fft_wl_G<W>(s, x, lad);
        else if (ph != 6) fft_wl_F<W>(s, x, lad, ph == 3);
        const int tid = opaque_tid();
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int i = tid + q * NTB;
            if (ph == 0) {
                const float4 g = Gn[i];
                o2[i] = x[q];
                x[q] = conjf2(cmul(x[q], make_float2(g.x, g.y)));
            } else if (ph == 1) {
                x[q] = cmul(conjf2(x[q]), mod[i]);
            } else if (ph == 2 || ph == 5) {
                h[q] = x[q];
            } else if (ph == 3) {
                const float4 g = Gn[i];
                const float2 og = cmul(x[q], make_float2(g.z, g.w));
                o2[n + i] = cmul(x[q], make_float2(h[q].x - fn * og.x, h[q].y - fn * og.y));
                x[q] = conjf2(og);
            } else if (ph == 4) {
                x[q] = conjf2(cmul(x[q], mod[i]));
            } else {
                const float4 g = Gn[i];
                const float2 e = o2[i];
                const float2 eg = cmul(e, make_float2(g.x, g.y));
                o2[i] = cmul(e, make_float2(fn * eg.x - h[q].x, fn * eg.y - h[q].y));
            }
        }
    }
}

// Two-level deterministic sum of per-pair rows (P complex each, as the two
// planes written by k_tpair_half: [0, P/2) even, [P/2, P) odd frequencies).
// Level 1: partial[slice] = sum of the rows of one slice (at most SLICE rows
// of one shot) in natural order, two frequencies per thread, eight rows in
// flight.
constexpr int SLICE = 32;
__global__ void __launch_bounds__(256) k_sum_slices(float4* __restrict__ partial,
                                                    const float2* __restrict__ rows,
                                                    const int* __restrict__ sl_beg,
                                                    const int* __restrict__ sl_end, int P2) {
    const int k = blockIdx.x * 256 + threadIdx.x;  // frequency pair (2k, 2k+1), P2 = P / 2
    if (k >= P2) return;
    const int b = sl_beg[blockIdx.y], e = sl_end[blockIdx.y];
    if (b == e) return;  // padding slice of a replayed launch: never read
    const size_t P = 2 * (size_t)P2;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    // each row is read once: non-temporal loads (7 % faster on the 2048^2
    // fBm workload's 2.9 GB of rows per launch)
    typedef float v2f __attribute__((ext_vector_type(2)));
    auto ld = [&](size_t i) {
        const v2f v = __builtin_nontemporal_load(reinterpret_cast<const v2f*>(rows + i));
        return make_float2(v.x, v.y);
    };
    int r = b;
    for (; r + 8 <= e; r += 8) {
        float2 ve[8], vo[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            ve[u] = ld((size_t)(r + u) * P + k);
            vo[u] = ld((size_t)(r + u) * P + P2 + k);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            acc.x += ve[u].x; acc.y += ve[u].y; acc.z += vo[u].x; acc.w += vo[u].y;
        }
    }
    for (; r < e; ++r) {
        const float2 ve = ld((size_t)r * P + k), vo = ld((size_t)r * P + P2 + k);
        acc.x += ve.x; acc.y += ve.y; acc.z += vo.x; acc.w += vo.y;
    }
    partial[(size_t)blockIdx.y * P2 + k] = acc;
}

// Level 2: S[shot] += scale * sum of the slices [beg, end) of one shot,
// written in the permuted order of a two-pass transform when pN1 > 0.
// seq: each slice is added to S on its own, in order (S += scale * slice),
// so that S depends only on the slices and not on how launches group them.
__global__ void __launch_bounds__(256) k_sum_final(float2* __restrict__ S,
                                                   const float2* __restrict__ partial,
                                                   const int* __restrict__ seg_shot,
                                                   const int* __restrict__ seg_beg,
                                                   const int* __restrict__ seg_end, int P,
                                                   float scale, int pN1, int pN2,
                                                   const float* __restrict__ seg_scale, int seq,
                                                   int wl) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= P) return;
    const int sg = blockIdx.y;
    if (seg_beg[sg] == seg_end[sg]) return;  // padding segment
    if (seg_scale) scale = seg_scale[sg];   // segments of several tracks
    // wl = 1 + log2 W: rows in the permuted layout of k_tpair_wl (planes of
    // n = P / 2: position k / 2 of plane k % 2 holds frequency
    // 2 wl_perm(k / 2) + k % 2)
    const int f = wl ? 2 * wl_perm(k >> 1, wl - 1) + (k & 1) : k;
    const int kd = pN1 > 0 ? (f % pN2) * pN1 + f / pN2 : f;
    float2* d = S + (size_t)seg_shot[sg] * P + kd;
    if (seq) {
        float2 a = *d;
        for (int r = seg_beg[sg]; r < seg_end[sg]; ++r) {
            const float2 v = partial[(size_t)r * P + k];
            const float2 t = make_float2(scale * v.x, scale * v.y);
            a.x += t.x;
            a.y += t.y;
        }
        *d = a;
        return;
    }
    float2 acc = make_float2(0.f, 0.f);
    for (int r = seg_beg[sg]; r < seg_end[sg]; ++r) {
        const float2 v = partial[(size_t)r * P + k];
        acc.x += v.x;
        acc.y += v.y;
    }
    d->x += scale * acc.x;
    d->y += scale * acc.y;
}

// natural-order copy of a two-pass spectrum: dst[k] = src[(k mod N2) N1 + k / N2]
__global__ void k_unpermute(float2* __restrict__ dst, const float2* __restrict__ src, int N1,
                            int N2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < N1 * N2) dst[k] = src[(k % N2) * N1 + k / N2];
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

static const float2* inter_table(erp_ctx* ctx, const FftPlan& p, hipStream_t s);

// The ladder table and the inter-pass tables of every two-pass size a run
// can use (up to the 2^17-point Bluestein transforms of nr = 32767): built
// once per context, so that no run allocates a table (1.9 MB in all).
void fft_init_twiddles(erp_ctx* ctx) {
    ctx->twiddle.ensure(sizeof(float2) * (size_t)LADN);
    hipLaunchKernelGGL(k_twiddle_init, dim3(LADN / 256), dim3(256), 0, ctx->stream,
                       ctx->twiddle.as<float2>());
    ERP_CHECK_LAUNCH("k_twiddle_init");
    for (int n = 2 * NPTS; n <= (1 << 17); n *= 2) inter_table(ctx, fft_plan(n), ctx->stream);
}

// inter-pass twiddle table of a two-pass plan (cached per size in the context)
static const float2* inter_table(erp_ctx* ctx, const FftPlan& p, hipStream_t s) {
    std::lock_guard<std::recursive_mutex> lk(ctx->cache_mu);
    for (auto& t : ctx->inter)
        if (t.first == p.N) return t.second->as<float2>();
    DevBuf* b = new DevBuf();
    b->ensure(sizeof(float2) * (size_t)p.N);
    hipLaunchKernelGGL(k_inter_init, dim3((p.N + 255) / 256), dim3(256), 0, s, b->as<float2>(),
                       p.N1, p.N);
    ERP_CHECK_LAUNCH("k_inter_init");
    ERP_HIP_TRY(hipStreamSynchronize(s));  // usable from other streams once published
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
    const float2* it = inter_table(ctx, p, s);
    dim3 gcol(p.N1 / CT, batch), grow(p.N2 / RT, batch);
    if (!inv) {
        hipLaunchKernelGGL(k_fft_cols, gcol, dim3(NT), 0, s, data, p.N1, p.N2, ilog2(p.N2), 0,
                           tw, it);
        ERP_CHECK_LAUNCH("k_fft_cols");
        hipLaunchKernelGGL(k_fft_rows, grow, dim3(NT), 0, s, data, p.N1, ilog2(p.N1), p.N2, 0,
                           tw, it);
        ERP_CHECK_LAUNCH("k_fft_rows");
    } else {
        hipLaunchKernelGGL(k_fft_rows, grow, dim3(NT), 0, s, data, p.N1, ilog2(p.N1), p.N2, 1,
                           tw, it);
        ERP_CHECK_LAUNCH("k_fft_rows");
        hipLaunchKernelGGL(k_fft_cols, gcol, dim3(NT), 0, s, data, p.N1, p.N2, ilog2(p.N2), 1,
                           tw, it);
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
// Y hold n rows of P; the work buffer holds max(ngroups, n) rows.
void target_spectra_fused(erp_ctx* ctx, const FftPlan& p, float2* A, float2* Y,
                          const int* pbase, const int* pwin, int nr, int n, const float2* G,
                          float2* work, int ngroups, const int* gbeg, const int* gend,
                          const int* gseg_shot, const int* gseg_beg, const int* gseg_end,
                          int nseg, float2* S, float scale, hipStream_t st) {
    const float2* tw = ctx->twiddle.as<float2>();
    const int P = p.N;
    if (!p.two_pass) {
        hipLaunchKernelGGL(k_tfft_small, dim3(n), dim3(NT), 0, st, A, P, p.log2N, pbase, pwin, nr,
                           G, work, tw);
        ERP_CHECK_LAUNCH("k_tfft_small");
        // groups are unused here: sum the per-pair products per shot segment
        hipLaunchKernelGGL(k_sum_rows, dim3(std::max(1, P / 256), nseg), dim3(256), 0, st, S,
                           work, gseg_shot, gseg_beg, gseg_end, P, scale, 0, 0);
        ERP_CHECK_LAUNCH("k_sum_rows");
        return;
    }
    const int CT = NPTS / p.N2, RT = NPTS / p.N1;
    const int l1 = ilog2(p.N1), l2 = ilog2(p.N2);
    if (nr % p.N1 != 0) throw ErpError{ERP_E_INVALID, "fused target chain needs nr % N1 == 0"};
    const float2* it = inter_table(ctx, p, st);
    hipLaunchKernelGGL(k_tfft_cols_fwd, dim3(p.N1 / CT, n), dim3(NT), 0, st, A, p.N1, p.N2, l2,
                       pbase, pwin, nr, tw, it);
    ERP_CHECK_LAUNCH("k_tfft_cols_fwd");
    hipLaunchKernelGGL(k_tfft_rows_conv, dim3(p.N2 / RT, n), dim3(NT), 0, st, A, Y, p.N1, l1,
                       p.N2, G, tw, it);
    ERP_CHECK_LAUNCH("k_tfft_rows_conv");
    hipLaunchKernelGGL(k_tfft_cols_shift, dim3(p.N1 / CT, n), dim3(NT), 0, st, Y, p.N1, p.N2, l2,
                       nr / p.N1, tw, it);
    ERP_CHECK_LAUNCH("k_tfft_cols_shift");
    hipLaunchKernelGGL(k_tfft_rows_mac, dim3(p.N2 / RT, ngroups), dim3(NT), 0, st, A, Y, p.N1, l1,
                       p.N2, gbeg, gend, work, tw);
    ERP_CHECK_LAUNCH("k_tfft_rows_mac");
    hipLaunchKernelGGL(k_sum_rows, dim3(std::max(1, P / 256), nseg), dim3(256), 0, st, S, work,
                       gseg_shot, gseg_beg, gseg_end, P, scale, 0, 0);
    ERP_CHECK_LAUNCH("k_sum_rows");
}

int target_half_slice() { return SLICE; }

// the wave-local chain from n = 8192 (one or two workgroups per CU, whose
// barriers idle it); below, several workgroups share a CU and hide the
// barriers of k_tpair_half, which is then 2-3 % faster (fBm geometry,
// 50 shots: n = 8192 13.4 against 14.7 ms; 4096 6.58 against 6.39)
bool target_wl(const FftPlan& p) { return p.N / 2 >= 8192 && p.N / 2 <= 16384; }

void wl_perm_table(const float2* Gnat, float2* Gwl, int P, hipStream_t st) {
    const int n = P / 2;
    hipLaunchKernelGGL(k_wl_perm_table, dim3((n + 255) / 256), dim3(256), 0, st,
                       reinterpret_cast<float4*>(Gwl), reinterpret_cast<const float4*>(Gnat), n,
                       clog2(n / 1024));
    ERP_CHECK_LAUNCH("k_wl_perm_table");
}

// P = 2n with n the next power of two of nr, from 1024 to 16384 (nr below
// n runs the cropped variant)
bool target_half_supported(const FftPlan& p, int nr) {
    const int n = p.N / 2;
    return p.N == 2 * next_pow2(nr) && n >= 1024 && n <= 16384;
}

void fft_natural_order(const FftPlan& p, const float2* src, float2* dst, hipStream_t st) {
    if (!p.two_pass) {
        ERP_HIP_TRY(hipMemcpyAsync(dst, src, sizeof(float2) * (size_t)p.N,
                                   hipMemcpyDeviceToDevice, st));
        return;
    }
    hipLaunchKernelGGL(k_unpermute, dim3((p.N + 255) / 256), dim3(256), 0, st, dst, src, p.N1,
                       p.N2);
    ERP_CHECK_LAUNCH("k_unpermute");
}

// Host driver of the one-workgroup-per-pair chain (see k_tpair_half): n
// pairs of A (rows of P) -> per-pair products in work (even/odd planes).
void target_pairs_half(erp_ctx* ctx, const FftPlan& p, const float2* A, const int* pbase,
                       const int* pwin, int nr, int n, const float2* Gnat, float2* work,
                       hipStream_t st, const float4* const* Gpair, const int* npairs_dev) {
    if (!target_half_supported(p, nr))
        throw ErpError{ERP_E_INVALID,
                       "per-pair LDS chain needs P = 2 next_pow2(nr) in [2048, 32768]"};
    const float2* lad = ctx->twiddle.as<float2>();
    const float4* G4 = reinterpret_cast<const float4*>(Gnat);
    float2* o4 = work;
    // n: the pairs, or (multi-track, Gpair set) the bound of a grid whose
    // pairs past *npairs_dev exit at once
    auto go = [&](auto kern, int nt) {
        hipLaunchKernelGGL(kern, dim3(n), dim3(nt), 0, st, A, p.N, pbase, pwin, G4, o4, lad,
                           Gpair, npairs_dev, nr);
    };
    const bool mt = Gpair != nullptr;
    const bool crop = nr < p.N / 2;
    // the 2 x 2 variants (multi-track, cropped) of one transform size
    auto pick = [&](auto k00, auto k10, auto k01, auto k11, int nt) {
        if (crop) mt ? go(k11, nt) : go(k01, nt);
        else mt ? go(k10, nt) : go(k00, nt);
    };
#define ERP_TPAIR_VARIANTS(NTB)                                                       \
    pick(k_tpair_half<NTB, false, false>, k_tpair_half<NTB, true, false>,            \
         k_tpair_half<NTB, false, true>, k_tpair_half<NTB, true, true>, NTB)
    if (target_wl(p)) {
        // the wave-local transforms (Gnat in their layout)
#define ERP_TPAIR_WL(W)                                                                \
    pick(k_tpair_wl<W, false, false>, k_tpair_wl<W, true, false>,                     \
         k_tpair_wl<W, false, true>, k_tpair_wl<W, true, true>, 64 * W)
        switch (p.N / 2) {
            case 1024: ERP_TPAIR_WL(1); break;
            case 2048: ERP_TPAIR_WL(2); break;
            case 4096: ERP_TPAIR_WL(4); break;
            case 8192: ERP_TPAIR_WL(8); break;
            default: ERP_TPAIR_WL(16); break;
        }
#undef ERP_TPAIR_WL
        ERP_CHECK_LAUNCH("k_tpair_wl");
        return;
    }
    switch (p.N / 2) {
        case 1024: ERP_TPAIR_VARIANTS(64); break;
        case 2048: ERP_TPAIR_VARIANTS(128); break;
        case 4096: ERP_TPAIR_VARIANTS(256); break;
        case 8192: ERP_TPAIR_VARIANTS(512); break;
        default: ERP_TPAIR_VARIANTS(1024); break;
    }
#undef ERP_TPAIR_VARIANTS
    ERP_CHECK_LAUNCH("k_tpair_half");
}

// Per-shot sums of the per-pair rows of work: slice sums into partial
// (nslices rows; slice i = pairs [sl_beg[i], sl_end[i]) of one shot, at most
// target_half_slice() pairs), then the slices [seg_beg, seg_end) of each
// shot into S, scaled and in the order fft_inverse expects.
void target_sum_half(const FftPlan& p, const float2* work, float2* partial, const int* sl_beg,
                     const int* sl_end, int nslices, const int* seg_shot, const int* seg_beg,
                     const int* seg_end, int nseg, float2* S, float scale, hipStream_t st,
                     const float* seg_scale, bool seq, bool wl) {
    const int P2 = p.N / 2;
    hipLaunchKernelGGL(k_sum_slices, dim3((P2 + 255) / 256, nslices), dim3(256), 0, st,
                       reinterpret_cast<float4*>(partial), work, sl_beg, sl_end, P2);
    ERP_CHECK_LAUNCH("k_sum_slices");
    hipLaunchKernelGGL(k_sum_final, dim3((p.N + 255) / 256, nseg), dim3(256), 0, st, S, partial,
                       seg_shot, seg_beg, seg_end, p.N, scale, p.two_pass ? p.N1 : 0,
                       p.two_pass ? p.N2 : 0, seg_scale, seq ? 1 : 0,
                       wl ? 1 + clog2(p.N / 2048) : 0);
    ERP_CHECK_LAUNCH("k_sum_final");
}
