// erp_focus.hip — SAR focusing of a radargram (simple_focusing.py:310-404).
//
// The reference focuses pixel (i, j) with an FFT product per pixel: the
// aperture window of traces around j, each column rolled by its range-bin
// shift, the row i of that window zero-padded to L = 2 n_sx, multiplied in
// the frequency domain by the zero-padded matched filter, transformed back,
// and one sample taken: index (j - az_st) + (L - w) / 2.  One sample of a
// circular convolution is a dot product, so each pixel here is
//     sum_{m < w} rdr[row_m, az_st + m] * filt[(n - m) mod L],
// with filt[t] = 0 for t >= n_sx.  One thread per pixel; the threads of a
// workgroup cover consecutive traces of one range bin, so for each m they
// read consecutive columns of one radargram row (coalesced), and the row's
// shifts and filter are shared by the workgroup through LDS.
#include "erp_internal.h"

#include <algorithm>

namespace {

constexpr int FNT = 256;
constexpr int FCAP = 4096;  // aperture samples staged in LDS (larger: read from global)

__global__ void __launch_bounds__(FNT) k_focus(const float2* __restrict__ rdr, int Nr, int Na,
                                               int row0, const int* __restrict__ row_hs,
                                               const int* __restrict__ row_nsx,
                                               const int64_t* __restrict__ row_off,
                                               const int* __restrict__ rows,
                                               const float2* __restrict__ filt,
                                               float2* __restrict__ out) {
    // out: nrows x Na, row q is range bin row0 + q
    __shared__ int s_rows[FCAP];
    __shared__ float2 s_filt[FCAP];
    const int q = blockIdx.y;
    const int i = row0 + q;
    const int hs = row_hs[q], nsx = row_nsx[q];
    const int64_t off = row_off[q];
    const bool staged = nsx <= FCAP;
    if (staged) {
        for (int t = threadIdx.x; t < nsx; t += FNT) {
            s_rows[t] = rows[off + t];
            s_filt[t] = filt[off + t];
        }
    }
    __syncthreads();
    const int j = blockIdx.x * FNT + threadIdx.x;
    if (j >= Na || i >= Nr) return;
    const int L = 2 * nsx;
    const int az_st = max(0, j - hs), az_en = min(Na, j + hs);
    const int w = az_en - az_st;
    const int n = (j - az_st) + (L - w) / 2;
    double ax = 0.0, ay = 0.0;
    for (int m = 0; m < w; ++m) {
        int t = n - m;
        if (t < 0) t += L;
        if (t >= nsx) continue;
        const int r = staged ? s_rows[m] : rows[off + m];
        const float2 f = staged ? s_filt[t] : filt[off + t];
        const float2 v = rdr[(size_t)r * Na + az_st + m];
        ax += (double)v.x * f.x - (double)v.y * f.y;
        ay += (double)v.x * f.y + (double)v.y * f.x;
    }
    out[(size_t)q * Na + j] = make_float2((float)ax, (float)ay);  // rows of this call only
}

}  // namespace

void focus_rows(const float2* rdr, int Nr, int Na, int row0, int nrows, const int* row_hs,
                const int* row_nsx, const int64_t* row_off, const int* rows, const float2* filt,
                float2* out, hipStream_t st) {
    if (nrows <= 0 || Na <= 0) return;
    hipLaunchKernelGGL(k_focus, dim3((Na + FNT - 1) / FNT, nrows), dim3(FNT), 0, st, rdr, Nr, Na,
                       row0, row_hs, row_nsx, row_off, rows, filt, out);
    ERP_CHECK_LAUNCH("k_focus");
}
