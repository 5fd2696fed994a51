/*
 * erp_oracle.h — CPU restatement of the EuropaRaypaths per-shot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (europaraypaths_amd/)
 * links, loads or calls this library; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg do, as the reference they compare with.
 *
 * precision = 64: every step in IEEE double (the parity oracle).
 * precision = 32: the reference's FP32 arithmetic restated on the CPU (the
 *                 "reference CPU path" that bench.py times as cpu_baseline).
 */
#ifndef ERP_ORACLE_H
#define ERP_ORACLE_H

#include <stdint.h>
#include "../include/erp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Full shot loop (src/CUDA/sim.cu:615-1008).  out: ns*nr*2 doubles,
 * shot-major interleaved.  parts (optional): ns*3*nr*2 doubles holding, per
 * shot, the raw surface phasor trace before range compression, the reflected
 * signal and the refracted signal.  nvalid (optional): illuminated facets per
 * shot.  Returns 0 or a negative ERP_E_* code. */
int erp_oracle_run(const erp_params* par, const float* soa, int64_t F,
                   const float* tgt, const uint8_t* ttype, int64_t T,
                   const erp_shot* shots, int64_t ns,
                   const erp_prism* prisms, int64_t nprisms,
                   int precision, int nthreads,
                   double* out, double* parts, int32_t* nvalid);

/* Per-facet intermediates for one shot (geometry_kernels.cu:75-140,426-517,
 * radar_kernels.cu:714-764).  out: F*12 doubles per facet:
 * [illum, Itd, Ith, Iph, Rth, fRe, fRfrC, G_fct, Psrc, bin, w, phase]. */
int erp_oracle_facets(const erp_params* par, const float* soa, int64_t F,
                      const erp_shot* shot, int precision, double* out);

/* Per-(facet, target) intermediates (geometry_kernels.cu:272-350,520-584,
 * radar_kernels.cu:767-892) for one shot and one target index.
 * out: F*10 doubles: [illum, Ttd, Tth, TargetTh, Tph, EI, EO, rngt,
 * Pray_in, Pray_out]. */
int erp_oracle_target_pairs(const erp_params* par, const float* soa, int64_t F,
                            const float* tgt, const uint8_t* ttype, int64_t T,
                            int64_t it, const erp_shot* shot,
                            const erp_prism* prisms, int64_t nprisms,
                            int precision, double* out);

/* Generic-length complex DFT in double (radix-2 for powers of two,
 * Bluestein otherwise).  in/out: n interleaved complex.  Unnormalised in
 * both directions (cuFFT convention). */
int erp_oracle_dft(const double* in, int64_t n, int inverse, double* out);

/* Convolutions of the reference for pinning (kind 0: linear chirp
 * convolution, 1: circular, 2: convolveComplex of x with y). */
int erp_oracle_conv(int kind, const double* x, const double* y, int nr, double dr,
                    double res, double rst, double smpl, double Bw, int precision,
                    double* out);

/* Planted mistakes for the pin tests: 0 restores the oracle; any other
 * value swaps one detail for a plausible slip (see the MUT_* list in
 * erp_oracle.cpp).  Returns the previous value.  Not thread-safe: set it
 * only between runs. */
int erp_oracle_set_mutation(int m);

/* Version string of the oracle build. */
const char* erp_oracle_version(void);

#ifdef __cplusplus
}
#endif

#endif
