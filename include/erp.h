/*
 * erp.h — C ABI of the MI355X (gfx950) per-shot raypath / phasor / range-
 * compression path of the EuropaRaypaths radar-sounder simulator.
 *
 * The reference has no library boundary: its path is reached through the
 * process boundary `sim <params.json> <facets.fct> <targets.txt> <outdir>`
 * (src/CUDA/sim.cu:16, 92-106, 187-224) and writes one text trace per shot
 * (sim.cu:995-998, file_io.cu:284-296).  Each entry point below replaces one
 * part of that process so that a ctypes caller can drive the path in-process:
 *
 *   erp_ctx_create / erp_ctx_destroy  - device set-up and tear-down that
 *                                        sim.cu:272-607 performs inline.
 *   erp_mesh_upload                    - loadFacetFile + the 12 SoA cudaMalloc/
 *                                        cudaMemcpy calls (file_io.cu:337-362,
 *                                        sim.cu:228-311).
 *   erp_targets_upload                 - loadTargetFile + host target arrays
 *                                        (file_io.cu:365-426, sim.cu:185-210).
 *   erp_run / erp_run_dev              - the shot loop sim.cu:615-1008 and every
 *                                        kernel it launches, for ns shots.
 *   erp_submit / erp_poll / erp_wait   - asynchronous form of erp_run used by
 *   erp_serve_slots                      the serving loop (no reference
 *                                        counterpart: the reference is
 *                                        synchronous, sim.cu:992-998).
 *   erp_last_error                     - replaces checkCUDAError's exit()
 *                                        (sim.cu:60-70): every call returns a
 *                                        status code and never exits.
 *
 * Conventions: 0 = success, negative = error (see ERP_E_*).  The caller owns
 * every host buffer; the library owns every device handle.  One context is
 * driven by one submitting thread.  Output traces are shot-major: shot s
 * occupies out[s*nr*2 .. (s+1)*nr*2) as interleaved (re, im) float32, the same
 * order as the reference's s%06d.txt rows (file_io.cu:291-293).
 */
#ifndef ERP_H
#define ERP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define ERP_OK             0
#define ERP_E_INVALID     -1  /* bad argument or unsupported mode combination */
#define ERP_E_HIP         -2  /* HIP runtime error (message in erp_last_error) */
#define ERP_E_NOMEM       -3  /* device allocation failed */
#define ERP_E_UNSUPPORTED -4  /* mode the reference itself cannot run */
#define ERP_E_NOTREADY    -5  /* erp_poll: job still running */
#define ERP_E_INTERNAL    -6  /* a device-side consistency check failed (message) */

/* ---- processing mode bits (erp_params.mode), file_io.cu:238-246 --------- */
#define ERP_MODE_CONVOLUTION     0x01u /* "convolution"        */
#define ERP_MODE_CONV_LINEAR     0x02u /* "convolution_linear" */
#define ERP_MODE_SPECULAR        0x04u /* "specular"           */
#define ERP_MODE_LOSSLESS        0x08u /* "lossless"           */
#define ERP_MODE_INCOHERENT      0x10u /* "incoherent"         */
#define ERP_MODE_DISABLE_SURFACE 0x20u /* "disable_surface"    */

/* ---- run flags (erp_run flags argument) -------------------------------- */
#define ERP_RUN_DETERMINISTIC 0x01u /* bitwise repeatable bins: each term rounded
                                       once to 64-bit fixed point (scaled from a
                                       per-shot / per-target bound) and summed
                                       with integer atomics; every method, and
                                       type-0 targets summed per shot over a
                                       bound pass.  Replaces the order-dependent
                                       float atomicAdd binning of
                                       radar_kernels.cu:754-758 and 821-825     */
/* Precision study of the surface echo and the subsurface-target
 * accumulation (non-specular method, not with ERP_RUN_DETERMINISTIC).
 * PHASE_*: arithmetic of Itd, Ttd, rngt = Itd + n Ttd, the bin coordinate
 * and the phase: FP32 as the reference computes them (radar_kernels.cu:
 * 681-686, 733-737, 811-813) or double-float; the default is FP64.  ACC_*:
 * the LDS bin window sums: plain FP32, or FP32 with the rounding error of
 * every addition recovered by TwoSum and summed beside it; the default is
 * FP64.  The facet records the target stage reads keep the FP64 Itd. */
#define ERP_RUN_PHASE_F32     0x02u
#define ERP_RUN_PHASE_DF      0x04u
#define ERP_RUN_RAW_SURFACE   0x08u /* output the raw surface phasor trace before
                                       range compression (the Psurf dump of
                                       debug_surface, sim.cu:782-788)           */
#define ERP_RUN_ACC_F32       0x10u
#define ERP_RUN_ACC_COMP      0x20u

/*
 * Simulation parameters after the derivations of parseSimulationParameters
 * (file_io.cu:122-249).  Every float field holds exactly the float value the
 * reference stores in SimulationParameters (file_io.cu:40-120); the host side
 * computes them with the reference's FP32 operation order.  Per-shot values
 * (source position, rx window start, gains) live in erp_shot.
 */
typedef struct erp_params {
    float lam;          /* c / f0                          file_io.cu:170 */
    float k;            /* 2 pi / lam                      file_io.cu:171 */
    float rng_res;      /* c / B                           file_io.cu:168 */
    float P;            /* transmit power                  file_io.cu:154 */
    float fs;           /* facet size                      file_io.cu:228 */
    float ks;           /* k * rms_height                  file_io.cu:179 */
    float eps_1, eps_2; /*                                 file_io.cu:182-183 */
    float c, c_1, c_2;  /* c, c/sqrt(eps)                  file_io.cu:42,187-188 */
    float nu_1, nu_2;   /* 376.7/sqrt(eps)                 file_io.cu:189-190 */
    float alpha1, alpha2;/* attenuation                    file_io.cu:196-206 */
    float aperture;     /* degrees                         file_io.cu:155 */
    float smpl;         /* rx sample rate                  file_io.cu:208 */
    float B;            /* bandwidth                       file_io.cu:153 */
    float f0;           /* centre frequency                file_io.cu:152 */
    float dr;           /* rx_window_m / nr                file_io.cu:213 */
    int32_t nr;         /* range bins                      file_io.cu:212 */
    int32_t rerad_funct;/* target reradiation function     file_io.cu:232 */
    int32_t pol;        /* always 0 (never parsed)         file_io.cu:70  */
    uint32_t mode;      /* ERP_MODE_* bits */
    double cos_ap;      /* cos of the float aperture threshold (pi/180)*aperture,
                           in double; the FP64 aperture predicate compares
                           against it (geometry_kernels.cu:95-96,183) */
    uint64_t seed;      /* incoherent-mode RNG seed (replaces clock64(),
                           radar_kernels.cu:689-711) */
    /* Second subsurface medium (an extension: the reference models one
     * medium eps_2, file_io.cu:81).  Below the horizontal interface z = z_int
     * the permittivity is eps_3; eps_3 <= 0 disables it.  The forced ray of
     * a subsurface target (compTargetRays, geometry_kernels.cu:272-350) then
     * has the optical length n_2 L_2 + n_3 L_3 over its parts above and
     * below the interface, and a ray that crosses it carries the
     * normal-incidence power transmission 1 - rho_23^2 each way,
     * rho_23 = (n_2 - n_3) / (n_2 + n_3).  With eps_3 = eps_2 every result
     * equals the single-medium one.  Plain method (non-specular, targets of
     * type != 0) only. */
    float eps_3;
    float z_int;
} erp_params;

/* One shot (source position).  sx, sy, sz carry the reference's float source
 * coordinates promoted to double (sim.cu:642-651); snx..snz already include
 * source_normal_multiplier; rst and the gains are the per-shot values of
 * sim.cu:624-633. */
typedef struct erp_shot {
    double sx, sy, sz;
    float snx, sny, snz;
    float rst;
    float grefl_lin, grefr_lin;
} erp_shot;

/* Attenuation prism: alpha already converted from conductivity
 * (file_io.cu:477-517). */
typedef struct erp_prism {
    float alpha, xmin, ymin, zmin, xmax, ymax, zmax;
} erp_prism;

typedef struct erp_ctx erp_ctx;
typedef struct erp_mesh erp_mesh;
typedef struct erp_targets erp_targets;
typedef struct erp_job erp_job;

/* Create a context on HIP device `device`.  Owns its stream, FFT twiddle
 * tables and the chirp-spectrum cache. */
int erp_ctx_create(int device, erp_ctx** out);
void erp_ctx_destroy(erp_ctx* ctx);
/* Last error message of this context (never NULL). */
const char* erp_last_error(const erp_ctx* ctx);

/* Upload a facet mesh: soa holds 12 consecutive float arrays of length F in
 * the order x, y, z, nx, ny, nz, ux, uy, uz, vx, vy, vz (file_io.cu:349,
 * sim.cu:228-239).  The mesh stays resident in HBM until released. */
int erp_mesh_upload(erp_ctx* ctx, const float* soa, int64_t F, erp_mesh** out);
void erp_mesh_release(erp_mesh* mesh);
/* Build a resident mesh on the device from a height grid: heights holds
 * ny rows of nx doubles (x fastest), facet (i, j) sits at (ox + i fs,
 * oy + j fs, heights[j nx + i]).  Normals come from central differences
 * and the tangents from cross products exactly as gen_normals computes them
 * (simple_surfaces.py:3-39), facets in make_surface's order
 * (simple_surfaces.py:147-181, no mask): the same mesh as erp_mesh_upload
 * of that SoA, with 8 bytes per facet crossing PCIe instead of 48 and no
 * host-side normal computation. */
int erp_mesh_from_heights(erp_ctx* ctx, const double* heights, int64_t nx, int64_t ny,
                          double fs, double ox, double oy, erp_mesh** out);

/* A synthetic terrain generated on the device.  The reference draws its
 * terrains on the host (gen_normals, simple_surfaces.py:3-39, over height
 * fields from its examples' scripts); this is the host generators of the
 * benchmark meshes (europaraypaths_amd/surfaces.py) on the GPU:
 *   ERP_TERRAIN_GAUSSIAN: white noise filtered by exp(-K^2 corr_len^2 / 8)
 *     (K the angular wavenumber on the fftfreq grid): a Gaussian field with
 *     autocorrelation exp(-r^2 / corr_len^2);
 *   ERP_TERRAIN_FBM: amplitude |k|^-(hurst + 1) (0 at k = 0) times a normal
 *     times exp(i phase), phase uniform in [0, 2 pi): fBm, PSD ~ |k|^-(2H+2).
 * Then the real part of the inverse 2-D DFT, mean removed, rescaled to RMS
 * `rms`, and the mesh of erp_mesh_from_heights over it.  `noise`, when not
 * NULL, holds the host's draws in the host generators' order (Gaussian: ny
 * nx normals; fBm: ny nx phases, then ny nx normals), so that the device
 * field is the host's to FP32 transform rounding; NULL draws them on the
 * device from `seed` (SplitMix64 counters, Box-Muller normals).  Any grid
 * size up to 65536 per side (Bluestein transforms). */
#define ERP_TERRAIN_GAUSSIAN 0
#define ERP_TERRAIN_FBM      1
typedef struct erp_terrain {
    int32_t kind;
    int32_t reserved;
    int64_t nx, ny;
    double fs, ox, oy;  /* facet spacing and grid origin, as erp_mesh_from_heights */
    double rms;         /* RMS height after the rescale */
    double corr_len;    /* ERP_TERRAIN_GAUSSIAN */
    double hurst;       /* ERP_TERRAIN_FBM */
    uint64_t seed;      /* device draws (noise NULL) */
    const double* noise;
} erp_terrain;
int erp_mesh_from_terrain(erp_ctx* ctx, const erp_terrain* terrain, erp_mesh** out);
/* Copy a resident mesh back as the 12 x F float32 SoA of erp_mesh_upload
 * (soa must hold 12 F floats). */
int erp_mesh_download(const erp_mesh* mesh, float* soa);

/* Upload T targets: xyzn holds 6 consecutive float arrays x, y, z, nx, ny, nz
 * of length T; type holds T target types (file_io.cu:365-426). */
int erp_targets_upload(erp_ctx* ctx, const float* xyzn, const uint8_t* type,
                       int64_t T, erp_targets** out);
void erp_targets_release(erp_targets* targets);

/* Run ns shots (sim.cu:615-1008).  targets may be NULL (no targets).
 * out: host buffer of ns*nr*2 floats (shot-major, interleaved re/im).
 * nvalid (optional, host, ns entries): illuminated facet count per shot; a
 * shot with 0 illuminated facets has an all-zero trace (the reference writes
 * no file for it, sim.cu:693-697).  Synchronous. */
int erp_run(erp_ctx* ctx, const erp_params* par, const erp_mesh* mesh,
            const erp_targets* targets, const erp_shot* shots, int64_t ns,
            const erp_prism* prisms, int64_t nprisms,
            float* out, int32_t* nvalid, uint32_t flags);

/* Same as erp_run with the output left in device memory: out_dev is a device
 * pointer to ns*nr*2 floats, nvalid_dev an optional device pointer to ns
 * int32.  Returns after the work is queued and complete on the context's
 * stream (synchronous). */
int erp_run_dev(erp_ctx* ctx, const erp_params* par, const erp_mesh* mesh,
                const erp_targets* targets, const erp_shot* shots, int64_t ns,
                const erp_prism* prisms, int64_t nprisms,
                void* out_dev, void* nvalid_dev, uint32_t flags);

/* Allocation-free form of erp_run_dev.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory while they run a job
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as erp_run_dev, but every work
 * buffer and table of the run is carved from a per-device scratch array of
 * ERP_SCRATCH_BYTES that the library holds from the first context on that
 * device until the process ends; the call allocates nothing.  Shots per
 * batch and pairs per target launch shrink until the buffers fit (slower
 * than erp_run_dev for large jobs); a job of which one shot cannot fit
 * returns ERP_E_NOMEM with the sizes in erp_last_error.  All calls on one
 * device share the array and would race if they overlapped, so the library
 * serialises them (a call waits for the previous one to finish); nothing of
 * one call's scratch contents is used by the next.  No stage timers. */
#define ERP_SCRATCH_BYTES ((int64_t)256 << 20)
int erp_run_dev_noalloc(erp_ctx* ctx, const erp_params* par, const erp_mesh* mesh,
                        const erp_targets* targets, const erp_shot* shots, int64_t ns,
                        const erp_prism* prisms, int64_t nprisms,
                        void* out_dev, void* nvalid_dev, uint32_t flags);

/* One track of erp_run_tracks: the arguments of erp_run_dev for one job
 * (no attenuation prisms). */
typedef struct erp_track {
    const erp_params* par;
    const erp_mesh* mesh;
    const erp_targets* targets;  /* or NULL */
    const erp_shot* shots;       /* host, ns entries */
    int64_t ns;
    void* out_dev;               /* device: ns x nr complex64, shot-major */
    void* nvalid_dev;            /* device: ns int32, or NULL */
} erp_track;

/* Several tracks as one run.  No counterpart in the reference, whose sim
 * runs one track per process (sim.cu:615-1008); it is for many small or
 * uneven jobs that do not fill the device one at a time.  Tracks of the
 * plain method (convolution and convolution_linear, not specular) with nr
 * in [513, 16384] and no target of type 0 are grouped by
 * (nr, second medium or not); the (track, shot) rows of a group run in
 * batches of ERP_TUNE_TRACK_ROWS rows, every launch covering all rows of
 * the batch with each row reading its own track's mesh, targets, parameters
 * and range kernels, and each batch's four launch sequences replayed from
 * HIP graphs captured once.  Every other track runs alone, as erp_run_dev.
 * flags: 0 or ERP_RUN_DETERMINISTIC for grouped tracks (other flags run
 * every track alone).  Each radargram is erp_run_dev's on the same inputs:
 * within the parity tolerance, and with ERP_RUN_DETERMINISTIC the same
 * values.  Synchronous. */
int erp_run_tracks(erp_ctx* ctx, const erp_track* tracks, int64_t ntracks, uint32_t flags);

/* Asynchronous form of erp_run: queue a job and return at once.  Jobs run
 * in worker slots (erp_serve_slots; 4 are started at the first submit), each
 * a host thread with its own stream and work buffers, so jobs submitted
 * while others run overlap on the device.  `par`, `shots` and `prisms` are
 * copied before return; `mesh`, `targets`, `out` and `nvalid` must stay
 * valid until erp_wait/erp_poll reports completion.  Every job must be
 * finished with erp_wait or erp_poll before erp_ctx_destroy.  Replaces the
 * reference's one-shot-at-a-time process loop (sim.cu:615-1008) for callers
 * that keep several simulations in flight. */
int erp_submit(erp_ctx* ctx, const erp_params* par, const erp_mesh* mesh,
               const erp_targets* targets, const erp_shot* shots, int64_t ns,
               const erp_prism* prisms, int64_t nprisms,
               float* out, int32_t* nvalid, uint32_t flags, erp_job** job);
/* Worker slots for erp_submit: starts slots up to n (1..16); never stops
 * any.  Each slot keeps its work buffers between jobs.  No counterpart in
 * the reference, which runs one simulation per process. */
int erp_serve_slots(erp_ctx* ctx, int n);
/* ERP_OK when done (job resources released), ERP_E_NOTREADY while running. */
int erp_poll(erp_job* job);
/* Block until the job is done; releases it. */
int erp_wait(erp_job* job);

/* ---- stage timers (tracing) ------------------------------------------------
 * Replaces the reference's gettimeofday shot timers (sim.cu:44-58, 1000-1026)
 * with hipEvent timers recorded on the context stream around every stage of
 * erp_run / erp_run_dev.  units = work items the stage processed:
 *   MASK: facet-shot predicate tests, SURFACE: illuminated facet-shot pairs,
 *   TARGET_ACCUM: facet-target-shot triples, TARGET_FFT: target-shot traces. */
#define ERP_STAGE_MASK          0  /* k_illum_mask + k_scan_blocks          */
#define ERP_STAGE_SURFACE       1  /* fused surface kernel                  */
#define ERP_STAGE_SURFACE_CONV  2  /* surface range compression             */
#define ERP_STAGE_TARGET_SELECT 3  /* in-aperture target lists              */
#define ERP_STAGE_TARGET_ACCUM  4  /* facet x target accumulation kernel    */
#define ERP_STAGE_TARGET_FFT    5  /* per-target spectral chain             */
#define ERP_STAGE_SHOT_FINAL    6  /* Bluestein DFT/crop/IDFT + combine     */
#define ERP_STAGE_OTHER         7  /* type-0 targets, direct sums           */
#define ERP_STAGE_TARGET_SUM    8  /* per-shot sums of per-target spectra   */
#define ERP_NSTAGES             9
/* on != 0 enables the timers and clears the counters. */
int erp_profile_enable(erp_ctx* ctx, int on);
/* Totals since the last enable: summed milliseconds, number of timed
 * launches and units for one stage. */
int erp_profile_read(erp_ctx* ctx, int stage, double* ms, int64_t* launches, double* units);

/* ---- batch shapes ------------------------------------------------------------
 * Caps on the batch shapes erp_run / erp_run_dev / erp_submit /
 * erp_run_tracks use on this context; 0 restores the default.  The first two
 * can only lower the defaults, so erp_run_footprint stays an upper bound.  No
 * counterpart in the reference (its sim runs every shot in its own loop
 * iteration, sim.cu:615-1008).
 *   ERP_TUNE_SHOTS_PER_BATCH : shots per batch (default and maximum 64)
 *   ERP_TUNE_PAIRS_PER_LAUNCH: (shot, target) pairs per target launch
 *                              (default: the pairs whose rows fit 6 GB,
 *                              at most 16384; with ERP_RUN_DETERMINISTIC
 *                              at least 32, so that the results do not
 *                              depend on it)
 *   ERP_TUNE_TRACK_ROWS      : (track, shot) rows per multi-track batch
 *                              (default 64, at most 256)
 *   ERP_TUNE_DIRECT_LAUNCH   : > 0: multi-track batches launch their kernels
 *                              directly instead of replaying HIP graphs
 *   ERP_TUNE_WINDOW_SHIFT    : tests only: shift the bin window of every
 *                              (shot, target) pair of the plain target kernel
 *                              up by this many bins, so that its lowest terms
 *                              fall outside it and the run fails with
 *                              ERP_E_INTERNAL (the check that a window bound
 *                              held; 0 = off)
 *   ERP_TUNE_SINGLE_SHOT     : > 0: the plain target kernel takes each
 *                              (shot, target) pair on its own instead of a
 *                              target with two consecutive shots (which
 *                              shares the facet -> target leg of their common
 *                              facets; the default) */
#define ERP_TUNE_SHOTS_PER_BATCH  1
#define ERP_TUNE_PAIRS_PER_LAUNCH 2
#define ERP_TUNE_TRACK_ROWS       3
#define ERP_TUNE_DIRECT_LAUNCH    4
#define ERP_TUNE_WINDOW_SHIFT     5
#define ERP_TUNE_SINGLE_SHOT      6
int erp_set_tuning(erp_ctx* ctx, int key, int64_t value);
/* The value set for key (0 = default). */
int erp_get_tuning(erp_ctx* ctx, int key, int64_t* value);

/* ---- SAR focusing (simple_focusing.py:310-404) ------------------------------
 * Focus range bins row0 .. row0 + nrows - 1 of an Nr x Na complex radargram
 * (rdr: Nr * Na * 2 floats, row-major, as compile_rdrgrm lays it out).  For
 * row q the caller gives the half aperture row_hs[q] and aperture length
 * row_nsx[q] (2 row_hs[q] <= row_nsx[q]) in traces, and at row_off[q] in the
 * ntab-long tables the radargram row each aperture sample reads
 * (rows[], in [0, Nr): the sample's shifted range bin modulo Nr) and the
 * matched filter (filt[], complex).  out receives nrows * Na * 2 floats:
 *     out[q][j] = sum_{m < w} rdr[rows[off + m]][az_st + m] filt[off + ((n - m) mod L)]
 * with az_st = max(0, j - hs), w = min(Na, j + hs) - az_st, L = 2 nsx,
 * n = (j - az_st) + (L - w) / 2 and filt zero past nsx: the sample of the
 * reference's per-pixel FFT product that it keeps (focus_rdrgrm's inner loop,
 * simple_focusing.py:357-394). */
int erp_focus(erp_ctx* ctx, const float* rdr, int64_t Nr, int64_t Na, int64_t row0,
              int64_t nrows, const int32_t* row_hs, const int32_t* row_nsx,
              const int64_t* row_off, const int32_t* rows, const float* filt, int64_t ntab,
              float* out);

/* Device bytes a run of (F facets, T targets, ns shots, nr bins) needs beyond
 * the resident mesh and targets; used for admission by footprint. */
int64_t erp_run_footprint(const erp_params* par, int64_t F, int64_t T, int64_t ns);

#ifdef __cplusplus
}
#endif

#endif /* ERP_H */
