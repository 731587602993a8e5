/*
 * scarf3d_srm.h - extensions of the MI355X SCARF3D SRM library beyond the
 * reference API of scarf3d.h. None of these has a counterpart in the
 * reference except where noted; all return 0 on success and a negative code
 * on failure (details from scarf_last_error).
 */
#ifndef SCARF3D_SRM_H_
#define SCARF3D_SRM_H_

#include "scarf3d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* field kernels */
enum scarf_kernel {
    SCARF_KERNEL_AUTO = 0,        /* mfma_bf16x3 on grids, direct_f32 on point lists (float lib); direct_f64 (double lib) */
    SCARF_KERNEL_DIRECT_F32 = 1,  /* FP32 cos on FP64-reduced phases, FP64-promoted sums */
    SCARF_KERNEL_DIRECT_F64 = 2,  /* FP64 throughout */
    SCARF_KERNEL_MFMA_BF16X3 = 3, /* grid GEMM, 3-term bf16 split, 6 MFMA products */
    SCARF_KERNEL_MFMA_F16X2 = 4,  /* grid GEMM, scaled 2-term fp16 split, 3 MFMA products */
    SCARF_KERNEL_MFMA_BF16X2 = 5  /* grid GEMM, 2-term bf16 split, 3 MFMA products (about 1e-5 sigma) */
};

/* Number of harmonics of the spectral sum. The reference fixes it at compile
 * time to 10000 (src/scarflib_srm.f90:90); that stays the default. */
int scarf_set_nharm(int nharm);
int scarf_get_nharm(void);

/* Field kernel for subsequent executes (environment SCARF_KERNEL=name also
 * works: auto, direct_f32, direct_f64, mfma_bf16x3, mfma_f16x2, mfma_bf16x2). */
int scarf_set_kernel(int kernel);
int scarf_get_kernel(void);

/* Last error message (empty string when the last call succeeded). */
const char *scarf_last_error(void);

/* scarf_execute writing a device-resident field (d_field: device pointer
 * with room for the local block or point list); no host copy of the field. */
int scarf_execute_device(const int seed, real *d_field, real stats[]);

/* Two halves of scarf_execute_device, so that the device work can be timed
 * with the harmonic table already resident: scarf_srm_draw samples the
 * harmonics of `seed` on the host and uploads them (host seconds in
 * *host_seconds); scarf_srm_run runs tables + field + statistics + rescale
 * + taper on the resident table. Returns the number of harmonics kept
 * (scarf_srm_draw) or 0. */
int scarf_srm_draw(const int seed, double *host_seconds);
int scarf_srm_run(real *d_field, real stats[]);

/* scarf_srm_run that allocates nothing. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The
 * operand tables, statistics partials, statistics and points of interest
 * live in a static device array of 64 MiB that each GPU holds for the life
 * of the process; the tables are built one chunk of harmonics at a time
 * (chunks are multiples of 128 harmonics and carry the field's running
 * sums, so the field equals scarf_srm_run's bit for bit). Takes the resident table of
 * scarf_srm_draw, and the MFMA kernels of structured grids only; uses the
 * tile chosen earlier for the grid (or the least-padding one) and tunes
 * nothing. Refuses, with an error, grids whose operand tables for 128
 * harmonics do not fit in the array. Every call on one GPU shares the array:
 * calls that overlap on one GPU race. Returns 0 or -1 (scarf_last_error). */
int scarf_srm_run_noalloc(real *d_field, real stats[]);

/* FP64 field of the resident table (FP64 kernel, FP64 output, FP64
 * epilogue) into d_field (device, doubles), whatever `real` is: the exact
 * reference against which the faster kernels' errors are measured.
 * stats_out (may be NULL) = {count, mean, variance, sqrt(variance)}. */
int scarf_srm_run_fp64(double *d_field, double stats_out[4]);

/* One realization of a Monte Carlo ensemble on a structured grid: the
 * arguments of scarf_cart_initialize plus the options an ensemble varies
 * (dh <= 0 means dh = ds; the model box defaults to the grid as in
 * src/scarflib.f90:157-158; no points of interest). */
struct scarf_member {
    int nd;
    int fs[3], fe[3];
    real ds, dh;
    int acf;
    real cl[3];
    real sigma, hurst;
    real alpha, beta, gamma;
    int rescale;
    int nharm;
    int seed;
};

/* Runs n realizations in one persistent grouped-GEMM launch (MFMA kernel
 * `kernel`, SCARF_KERNEL_AUTO = mfma_bf16x3) replayed from a hipGraph
 * captured per ensemble shape. d_fields[m]: device output of member m.
 * stats: n x 8 values laid out like scarf_execute's. Harmonics are drawn on
 * the host in parallel, one independent stream per seed. */
int scarf_ensemble_execute(int n, const struct scarf_member *members, real *const *d_fields, real *stats,
                           int kernel);

/* Device milliseconds of the last ensemble (graph launch to completion). */
double scarf_ensemble_last_ms(void);

/* Block tile of the last MFMA run: out = {x points, j values, k values}
 * (j x k = 256 grid columns). The tile is picked per grid shape by a tuner
 * that times the candidates on a short synthetic harmonic record (SCARF_TUNE=0:
 * least padding instead; SCARF_TUNE_CACHE=<file> keeps the choices across
 * processes; SCARF_TILE=x,j,k fixes the tile of single-field runs, e.g.
 * SCARF_TILE=128,32,8); it changes the speed only, never the field. Returns
 * 3, or 0 before any MFMA run. */
int scarf_last_tile(int out[3]);

/* Device milliseconds of the last run: {tables, field, statistics,
 * epilogue, total}; returns the number written (<= n). */
int scarf_last_timings(double *ms, int n);

/* HIP stream the library launches on (hipStream_t as void*). */
void *scarf_stream(void);

/* Setup values of the current configuration (for checks):
 * out[0..8] = {ndim, kmin, kmax, kc, info1, info2, scaling, intg(kmin,kmax),
 * intg()}, out[9..17] = rotation matrix (column-major), out[18..20] = bar. */
int scarf_setup_values(double *out, int n);

/* Sets one parameter of the initialized object to `value` as given, with the
 * semantics of a present Fortran OPTIONAL argument of scarf_initialize
 * (src/scarflib.f90:163-175, :276-288): unlike struct scarf_opt, a value of
 * 0 or below counts (e.g. mute = 0). Names: method, hurst, dh, ds, mute,
 * taper, rescale, pad, alpha, beta, gamma, nc1..nc3, fc1..fc3. Used by the
 * Fortran module m_scarflib. Two more names set how the host draw runs and
 * never change the harmonics drawn: draw_threads (workers; 0 = environment
 * SCARF_DRAW_THREADS or min(16, hardware threads)) and draw_block (stream
 * pairs per block; 0 = 2^18). */
int scarf_set_option(const char *name, double value);

/* Parameters held after initialize (defaults applied), for checks:
 * {nd, structured, fs[3], fe[3], ds, dh, acf, cl[3], sigma, hurst, nc[3],
 * fc[3], alpha, beta, gamma, mute, taper, rescale, pad, method, npoi, npts};
 * returns the number written (<= n). */
int scarf_get_config(double *out, int n);

/* Harmonics of `seed` as drawn on the host: harm (nharm x 5, row-major
 * {v1, v2, v3, a, b}) and used[nharm] (0 = dropped by u > kc). Returns the
 * number of uniforms consumed. */
long scarf_draw_harmonics(const int seed, double *harm, unsigned char *used);

#ifdef __cplusplus
}
#endif

#endif
