/*
 * oracle.h — TEST INFRASTRUCTURE. Declarations of the CPU oracle for the SRM
 * hot path of SCARF3D (src/scarflib_srm.f90). Only tests/, the smoke check
 * and the bench's CPU-baseline leg may load the oracle, and only as the
 * checker: the product (scarf3d_amd) never links or calls it.
 *
 * Every function restates one reference routine; the .c files cite
 * file:line for each. "_f" entry points follow the single-precision build of
 * the reference (real = float, libscarf3df), "_d" the DOUBLE_PREC build
 * (real = double, libscarf3d).
 */
#ifndef SCARF_ORACLE_H
#define SCARF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double (*orc_integrand)(double x, void *ctx);

/* QUADPACK (quadpack_oracle.c) */
double orc_dqng(orc_integrand f, void *ctx, double a, double b, double epsabs,
                double epsrel, double *abserr, int *neval, int *ier);
double orc_dqagi(orc_integrand f, void *ctx, double bound, int inf, double epsabs,
                 double epsrel, int limit, double *abserr, int *neval, int *ier, int *last);

/* TRNG4 yarn2 + uniform01_dist<double> restatement (srm_oracle.c) */
typedef struct { int32_t r0, r1; } orc_yarn2;
void orc_yarn2_seed(orc_yarn2 *g, unsigned long s);
int32_t orc_yarn2_next(orc_yarn2 *g);
double orc_uniform01(orc_yarn2 *g);
void orc_yarn2_stream(long seed, int n, int32_t *out);
void orc_uniform_stream(long seed, int n, double *out);
int64_t orc_powmod(int64_t g, int64_t e, int64_t m);

/* Resolved SRM parameters (the values the Fortran object holds after
 * initialize, src/scarflib.f90:97-296). Real-valued members hold values of
 * the library's `real` type (float-representable for the _f variants). */
typedef struct {
    int nd;                  /* problem dimension passed at initialize (poi stride) */
    int structured;          /* 1 = Cartesian grid, 0 = point list */
    int fs[3], fe[3];        /* structured: first/last global index (1-based) */
    double ds, dh;           /* grid step, minimum grid step */
    int acf;                 /* 0 von Karman, 1 Gaussian */
    double cl[3];            /* correlation lengths (cl[2] = 0 -> 2D) */
    double sigma, hurst;
    double nc[3], fc[3];     /* near / far corner of the model */
    double alpha, beta, gamma;
    double mute, taper;
    int rescale;
    int nharm;               /* harmonics in the spectral sum (reference: 10000) */
    int npoi;
    const double *poi;       /* (nd, npoi) column-major */
    int npts;                /* unstructured: number of points */
    const double *x, *y, *z; /* unstructured coordinates */
} orc_params;

/* Setup of scarf3d_structured_srm / scarf3d_unstructured_srm
 * (src/scarflib_srm.f90:617-720 / :214-317). */
typedef struct {
    int ndim;
    double kmin, kmax, kc;
    double info1, info2;
    double scaling;          /* sigma/sqrt(nharm)*sqrt(intg(kmin,kmax)/intg()) */
    double intg_band, intg_all;
    double matrix[9];        /* column-major M(r,c) = matrix[r + 3c] */
    double bar[3];
    double nu, cnst;
} orc_setup;

double orc_intg_raw_f(int acf, int ndim, double nu, int finite, double k0, double k1,
                      double *abserr, int *neval, int *ier);
double orc_intg_raw_d(int acf, int ndim, double nu, int finite, double k0, double k1,
                      double *abserr, int *neval, int *ier);
int orc_setup_f(const orc_params *p, orc_setup *s);
int orc_setup_d(const orc_params *p, orc_setup *s);

/* Harmonic draw (src/scarflib_srm.f90:731-778). harm is (nharm, 5) row-major
 * {v1, v2, v3, a, b}; used[l] = 0 for harmonics skipped by u > kc. Returns the
 * number of uniforms consumed. */
long orc_draw_f(const orc_params *p, const orc_setup *s, int seed, double *harm, uint8_t *used);
long orc_draw_d(const orc_params *p, const orc_setup *s, int seed, double *harm, uint8_t *used);

/* FP64 direct sum over a harmonic table (src/scarflib_srm.f90:794-808 and
 * :399-405, then the scale :843-851 / :431-436). Returns scaling*sum. The
 * structured variant evaluates z-slabs [k0, k1) (0-based, local) only. */
void orc_field_structured(const orc_params *p, const orc_setup *s, const double *harm,
                          const uint8_t *used, int k0, int k1, double *out);
void orc_field_unstructured(const orc_params *p, const orc_setup *s, const double *harm,
                            const uint8_t *used, double *out);
/* FP64 field at n selected structured points given by 0-based local (i,j,k). */
void orc_field_structured_at(const orc_params *p, const orc_setup *s, const double *harm,
                             const uint8_t *used, long n, const int *ii, const int *jj,
                             const int *kk, double *out);

/* Statistics (src/scarflib_common.f90:652-875). */
double orc_mean_f(const float *r, long n);
double orc_mean_d(const double *r, long n);
double orc_variance_f(const float *r, long n);
double orc_variance_d(const double *r, long n);
void orc_parallel_variance(const double *varset, const double *avgset, const long *nset, int nsets,
                           double *var, double *avg);

/* Taper / mute (src/scarflib_common.f90:881-994), one POI. */
void orc_taper_structured_f(float *field, const int *fs, const int *fe, double dh,
                            const double *poi3, double mute, double taper);
void orc_taper_structured_d(double *field, const int *fs, const int *fe, double dh,
                            const double *poi3, double mute, double taper);
void orc_taper_unstructured_f(float *field, long n, const double *x, const double *y,
                              const double *z, const double *poi3, double mute, double taper);
void orc_taper_unstructured_d(double *field, long n, const double *x, const double *y,
                              const double *z, const double *poi3, double mute, double taper);

/* Whole scarf3d_*_srm pipeline with the FP64 field: setup, draw, FP64 direct
 * sum, statistics, rescale, taper. field is `real` (float for _f). stats[8]
 * follows src/scarflib.f90:363-377 (stats(5:6) timing left at 0). */
int orc_srm_f(const orc_params *p, int seed, float *field, double *stats);
int orc_srm_d(const orc_params *p, int seed, double *field, double *stats);

/* CPU baseline: the reference's own loop order (harmonic outer, points
 * inner, src/scarflib_srm.f90:725-851) at `real` precision over the z-slabs
 * [k0, k1), using `threads` OpenMP threads (each owning z-slabs, like MPI
 * ranks). Accumulates into field (zeroed by the caller) and applies the
 * scaling pass at the end. */
void orc_baseline_structured_f(const orc_params *p, const orc_setup *s, const double *harm,
                               const uint8_t *used, int k0, int k1, float *field, int threads);
void orc_baseline_structured_d(const orc_params *p, const orc_setup *s, const double *harm,
                               const uint8_t *used, int k0, int k1, double *field, int threads);
/* the same on a point list: points [i0, i1), float or double precision */
void orc_baseline_unstructured_f(const orc_params *p, const orc_setup *s, const double *harm,
                                 const uint8_t *used, long i0, long i1, float *field, int threads);
void orc_baseline_unstructured_d(const orc_params *p, const orc_setup *s, const double *harm,
                                 const uint8_t *used, long i0, long i1, double *field, int threads);

#ifdef __cplusplus
}
#endif
#endif
