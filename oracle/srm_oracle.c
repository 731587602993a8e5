/*
 * srm_oracle.c — TEST INFRASTRUCTURE (the oracle). Only tests/, the smoke
 * check and the bench's CPU-baseline leg load it, and only as the checker.
 * The product (scarf3d_amd) never links or calls it.
 *
 * Plain-C restatement of SCARF3D's spectral representation method (SRM):
 *   - TRNG4 trng::yarn2 + trng::uniform01_dist<double> (third-party, not in
 *     the reference tree; src/prng.cpp:36-38, :60, :147-165 bind it). TRNG
 *     >= 4.20 per the reference manual; restated from its published design
 *     (Bauke & Mertens, Phys. Rev. E 75, 066701, 2007): an order-2 multiple
 *     recursive generator modulo m = 2^31-1 with L'Ecuyer's coefficients
 *     (1498809829, 1160990996), whose output is g^r mod m, g = 123567893;
 *     seed s -> state (s mod m, 1); uniform01 = r / m from one draw.
 *     Same-seed bit parity with a TRNG build: parity unpinned (TRNG absent).
 *     Pinned here only by the number theory the design fixes (primitive
 *     root g, full-period characteristic polynomial, seed reduction mod m).
 *   - PSDFs, srm_fn, intg            (src/psdf.f90:50-90, scarflib_srm.f90:935-995)
 *   - setup (bounds, rotation, bar)  (src/scarflib_srm.f90:617-720, :214-317)
 *   - harmonic draw                  (src/scarflib_srm.f90:731-778, :328-375)
 *   - FP64 direct sum + scale        (src/scarflib_srm.f90:794-851, :397-436)
 *   - statistics / rescale / taper   (src/scarflib_common.f90:652-994,
 *                                     src/scarflib_srm.f90:877-925, :451-495)
 * The precision-dependent parts are written once in srm_oracle_t.h and
 * instantiated for real = float (_f) and real = double (_d).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* ---- yarn2 ------------------------------------------------------------ */
#define Y2_M 2147483647LL
#define Y2_A0 1498809829LL
#define Y2_A1 1160990996LL
#define Y2_G 123567893LL

int64_t orc_powmod(int64_t g, int64_t e, int64_t m)
{
    int64_t r = 1 % m, b = g % m;
    if (b < 0) b += m;
    while (e > 0) {
        if (e & 1) r = (int64_t)((__int128)r * b % m);
        b = (int64_t)((__int128)b * b % m);
        e >>= 1;
    }
    return r;
}

void orc_yarn2_seed(orc_yarn2 *g, unsigned long s)
{
    int64_t t = (int64_t)s;
    t %= Y2_M;
    if (t < 0) t += Y2_M;
    g->r0 = (int32_t)t;
    g->r1 = 1;
}

static void y2_step(orc_yarn2 *g)
{
    const uint64_t t = (uint64_t)Y2_A0 * (uint64_t)g->r0 + (uint64_t)Y2_A1 * (uint64_t)g->r1;
    g->r1 = g->r0;
    g->r0 = (int32_t)(t % (uint64_t)Y2_M);
}

int32_t orc_yarn2_next(orc_yarn2 *g)
{
    y2_step(g);
    if (g->r0 == 0) return 0;
    return (int32_t)orc_powmod(Y2_G, g->r0, Y2_M);
}

double orc_uniform01(orc_yarn2 *g)
{
    return (double)orc_yarn2_next(g) / (double)Y2_M;
}

void orc_yarn2_stream(long seed, int n, int32_t *out)
{
    orc_yarn2 g;
    orc_yarn2_seed(&g, (unsigned long)seed);
    for (int i = 0; i < n; ++i) out[i] = orc_yarn2_next(&g);
}

void orc_uniform_stream(long seed, int n, double *out)
{
    orc_yarn2 g;
    orc_yarn2_seed(&g, (unsigned long)seed);
    for (int i = 0; i < n; ++i) out[i] = orc_uniform01(&g);
}

/* ---- FP64 field (the oracle's truth) ------------------------------------ */
/* arg exactly as src/scarflib_srm.f90:802-804 (:401-403), in double */
static inline double srm_arg(const orc_setup *s, const double *h, double x, double y, double z)
{
    const double *m = s->matrix;
    return h[0] * ((m[0] * x + m[3] * y + m[6] * z) + s->bar[0]) +
           h[1] * ((m[1] * x + m[4] * y + m[7] * z) + s->bar[1]) +
           h[2] * ((m[2] * x + m[5] * y + m[8] * z) + s->bar[2]);
}

static double srm_point(const orc_params *p, const orc_setup *s, const double *harm,
                        const uint8_t *used, double x, double y, double z)
{
    double acc = 0.0;
    for (int l = 0; l < p->nharm; ++l) {
        if (!used[l]) continue;
        const double *h = harm + 5 * l;
        const double arg = srm_arg(s, h, x, y, z);
        acc = acc + h[3] * sin(arg) + h[4] * cos(arg);
    }
    return acc * s->scaling;
}

void orc_field_structured(const orc_params *p, const orc_setup *s, const double *harm,
                          const uint8_t *used, int k0, int k1, double *out)
{
    const int nx = p->fe[0] - p->fs[0] + 1, ny = p->fe[1] - p->fs[1] + 1;
    const long rows = (long)(k1 - k0) * ny;
#pragma omp parallel for schedule(dynamic, 1)
    for (long r = 0; r < rows; ++r) {
        const int k = k0 + (int)(r / ny), j = (int)(r % ny);
        const double z = (double)(k + p->fs[2] - 1) * p->ds;
        const double y = (double)(j + p->fs[1] - 1) * p->ds;
        for (int i = 0; i < nx; ++i) {
            const double x = (double)(i + p->fs[0] - 1) * p->ds;
            out[(size_t)(k - k0) * nx * ny + (size_t)j * nx + i] = srm_point(p, s, harm, used, x, y, z);
        }
    }
}

void orc_field_structured_at(const orc_params *p, const orc_setup *s, const double *harm,
                             const uint8_t *used, long n, const int *ii, const int *jj,
                             const int *kk, double *out)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (long q = 0; q < n; ++q) {
        const double x = (double)(ii[q] + p->fs[0] - 1) * p->ds;
        const double y = (double)(jj[q] + p->fs[1] - 1) * p->ds;
        const double z = (double)(kk[q] + p->fs[2] - 1) * p->ds;
        out[q] = srm_point(p, s, harm, used, x, y, z);
    }
}

void orc_field_unstructured(const orc_params *p, const orc_setup *s, const double *harm,
                            const uint8_t *used, double *out)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < p->npts; ++i) {
        out[i] = srm_point(p, s, harm, used, p->x[i], p->y[i], p->z ? p->z[i] : 0.0);
    }
}

/* Chan et al. merge of per-set variance/mean (src/scarflib_common.f90:823-875) */
void orc_parallel_variance(const double *varset, const double *avgset, const long *nset, int nsets,
                           double *var, double *avg)
{
    double sig = varset[0], mu = avgset[0], n = (double)nset[0];
    for (int i = 1; i < nsets; ++i) {
        const double ni = (double)nset[i];
        const double delta = avgset[i] - mu;
        const double m1 = sig * (n - 1.0);
        const double m2 = varset[i] * (ni - 1.0);
        const double m = m1 + m2 + delta * delta * n * ni / (n + ni);
        mu = (mu * n + avgset[i] * ni) / (n + ni);
        n = n + ni;
        sig = m / (n - 1.0);
    }
    *var = sig;
    *avg = mu;
}

/* ---- precision-templated parts ------------------------------------------ */
#define REAL float
#define SFX(name) name##_f
#define RSQRT sqrtf
#define RLOG logf
#define RSIN sinf
#define RCOS cosf
#define RACOS acosf
#define RPOW powf
#define REXP expf
#include "srm_oracle_t.h"
#undef REAL
#undef SFX
#undef RSQRT
#undef RLOG
#undef RSIN
#undef RCOS
#undef RACOS
#undef RPOW
#undef REXP

#define REAL double
#define SFX(name) name##_d
#define RSQRT sqrt
#define RLOG log
#define RSIN sin
#define RCOS cos
#define RACOS acos
#define RPOW pow
#define REXP exp
#include "srm_oracle_t.h"
