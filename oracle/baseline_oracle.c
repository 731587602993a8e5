/*
 * baseline_oracle.c — TEST INFRASTRUCTURE (the oracle's CPU-baseline leg).
 *
 * The reference's own field loop restated for timing on the host: harmonic
 * outer, points inner, real precision, libm sin/cos, then the scaling pass
 * (src/scarflib_srm.f90:725-851). Each OpenMP thread owns a contiguous block
 * of z-slabs and runs the whole harmonic loop over it, which is what each MPI
 * rank of the reference does with its block (every rank redraws the same
 * harmonics, :627-778). Compiled -O3 for an AVX2/FMA host like the
 * reference's -O3 builds; only timed, never used as the parity truth.
 */
#include <math.h>
#include <stddef.h>
#include "oracle.h"

/* CPU baseline: the reference's loop nest at real precision (:794-808, :843-851) */
void orc_baseline_structured_f(const orc_params *p, const orc_setup *s, const double *harm,
                                  const uint8_t *used, int k0, int k1, float *field, int threads)
{
    const int nx = p->fe[0] - p->fs[0] + 1, ny = p->fe[1] - p->fs[1] + 1;
    const float ds = (float)p->ds;
    float m[9], bar[3];
    for (int i = 0; i < 9; ++i) m[i] = (float)s->matrix[i];
    for (int i = 0; i < 3; ++i) bar[i] = (float)s->bar[i];
    const float scaling = (float)s->scaling;
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(static, 1)
    for (int t = 0; t < threads; ++t) {
        /* each thread owns a contiguous block of z-slabs, like an MPI rank */
        const int span = k1 - k0;
        const int kb = k0 + (int)((long)span * t / threads), ke = k0 + (int)((long)span * (t + 1) / threads);
        for (int l = 0; l < p->nharm; ++l) {
            if (!used[l]) continue;
            const double *h = harm + 5 * l;
            const float v1 = (float)h[0], v2 = (float)h[1], v3 = (float)h[2], a = (float)h[3], b = (float)h[4];
            for (int k = kb; k < ke; ++k) {
                const float z = (float)(k + p->fs[2] - 1) * ds;
                for (int j = 0; j < ny; ++j) {
                    const float y = (float)(j + p->fs[1] - 1) * ds;
                    float *row = field + (size_t)(k - k0) * nx * ny + (size_t)j * nx;
                    for (int i = 0; i < nx; ++i) {
                        const float x = (float)(i + p->fs[0] - 1) * ds;
                        const float arg = v1 * ((m[0] * x + m[3] * y + m[6] * z) + bar[0]) +
                                         v2 * ((m[1] * x + m[4] * y + m[7] * z) + bar[1]) +
                                         v3 * ((m[2] * x + m[5] * y + m[8] * z) + bar[2]);
                        row[i] = row[i] + a * sinf(arg) + b * cosf(arg);
                    }
                }
            }
        }
        for (int k = kb; k < ke; ++k) {
            float *slab = field + (size_t)(k - k0) * nx * ny;
            for (long q = 0; q < (long)nx * ny; ++q) slab[q] = slab[q] * scaling;
        }
    }
}

/* CPU baseline: the reference's loop nest at real precision (:794-808, :843-851) */
void orc_baseline_structured_d(const orc_params *p, const orc_setup *s, const double *harm,
                                  const uint8_t *used, int k0, int k1, double *field, int threads)
{
    const int nx = p->fe[0] - p->fs[0] + 1, ny = p->fe[1] - p->fs[1] + 1;
    const double ds = (double)p->ds;
    double m[9], bar[3];
    for (int i = 0; i < 9; ++i) m[i] = (double)s->matrix[i];
    for (int i = 0; i < 3; ++i) bar[i] = (double)s->bar[i];
    const double scaling = (double)s->scaling;
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(static, 1)
    for (int t = 0; t < threads; ++t) {
        /* each thread owns a contiguous block of z-slabs, like an MPI rank */
        const int span = k1 - k0;
        const int kb = k0 + (int)((long)span * t / threads), ke = k0 + (int)((long)span * (t + 1) / threads);
        for (int l = 0; l < p->nharm; ++l) {
            if (!used[l]) continue;
            const double *h = harm + 5 * l;
            const double v1 = (double)h[0], v2 = (double)h[1], v3 = (double)h[2], a = (double)h[3], b = (double)h[4];
            for (int k = kb; k < ke; ++k) {
                const double z = (double)(k + p->fs[2] - 1) * ds;
                for (int j = 0; j < ny; ++j) {
                    const double y = (double)(j + p->fs[1] - 1) * ds;
                    double *row = field + (size_t)(k - k0) * nx * ny + (size_t)j * nx;
                    for (int i = 0; i < nx; ++i) {
                        const double x = (double)(i + p->fs[0] - 1) * ds;
                        const double arg = v1 * ((m[0] * x + m[3] * y + m[6] * z) + bar[0]) +
                                         v2 * ((m[1] * x + m[4] * y + m[7] * z) + bar[1]) +
                                         v3 * ((m[2] * x + m[5] * y + m[8] * z) + bar[2]);
                        row[i] = row[i] + a * sin(arg) + b * cos(arg);
                    }
                }
            }
        }
        for (int k = kb; k < ke; ++k) {
            double *slab = field + (size_t)(k - k0) * nx * ny;
            for (long q = 0; q < (long)nx * ny; ++q) slab[q] = slab[q] * scaling;
        }
    }
}

/* CPU baseline on a point list: the reference's unstructured loop nest at
 * float precision (:399-405, :431-436), each thread owning a contiguous range
 * of points [i0, i1) like an MPI rank. */
void orc_baseline_unstructured_f(const orc_params *p, const orc_setup *s, const double *harm,
                                 const uint8_t *used, long i0, long i1, float *field, int threads)
{
    float m[9], bar[3];
    for (int i = 0; i < 9; ++i) m[i] = (float)s->matrix[i];
    for (int i = 0; i < 3; ++i) bar[i] = (float)s->bar[i];
    const float scaling = (float)s->scaling;
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(static, 1)
    for (int t = 0; t < threads; ++t) {
        const long span = i1 - i0;
        const long pb = i0 + span * t / threads, pe = i0 + span * (t + 1) / threads;
        for (int l = 0; l < p->nharm; ++l) {
            if (!used[l]) continue;
            const double *h = harm + 5 * l;
            const float v1 = (float)h[0], v2 = (float)h[1], v3 = (float)h[2], a = (float)h[3], b = (float)h[4];
            for (long q = pb; q < pe; ++q) {
                const float x = (float)p->x[q], y = (float)p->y[q], z = (float)p->z[q];
                const float arg = v1 * ((m[0] * x + m[3] * y + m[6] * z) + bar[0]) +
                                 v2 * ((m[1] * x + m[4] * y + m[7] * z) + bar[1]) +
                                 v3 * ((m[2] * x + m[5] * y + m[8] * z) + bar[2]);
                field[q - i0] = field[q - i0] + a * sinf(arg) + b * cosf(arg);
            }
        }
        for (long q = pb; q < pe; ++q) field[q - i0] = field[q - i0] * scaling;
    }
}

/* The same at double precision (the DOUBLE_PREC build's loop). */
void orc_baseline_unstructured_d(const orc_params *p, const orc_setup *s, const double *harm,
                                 const uint8_t *used, long i0, long i1, double *field, int threads)
{
    const double *m = s->matrix, *bar = s->bar;
    const double scaling = s->scaling;
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(static, 1)
    for (int t = 0; t < threads; ++t) {
        const long span = i1 - i0;
        const long pb = i0 + span * t / threads, pe = i0 + span * (t + 1) / threads;
        for (int l = 0; l < p->nharm; ++l) {
            if (!used[l]) continue;
            const double *h = harm + 5 * l;
            const double v1 = h[0], v2 = h[1], v3 = h[2], a = h[3], b = h[4];
            for (long q = pb; q < pe; ++q) {
                const double x = p->x[q], y = p->y[q], z = p->z[q];
                const double arg = v1 * ((m[0] * x + m[3] * y + m[6] * z) + bar[0]) +
                                   v2 * ((m[1] * x + m[4] * y + m[7] * z) + bar[1]) +
                                   v3 * ((m[2] * x + m[5] * y + m[8] * z) + bar[2]);
                field[q - i0] = field[q - i0] + a * sin(arg) + b * cos(arg);
            }
        }
        for (long q = pb; q < pe; ++q) field[q - i0] = field[q - i0] * scaling;
    }
}
