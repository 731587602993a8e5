/*
 * srm_oracle_t.h — TEST INFRASTRUCTURE. Precision-templated half of the SRM
 * oracle; included twice by srm_oracle.c with REAL = float (suffix _f, the
 * reference's default single-precision build) and REAL = double (suffix _d,
 * the DOUBLE_PREC build). Expressions keep the reference's evaluation order
 * (Fortran left-to-right for * and /), and the file is compiled with
 * -ffp-contract=off so that no FMA changes a rounding.
 */

/* pi as the reference declares it: REAL(f_real), PARAMETER :: pi (src/scarflib_common.f90:94) */
#define RPI ((REAL)3.141592653589793)

typedef struct {
    int ndim, acf;
    REAL nu;
} SFX(psd_ctx);

/* fn_vk / fn_gs (src/psdf.f90:50-67, :73-90), evaluated at real precision */
static REAL SFX(psd)(const SFX(psd_ctx) * c, REAL x)
{
    if (c->acf == 0) return (REAL)1 / RPOW((REAL)1 + x * x, c->nu);
    return REXP(-((REAL)0.25 * (x * x)));
}

/* srm_fn (src/scarflib_srm.f90:975-995): x * fun(real(x, f_real)) [* x in 3D] */
static double SFX(srm_fn)(double x, void *ctx)
{
    const SFX(psd_ctx) *c = (const SFX(psd_ctx) *)ctx;
    double r = x * (double)SFX(psd)(c, (REAL)x);
    if (c->ndim == 3) r = r * x;
    return r;
}

/* intg (src/scarflib_srm.f90:935-969): DQNG on [k0, k1] or DQAGI on (0, inf) */
static REAL SFX(intg)(SFX(psd_ctx) * c, int finite, REAL k0, REAL k1)
{
    double abserr;
    int neval, ier, last;
    double solv;
    if (finite)
        solv = orc_dqng(SFX(srm_fn), c, (double)k0, (double)k1, 1.0e-8, 1.0e-5, &abserr, &neval, &ier);
    else
        solv = orc_dqagi(SFX(srm_fn), c, 0.0, 1, 1.0e-8, 1.0e-5, 500, &abserr, &neval, &ier, &last);
    return (REAL)solv;
}

/* raw FP64 QUADPACK result for the SRM integrand (before the real rounding
 * of intg), for bit comparison with the reference QUADPACK */
double SFX(orc_intg_raw)(int acf, int ndim, double nu, int finite, double k0, double k1,
                         double *abserr, int *neval, int *ier)
{
    SFX(psd_ctx) c;
    c.ndim = ndim;
    c.acf = acf;
    c.nu = (REAL)nu;
    int last;
    if (finite) return orc_dqng(SFX(srm_fn), &c, k0, k1, 1.0e-8, 1.0e-5, abserr, neval, ier);
    return orc_dqagi(SFX(srm_fn), &c, 0.0, 1, 1.0e-8, 1.0e-5, 500, abserr, neval, ier, &last);
}

int SFX(orc_setup)(const orc_params *p, orc_setup *s)
{
    memset(s, 0, sizeof *s);
    const int ndim = ((REAL)p->cl[2] > (REAL)0) ? 3 : 2;
    REAL cl[3] = {(REAL)p->cl[0], (REAL)p->cl[1], (REAL)p->cl[2]};
    REAL span[3];
    for (int i = 0; i < 3; ++i) span[i] = (REAL)p->fc[i] - (REAL)p->nc[i];
    const REAL ds = (REAL)p->ds, dh = (REAL)p->dh;

    /* wavenumber bounds (:644-652) */
    REAL mx = cl[0] / span[0], mn = cl[0];
    for (int i = 1; i < ndim; ++i) {
        const REAL q = cl[i] / span[i];
        if (q > mx) mx = q;
        if (cl[i] < mn) mn = cl[i];
    }
    const REAL kmin = (REAL)2 * RPI * mx;
    const REAL kmax = RPI / dh * mn;
    const REAL kc = RPI / ds * mn;

    /* warning flags (:660-662) */
    REAL info1 = 0, info2 = 0;
    for (int i = 0; i < ndim; ++i) {
        if ((REAL)2 * RPI / span[i] >= (REAL)1 / cl[i]) info1 = 1;
        if (ds > cl[i] / (REAL)2) info2 = 1;
    }

    REAL scaling = (REAL)p->sigma / RSQRT((REAL)p->nharm);

    /* rotation matrix, columns as written at :689-691, angles in degrees */
    const REAL ca = RCOS((REAL)p->alpha * RPI / (REAL)180), sa = RSIN((REAL)p->alpha * RPI / (REAL)180);
    const REAL cb = RCOS((REAL)p->beta * RPI / (REAL)180), sb = RSIN((REAL)p->beta * RPI / (REAL)180);
    const REAL cg = RCOS((REAL)p->gamma * RPI / (REAL)180), sg = RSIN((REAL)p->gamma * RPI / (REAL)180);
    REAL m[9];
    m[0] = ca * cb;
    m[1] = sa * cb;
    m[2] = -sb;
    m[3] = ca * sb * sg - sa * cg;
    m[4] = sa * sb * sg + ca * cg;
    m[5] = cb * sg;
    m[6] = ca * sb * cg + sa * sg;
    m[7] = sa * sb * cg - ca * sg;
    m[8] = cb * cg;

    /* bar = span/2 - M*(span/2) (:694-700) */
    REAL bar[3], obar[3];
    for (int i = 0; i < 3; ++i) bar[i] = ((REAL)p->fc[i] - (REAL)p->nc[i]) / (REAL)2;
    for (int r = 0; r < 3; ++r) obar[r] = m[r] * bar[0] + m[r + 3] * bar[1] + m[r + 6] * bar[2];
    for (int i = 0; i < 3; ++i) bar[i] = bar[i] - obar[i];

    /* spectrum (:703-716). 2D Gaussian: the reference leaves const unset; 1 is used. */
    SFX(psd_ctx) c;
    c.ndim = ndim;
    c.acf = p->acf;
    c.nu = (REAL)p->hurst + (REAL)ndim / (REAL)2;
    REAL cnst = 1;
    if (p->acf == 1 && ndim == 3) cnst = (REAL)1 / RSQRT(RPI);
    if (p->acf != 0 && p->acf != 1) return -1;

    /* normalisation (:720) */
    const REAL band = SFX(intg)(&c, 1, kmin, kmax);
    const REAL all = SFX(intg)(&c, 0, 0, 0);
    scaling = scaling * RSQRT(band / all);

    s->ndim = ndim;
    s->kmin = kmin;
    s->kmax = kmax;
    s->kc = kc;
    s->info1 = info1;
    s->info2 = info2;
    s->scaling = scaling;
    s->intg_band = band;
    s->intg_all = all;
    for (int i = 0; i < 9; ++i) s->matrix[i] = m[i];
    for (int i = 0; i < 3; ++i) s->bar[i] = bar[i];
    s->nu = c.nu;
    s->cnst = cnst;
    return 0;
}

/* srng (src/prng.cpp:152-165): two uniform01 draws rounded to real */
static inline void SFX(srng)(orc_yarn2 *g, REAL r[2])
{
    r[0] = (REAL)orc_uniform01(g);
    r[1] = (REAL)orc_uniform01(g);
}

long SFX(orc_draw)(const orc_params *p, const orc_setup *s, int seed, double *harm, uint8_t *used)
{
    orc_yarn2 g;
    orc_yarn2_seed(&g, (unsigned long)seed); /* set_seed: unsigned long myseed = seed */
    SFX(psd_ctx) c;
    c.ndim = s->ndim;
    c.acf = p->acf;
    c.nu = (REAL)s->nu;
    const REAL kmin = (REAL)s->kmin, kmax = (REAL)s->kmax, kc = (REAL)s->kc, cnst = (REAL)s->cnst;
    const REAL cl0 = (REAL)p->cl[0], cl1 = (REAL)p->cl[1], cl2 = (REAL)p->cl[2];
    long ndraw = 0;
    REAL r[2];
    for (int l = 0; l < p->nharm; ++l) {
        REAL u;
        for (;;) {
            SFX(srng)(&g, r);
            ndraw += 2;
            u = r[0] * (kmax - kmin) + kmin;
            const REAL d = (REAL)(SFX(srm_fn)((double)u, &c) * (double)cnst);
            if (r[1] < d) break;
            if (ndraw > 4000000000L) return -1; /* the reference would spin forever here */
        }
        SFX(srng)(&g, r);
        ndraw += 2;
        double *h = harm + 5 * l;
        if (u > kc) { /* skipped, but the harmonic still counts in scaling (Q1) */
            used[l] = 0;
            h[0] = h[1] = h[2] = h[3] = h[4] = 0.0;
            continue;
        }
        const REAL phi = r[0] * (REAL)2 * RPI;
        REAL v1, v2, v3;
        if (s->ndim == 3) {
            const REAL theta = RACOS((REAL)1 - (REAL)2 * r[1]);
            v1 = u * RSIN(phi) * RSIN(theta) / cl0;
            v2 = u * RCOS(phi) * RSIN(theta) / cl1;
            v3 = u * RCOS(theta) / cl2;
        } else {
            v1 = u * RCOS(phi) / cl0;
            v2 = u * RSIN(phi) / cl1;
            v3 = 0;
        }
        SFX(srng)(&g, r);
        const REAL a = RSQRT(-((REAL)2 * RLOG(r[0]))) * RCOS((REAL)2 * RPI * r[1]);
        SFX(srng)(&g, r);
        const REAL b = RSQRT(-((REAL)2 * RLOG(r[0]))) * RCOS((REAL)2 * RPI * r[1]);
        ndraw += 4;
        h[0] = v1;
        h[1] = v2;
        h[2] = v3;
        h[3] = a;
        h[4] = b;
        used[l] = 1;
    }
    (void)cl2;
    return ndraw;
}

/* mean_1d/3d (src/scarflib_common.f90:743-817): running-average recurrence */
double SFX(orc_mean)(const REAL *r, long n)
{
    double v = 0.0, c = 1.0;
    for (long i = 0; i < n; ++i) {
        v = v + ((double)r[i] - v) / c;
        c = c + 1.0;
    }
    return (REAL)v;
}

/* variance_1d/3d (:652-737): two-pass, compensated, FP64 inside */
double SFX(orc_variance)(const REAL *r, long n)
{
    const double mu = SFX(orc_mean)(r, n);
    double s1 = 0.0, s2 = 0.0;
    for (long i = 0; i < n; ++i) {
        const double x = (double)r[i] - mu;
        s1 = s1 + x;
        s2 = s2 + x * x;
    }
    s1 = (s1 * s1) / (double)n;
    const double v = (s2 - s1) / (double)(n - 1);
    return (REAL)v;
}

/* tapering_structured (:881-941), global index (i-1)*dh */
void SFX(orc_taper_structured)(REAL *field, const int *fs, const int *fe, double dh_,
                               const double *poi3, double mute_, double taper_)
{
    const REAL dh = (REAL)dh_, mute = (REAL)mute_, taper = (REAL)taper_;
    const REAL p0 = (REAL)poi3[0], p1 = (REAL)poi3[1], p2 = (REAL)poi3[2];
    const REAL dsr = mute + taper;
    const long nx = fe[0] - fs[0] + 1, ny = fe[1] - fs[1] + 1;
    for (int k = fs[2]; k <= fe[2]; ++k) {
        REAL dz = (REAL)(k - 1) * dh - p2;
        dz = dz * dz;
        for (int j = fs[1]; j <= fe[1]; ++j) {
            REAL dy = (REAL)(j - 1) * dh - p1;
            dy = dy * dy;
            for (int i = fs[0]; i <= fe[0]; ++i) {
                REAL dx = (REAL)(i - 1) * dh - p0;
                dx = dx * dx;
                const REAL d = RSQRT(dx + dy + dz);
                REAL *f = field + (size_t)(k - fs[2]) * nx * ny + (size_t)(j - fs[1]) * nx + (i - fs[0]);
                if (d <= mute) {
                    *f = 0;
                } else if (d > mute && d <= dsr) {
                    const REAL t = (d - mute) / taper;
                    *f = *f * ((REAL)0.5 - (REAL)0.5 * RCOS(t * RPI));
                }
            }
        }
    }
}

/* tapering_unstructured (:947-994) */
void SFX(orc_taper_unstructured)(REAL *field, long n, const double *x, const double *y,
                                 const double *z, const double *poi3, double mute_, double taper_)
{
    const REAL mute = (REAL)mute_, taper = (REAL)taper_;
    const REAL p0 = (REAL)poi3[0], p1 = (REAL)poi3[1], p2 = (REAL)poi3[2];
    const REAL dsr = mute + taper;
    for (long i = 0; i < n; ++i) {
        const REAL ax = (REAL)x[i] - p0, ay = (REAL)y[i] - p1, az = (REAL)(z ? z[i] : 0.0) - p2;
        const REAL d = RSQRT(ax * ax + ay * ay + az * az);
        if (d <= mute) {
            field[i] = 0;
        } else if (d > mute && d <= dsr) {
            const REAL t = (d - mute) / taper;
            field[i] = field[i] * ((REAL)0.5 - (REAL)0.5 * RCOS(t * RPI));
        }
    }
}

/* Whole scarf3d_*_srm with the FP64 field (:617-929 / :214-499) */
int SFX(orc_srm)(const orc_params *p, int seed, REAL *field, double *stats)
{
    orc_setup s;
    for (int i = 0; i < 8; ++i) stats[i] = 0.0;
    if (SFX(orc_setup)(p, &s) != 0) return -1;
    double *harm = (double *)malloc(sizeof(double) * 5 * (size_t)p->nharm);
    uint8_t *used = (uint8_t *)malloc((size_t)p->nharm);
    SFX(orc_draw)(p, &s, seed, harm, used);
    long n;
    double *f64;
    if (p->structured) {
        const int nz = p->fe[2] - p->fs[2] + 1;
        n = (long)(p->fe[0] - p->fs[0] + 1) * (p->fe[1] - p->fs[1] + 1) * nz;
        f64 = (double *)malloc(sizeof(double) * (size_t)n);
        orc_field_structured(p, &s, harm, used, 0, nz, f64);
    } else {
        n = p->npts;
        f64 = (double *)malloc(sizeof(double) * (size_t)n);
        orc_field_unstructured(p, &s, harm, used, f64);
    }
    for (long i = 0; i < n; ++i) field[i] = (REAL)f64[i];
    free(f64);
    free(harm);
    free(used);

    const REAL var = (REAL)SFX(orc_variance)(field, n);
    const REAL mu = (REAL)SFX(orc_mean)(field, n);
    REAL info3 = RSQRT(var), info4 = mu;
    if (p->rescale == 1) {
        const REAL sc = (REAL)p->sigma / info3;
        for (long i = 0; i < n; ++i) field[i] = field[i] * sc - info4;
        info3 = (REAL)p->sigma;
        info4 = 0;
    }
    for (int q = 0; q < p->npoi; ++q) {
        double pt[3] = {0.0, 0.0, 0.0};
        const int stride = p->nd > 0 ? p->nd : s.ndim;
        for (int d = 0; d < s.ndim; ++d) pt[d] = p->poi[(size_t)q * stride + d];
        if (p->structured)
            SFX(orc_taper_structured)(field, p->fs, p->fe, p->ds, pt, p->mute, p->taper);
        else
            SFX(orc_taper_unstructured)(field, n, p->x, p->y, p->z, pt, p->mute, p->taper);
    }
    stats[0] = s.info1;
    stats[1] = s.info2;
    stats[2] = info3;
    stats[3] = info4;
    return 0;
}

#undef RPI
