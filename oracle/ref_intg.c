/*
 * ref_intg.c — TEST INFRASTRUCTURE. Harness that drives the reference's own
 * QUADPACK (src/quadpack.f90, compiled from where it lies by oracle/Makefile
 * into oracle/_ref/) with the SRM integrand, so that the oracle's C
 * restatement of DQNG / DQAGI can be checked bit for bit.
 *
 * The integrand restates srm_fn (src/scarflib_srm.f90:975-995) over the PSDFs
 * of src/psdf.f90:50-90 and is passed to the Fortran routines by reference,
 * as the reference passes srm_fn (src/scarflib_srm.f90:962, :964).
 */
#include <math.h>

extern void dqng_(double (*f)(double *), double *a, double *b, double *epsabs, double *epsrel,
                  double *result, double *abserr, int *neval, int *ier);
extern void dqagi_(double (*f)(double *), double *bound, int *inf, double *epsabs, double *epsrel,
                   double *result, double *abserr, int *neval, int *ier, int *limit, int *lenw,
                   int *last, int *iwork, double *work);

static int g_ndim, g_acf, g_single;
static double g_nu;

static double srm_fn_ref(double *xp)
{
    const double x = *xp;
    double fv;
    if (g_single) {
        const float xf = (float)x, nuf = (float)g_nu;
        fv = (g_acf == 0) ? (double)(1.0f / powf(1.0f + xf * xf, nuf)) : (double)expf(-(0.25f * (xf * xf)));
    } else {
        fv = (g_acf == 0) ? 1.0 / pow(1.0 + x * x, g_nu) : exp(-(0.25 * (x * x)));
    }
    double r = x * fv;
    if (g_ndim == 3) r = r * x;
    return r;
}

/* finite != 0: DQNG on [k0, k1] (epsabs 1e-8, epsrel 1e-5); else DQAGI on
 * (0, inf) with limit 500 and the work sizes of src/scarflib_srm.f90:950-955. */
double ref_intg(int acf, int ndim, double nu, int single, int finite, double k0, double k1,
                double *abserr, int *neval, int *ier)
{
    g_acf = acf;
    g_ndim = ndim;
    g_nu = nu;
    g_single = single;
    double epsabs = 1.0e-8, epsrel = 1.0e-5, result = 0.0;
    if (finite) {
        dqng_(srm_fn_ref, &k0, &k1, &epsabs, &epsrel, &result, abserr, neval, ier);
    } else {
        enum { limlst = 50, limit = 500, leniw = 2 * limit + limlst, maxp1 = 21, lenw = leniw * 2 + maxp1 * 25 };
        static int iwork[leniw];
        static double work[lenw];
        double bound = 0.0;
        int inf = 1, lim = limit, lw = lenw, last = 0;
        dqagi_(srm_fn_ref, &bound, &inf, &epsabs, &epsrel, &result, abserr, neval, ier, &lim, &lw,
               &last, iwork, work);
    }
    return result;
}
