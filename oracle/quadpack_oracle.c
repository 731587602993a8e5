/*
 * quadpack_oracle.c — TEST INFRASTRUCTURE (the oracle). Only tests/, the
 * smoke check and the bench's CPU-baseline leg load this code, and only as a
 * checker. The product never links it.
 *
 * Plain-C restatement of the five QUADPACK routines the reference's SRM uses
 * to normalise the spectrum (src/scarflib_srm.f90:959-965):
 *
 *   dqng   - non-adaptive 21/43/87-point Gauss-Kronrod-Patterson rule
 *            (src/quadpack.f90:8137-8528)
 *   dqagi  - adaptive integration on a semi-infinite range
 *            (src/quadpack.f90:1177-1370) -> dqagie (:712-1176), which calls
 *            dqk15i (:6589-6783), dqpsrt (:8529-8662), dqelg (:6230-6416).
 *
 * The arithmetic order of every expression follows the Fortran statement it
 * restates, so that the returned values (not just the integrals) agree with
 * the reference bit for bit. Built with -ffp-contract=off. Machine constants:
 * EPSILON(1d0) = DBL_EPSILON, TINY(1d0) = DBL_MIN, HUGE(1d0) = DBL_MAX.
 *
 * Pinned by: the reference QUADPACK compiled from its own source into
 * oracle/_ref (bit-for-bit comparison in tests/test_oracle_quadpack.py), and
 * by closed-form integrals of the spectra (Beta / erf forms).
 */
#include <float.h>
#include <math.h>
#include "gk_rules.h"
#include "oracle.h"

/* ---- dqng ------------------------------------------------------------- */
double orc_dqng(orc_integrand f, void *ctx, double a, double b, double epsabs,
                double epsrel, double *abserr, int *neval, int *ier)
{
    const double epmach = DBL_EPSILON, uflow = DBL_MIN;
    double result = 0.0, err = 0.0;
    double res10 = 0, res21 = 0, res43 = 0, res87 = 0, resabs = 0, resasc = 0, reskh;
    double fv1[5], fv2[5], fv3[5], fv4[5], savfun[21];
    int ipx = 0;

    *neval = 0;
    *ier = 6;
    *abserr = 0.0;
    if (epsabs <= 0.0 && epsrel < fmax(50.0 * epmach, 0.5e-28))
        return 0.0;

    const double hlgth = 0.5 * (b - a);
    const double dhlgth = fabs(hlgth);
    const double centr = 0.5 * (b + a);
    const double fcentr = f(centr, ctx);
    *neval = 21;
    *ier = 1;

    for (int l = 1; l <= 3; ++l) {
        if (l == 1) {
            /* 10- and 21-point results */
            res10 = 0.0;
            res21 = gkp_w21b[5] * fcentr;
            resabs = gkp_w21b[5] * fabs(fcentr);
            for (int k = 0; k < 5; ++k) {
                const double absc = hlgth * gkp_x1[k];
                const double fval1 = f(centr + absc, ctx);
                const double fval2 = f(centr - absc, ctx);
                const double fval = fval1 + fval2;
                res10 = res10 + gkp_w10[k] * fval;
                res21 = res21 + gkp_w21a[k] * fval;
                resabs = resabs + gkp_w21a[k] * (fabs(fval1) + fabs(fval2));
                savfun[k] = fval;
                fv1[k] = fval1;
                fv2[k] = fval2;
            }
            ipx = 5;
            for (int k = 0; k < 5; ++k) {
                ++ipx;
                const double absc = hlgth * gkp_x2[k];
                const double fval1 = f(centr + absc, ctx);
                const double fval2 = f(centr - absc, ctx);
                const double fval = fval1 + fval2;
                res21 = res21 + gkp_w21b[k] * fval;
                resabs = resabs + gkp_w21b[k] * (fabs(fval1) + fabs(fval2));
                savfun[ipx - 1] = fval;
                fv3[k] = fval1;
                fv4[k] = fval2;
            }
            result = res21 * hlgth;
            resabs = resabs * dhlgth;
            reskh = 0.5 * res21;
            resasc = gkp_w21b[5] * fabs(fcentr - reskh);
            for (int k = 0; k < 5; ++k) {
                resasc = resasc + gkp_w21a[k] * (fabs(fv1[k] - reskh) + fabs(fv2[k] - reskh))
                                + gkp_w21b[k] * (fabs(fv3[k] - reskh) + fabs(fv4[k] - reskh));
            }
            err = fabs((res21 - res10) * hlgth);
            resasc = resasc * dhlgth;
        } else if (l == 2) {
            /* 43-point result */
            res43 = gkp_w43b[11] * fcentr;
            *neval = 43;
            for (int k = 0; k < 10; ++k)
                res43 = res43 + savfun[k] * gkp_w43a[k];
            for (int k = 0; k < 11; ++k) {
                ++ipx;
                const double absc = hlgth * gkp_x3[k];
                const double fval = f(absc + centr, ctx) + f(centr - absc, ctx);
                res43 = res43 + fval * gkp_w43b[k];
                savfun[ipx - 1] = fval;
            }
            result = res43 * hlgth;
            err = fabs((res43 - res21) * hlgth);
        } else {
            /* 87-point result */
            res87 = gkp_w87b[22] * fcentr;
            *neval = 87;
            for (int k = 0; k < 21; ++k)
                res87 = res87 + savfun[k] * gkp_w87a[k];
            for (int k = 0; k < 22; ++k) {
                const double absc = hlgth * gkp_x4[k];
                res87 = res87 + gkp_w87b[k] * (f(absc + centr, ctx) + f(centr - absc, ctx));
            }
            result = res87 * hlgth;
            err = fabs((res87 - res43) * hlgth);
        }
        if (resasc != 0.0 && err != 0.0)
            err = resasc * fmin(1.0, pow(200.0 * err / resasc, 1.5));
        if (resabs > uflow / (50.0 * epmach))
            err = fmax((epmach * 50.0) * resabs, err);
        if (err <= fmax(epsabs, epsrel * fabs(result))) {
            *ier = 0;
            break;
        }
    }
    *abserr = err;
    return result;
}

/* ---- dqk15i: 15-point transformed Kronrod rule on (a,b) within (0,1] ---- */
static void dqk15i(orc_integrand f, void *ctx, double boun, int inf, double a, double b,
                   double *result, double *abserr, double *resabs, double *resasc)
{
    const double epmach = DBL_EPSILON, uflow = DBL_MIN;
    const double dinf = (double)(inf < 1 ? inf : 1);
    double fv1[7], fv2[7];
    const double centr = 0.5 * (a + b);
    const double hlgth = 0.5 * (b - a);
    const double tabsc1c = boun + dinf * (1.0 - centr) / centr;
    double fval1 = f(tabsc1c, ctx);
    if (inf == 2) fval1 = fval1 + f(-tabsc1c, ctx);
    const double fc = (fval1 / centr) / centr;
    double resg = gkp_k15_wg[7] * fc;
    double resk = gkp_k15_wgk[7] * fc;
    double rabs = fabs(resk);
    for (int j = 0; j < 7; ++j) {
        const double absc = hlgth * gkp_k15_xgk[j];
        const double absc1 = centr - absc;
        const double absc2 = centr + absc;
        const double tabsc1 = boun + dinf * (1.0 - absc1) / absc1;
        const double tabsc2 = boun + dinf * (1.0 - absc2) / absc2;
        double f1 = f(tabsc1, ctx);
        double f2 = f(tabsc2, ctx);
        if (inf == 2) f1 = f1 + f(-tabsc1, ctx);
        if (inf == 2) f2 = f2 + f(-tabsc2, ctx);
        f1 = (f1 / absc1) / absc1;
        f2 = (f2 / absc2) / absc2;
        fv1[j] = f1;
        fv2[j] = f2;
        const double fsum = f1 + f2;
        resg = resg + gkp_k15_wg[j] * fsum;
        resk = resk + gkp_k15_wgk[j] * fsum;
        rabs = rabs + gkp_k15_wgk[j] * (fabs(f1) + fabs(f2));
    }
    const double reskh = resk * 0.5;
    double rasc = gkp_k15_wgk[7] * fabs(fc - reskh);
    for (int j = 0; j < 7; ++j)
        rasc = rasc + gkp_k15_wgk[j] * (fabs(fv1[j] - reskh) + fabs(fv2[j] - reskh));
    *result = resk * hlgth;
    rasc = rasc * hlgth;
    rabs = rabs * hlgth;
    double err = fabs((resk - resg) * hlgth);
    if (rasc != 0.0 && err != 0.0)
        err = rasc * fmin(1.0, pow(200.0 * err / rasc, 1.5));
    if (rabs > uflow / (50.0 * epmach))
        err = fmax((epmach * 50.0) * rabs, err);
    *abserr = err;
    *resabs = rabs;
    *resasc = rasc;
}

/* ---- dqpsrt: keep the descending ordering of error estimates (1-based) ---- */
static void dqpsrt(int limit, int last, int *maxerr, double *ermax, const double *elist,
                   int *iord, int *nrmax)
{
    int i, k, jbnd, jupbn, ibeg, isucc;
    double errmax, errmin;
    if (last <= 2) {
        iord[1] = 1;
        iord[2] = 2;
        goto done;
    }
    errmax = elist[*maxerr];
    {
        const int ido = *nrmax - 1;
        for (i = 1; i <= ido; ++i) {
            isucc = iord[*nrmax - 1];
            if (errmax <= elist[isucc]) break;
            iord[*nrmax] = isucc;
            *nrmax = *nrmax - 1;
        }
    }
    jupbn = last;
    if (last > (limit / 2 + 2)) jupbn = limit + 3 - last;
    errmin = elist[last];
    jbnd = jupbn - 1;
    ibeg = *nrmax + 1;
    for (i = ibeg; i <= jbnd; ++i) {
        isucc = iord[i];
        if (errmax >= elist[isucc]) goto insert_max;
        iord[i - 1] = isucc;
    }
    iord[jbnd] = *maxerr;
    iord[jupbn] = last;
    goto done;
insert_max:
    iord[i - 1] = *maxerr;
    k = jbnd;
    for (int j = i; j <= jbnd; ++j) {
        isucc = iord[k];
        if (errmin < elist[isucc]) {
            iord[k + 1] = last;
            goto done;
        }
        iord[k + 1] = isucc;
        k = k - 1;
    }
    iord[i] = last;
done:
    *maxerr = iord[*nrmax];
    *ermax = elist[*maxerr];
}

/* ---- dqelg: epsilon algorithm (1-based epstab[1..52], res3la[1..3]) ---- */
static void dqelg(int *n, double *epstab, double *result, double *abserr, double *res3la,
                  int *nres)
{
    const double epmach = DBL_EPSILON, oflow = DBL_MAX;
    *nres = *nres + 1;
    *abserr = oflow;
    *result = epstab[*n];
    if (*n < 3) goto finish;
    {
        const int limexp = 50;
        epstab[*n + 2] = epstab[*n];
        const int newelm = (*n - 1) / 2;
        epstab[*n] = oflow;
        const int num = *n;
        int k1 = *n;
        for (int i = 1; i <= newelm; ++i) {
            const int k2 = k1 - 1;
            const int k3 = k1 - 2;
            double res = epstab[k1 + 2];
            const double e0 = epstab[k3];
            const double e1 = epstab[k2];
            const double e2 = res;
            const double e1abs = fabs(e1);
            const double delta2 = e2 - e1;
            const double err2 = fabs(delta2);
            const double tol2 = fmax(fabs(e2), e1abs) * epmach;
            const double delta3 = e1 - e0;
            const double err3 = fabs(delta3);
            const double tol3 = fmax(e1abs, fabs(e0)) * epmach;
            if (!(err2 > tol2 || err3 > tol3)) {
                /* e0, e1 and e2 are equal to within machine accuracy */
                *result = res;
                *abserr = err2 + err3;
                goto finish;
            }
            const double e3 = epstab[k1];
            epstab[k1] = e1;
            const double delta1 = e1 - e3;
            const double err1 = fabs(delta1);
            const double tol1 = fmax(e1abs, fabs(e3)) * epmach;
            double ss;
            if (err1 <= tol1 || err2 <= tol2 || err3 <= tol3) goto truncate;
            ss = 1.0 / delta1 + 1.0 / delta2 - 1.0 / delta3;
            if (!(fabs(ss * e1) > 0.1e-03)) goto truncate;
            res = e1 + 1.0 / ss;
            epstab[k1] = res;
            k1 = k1 - 2;
            {
                const double error = err2 + fabs(res - e2) + err3;
                if (error <= *abserr) {
                    *abserr = error;
                    *result = res;
                }
            }
            continue;
        truncate:
            *n = i + i - 1;
            break;
        }
        if (*n == limexp) *n = 2 * (limexp / 2) - 1;
        int ib = 1;
        if ((num / 2) * 2 == num) ib = 2;
        const int ie = newelm + 1;
        for (int i = 1; i <= ie; ++i) {
            const int ib2 = ib + 2;
            epstab[ib] = epstab[ib2];
            ib = ib2;
        }
        if (num != *n) {
            int indx = num - *n + 1;
            for (int i = 1; i <= *n; ++i) {
                epstab[i] = epstab[indx];
                ++indx;
            }
        }
        if (*nres < 4) {
            res3la[*nres] = *result;
            *abserr = oflow;
        } else {
            *abserr = fabs(*result - res3la[3]) + fabs(*result - res3la[2]) + fabs(*result - res3la[1]);
            res3la[1] = res3la[2];
            res3la[2] = res3la[3];
            res3la[3] = *result;
        }
    }
finish:
    *abserr = fmax(*abserr, 5.0 * epmach * fabs(*result));
}

/* ---- dqagie / dqagi ----------------------------------------------------- */
#define ORC_QP_MAXLIMIT 1000
static double dqagie(orc_integrand f, void *ctx, double bound, int inf, double epsabs,
                     double epsrel, int limit, double *abserr_out, int *neval, int *ier,
                     int *last_out)
{
    const double epmach = DBL_EPSILON;
    double alist[ORC_QP_MAXLIMIT + 2], blist[ORC_QP_MAXLIMIT + 2];
    double rlist[ORC_QP_MAXLIMIT + 2], elist[ORC_QP_MAXLIMIT + 2];
    int iord[ORC_QP_MAXLIMIT + 2];
    double rlist2[53], res3la[4];
    double result = 0.0, abserr = 0.0;
    double area = 0, errsum = 0, errmax = 0, erlarg = 0, ertest = 0, small = 0, correc = 0;
    double area1, area2, area12, error1, error2, erro12, defab1, defab2, resabs, defabs;
    double a1, a2, b1, b2, erlast, errbnd, dres, boun, reseps, abseps;
    const double uflow = DBL_MIN, oflow = DBL_MAX;
    int last = 0, maxerr = 0, nrmax = 0, nres = 0, ktmin = 0, numrl2 = 0, ierro = 0;
    int iroff1 = 0, iroff2 = 0, iroff3 = 0, ksgn, extrap = 0, noext = 0, k;

    *ier = 0;
    *neval = 0;
    alist[1] = 0.0;
    blist[1] = 1.0;
    rlist[1] = 0.0;
    elist[1] = 0.0;
    iord[1] = 0;
    if (epsabs <= 0.0 && epsrel < fmax(50.0 * epmach, 0.5e-28)) *ier = 6;
    if (*ier == 6) {
        *abserr_out = 0.0;
        *last_out = 0;
        return 0.0;
    }
    boun = bound;
    if (inf == 2) boun = 0.0;
    dqk15i(f, ctx, boun, inf, 0.0, 1.0, &result, &abserr, &defabs, &resabs);
    last = 1;
    rlist[1] = result;
    elist[1] = abserr;
    iord[1] = 1;
    dres = fabs(result);
    errbnd = fmax(epsabs, epsrel * dres);
    if (abserr <= 100.0 * epmach * defabs && abserr > errbnd) *ier = 2;
    if (limit == 1) *ier = 1;
    if (*ier != 0 || (abserr <= errbnd && abserr != resabs) || abserr == 0.0) goto L130;

    rlist2[1] = result;
    errmax = abserr;
    maxerr = 1;
    area = result;
    errsum = abserr;
    abserr = oflow;
    nrmax = 1;
    nres = 0;
    ktmin = 0;
    numrl2 = 2;
    extrap = 0;
    noext = 0;
    ierro = 0;
    iroff1 = iroff2 = iroff3 = 0;
    ksgn = -1;
    if (dres >= (1.0 - 50.0 * epmach) * defabs) ksgn = 1;

    for (last = 2; last <= limit; ++last) {
        a1 = alist[maxerr];
        b1 = 0.5 * (alist[maxerr] + blist[maxerr]);
        a2 = b1;
        b2 = blist[maxerr];
        erlast = errmax;
        dqk15i(f, ctx, boun, inf, a1, b1, &area1, &error1, &resabs, &defab1);
        dqk15i(f, ctx, boun, inf, a2, b2, &area2, &error2, &resabs, &defab2);
        area12 = area1 + area2;
        erro12 = error1 + error2;
        errsum = errsum + erro12 - errmax;
        area = area + area12 - rlist[maxerr];
        if (!(defab1 == error1 || defab2 == error2)) {
            if (!(fabs(rlist[maxerr] - area12) > 0.1e-04 * fabs(area12) || erro12 < 0.99 * errmax)) {
                if (extrap) iroff2 = iroff2 + 1;
                if (!extrap) iroff1 = iroff1 + 1;
            }
            if (last > 10 && erro12 > errmax) iroff3 = iroff3 + 1;
        }
        rlist[maxerr] = area1;
        rlist[last] = area2;
        errbnd = fmax(epsabs, epsrel * fabs(area));
        if (iroff1 + iroff2 >= 10 || iroff3 >= 20) *ier = 2;
        if (iroff2 >= 5) ierro = 3;
        if (last == limit) *ier = 1;
        if (fmax(fabs(a1), fabs(b2)) <= (1.0 + 100.0 * epmach) * (fabs(a2) + 1000.0 * uflow)) *ier = 4;
        if (!(error2 > error1)) {
            alist[last] = a2;
            blist[maxerr] = b1;
            blist[last] = b2;
            elist[maxerr] = error1;
            elist[last] = error2;
        } else {
            alist[maxerr] = a2;
            alist[last] = a1;
            blist[last] = b1;
            rlist[maxerr] = area2;
            rlist[last] = area1;
            elist[maxerr] = error2;
            elist[last] = error1;
        }
        dqpsrt(limit, last, &maxerr, &errmax, elist, iord, &nrmax);
        if (errsum <= errbnd) goto L115;
        if (*ier != 0) goto L100;
        if (last == 2) {
            small = 0.375;
            erlarg = errsum;
            ertest = errbnd;
            rlist2[2] = area;
            continue;
        }
        if (noext) continue;
        erlarg = erlarg - erlast;
        if (fabs(b1 - a1) > small) erlarg = erlarg + erro12;
        if (!extrap) {
            if (fabs(blist[maxerr] - alist[maxerr]) > small) continue;
            extrap = 1;
            nrmax = 2;
        }
        if (!(ierro == 3 || erlarg <= ertest)) {
            const int id = nrmax;
            int jupbnd = last;
            int leave = 0;
            if (last > (2 + limit / 2)) jupbnd = limit + 3 - last;
            for (k = id; k <= jupbnd; ++k) {
                maxerr = iord[nrmax];
                errmax = elist[maxerr];
                if (fabs(blist[maxerr] - alist[maxerr]) > small) {
                    leave = 1;
                    break;
                }
                nrmax = nrmax + 1;
            }
            if (leave) continue;
        }
        /* perform extrapolation */
        numrl2 = numrl2 + 1;
        rlist2[numrl2] = area;
        dqelg(&numrl2, rlist2, &reseps, &abseps, res3la, &nres);
        ktmin = ktmin + 1;
        if (ktmin > 5 && abserr < 0.1e-02 * errsum) *ier = 5;
        if (abseps < abserr) {
            ktmin = 0;
            abserr = abseps;
            result = reseps;
            correc = erlarg;
            ertest = fmax(epsabs, epsrel * fabs(reseps));
            if (abserr <= ertest) goto L100;
        }
        if (numrl2 == 1) noext = 1;
        if (*ier == 5) goto L100;
        maxerr = iord[1];
        errmax = elist[maxerr];
        nrmax = 1;
        extrap = 0;
        small = small * 0.5;
        erlarg = errsum;
    }
L100:
    if (abserr == oflow) goto L115;
    if ((*ier + ierro) == 0) goto L110;
    if (ierro == 3) abserr = abserr + correc;
    if (*ier == 0) *ier = 3;
    if (result != 0.0 && area != 0.0) {
        if (abserr / fabs(result) > errsum / fabs(area)) goto L115;
        goto L110;
    }
    if (abserr > errsum) goto L115;
    if (area == 0.0) goto L130;
L110:
    if (ksgn == -1 && fmax(fabs(result), fabs(area)) <= defabs * 0.1e-01) goto L130;
    if (0.1e-01 > (result / area) || (result / area) > 0.1e+03 || errsum > fabs(area)) *ier = 6;
    goto L130;
L115:
    result = 0.0;
    for (k = 1; k <= last; ++k) result = result + rlist[k];
    abserr = errsum;
L130:
    *neval = 30 * last - 15;
    if (inf == 2) *neval = 2 * *neval;
    if (*ier > 2) *ier = *ier - 1;
    *abserr_out = abserr;
    *last_out = last;
    return result;
}

double orc_dqagi(orc_integrand f, void *ctx, double bound, int inf, double epsabs, double epsrel,
                 int limit, double *abserr, int *neval, int *ier, int *last)
{
    *ier = 6;
    *neval = 0;
    *last = 0;
    *abserr = 0.0;
    if (limit < 1 || limit > ORC_QP_MAXLIMIT) return 0.0;
    return dqagie(f, ctx, bound, inf, epsabs, epsrel, limit, abserr, neval, ier, last);
}
