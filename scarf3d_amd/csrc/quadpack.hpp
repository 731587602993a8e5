// QUADPACK integrators used to normalise the SRM spectrum.
//
// The reference integrates its PSD with two QUADPACK routines
// (src/scarflib_srm.f90:959-965): DQNG on the finite band [kmin, kmax]
// (src/quadpack.f90:8137-8528) and DQAGI on (0, inf) (src/quadpack.f90:1177,
// i.e. DQAGIE :712-1176 with DQK15I :6589, DQPSRT :8529 and DQELG :6230).
// The SRM scaling is sqrt(band / all), and a relative error e in it moves
// every field value by e * |f|, so the values QUADPACK *returns* (not the
// exact integrals) are what has to be reproduced. Every floating-point
// expression below keeps the operand order of the Fortran statement it
// restates; the translation unit is compiled with -ffp-contract=off.
#pragma once
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "gk_tables.hpp"

namespace scarf::quad {

struct Result {
    double value = 0.0;
    double abserr = 0.0;
    int neval = 0;
    int ier = 0;
    int last = 0;  // subintervals used (DQAGI only)
};

// DQNG: 21-, 43-, then 87-point rule until the error estimate meets
// max(epsabs, epsrel*|result|).
template <class F>
Result dqng(F &&f, double a, double b, double epsabs, double epsrel)
{
    constexpr double eps = DBL_EPSILON, tiny = DBL_MIN;
    Result out;
    out.ier = 6;
    if (epsabs <= 0.0 && epsrel < std::max(50.0 * eps, 0.5e-28)) return out;

    const double half = 0.5 * (b - a);
    const double abshalf = std::fabs(half);
    const double mid = 0.5 * (b + a);
    const double fmid = f(mid);
    out.neval = 21;
    out.ier = 1;

    double saved[21];  // f(mid+x)+f(mid-x) at the 21-point abscissae, reused by later rules
    double lo1[5], lo2[5], hi1[5], hi2[5];
    double r10 = 0.0, r21 = 0.0, r43 = 0.0, r87 = 0.0, rabs = 0.0, rasc = 0.0;
    double err = 0.0, res = 0.0;

    for (int rule = 0; rule < 3; ++rule) {
        if (rule == 0) {
            r21 = kKronrod21WeightsNew[5] * fmid;
            rabs = kKronrod21WeightsNew[5] * std::fabs(fmid);
            for (int k = 0; k < 5; ++k) {
                const double dx = half * kGauss10Nodes[k];
                const double fp = f(mid + dx), fm = f(mid - dx), s = fp + fm;
                r10 = r10 + kGauss10Weights[k] * s;
                r21 = r21 + kKronrod21WeightsOld[k] * s;
                rabs = rabs + kKronrod21WeightsOld[k] * (std::fabs(fp) + std::fabs(fm));
                saved[k] = s;
                lo1[k] = fp;
                lo2[k] = fm;
            }
            for (int k = 0; k < 5; ++k) {
                const double dx = half * kKronrod21Nodes[k];
                const double fp = f(mid + dx), fm = f(mid - dx), s = fp + fm;
                r21 = r21 + kKronrod21WeightsNew[k] * s;
                rabs = rabs + kKronrod21WeightsNew[k] * (std::fabs(fp) + std::fabs(fm));
                saved[5 + k] = s;
                hi1[k] = fp;
                hi2[k] = fm;
            }
            res = r21 * half;
            rabs = rabs * abshalf;
            const double mean = 0.5 * r21;
            rasc = kKronrod21WeightsNew[5] * std::fabs(fmid - mean);
            for (int k = 0; k < 5; ++k)
                rasc = rasc + kKronrod21WeightsOld[k] * (std::fabs(lo1[k] - mean) + std::fabs(lo2[k] - mean)) +
                       kKronrod21WeightsNew[k] * (std::fabs(hi1[k] - mean) + std::fabs(hi2[k] - mean));
            err = std::fabs((r21 - r10) * half);
            rasc = rasc * abshalf;
        } else if (rule == 1) {
            r43 = kPatterson43WeightsNew[11] * fmid;
            out.neval = 43;
            for (int k = 0; k < 10; ++k) r43 = r43 + saved[k] * kPatterson43WeightsOld[k];
            for (int k = 0; k < 11; ++k) {
                const double dx = half * kPatterson43Nodes[k];
                const double s = f(dx + mid) + f(mid - dx);
                r43 = r43 + s * kPatterson43WeightsNew[k];
                saved[10 + k] = s;
            }
            res = r43 * half;
            err = std::fabs((r43 - r21) * half);
        } else {
            r87 = kPatterson87WeightsNew[22] * fmid;
            out.neval = 87;
            for (int k = 0; k < 21; ++k) r87 = r87 + saved[k] * kPatterson87WeightsOld[k];
            for (int k = 0; k < 22; ++k) {
                const double dx = half * kPatterson87Nodes[k];
                r87 = r87 + kPatterson87WeightsNew[k] * (f(dx + mid) + f(mid - dx));
            }
            res = r87 * half;
            err = std::fabs((r87 - r43) * half);
        }
        if (rasc != 0.0 && err != 0.0) err = rasc * std::min(1.0, std::pow(200.0 * err / rasc, 1.5));
        if (rabs > tiny / (50.0 * eps)) err = std::max((eps * 50.0) * rabs, err);
        if (err <= std::max(epsabs, epsrel * std::fabs(res))) {
            out.ier = 0;
            break;
        }
    }
    out.value = res;
    out.abserr = err;
    return out;
}

namespace detail {

struct K15 {
    double area, err, resabs, resasc;
};

// DQK15I: 15-point Kronrod rule on (a, b) within (0, 1] after the map
// x = bound + (1 - t) / t (inf = 2 adds the mirrored half line).
template <class F>
K15 k15_transformed(F &f, double bound, int inf, double a, double b)
{
    constexpr double eps = DBL_EPSILON, tiny = DBL_MIN;
    const double dinf = static_cast<double>(std::min(1, inf));
    const double mid = 0.5 * (a + b);
    const double half = 0.5 * (b - a);
    const auto mapped = [&](double t) { return bound + dinf * (1.0 - t) / t; };
    const double xm = mapped(mid);
    double fm = f(xm);
    if (inf == 2) fm = fm + f(-xm);
    const double fc = (fm / mid) / mid;
    double rg = kGauss7WeightsInterleaved[7] * fc;
    double rk = kKronrod15Weights[7] * fc;
    double rabs = std::fabs(rk);
    double left[7], right[7];
    for (int j = 0; j < 7; ++j) {
        const double dx = half * kKronrod15Nodes[j];
        const double t1 = mid - dx, t2 = mid + dx;
        const double x1 = mapped(t1), x2 = mapped(t2);
        double f1 = f(x1), f2 = f(x2);
        if (inf == 2) f1 = f1 + f(-x1);
        if (inf == 2) f2 = f2 + f(-x2);
        f1 = (f1 / t1) / t1;
        f2 = (f2 / t2) / t2;
        left[j] = f1;
        right[j] = f2;
        const double s = f1 + f2;
        rg = rg + kGauss7WeightsInterleaved[j] * s;
        rk = rk + kKronrod15Weights[j] * s;
        rabs = rabs + kKronrod15Weights[j] * (std::fabs(f1) + std::fabs(f2));
    }
    const double mean = rk * 0.5;
    double rasc = kKronrod15Weights[7] * std::fabs(fc - mean);
    for (int j = 0; j < 7; ++j)
        rasc = rasc + kKronrod15Weights[j] * (std::fabs(left[j] - mean) + std::fabs(right[j] - mean));
    K15 r;
    r.area = rk * half;
    rasc = rasc * half;
    rabs = rabs * half;
    double err = std::fabs((rk - rg) * half);
    if (rasc != 0.0 && err != 0.0) err = rasc * std::min(1.0, std::pow(200.0 * err / rasc, 1.5));
    if (rabs > tiny / (50.0 * eps)) err = std::max((eps * 50.0) * rabs, err);
    r.err = err;
    r.resabs = rabs;
    r.resasc = rasc;
    return r;
}

// Interval bookkeeping of DQAGIE: 1-based arrays, `order` keeps the
// descending error order maintained by DQPSRT.
struct Intervals {
    explicit Intervals(int limit)
        : lo(limit + 2), hi(limit + 2), area(limit + 2), err(limit + 2), order(limit + 2) {}
    std::vector<double> lo, hi, area, err;
    std::vector<int> order;
};

// DQPSRT: re-establish the descending ordering of error estimates after a
// bisection; updates maxerr/ermax and nrmax.
void sort_errors(int limit, int last, int &maxerr, double &ermax, Intervals &iv, int &nrmax);

// DQELG: epsilon-algorithm extrapolation over table[1..n] (n may shrink).
void epsilon_extrapolate(int &n, double *table, double &result, double &abserr, double *last3,
                         int &nres);

}  // namespace detail

// DQAGI: adaptive integration of f over (bound, +inf) (inf = 1),
// (-inf, bound) (inf = -1) or (-inf, +inf) (inf = 2).
template <class F>
Result dqagi(F &&f, double bound, int inf, double epsabs, double epsrel, int limit)
{
    constexpr double eps = DBL_EPSILON, tiny = DBL_MIN, huge = DBL_MAX;
    Result out;
    out.ier = 6;
    if (limit < 1) return out;
    out.ier = 0;
    if (epsabs <= 0.0 && epsrel < std::max(50.0 * eps, 0.5e-28)) {
        out.ier = 6;
        return out;
    }
    detail::Intervals iv(limit);
    double table[53] = {0};
    double last3[4] = {0};
    const double b0 = (inf == 2) ? 0.0 : bound;

    detail::K15 first = detail::k15_transformed(f, b0, inf, 0.0, 1.0);
    double result = first.area, abserr = first.err;
    const double defabs = first.resabs;
    double resabs = first.resasc;
    int last = 1;
    iv.lo[1] = 0.0;
    iv.hi[1] = 1.0;
    iv.area[1] = result;
    iv.err[1] = abserr;
    iv.order[1] = 1;
    const double dres = std::fabs(result);
    double errbnd = std::max(epsabs, epsrel * dres);
    int ier = 0;
    if (abserr <= 100.0 * eps * defabs && abserr > errbnd) ier = 2;
    if (limit == 1) ier = 1;

    double area = 0, errsum = 0, errmax = 0, erlarg = 0, ertest = 0, small = 0, correc = 0;
    int ksgn = -1;
    enum class Exit { Done, Check, Sum, Final } exit = Exit::Done;

    if (ier != 0 || (abserr <= errbnd && abserr != resabs) || abserr == 0.0) {
        exit = Exit::Done;
    } else {
        table[1] = result;
        errmax = abserr;
        int maxerr = 1;
        area = result;
        errsum = abserr;
        abserr = huge;
        int nrmax = 1, nres = 0, ktmin = 0, numrl2 = 2, ierro = 0;
        int iroff1 = 0, iroff2 = 0, iroff3 = 0;
        bool extrap = false, noext = false;
        if (dres >= (1.0 - 50.0 * eps) * defabs) ksgn = 1;

        exit = Exit::Check;  // reached only through a break below
        for (last = 2; last <= limit; ++last) {
            const double a1 = iv.lo[maxerr];
            const double b1 = 0.5 * (iv.lo[maxerr] + iv.hi[maxerr]);
            const double a2 = b1;
            const double b2 = iv.hi[maxerr];
            const double erlast = errmax;
            const detail::K15 left = detail::k15_transformed(f, b0, inf, a1, b1);
            const detail::K15 right = detail::k15_transformed(f, b0, inf, a2, b2);
            resabs = right.resabs;
            const double area12 = left.area + right.area;
            const double erro12 = left.err + right.err;
            errsum = errsum + erro12 - errmax;
            area = area + area12 - iv.area[maxerr];
            if (!(left.resasc == left.err || right.resasc == right.err)) {
                if (!(std::fabs(iv.area[maxerr] - area12) > 0.1e-04 * std::fabs(area12) ||
                      erro12 < 0.99 * errmax)) {
                    if (extrap) ++iroff2;
                    if (!extrap) ++iroff1;
                }
                if (last > 10 && erro12 > errmax) ++iroff3;
            }
            iv.area[maxerr] = left.area;
            iv.area[last] = right.area;
            errbnd = std::max(epsabs, epsrel * std::fabs(area));
            if (iroff1 + iroff2 >= 10 || iroff3 >= 20) ier = 2;
            if (iroff2 >= 5) ierro = 3;
            if (last == limit) ier = 1;
            if (std::max(std::fabs(a1), std::fabs(b2)) <= (1.0 + 100.0 * eps) * (std::fabs(a2) + 1000.0 * tiny))
                ier = 4;
            if (!(right.err > left.err)) {
                iv.lo[last] = a2;
                iv.hi[maxerr] = b1;
                iv.hi[last] = b2;
                iv.err[maxerr] = left.err;
                iv.err[last] = right.err;
            } else {
                iv.lo[maxerr] = a2;
                iv.lo[last] = a1;
                iv.hi[last] = b1;
                iv.area[maxerr] = right.area;
                iv.area[last] = left.area;
                iv.err[maxerr] = right.err;
                iv.err[last] = left.err;
            }
            detail::sort_errors(limit, last, maxerr, errmax, iv, nrmax);
            if (errsum <= errbnd) {
                exit = Exit::Sum;
                break;
            }
            if (ier != 0) break;
            if (last == 2) {
                small = 0.375;
                erlarg = errsum;
                ertest = errbnd;
                table[2] = area;
                continue;
            }
            if (noext) continue;
            erlarg = erlarg - erlast;
            if (std::fabs(b1 - a1) > small) erlarg = erlarg + erro12;
            if (!extrap) {
                if (std::fabs(iv.hi[maxerr] - iv.lo[maxerr]) > small) continue;
                extrap = true;
                nrmax = 2;
            }
            if (!(ierro == 3 || erlarg <= ertest)) {
                const int first_k = nrmax;
                int upper = last;
                if (last > (2 + limit / 2)) upper = limit + 3 - last;
                bool large_left = false;
                for (int k = first_k; k <= upper; ++k) {
                    maxerr = iv.order[nrmax];
                    errmax = iv.err[maxerr];
                    if (std::fabs(iv.hi[maxerr] - iv.lo[maxerr]) > small) {
                        large_left = true;
                        break;
                    }
                    ++nrmax;
                }
                if (large_left) continue;
            }
            ++numrl2;
            table[numrl2] = area;
            double reseps, abseps;
            detail::epsilon_extrapolate(numrl2, table, reseps, abseps, last3, nres);
            ++ktmin;
            if (ktmin > 5 && abserr < 0.1e-02 * errsum) ier = 5;
            if (abseps < abserr) {
                ktmin = 0;
                abserr = abseps;
                result = reseps;
                correc = erlarg;
                ertest = std::max(epsabs, epsrel * std::fabs(reseps));
                if (abserr <= ertest) break;
            }
            if (numrl2 == 1) noext = true;
            if (ier == 5) break;
            maxerr = iv.order[1];
            errmax = iv.err[maxerr];
            nrmax = 1;
            extrap = false;
            small = small * 0.5;
            erlarg = errsum;
        }

        if (exit == Exit::Check) {
            // label 100 of DQAGIE: decide between the extrapolated result and the sum
            if (abserr == huge) {
                exit = Exit::Sum;
            } else if ((ier + ierro) == 0) {
                exit = Exit::Final;
            } else {
                if (ierro == 3) abserr = abserr + correc;
                if (ier == 0) ier = 3;
                if (result != 0.0 && area != 0.0) {
                    exit = (abserr / std::fabs(result) > errsum / std::fabs(area)) ? Exit::Sum : Exit::Final;
                } else if (abserr > errsum) {
                    exit = Exit::Sum;
                } else if (area == 0.0) {
                    exit = Exit::Done;
                } else {
                    exit = Exit::Final;
                }
            }
        }
        if (exit == Exit::Final) {
            if (!(ksgn == -1 && std::max(std::fabs(result), std::fabs(area)) <= defabs * 0.1e-01)) {
                if (0.1e-01 > (result / area) || (result / area) > 0.1e+03 || errsum > std::fabs(area)) ier = 6;
            }
        } else if (exit == Exit::Sum) {
            result = 0.0;
            for (int k = 1; k <= last; ++k) result = result + iv.area[k];
            abserr = errsum;
        }
    }
    out.neval = 30 * last - 15;
    if (inf == 2) out.neval = 2 * out.neval;
    if (ier > 2) ier = ier - 1;
    out.ier = ier;
    out.value = result;
    out.abserr = abserr;
    out.last = last;
    return out;
}

}  // namespace scarf::quad
