// Non-template parts of the QUADPACK restatement (see quadpack.hpp).
#include "quadpack.hpp"

namespace scarf::quad::detail {

// DQPSRT (src/quadpack.f90:8529-8662)
void sort_errors(int limit, int last, int &maxerr, double &ermax, Intervals &iv, int &nrmax)
{
    std::vector<int> &ord = iv.order;
    const std::vector<double> &e = iv.err;
    if (last <= 2) {
        ord[1] = 1;
        ord[2] = 2;
    } else {
        const double emax = e[maxerr];
        // move the bisected interval's successor up while it is larger
        const int steps = nrmax - 1;
        for (int s = 0; s < steps; ++s) {
            const int succ = ord[nrmax - 1];
            if (emax <= e[succ]) break;
            ord[nrmax] = succ;
            --nrmax;
        }
        int top = last;
        if (last > (limit / 2 + 2)) top = limit + 3 - last;
        const double emin = e[last];
        const int bnd = top - 1;
        int i = nrmax + 1;
        for (; i <= bnd; ++i) {
            const int succ = ord[i];
            if (emax >= e[succ]) break;
            ord[i - 1] = succ;
        }
        if (i > bnd) {
            ord[bnd] = maxerr;
            ord[top] = last;
        } else {
            ord[i - 1] = maxerr;
            int k = bnd;
            bool placed = false;
            for (int j = i; j <= bnd; ++j) {
                const int succ = ord[k];
                if (emin < e[succ]) {
                    ord[k + 1] = last;
                    placed = true;
                    break;
                }
                ord[k + 1] = succ;
                --k;
            }
            if (!placed) ord[i] = last;
        }
    }
    maxerr = ord[nrmax];
    ermax = e[maxerr];
}

// DQELG (src/quadpack.f90:6230-6416): Wynn's epsilon algorithm on
// table[1..n] (table has 52 usable entries plus 2 of headroom); last3 holds
// the three previous results (1-based).
void epsilon_extrapolate(int &n, double *table, double &result, double &abserr, double *last3,
                         int &nres)
{
    constexpr double eps = DBL_EPSILON, huge = DBL_MAX;
    ++nres;
    abserr = huge;
    result = table[n];
    if (n >= 3) {
        constexpr int limexp = 50;
        table[n + 2] = table[n];
        const int newelm = (n - 1) / 2;
        table[n] = huge;
        const int num = n;
        int k1 = n;
        bool converged = false;
        for (int i = 1; i <= newelm; ++i) {
            const int k2 = k1 - 1, k3 = k1 - 2;
            double res = table[k1 + 2];
            const double e0 = table[k3], e1 = table[k2], e2 = res;
            const double e1abs = std::fabs(e1);
            const double delta2 = e2 - e1, err2 = std::fabs(delta2);
            const double tol2 = std::max(std::fabs(e2), e1abs) * eps;
            const double delta3 = e1 - e0, err3 = std::fabs(delta3);
            const double tol3 = std::max(e1abs, std::fabs(e0)) * eps;
            if (!(err2 > tol2 || err3 > tol3)) {
                result = res;
                abserr = err2 + err3;
                converged = true;
                break;
            }
            const double e3 = table[k1];
            table[k1] = e1;
            const double delta1 = e1 - e3, err1 = std::fabs(delta1);
            const double tol1 = std::max(e1abs, std::fabs(e3)) * eps;
            bool irregular = (err1 <= tol1 || err2 <= tol2 || err3 <= tol3);
            double ss = 0.0;
            if (!irregular) {
                ss = 1.0 / delta1 + 1.0 / delta2 - 1.0 / delta3;
                irregular = !(std::fabs(ss * e1) > 0.1e-03);
            }
            if (irregular) {
                n = i + i - 1;
                break;
            }
            res = e1 + 1.0 / ss;
            table[k1] = res;
            k1 = k1 - 2;
            const double error = err2 + std::fabs(res - e2) + err3;
            if (error <= abserr) {
                abserr = error;
                result = res;
            }
        }
        if (!converged) {
            if (n == limexp) n = 2 * (limexp / 2) - 1;
            int ib = ((num / 2) * 2 == num) ? 2 : 1;
            const int ie = newelm + 1;
            for (int i = 1; i <= ie; ++i) {
                table[ib] = table[ib + 2];
                ib += 2;
            }
            if (num != n) {
                int idx = num - n + 1;
                for (int i = 1; i <= n; ++i) table[i] = table[idx++];
            }
            if (nres < 4) {
                last3[nres] = result;
                abserr = huge;
            } else {
                abserr = std::fabs(result - last3[3]) + std::fabs(result - last3[2]) + std::fabs(result - last3[1]);
                last3[1] = last3[2];
                last3[2] = last3[3];
                last3[3] = result;
            }
        }
    }
    abserr = std::max(abserr, 5.0 * eps * std::fabs(result));
}

}  // namespace scarf::quad::detail
