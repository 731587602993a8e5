// FIM constants on the host, templated on the library's `real` so that the
// internal grid size, offsets and spectrum constants round as in the
// reference's single- or double-precision build. Compiled with
// -ffp-contract=off: npts and off_axis decide which internal point every
// output point takes.
#include <cmath>
#include <string>

#include "fim.hpp"

namespace scarf {
namespace {

template <class Real>
constexpr Real kPi = static_cast<Real>(3.141592653589793);  // src/scarflib_common.f90:94

// Fortran NINT: nearest integer, halfway cases away from zero
template <class Real>
long nint(Real x)
{
    return std::lround(x);
}

}  // namespace

// scarf3d_structured_fim / scarf3d_unstructured_fim with DIP off (the
// default build, Makefile.inc: DIP = no): no rotation, bar = 0.
//   npts, offset, off_axis: src/scarflib_fim.f90:805-835
//   info(1:2): :846-849
//   dk, kc, filter, nu, const: compute_spectrum :2440-2482
//   scaling: :951
//   points of interest on the internal grid: :1049-1058
template <class Real>
int fim_setup(const SrmConfig &cfg, FimSetup<Real> &out, std::string &err)
{
    const Real pi = kPi<Real>;
    const Real dh = static_cast<Real>(cfg.dh), ds = static_cast<Real>(cfg.ds);
    const Real cl[3] = {static_cast<Real>(cfg.cl[0]), static_cast<Real>(cfg.cl[1]), static_cast<Real>(cfg.cl[2])};
    const int n = cl[2] > Real(0) ? 3 : 2;
    if (cfg.acf != 0 && cfg.acf != 1) {
        err = "acf must be 0 (von Karman) or 1 (Gaussian)";
        return -1;
    }
    if (cfg.acf == 0 && !(cfg.hurst > 0)) {
        err = "von Karman spectrum needs hurst > 0";
        return -1;
    }
    if (!(dh > Real(0)) || !(ds > Real(0))) {
        err = "grid steps must be positive";
        return -1;
    }
    for (int i = 0; i < n; ++i)
        if (!(cl[i] > Real(0))) {
            err = "correlation lengths must be positive";
            return -1;
        }

    const Real half = dh * Real(0.5);
    for (int i = 0; i < 3; ++i) {
        const Real lo = static_cast<Real>(cfg.nc[i]), hi = static_cast<Real>(cfg.fc[i]);
        long np = nint<Real>((((hi + half) - lo) + half) / dh) + 1;
        long offset = nint<Real>(cl[i] / dh);
        if (cfg.pad != 1) offset = 0;
        np += 2 * offset;
        if (np % 2 != 0) np += 1;
        Real off = lo - (static_cast<Real>(offset) + Real(0.5)) * dh;
        if (n == 2 && i == 2) {
            off = -off;
            np = 1;
        }
        if (np < 1 || np > (1L << 24)) {
            err = "internal FIM grid size out of range along axis " + std::to_string(i + 1);
            return -1;
        }
        out.npts[i] = static_cast<int>(np);
        out.off_axis[i] = off;
    }
    out.info1 = 0;
    out.info2 = 0;
    for (int i = 0; i < n; ++i) {
        if (out.npts[i] <= nint<Real>(((Real(2) * pi) * cl[i]) / dh)) out.info1 = 1;
        if (dh > cl[i] / Real(2)) out.info2 = 1;
    }

    out.ndim = n;
    out.dh = dh;
    out.ds = ds;
    for (int i = 0; i < 3; ++i) {
        out.dk[i] = (Real(2) * pi) / (static_cast<Real>(out.npts[i]) * dh);
        out.cl[i] = cl[i];
    }
    out.kc = pi / ds;
    out.filter = ds != dh;
    out.acf = cfg.acf;
    const Real hurst = static_cast<Real>(cfg.hurst), sigma = static_cast<Real>(cfg.sigma);
    out.nu = hurst + static_cast<Real>(n) / Real(2);
    Real prod = cl[0];
    for (int i = 1; i < n; ++i) prod = prod * cl[i];
    const Real pin = std::pow(pi, static_cast<Real>(n) / Real(2));
    if (cfg.acf == 0)
        out.cnst = static_cast<Real>(1 << n) * pin * std::tgamma(out.nu) * (sigma * sigma) * prod / std::tgamma(hurst);
    else
        out.cnst = pin * (sigma * sigma) * prod;
    out.sigma = sigma;

    Real dhn = dh;
    for (int i = 1; i < n; ++i) dhn = dhn * dh;  // dh**n
    out.scaling = Real(1) / std::sqrt(static_cast<Real>(out.npts[0]) * static_cast<Real>(out.npts[1]) *
                                      static_cast<Real>(out.npts[2]) * dhn);

    // pt(k) = sum_j matrix(k, j) (poi(j) - off_axis(j)) + bar(k) with the
    // identity matrix and bar = 0: poi(k) - off_axis(k) for k <= n, 0 beyond
    out.pt.assign(size_t(cfg.npoi) * 3, Real(0));
    for (int q = 0; q < cfg.npoi; ++q)
        for (int k = 0; k < n && k < cfg.nd; ++k)
            out.pt[size_t(q) * 3 + k] = static_cast<Real>(cfg.poi[size_t(q) * cfg.nd + k]) - out.off_axis[k];
    return 0;
}

template int fim_setup<float>(const SrmConfig &, FimSetup<float> &, std::string &);
template int fim_setup<double>(const SrmConfig &, FimSetup<double> &, std::string &);

}  // namespace scarf
