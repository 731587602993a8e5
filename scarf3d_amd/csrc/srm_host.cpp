// SRM setup and harmonic draw on the host, templated on the library's `real`
// (float for libscarf3df, double for libscarf3d) so that every rounding of
// the reference's single- or double-precision build is reproduced. Compiled
// with -ffp-contract=off: an FMA could flip an accept/reject decision and
// shift every later harmonic.
#include <algorithm>
#include <cmath>

#include "quadpack.hpp"
#include "srm_host.hpp"
#include "yarn2.hpp"

namespace scarf {
namespace {

template <class Real>
constexpr Real kPi = static_cast<Real>(3.141592653589793);  // src/scarflib_common.f90:94

// PSD shapes at `real` precision: fn_vk / fn_gs (src/psdf.f90:50-90)
template <class Real>
struct Spectrum {
    int acf = 0;
    int ndim = 3;
    Real nu = 0;

    Real psd(Real x) const
    {
        if (acf == 0) return Real(1) / std::pow(Real(1) + x * x, nu);
        return std::exp(-(Real(0.25) * (x * x)));
    }
    // srm_fn (src/scarflib_srm.f90:975-995): FP64 integrand around a `real` PSD
    double operator()(double x) const
    {
        double r = x * static_cast<double>(psd(static_cast<Real>(x)));
        if (ndim == 3) r = r * x;
        return r;
    }
};

// intg (src/scarflib_srm.f90:935-969)
template <class Real>
Real integrate(const Spectrum<Real> &s, bool finite, Real k0, Real k1)
{
    const quad::Result r = finite ? quad::dqng(s, static_cast<double>(k0), static_cast<double>(k1), 1.0e-8, 1.0e-5)
                                  : quad::dqagi(s, 0.0, 1, 1.0e-8, 1.0e-5, 500);
    return static_cast<Real>(r.value);
}

// srng (src/prng.cpp:152-165): two uniforms rounded to `real`
template <class Real>
inline void two_uniforms(Yarn2 &g, Real r[2])
{
    r[0] = static_cast<Real>(g.uniform01());
    r[1] = static_cast<Real>(g.uniform01());
}

}  // namespace

template <class Real>
int srm_setup(const SrmConfig &cfg, SrmSetup &out, std::string &err)
{
    out = SrmSetup{};
    const Real pi = kPi<Real>;
    const int ndim = (static_cast<Real>(cfg.cl[2]) > Real(0)) ? 3 : 2;
    const Real cl[3] = {static_cast<Real>(cfg.cl[0]), static_cast<Real>(cfg.cl[1]), static_cast<Real>(cfg.cl[2])};
    Real span[3];
    for (int i = 0; i < 3; ++i) span[i] = static_cast<Real>(cfg.fc[i]) - static_cast<Real>(cfg.nc[i]);
    const Real ds = static_cast<Real>(cfg.ds), dh = static_cast<Real>(cfg.dh);
    if (cfg.acf != 0 && cfg.acf != 1) {
        err = "acf must be 0 (von Karman) or 1 (Gaussian): the user-defined PSD (acf=2) has no body";
        return -1;
    }
    if (cfg.acf == 0 && !(static_cast<Real>(cfg.hurst) > Real(0))) {
        err = "von Karman spectrum needs hurst > 0 (hurst = 0 makes the spectral integral diverge)";
        return -1;
    }
    for (int i = 0; i < ndim; ++i) {
        if (!(cl[i] > Real(0)) || !(span[i] > Real(0))) {
            err = "correlation lengths and model extents must be positive";
            return -1;
        }
    }
    if (!(ds > Real(0)) || !(dh > Real(0)) || cfg.nharm < 1) {
        err = "ds, dh and nharm must be positive";
        return -1;
    }

    // band limits (src/scarflib_srm.f90:644-652)
    Real largest = cl[0] / span[0], shortest = cl[0];
    for (int i = 1; i < ndim; ++i) {
        largest = std::max(largest, cl[i] / span[i]);
        shortest = std::min(shortest, cl[i]);
    }
    const Real kmin = Real(2) * pi * largest;
    const Real kmax = pi / dh * shortest;
    const Real kc = pi / ds * shortest;

    // representation warnings (:660-662)
    Real info1 = 0, info2 = 0;
    for (int i = 0; i < ndim; ++i) {
        if (Real(2) * pi / span[i] >= Real(1) / cl[i]) info1 = 1;
        if (ds > cl[i] / Real(2)) info2 = 1;
    }

    // rotation about z (alpha), y (beta), x (gamma) in degrees (:677-691)
    const auto rad = [&](double deg) { return static_cast<Real>(deg) * pi / Real(180); };
    const Real ca = std::cos(rad(cfg.alpha)), sa = std::sin(rad(cfg.alpha));
    const Real cb = std::cos(rad(cfg.beta)), sb = std::sin(rad(cfg.beta));
    const Real cg = std::cos(rad(cfg.gamma)), sg = std::sin(rad(cfg.gamma));
    const Real m[9] = {ca * cb,           sa * cb,           -sb,
                       ca * sb * sg - sa * cg, sa * sb * sg + ca * cg, cb * sg,
                       ca * sb * cg + sa * sg, sa * sb * cg - ca * sg, cb * cg};

    // rotate about the half-span point (:694-700)
    Real half[3], bar[3];
    for (int i = 0; i < 3; ++i) half[i] = (static_cast<Real>(cfg.fc[i]) - static_cast<Real>(cfg.nc[i])) / Real(2);
    for (int r = 0; r < 3; ++r) bar[r] = half[r] - (m[r] * half[0] + m[r + 3] * half[1] + m[r + 6] * half[2]);

    Spectrum<Real> spec;
    spec.acf = cfg.acf;
    spec.ndim = ndim;
    spec.nu = static_cast<Real>(cfg.hurst) + static_cast<Real>(ndim) / Real(2);
    // majorant constant (:706-716); the 2D Gaussian value is unset upstream, 1 bounds it
    Real cnst = Real(1);
    if (cfg.acf == 1 && ndim == 3) cnst = Real(1) / std::sqrt(pi);

    // discrete-to-continuous normalisation (:665, :720)
    Real scaling = static_cast<Real>(cfg.sigma) / std::sqrt(static_cast<Real>(cfg.nharm));
    const Real band = integrate(spec, true, kmin, kmax);
    const Real all = integrate(spec, false, Real(0), Real(0));
    scaling = scaling * std::sqrt(band / all);

    // The rejection loop of the draw only ends if srm_fn*const reaches a
    // usable acceptance somewhere in [kmin, kmax]; upstream it spins forever
    // otherwise (e.g. a Gaussian spectrum on a box much smaller than cl).
    double peak = 0.0;
    for (int q = 0; q <= 4096; ++q) {
        const Real u = kmin + (kmax - kmin) * static_cast<Real>(q) / Real(4096);
        peak = std::max(peak, spec(static_cast<double>(u)) * static_cast<double>(cnst));
    }
    if (!(kmax > kmin) || !(peak > 1.0e-7)) {
        err = "the spectrum has no usable power in the sampled band [kmin, kmax] (peak acceptance " +
              std::to_string(peak) + "); the harmonic draw would not terminate";
        return -1;
    }

    out.ndim = ndim;
    out.kmin = kmin;
    out.kmax = kmax;
    out.kc = kc;
    out.info1 = info1;
    out.info2 = info2;
    out.scaling = scaling;
    out.band = band;
    out.all = all;
    for (int i = 0; i < 9; ++i) out.matrix[i] = m[i];
    for (int i = 0; i < 3; ++i) out.bar[i] = bar[i];
    out.nu = spec.nu;
    out.cnst = cnst;
    return 0;
}

// Harmonic draw (src/scarflib_srm.f90:731-778): rejection sampling of the
// radial wavenumber against srm_fn*const, a uniform direction, and a, b from
// the cosine branch of Box-Muller with one srng call each.
template <class Real>
HarmonicDraw srm_draw(const SrmConfig &cfg, const SrmSetup &s, int seed)
{
    HarmonicDraw d;
    const int n = cfg.nharm;
    d.v1.assign(n, 0.0);
    d.v2.assign(n, 0.0);
    d.v3.assign(n, 0.0);
    d.a.assign(n, 0.0);
    d.b.assign(n, 0.0);
    d.used.assign(n, 0);

    Yarn2 g;
    g.seed(static_cast<unsigned long>(seed));  // set_seed: unsigned long myseed = seed
    Spectrum<Real> spec;
    spec.acf = cfg.acf;
    spec.ndim = s.ndim;
    spec.nu = static_cast<Real>(s.nu);
    const Real pi = kPi<Real>;
    const Real kmin = static_cast<Real>(s.kmin), kmax = static_cast<Real>(s.kmax), kc = static_cast<Real>(s.kc);
    const Real cnst = static_cast<Real>(s.cnst);
    const Real cl0 = static_cast<Real>(cfg.cl[0]), cl1 = static_cast<Real>(cfg.cl[1]), cl2 = static_cast<Real>(cfg.cl[2]);
    long draws = 0;
    Real r[2];
    for (int l = 0; l < n; ++l) {
        Real u;
        for (;;) {
            two_uniforms(g, r);
            draws += 2;
            u = r[0] * (kmax - kmin) + kmin;
            const Real accept = static_cast<Real>(spec(static_cast<double>(u)) * static_cast<double>(cnst));
            if (r[1] < accept) break;
        }
        two_uniforms(g, r);
        draws += 2;
        if (u > kc) continue;  // dropped, still counted by the scaling
        const Real phi = r[0] * Real(2) * pi;
        Real v1, v2, v3;
        if (s.ndim == 3) {
            const Real theta = std::acos(Real(1) - Real(2) * r[1]);
            v1 = u * std::sin(phi) * std::sin(theta) / cl0;
            v2 = u * std::cos(phi) * std::sin(theta) / cl1;
            v3 = u * std::cos(theta) / cl2;
        } else {
            v1 = u * std::cos(phi) / cl0;
            v2 = u * std::sin(phi) / cl1;
            v3 = 0;
        }
        two_uniforms(g, r);
        const Real a = std::sqrt(-(Real(2) * std::log(r[0]))) * std::cos(Real(2) * pi * r[1]);
        two_uniforms(g, r);
        const Real b = std::sqrt(-(Real(2) * std::log(r[0]))) * std::cos(Real(2) * pi * r[1]);
        draws += 4;
        d.v1[l] = v1;
        d.v2[l] = v2;
        d.v3[l] = v3;
        d.a[l] = a;
        d.b[l] = b;
        d.used[l] = 1;
        ++d.n_used;
    }
    d.uniforms = draws;
    return d;
}

HarmonicTable build_table(const SrmSetup &s, const HarmonicDraw &d)
{
    HarmonicTable t;
    const int n = d.n_used;
    t.wx.reserve(n);
    t.wy.reserve(n);
    t.wz.reserve(n);
    t.phase.reserve(n);
    t.amp.reserve(n);
    const double *m = s.matrix;
    for (size_t l = 0; l < d.used.size(); ++l) {
        if (!d.used[l]) continue;
        const double v[3] = {d.v1[l], d.v2[l], d.v3[l]};
        // w = M^T v: theta = sum_r v_r ((M p)_r + bar_r) = (M^T v) . p + v . bar
        t.wx.push_back(v[0] * m[0] + v[1] * m[1] + v[2] * m[2]);
        t.wy.push_back(v[0] * m[3] + v[1] * m[4] + v[2] * m[5]);
        t.wz.push_back(v[0] * m[6] + v[1] * m[7] + v[2] * m[8]);
        const double vbar = v[0] * s.bar[0] + v[1] * s.bar[1] + v[2] * s.bar[2];
        t.phase.push_back(vbar - std::atan2(d.a[l], d.b[l]));
        t.amp.push_back(std::hypot(d.a[l], d.b[l]));
    }
    return t;
}

template int srm_setup<float>(const SrmConfig &, SrmSetup &, std::string &);
template int srm_setup<double>(const SrmConfig &, SrmSetup &, std::string &);
template HarmonicDraw srm_draw<float>(const SrmConfig &, const SrmSetup &, int);
template HarmonicDraw srm_draw<double>(const SrmConfig &, const SrmSetup &, int);

}  // namespace scarf
