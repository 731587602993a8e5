// SRM setup and harmonic draw on the host, templated on the library's `real`
// (float for libscarf3df, double for libscarf3d) so that every rounding of
// the reference's single- or double-precision build is reproduced. Compiled
// with -ffp-contract=off: an FMA could flip an accept/reject decision and
// shift every later harmonic.
#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <thread>

#include "quadpack.hpp"
#include "srm_host.hpp"
#include "worker_pool.hpp"
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

namespace {

// Pairs of uniforms (one srng call each) from index `start` of a seed's
// stream, with the rejection test of every pair evaluated in advance. Whether
// a pair is accepted depends on its two values only, never on where the
// draw loop stands, so the test can run ahead of the loop, in parallel.
template <class Real>
struct PairBlock {
    long start = 0, end = 0;      // pair indices [start, end)
    std::vector<int64_t> raw;     // 2 generator outputs per pair
    std::vector<uint8_t> accept;  // rejection test of the pair as a trial
};

template <class Real>
struct DrawContext {
    Spectrum<Real> spec;
    Real kmin, kmax, kc, cnst;
    unsigned long seed;

    // the trial of src/scarflib_srm.f90:737-743 on one srng pair
    Real radial(Real r0) const { return r0 * (kmax - kmin) + kmin; }
    Real acceptance(Real u) const
    {
        return static_cast<Real>(spec(static_cast<double>(u)) * static_cast<double>(cnst));
    }
    bool accepts(Real r0, Real r1) const { return r1 < acceptance(radial(r0)); }
    static Real uniform(int64_t raw) { return static_cast<Real>(Yarn2::to_uniform01(raw)); }

    // Squeeze: bound[b] is at least the acceptance of every r0 in
    // [b/kBins, (b+1)/kBins], so a trial with r1 >= bound[b] is rejected
    // without evaluating the spectrum; the others take the exact test. The
    // decisions are the exact test's. srm_fn * const = u^d psd(u) has one
    // maximum, at u^2 = d / (2 nu - d) (von Karman) or 2d (Gaussian), with
    // d = 2 in 3D and 1 in 2D; on a bin away from it the exact function is
    // monotone, so its larger end bounds the bin, and the computed
    // acceptance (a few roundings of `real` and one libm pow or exp, well
    // under 1e-5 relative) stays below that end's value times 1 + 1e-4.
    // Bins near the maximum have no bound.
    static constexpr int kBins = 4096;
    std::vector<Real> bound;
    void make_bounds(int ndim)
    {
        const double d = ndim == 3 ? 2.0 : 1.0;
        const double nu = static_cast<double>(spec.nu);
        const double up2 = spec.acf == 0 ? (2.0 * nu > d ? d / (2.0 * nu - d) : -1.0) : 2.0 * d;
        const double up = up2 > 0.0 ? std::sqrt(up2) : 0.0;
        bound.assign(kBins, std::numeric_limits<Real>::infinity());
        for (int b = 0; b < kBins; ++b) {
            const Real ulo = radial(static_cast<Real>(b) / static_cast<Real>(kBins));
            const Real uhi = radial(static_cast<Real>(b + 1) / static_cast<Real>(kBins));
            Real end;
            if (static_cast<double>(ulo) >= up * (1.0 + 1e-3)) end = acceptance(ulo);       // decreasing
            else if (static_cast<double>(uhi) <= up * (1.0 - 1e-3)) end = acceptance(uhi);  // increasing
            else continue;
            if (std::isfinite(static_cast<double>(end)) && end >= Real(0))
                bound[b] = static_cast<Real>(static_cast<double>(end) * (1.0 + 1e-4));
        }
    }
    bool accepts_fast(Real r0, Real r1) const
    {
        const int b = std::min(static_cast<int>(r0 * static_cast<Real>(kBins)), kBins - 1);
        if (!(r1 < bound[b])) return false;
        return accepts(r0, r1);
    }

    void fill_range(PairBlock<Real> &blk, long first, long last) const
    {
        Yarn2 g;
        g.seed(seed);
        g.discard(2 * static_cast<uint64_t>(first));
        for (long p = first; p < last; ++p) {
            const int64_t a = g(), b = g();
            const size_t i = static_cast<size_t>(p - blk.start);
            blk.raw[2 * i] = a;
            blk.raw[2 * i + 1] = b;
            blk.accept[i] = accepts_fast(uniform(a), uniform(b)) ? 1 : 0;
        }
    }
};


}  // namespace

int default_draw_threads()
{
    if (const char *e = std::getenv("SCARF_DRAW_THREADS")) {
        const int v = std::atoi(e);
        if (v >= 1) return std::min(v, 256);
    }
    const unsigned hw = std::thread::hardware_concurrency();
    return static_cast<int>(std::max(1u, std::min(hw, 16u)));
}

namespace {

// A worker pool kept for the process, so that a draw does not start and join
// its threads on every call. One draw at a time uses it; a draw that finds it
// busy (the library is not re-entrant, but a caller may still overlap calls)
// makes a pool of its own.
struct SharedPool {
    std::mutex mu;
    std::unique_ptr<WorkerPool> pool;
};
SharedPool &shared_pool()
{
    static SharedPool *p = new SharedPool;  // never destroyed: its threads live until the process ends
    return *p;
}

// Harmonic draw (src/scarflib_srm.f90:731-778): rejection sampling of the
// radial wavenumber against srm_fn*const, a uniform direction, and a, b from
// the cosine branch of Box-Muller with one srng call each.
//
// The loop consumes the seed's stream pair by pair: trials until the first
// accepted pair q, then pair q+1 (direction) and, for a kept harmonic
// (u <= kc), pairs q+2 and q+3 (a and b). Here the stream is produced and
// every pair's trial evaluated ahead of that walk, in blocks, by `threads`
// workers that each jump to their sub-range (Yarn2::discard); the walk then
// only looks up the next accepted pair. The pairs consumed, and so every
// harmonic, are those of the one-pair-at-a-time loop.
template <class Real>
class Sampler final : public HarmonicSampler {
public:
    Sampler(const SrmConfig &cfg, const SrmSetup &s, int seed, int threads, long block_pairs)
        : n_(cfg.nharm), ndim_(s.ndim)
    {
        d_.v1.assign(n_, 0.0);
        d_.v2.assign(n_, 0.0);
        d_.v3.assign(n_, 0.0);
        d_.a.assign(n_, 0.0);
        d_.b.assign(n_, 0.0);
        d_.used.assign(n_, 0);
        ctx_.spec.acf = cfg.acf;
        ctx_.spec.ndim = s.ndim;
        ctx_.spec.nu = static_cast<Real>(s.nu);
        ctx_.kmin = static_cast<Real>(s.kmin);
        ctx_.kmax = static_cast<Real>(s.kmax);
        ctx_.kc = static_cast<Real>(s.kc);
        ctx_.cnst = static_cast<Real>(s.cnst);
        ctx_.seed = static_cast<unsigned long>(seed);  // set_seed: unsigned long myseed = seed
        ctx_.make_bounds(s.ndim);
        cl_[0] = static_cast<Real>(cfg.cl[0]);
        cl_[1] = static_cast<Real>(cfg.cl[1]);
        cl_[2] = static_cast<Real>(cfg.cl[2]);
        if (threads <= 0) threads = default_draw_threads();
        max_block_ = block_pairs > 0 ? block_pairs : (1L << 18);
        next_size_ = std::min(max_block_, std::max(64L, 16L * n_));  // first block: a few trials per harmonic
        if (threads > 1) {
            SharedPool &sp = shared_pool();
            lock_ = std::unique_lock<std::mutex>(sp.mu, std::try_to_lock);
            if (lock_.owns_lock()) {
                if (!sp.pool || sp.pool->size() != threads) sp.pool.reset(new WorkerPool(threads));
                pool_ = sp.pool.get();
            } else {
                own_.reset(new WorkerPool(threads));
                pool_ = own_.get();
            }
        }
    }

    bool finished() const override { return l_ >= n_; }
    const HarmonicDraw &draw() const override { return d_; }
    int position() const override { return l_; }

    int draw_until(int kept) override
    {
        const Real pi = kPi<Real>;
        Real r[2];
        while (l_ < n_ && d_.n_used < kept) {
            const int l = l_++;
            long q = -1;
            while (q < 0) {
                if (pos_ < blk_.start || pos_ >= blk_.end) fill(pos_);
                const uint8_t *f = blk_.accept.data();
                for (long p = pos_; p < blk_.end; ++p)
                    if (f[p - blk_.start]) {
                        q = p;
                        break;
                    }
                if (q < 0) pos_ = blk_.end;
            }
            if (q + 3 >= blk_.end) fill(q);  // the trial and the (up to) three pairs after it
            pair(q, r);
            const Real u = ctx_.radial(r[0]);
            pair(q + 1, r);
            if (u > ctx_.kc) {  // dropped, still counted by the scaling
                pos_ = q + 2;
                d_.uniforms = 2 * pos_;
                continue;
            }
            const Real phi = r[0] * Real(2) * pi;
            Real v1, v2, v3;
            if (ndim_ == 3) {
                const Real theta = std::acos(Real(1) - Real(2) * r[1]);
                v1 = u * std::sin(phi) * std::sin(theta) / cl_[0];
                v2 = u * std::cos(phi) * std::sin(theta) / cl_[1];
                v3 = u * std::cos(theta) / cl_[2];
            } else {
                v1 = u * std::cos(phi) / cl_[0];
                v2 = u * std::sin(phi) / cl_[1];
                v3 = 0;
            }
            pair(q + 2, r);
            const Real a = std::sqrt(-(Real(2) * std::log(r[0]))) * std::cos(Real(2) * pi * r[1]);
            pair(q + 3, r);
            const Real b = std::sqrt(-(Real(2) * std::log(r[0]))) * std::cos(Real(2) * pi * r[1]);
            pos_ = q + 4;
            d_.v1[l] = v1;
            d_.v2[l] = v2;
            d_.v3[l] = v3;
            d_.a[l] = a;
            d_.b[l] = b;
            d_.used[l] = 1;
            ++d_.n_used;
            d_.uniforms = 2 * pos_;
        }
        return d_.n_used;
    }

    HarmonicDraw take() { return std::move(d_); }

private:
    // make pairs [first, first + >= 4) resident
    void fill(long first)
    {
        const long count = std::max(next_size_, 4L);
        next_size_ = std::min(max_block_, 2 * count);
        blk_.start = first;
        blk_.end = first + count;
        blk_.raw.resize(2 * static_cast<size_t>(count));
        blk_.accept.resize(static_cast<size_t>(count));
        const int nt = (pool_ && count >= 4096) ? pool_->size() : 1;
        if (nt == 1) {
            ctx_.fill_range(blk_, first, first + count);
            return;
        }
        const long chunk = (count + nt - 1) / nt;
        const std::function<void(int)> job = [&](int t) {
            const long lo = first + t * chunk, hi = std::min(first + count, lo + chunk);
            if (lo < hi) ctx_.fill_range(blk_, lo, hi);
        };
        pool_->run(job);
    }
    void pair(long p, Real r[2]) const
    {
        const size_t i = static_cast<size_t>(p - blk_.start);
        r[0] = DrawContext<Real>::uniform(blk_.raw[2 * i]);
        r[1] = DrawContext<Real>::uniform(blk_.raw[2 * i + 1]);
    }

    const int n_, ndim_;
    Real cl_[3];
    DrawContext<Real> ctx_;
    HarmonicDraw d_;
    PairBlock<Real> blk_;
    long max_block_ = 0, next_size_ = 0;
    long pos_ = 0;  // next pair the loop would read
    int l_ = 0;     // next harmonic index
    std::unique_lock<std::mutex> lock_;
    std::unique_ptr<WorkerPool> own_;
    WorkerPool *pool_ = nullptr;
};

}  // namespace

template <class Real>
std::unique_ptr<HarmonicSampler> HarmonicSampler::create(const SrmConfig &cfg, const SrmSetup &setup, int seed,
                                                         int threads, long block_pairs)
{
    return std::unique_ptr<HarmonicSampler>(new Sampler<Real>(cfg, setup, seed, threads, block_pairs));
}

template <class Real>
HarmonicDraw srm_draw(const SrmConfig &cfg, const SrmSetup &s, int seed, int threads, long block_pairs)
{
    Sampler<Real> w(cfg, s, seed, threads, block_pairs);
    w.draw_until(cfg.nharm);
    return w.take();
}

void append_table(const SrmSetup &s, const HarmonicDraw &d, int l0, int l1, HarmonicTable &t)
{
    const double *m = s.matrix;
    for (int l = l0; l < l1; ++l) {
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
    append_table(s, d, 0, static_cast<int>(d.used.size()), t);
    return t;
}

template int srm_setup<float>(const SrmConfig &, SrmSetup &, std::string &);
template int srm_setup<double>(const SrmConfig &, SrmSetup &, std::string &);
template HarmonicDraw srm_draw<float>(const SrmConfig &, const SrmSetup &, int, int, long);
template HarmonicDraw srm_draw<double>(const SrmConfig &, const SrmSetup &, int, int, long);
template std::unique_ptr<HarmonicSampler> HarmonicSampler::create<float>(const SrmConfig &, const SrmSetup &, int,
                                                                         int, long);
template std::unique_ptr<HarmonicSampler> HarmonicSampler::create<double>(const SrmConfig &, const SrmSetup &, int,
                                                                          int, long);

}  // namespace scarf
