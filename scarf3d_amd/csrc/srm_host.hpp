// Host half of the SRM: configuration, setup (wavenumber band, rotation,
// spectral normalisation) and the harmonic draw. Everything that is
// O(points x harmonics) runs on the GPU (srm_device.hpp); this file produces
// the harmonic table the kernels consume.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace scarf {

// Parameters held by the object layer after initialize
// (scarf_obj, src/scarflib.f90:67-87, defaults :132-189 / :229-294).
// Real-valued members hold values of the library's `real` type.
struct SrmConfig {
    int nd = 3;               // dimension given at initialize
    bool structured = true;
    int fs[3] = {1, 1, 1};    // first/last global index, 1-based
    int fe[3] = {1, 1, 1};
    double ds = 0, dh = 0;
    int acf = 0;              // 0 von Karman, 1 Gaussian
    double cl[3] = {0, 0, 0};
    double sigma = 0, hurst = 0;
    double nc[3] = {0, 0, 0}, fc[3] = {0, 0, 0};
    double alpha = 0, beta = 0, gamma = 0;
    double mute = -999, taper = -999;
    int rescale = 0, pad = 0, method = 0;
    int nharm = 10000;        // compile-time PARAMETER in the reference (src/scarflib_srm.f90:90)
    std::vector<double> poi;  // (nd, npoi) column-major, copied at initialize
    int npoi = 0;
    long npts = 0;            // unstructured point count
};

// Derived quantities of scarf3d_*_srm (src/scarflib_srm.f90:617-720).
struct SrmSetup {
    int ndim = 3;
    double kmin = 0, kmax = 0, kc = 0;
    double info1 = 0, info2 = 0;
    double scaling = 0;
    double band = 0, all = 0;  // intg(kmin, kmax), intg()
    double matrix[9] = {};     // column-major 3x3 rotation
    double bar[3] = {};
    double nu = 0, cnst = 1;
};

// Harmonics drawn for one seed. v/a/b hold `real` values widened to double;
// used[l] == 0 marks harmonics dropped by u > kc (they still count in the
// 1/sqrt(nharm) scaling).
struct HarmonicDraw {
    std::vector<double> v1, v2, v3, a, b;
    std::vector<uint8_t> used;
    long uniforms = 0;   // number of uniform01 values consumed
    int n_used = 0;
};

// Per-harmonic record the device consumes, for used harmonics only:
//   theta_n(p) = w_n . p + phase_n  with  w_n = M^T v_n  (rad per m)
//   f(p) = scaling * sum_n amp_n cos(theta_n(p))
// where phase_n = v_n . bar - atan2(a_n, b_n) and amp_n = hypot(a_n, b_n), so
// that a sin(t) + b cos(t) = amp cos(t - atan2(a, b)). Stored structure of
// arrays, FP64.
struct HarmonicTable {
    std::vector<double> wx, wy, wz, phase, amp;
    int count() const { return static_cast<int>(amp.size()); }
};

template <class Real>
int srm_setup(const SrmConfig &cfg, SrmSetup &out, std::string &err);

// The harmonic draw of one seed as a resumable walk: draw_until(m) runs the
// reference's draw loop until m harmonics have been kept (or every harmonic
// has been drawn), so that a caller can hand the first harmonics to the
// device while the rest are still drawn. Where the walk is cut never
// changes the harmonics: srm_draw is one draw_until to the end.
// threads <= 0: default_draw_threads(); block_pairs <= 0: 2^18 pairs per
// block at most (smaller blocks only change how the work is cut, never the
// harmonics).
class HarmonicSampler {
public:
    template <class Real>
    static std::unique_ptr<HarmonicSampler> create(const SrmConfig &cfg, const SrmSetup &setup, int seed,
                                                   int threads = 0, long block_pairs = 0);
    virtual ~HarmonicSampler() = default;
    // returns the number of kept harmonics so far
    virtual int draw_until(int kept) = 0;
    // every harmonic 1..nharm has been drawn
    virtual bool finished() const = 0;
    // harmonics drawn so far: v/a/b/used for indices [0, position())
    virtual const HarmonicDraw &draw() const = 0;
    virtual int position() const = 0;
};

template <class Real>
HarmonicDraw srm_draw(const SrmConfig &cfg, const SrmSetup &setup, int seed, int threads = 0, long block_pairs = 0);

// SCARF_DRAW_THREADS if set, else min(16, hardware threads)
int default_draw_threads();

HarmonicTable build_table(const SrmSetup &setup, const HarmonicDraw &draw);
// the kept harmonics among indices [l0, l1) appended to t
void append_table(const SrmSetup &setup, const HarmonicDraw &draw, int l0, int l1, HarmonicTable &t);

}  // namespace scarf
