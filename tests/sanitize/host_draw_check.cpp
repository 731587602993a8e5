// Host half of the library (setup, QUADPACK, the resumable parallel draw,
// the device record) built with AddressSanitizer and UndefinedBehavior-
// Sanitizer by tests/test_sanitize.py: g++ on the product's own sources, no
// GPU code. Runs the setup and the draw of a few configurations (both `real`
// types, several thread counts and walk cuts) and prints every harmonic, so
// the test can check the sanitized build against the library.
//   usage: host_draw_check     (prints "case <i> <uniforms> <kept>" then
//                               "h <l> <v1> <v2> <v3> <a> <b> <used>" lines)
#include <cstdio>
#include <string>
#include <vector>

#include "srm_host.hpp"

using scarf::SrmConfig;

namespace {

SrmConfig grid(int n, double ds, double dh, int acf, double cl, double hurst, int nharm, double alpha = 0)
{
    SrmConfig c;
    c.nd = 3;
    c.structured = true;
    for (int i = 0; i < 3; ++i) {
        c.fs[i] = 1;
        c.fe[i] = n;
        c.cl[i] = cl;
        c.nc[i] = 0;
        c.fc[i] = (n - 1) * ds;
    }
    c.ds = ds;
    c.dh = dh;
    c.acf = acf;
    c.sigma = 0.05;
    c.hurst = hurst;
    c.alpha = alpha;
    c.method = 1;
    c.nharm = nharm;
    return c;
}

template <class Real>
int run_case(int id, const SrmConfig &c, int seed, int threads, int cut)
{
    scarf::SrmSetup s;
    std::string err;
    if (scarf::srm_setup<Real>(c, s, err) != 0) {
        std::fprintf(stderr, "case %d: %s\n", id, err.c_str());
        return 1;
    }
    std::unique_ptr<scarf::HarmonicSampler> w = scarf::HarmonicSampler::create<Real>(c, s, seed, threads, 4099);
    for (int kept = cut; !w->finished(); kept += cut) w->draw_until(kept);
    const scarf::HarmonicDraw &d = w->draw();
    const scarf::HarmonicTable t = scarf::build_table(s, d);
    std::printf("case %d %ld %d %d\n", id, d.uniforms, d.n_used, t.count());
    for (int l = 0; l < c.nharm; ++l)
        std::printf("h %d %.17g %.17g %.17g %.17g %.17g %d\n", l, d.v1[l], d.v2[l], d.v3[l], d.a[l], d.b[l],
                    int(d.used[l]));
    return 0;
}

}  // namespace

int main()
{
    int rc = 0;
    rc |= run_case<float>(0, grid(64, 20.0, 20.0, 0, 5000.0, 0.1, 700), 1234, 4, 128);
    rc |= run_case<double>(1, grid(64, 20.0, 8.0, 0, 3000.0, 0.3, 600, 30.0), 77, 3, 7);
    rc |= run_case<float>(2, grid(40, 50.0, 50.0, 1, 300.0, 0.0, 500), 5, 1, 1000);
    // an error path: von Karman without a Hurst exponent
    scarf::SrmSetup s;
    std::string err;
    if (scarf::srm_setup<float>(grid(16, 50.0, 50.0, 0, 2000.0, 0.0, 64), s, err) == 0 || err.empty()) rc |= 2;
    return rc;
}
