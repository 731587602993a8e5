// C++ drop-in driver: the SRM and FIM blocks of the reference's C++ example
// (examples/c++/driver.cpp:160-240, :270-300, :410-440, :540-600) on one
// process and without MPI: Scarf3D::Initialize<srm> (the constructor sets
// method = 1 and initializes; the destructor finalizes and resets
// Scarf3D::options) and Scarf3D::Initialize<fim> (method left at its
// default 0, FIM), compiled against include/scarf3d.h.
//
//   1. SRM, a 2D Cartesian grid (nx x ny) with Hurst 0.25;
//   2. SRM, a 3D Cartesian grid with Hurst 0.25, alpha = 30, two points of
//      interest, taper 2500 m and mute 500 m;
//   3. SRM, the 3D grid points as a point list;
//   4. FIM, the 3D grid with the options of 2 (alpha is not used by FIM);
//   5. FIM, the 2D grid points as a 2D point list.
//
// Fields go to <outdir>/cpp_2d.bin, cpp_3d.bin (S.io, the whole volume),
// cpp_3d_xslice.bin (S.io, a slice), cpp_points.bin, cpp_fim_3d.bin and
// cpp_fim_points2d.bin (raw); stats(1:6) are printed as
// "stats <name> s0 .. s5". Usage: srm_driver nx ny nz outdir
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "scarf3d.h"

namespace {

void print_stats(const char *name, const real *stats)
{
    std::printf("stats %s", name);
    for (int i = 0; i < 6; ++i) std::printf(" %.17g", static_cast<double>(stats[i]));
    std::printf("\n");
}

bool write_raw(const std::string &path, const std::vector<real> &v)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(v.data(), sizeof(real), v.size(), f) == v.size();
    return std::fclose(f) == 0 && ok;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s nx ny nz outdir\n", argv[0]);
        return 2;
    }
    const int n[3] = {std::atoi(argv[1]), std::atoi(argv[2]), std::atoi(argv[3])};
    const std::string outdir = argv[4];
    const real ds = 50.;
    const int acf = 0;
    const real cl[3] = {2000., 2000., 2000.};
    const real sigma = 0.05;
    const int seed = 1235;
    const real hurst = 0.25, alpha = 30, taper = 2500, mute = 500;
    const int npoi = 2;
    const int fs[3] = {1, 1, 1};
    const int fe[3] = {n[0], n[1], n[2]};
    real poi[6] = {(4 * n[0] / 5) * ds, (n[1] / 2) * ds, (2 * n[2] / 5) * ds,
                   (3 * n[0] / 10) * ds, (n[1] / 10) * ds, (9 * n[2] / 10) * ds};
    real stats[8];

    {  // 1. 2D grid
        const int nd = 2;
        std::vector<real> field(static_cast<size_t>(n[0]) * n[1]);
        Scarf3D::options.hurst = hurst;
        Scarf3D::Initialize<srm> S(nd, fs, fe, ds, acf, cl, sigma);
        S.execute(seed, field.data(), stats);
        print_stats("2d", stats);
        S.io(nd, n, field.data(), (outdir + "/cpp_2d.bin").c_str());
    }
    const long npts = static_cast<long>(n[0]) * n[1] * n[2];
    std::vector<real> field(npts);
    {  // 2. 3D grid with rotation and points of interest
        const int nd = 3;
        Scarf3D::options.hurst = hurst;
        Scarf3D::options.alpha = alpha;
        Scarf3D::options.npoi = npoi;
        Scarf3D::options.poi = poi;
        Scarf3D::options.taper = taper;
        Scarf3D::options.mute = mute;
        Scarf3D::Initialize<srm> S(nd, fs, fe, ds, acf, cl, sigma);
        S.execute(seed, field.data(), stats);
        print_stats("3d", stats);
        S.io(nd, n, field.data(), (outdir + "/cpp_3d.bin").c_str());
        S.io(n, "x", n[0] / 2, field.data(), (outdir + "/cpp_3d_xslice.bin").c_str());
    }
    {  // 3. the 3D grid points as a list
        std::vector<real> x(npts), y(npts), z(npts);
        long c = 0;
        for (int k = 0; k < n[2]; ++k)
            for (int j = 0; j < n[1]; ++j)
                for (int i = 0; i < n[0]; ++i, ++c) {
                    x[c] = (i + fs[0] - 1) * ds;
                    y[c] = (j + fs[1] - 1) * ds;
                    z[c] = (k + fs[2] - 1) * ds;
                }
        Scarf3D::options.hurst = hurst;
        Scarf3D::Initialize<srm> S(3, static_cast<int>(npts), x.data(), y.data(), z.data(), ds, acf, cl, sigma);
        S.execute(seed, field.data(), stats);
        print_stats("points", stats);
        if (!write_raw(outdir + "/cpp_points.bin", field)) return 1;
    }
    {  // 4. FIM on the 3D grid, the options of block 2
        const int nd = 3;
        Scarf3D::options.hurst = hurst;
        Scarf3D::options.alpha = alpha;
        Scarf3D::options.npoi = npoi;
        Scarf3D::options.poi = poi;
        Scarf3D::options.taper = taper;
        Scarf3D::options.mute = mute;
        Scarf3D::Initialize<fim> S(nd, fs, fe, ds, acf, cl, sigma);
        S.execute(seed, field.data(), stats);
        print_stats("fim_3d", stats);
        if (!write_raw(outdir + "/cpp_fim_3d.bin", field)) return 1;
    }
    {  // 5. FIM on the 2D grid points as a 2D point list
        const long n2 = static_cast<long>(n[0]) * n[1];
        std::vector<real> x(n2), y(n2), f2(n2);
        long c = 0;
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i, ++c) {
                x[c] = (i + fs[0] - 1) * ds;
                y[c] = (j + fs[1] - 1) * ds;
            }
        Scarf3D::options.hurst = hurst;
        Scarf3D::Initialize<fim> S(2, static_cast<int>(n2), x.data(), y.data(), ds, acf, cl, sigma);
        S.execute(seed, f2.data(), stats);
        print_stats("fim_points2d", stats);
        if (!write_raw(outdir + "/cpp_fim_points2d.bin", f2)) return 1;
    }
    return 0;
}
