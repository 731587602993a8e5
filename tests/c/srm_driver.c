/*
 * C drop-in driver: the SRM blocks of the reference's C example
 * (examples/c/driver.c:526-612), on one process and without MPI, compiled
 * against include/scarf3d.h and linked with libscarf3df / libscarf3d.
 * options.method is set to 1: the example leaves it commented out and so
 * runs FIM there (a quirk of the example, not of the library).
 *
 *   1. a 3D Cartesian grid with the example's parameters (ds = 50 m, von
 *      Karman, cl = 2000 m, sigma = 0.05, Hurst 0.25, seed 1235), rotated by
 *      alpha = 30, beta = -10, with two points of interest, mute 500 m and
 *      taper 2500 m;
 *   2. the same grid points given as a point list (scarf_nocart_initialize);
 *   3. FIM on the grid as the example's own 3D Cartesian FIM block runs it
 *      (examples/c/driver.c:405-430: method left at its default 0, hurst
 *      the only option).
 *
 * Each field goes to <outdir>/c_cart.bin, <outdir>/c_nocart.bin and
 * <outdir>/c_fim_cart.bin through scarf_io_one (the point list as raw
 * values); a z slice through
 * scarf_io_slice; stats(1:6) are printed as
 * "stats <name> s0 .. s5". Usage: srm_driver nx ny nz outdir
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "scarf3d.h"

static void print_stats(const char *name, const real *stats)
{
    printf("stats %s", name);
    for (int i = 0; i < 6; ++i) printf(" %.17g", (double)stats[i]);
    printf("\n");
}

int main(int argc, char **argv)
{
    if (argc != 5) {
        fprintf(stderr, "usage: %s nx ny nz outdir\n", argv[0]);
        return 2;
    }
    const int n[3] = {atoi(argv[1]), atoi(argv[2]), atoi(argv[3])};
    const char *outdir = argv[4];
    const real ds = 50.;
    const int acf = 0;
    const real cl[3] = {2000., 2000., 2000.};
    const real sigma = 0.05;
    const int seed = 1235;
    const real hurst = 0.25, alpha = 30, beta = -10, taper = 2500, mute = 500;
    const int npoi = 2;
    const int nd = 3;
    const int fs[3] = {1, 1, 1};
    const int fe[3] = {n[0], n[1], n[2]};
    /* two points of interest inside the grid, (nd, npoi) column-major */
    real poi[6] = {(4 * n[0] / 5) * ds, (n[1] / 2) * ds, (2 * n[2] / 5) * ds,
                   (3 * n[0] / 10) * ds, (n[1] / 10) * ds, (9 * n[2] / 10) * ds};
    const long npts = (long)n[0] * n[1] * n[2];
    real *field = (real *)malloc(sizeof(real) * npts);
    real *x = (real *)malloc(sizeof(real) * npts);
    real *y = (real *)malloc(sizeof(real) * npts);
    real *z = (real *)malloc(sizeof(real) * npts);
    real stats[8];
    char path[4096];
    struct scarf_opt options;
    if (!field || !x || !y || !z) return 1;

    /* 1. Cartesian grid */
    scarf_opt_init(&options);
    options.method = 1;
    options.hurst = hurst;
    options.alpha = alpha;
    options.beta = beta;
    options.npoi = npoi;
    options.poi = poi;
    options.taper = taper;
    options.mute = mute;
    scarf_cart_initialize(nd, fs, fe, ds, acf, cl, sigma, &options);
    memset(field, 0, sizeof(real) * npts);
    scarf_execute(seed, field, stats);
    print_stats("cart", stats);
    snprintf(path, sizeof path, "%s/c_cart.bin", outdir);
    scarf_io_one(nd, n, field, path, NULL);
    snprintf(path, sizeof path, "%s/c_cart_zslice.bin", outdir);
    scarf_io_slice(n, "z", n[2] / 2, field, path);
    scarf_finalize();

    /* 2. the same points as a list (examples/c/driver.c:383-394) */
    long c = 0;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                x[c] = (i + fs[0] - 1) * ds;
                y[c] = (j + fs[1] - 1) * ds;
                z[c] = (k + fs[2] - 1) * ds;
                ++c;
            }
    scarf_opt_init(&options);
    options.method = 1;
    options.hurst = hurst;
    scarf_nocart_initialize(nd, (int)npts, x, y, z, ds, acf, cl, sigma, &options);
    memset(field, 0, sizeof(real) * npts);
    scarf_execute(seed, field, stats);
    print_stats("nocart", stats);
    scarf_finalize();
    /* the reference example writes no point-list field: raw values here */
    snprintf(path, sizeof path, "%s/c_nocart.bin", outdir);
    FILE *f = fopen(path, "wb");
    if (!f || fwrite(field, sizeof(real), (size_t)npts, f) != (size_t)npts) return 1;
    fclose(f);

    /* 3. FIM (examples/c/driver.c:405-430) */
    scarf_opt_init(&options);
    options.hurst = hurst;
    scarf_cart_initialize(nd, fs, fe, ds, acf, cl, sigma, &options);
    memset(field, 0, sizeof(real) * npts);
    scarf_execute(seed, field, stats);
    print_stats("fim_cart", stats);
    snprintf(path, sizeof path, "%s/c_fim_cart.bin", outdir);
    scarf_io_one(nd, n, field, path, NULL);
    scarf_finalize();

    free(field);
    free(x);
    free(y);
    free(z);
    return 0;
}
