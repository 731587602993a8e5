/*
 * scarf3d.h - public C/C++ interface of the MI355X SCARF3D library (SRM and FIM).
 *
 * Drop-in for the reference's src/scarf3d.h:36-162: the same `real` rule,
 * the same struct scarf_opt layout and the same entry points, exported with C
 * linkage by libscarf3df.so (real = float) and libscarf3d.so (real = double,
 * built with -DDOUBLE_PREC), matching the two reference libraries
 * (src/Makefile:9-13). Extensions live in scarf3d_srm.h.
 *
 * Entry points and the reference code each one replaces:
 *   scarf_opt_init           src/c_binding.c:62-80     zero/NULL all options
 *   scarf_cart_initialize    src/c_binding.c:83-108    -> f_binding.f90:57-113 -> scarflib.f90:97-189
 *   scarf_nocart_initialize  src/c_binding.c:110-136   -> f_binding.f90:119-175 -> scarflib.f90:195-296
 *   scarf_execute            src/c_binding.c:138-143   -> f_binding.f90:181-217 -> scarflib.f90:302-426
 *   scarf_finalize           src/c_binding.c:145-150   -> f_binding.f90:223-240 -> scarflib.f90:432-455
 *   scarf_io_one             src/c_binding.c:154-158   -> f_binding.f90:246-292 (single writer here)
 *   scarf_io_slice           src/c_binding.c:160-164   -> f_binding.f90:298-338 (single writer here)
 *
 * Optional members of scarf_opt count as given when: method == 1, hurst > 0,
 * dh > 0 (structured) / ds > 0 (unstructured), poi != NULL with npoi > 0,
 * mute > 0, taper > 0, rescale == 1, pad == 1, nc and fc both non-NULL,
 * alpha/beta/gamma != 0 (src/c_binding.c:89-103, :116-130). method 0 (the
 * default) selects FIM, method 1 SRM. FIM runs as the reference's default
 * build (DIP off) runs it on one process: the rotation angles are ignored
 * (src/scarflib_fim.f90:614-1104, :141-612).
 */
#ifndef SCARF3D_H_
#define SCARF3D_H_

#ifdef DOUBLE_PREC
typedef double real;
#else
typedef float real;
#endif

#ifdef __cplusplus
extern "C" {
#endif

enum algorithm { fim, srm };

struct scarf_opt {
    int method;
    real hurst;
    real dh;
    real ds;
    real *poi;
    int npoi;
    real mute;
    real taper;
    int rescale;
    int pad;
    real *nc;
    real *fc;
    real alpha;
    real beta;
    real gamma;
};

void scarf_opt_init(struct scarf_opt *var);

void scarf_cart_initialize(const int nd, const int fs[], const int fe[], const real ds, const int acf,
                           const real cl[], const real sigma, struct scarf_opt *var);

void scarf_nocart_initialize(const int nd, const int npts, const real *x, const real *y, const real *z,
                             const real dh, const int acf, const real cl[], const real sigma,
                             struct scarf_opt *var);

/* stats[8]: {low-k flag, high-k flag, discrete std, mean, host s, device s, 0, 0} */
void scarf_execute(const int seed, real *field, real stats[]);

void scarf_finalize(void);

void scarf_io_one(const int nd, const int *npts, const real *field, const char fname[], const int *nwriters);

void scarf_io_slice(const int *npts, const char axis[], const int plane, const real *field, const char fname[]);

#ifdef __cplusplus
}

namespace Scarf3D {

inline scarf_opt options{0, 0, 0, 0, nullptr, 0, 0, 0, 0, 0, nullptr, nullptr, 0, 0, 0};

template <algorithm method>
class Initialize {
public:
    Initialize(const int nd, const int fs[], const int fe[], const real ds, const int acf, const real cl[],
               const real sigma)
    {
        if (method == srm) options.method = 1;
        scarf_cart_initialize(nd, fs, fe, ds, acf, cl, sigma, &options);
    }

    Initialize(const int nd, const int npts, const real *x, const real *y, const real *z, const real dh,
               const int acf, const real cl[], const real sigma)
    {
        if (method == srm) options.method = 1;
        scarf_nocart_initialize(nd, npts, x, y, z, dh, acf, cl, sigma, &options);
    }

    Initialize(const int nd, const int npts, const real *x, const real *y, const real dh, const int acf,
               const real cl[], const real sigma)
    {
        if (method == srm) options.method = 1;
        scarf_nocart_initialize(nd, npts, x, y, nullptr, dh, acf, cl, sigma, &options);
    }

    ~Initialize()
    {
        scarf_finalize();
        options = {0, 0, 0, 0, nullptr, 0, 0, 0, 0, 0, nullptr, nullptr, 0, 0, 0};
    }

    void execute(const int seed, real *field, real stats[]) { scarf_execute(seed, field, stats); }

    void io(const int nd, const int npts[], const real *field, const char fname[], const int *nwriters = nullptr)
    {
        scarf_io_one(nd, npts, field, fname, nwriters);
    }

    void io(const int npts[], const char axis[], const int plane, const real *field, const char fname[])
    {
        scarf_io_slice(npts, axis, plane, field, fname);
    }
};

}  // namespace Scarf3D
#endif

#endif
