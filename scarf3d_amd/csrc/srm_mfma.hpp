// Structured-grid SRM as a real GEMM on the MFMA units.
//
// On a grid every harmonic factors per axis:
//   amp_n cos(tx + ty + tz) = Re( e^{i tx} * [amp_n e^{i tz}] * e^{i ty} )
// so with X[n,i] = e^{i tx[n,i]}, Y[n,j] = e^{i ty[n,j]}, Z[n,k] = amp_n e^{i tz[n,k]}
//   F[(j,k), i] = sum_n  Re(Z Y)[n,(j,k)] cos tx[n,i] - Im(Z Y)[n,(j,k)] sin tx[n,i]
// i.e. a GEMM with inner dimension 2N between a generated operand G = Z*Y
// (one row per grid column (j,k)) and the x-table X (one column per x).
// Both operands are split into 16-bit parts whose partial products are
// accumulated in FP32 (hi*hi + hi*lo + lo*hi [+ ...]) and promoted into a
// second FP32 accumulator every 128 harmonics.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "srm_device.hpp"

namespace scarf {

struct MfmaPlan {
    int kernel = 0;
    int parts = 0;          // 16-bit parts per operand value
    int np_ = 0;            // padded harmonic count
    int stages = 0;         // np_ / harmonics per stage
    int nx_pad = 0;         // x extent padded to the block tile
    int tiles_x = 0;        // block tiles along x
    int tj = 0, tk = 0;     // block column tile: tj values of j by tk values of k
    int tiles_j = 0, tiles_k = 0;
    long blocks = 0;
    int nparts = 0;         // statistics partials (one per block)
    size_t x_bytes = 0, y_bytes = 0, z_bytes = 0, table_bytes = 0;
};

bool mfma_plan(const Geometry &g, int np_, int kernel, MfmaPlan &plan, std::string &err);
bool mfma_tables(const MfmaPlan &plan, const Geometry &g, const double *d_harm, int np_, void *d_tab,
                 hipStream_t stream, std::string &err);
template <class Real>
bool mfma_field(const MfmaPlan &plan, const Geometry &g, const void *d_tab, double scaling, Real *d_field,
                double *d_part, hipStream_t stream, std::string &err);

}  // namespace scarf
