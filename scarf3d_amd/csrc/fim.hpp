// Fourier integral method (FIM, method 0) on one GPU.
//
// The reference computes a random-phase amplitude spectrum on an internal
// (FFT) grid that covers the model plus optional padding, makes it Hermitian
// on the kx = 0 and kx = Nyquist planes, inverse-transforms it along z, y and
// x, scales it, takes its statistics, rescales and tapers it on the internal
// grid, and maps it onto the caller's structured grid or point list by
// nearest neighbour (src/scarflib_fim.f90:141-612 point lists, :614-1104
// structured grids). Here one process holds the whole internal grid, as one
// MPI rank of the reference does, and every step is a device pass:
//   fim_spectrum   uniforms of the global stream + amplitude, half spectrum
//   fim_symmetry   the conjugate pairs of the kx = 0 and Nyquist planes
//                  (after it both planes are exactly Hermitian)
//   hipFFT         one 3D (2D) c2r, equal to the reference's backward c2c
//                  along z and y followed by c2r along x
//   fim_moments    FP64 moments of the scaled field; fim_finalize
//   fim_epilogue   scaling, optional rescale and taper on the internal grid
//   fim_map_*      nearest-neighbour values at the caller's points
// Host-side constants (internal grid size, offsets, spectrum constants) are
// computed at the library's `real` precision with the reference's roundings
// (fim_setup, compiled without FMA contraction).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "srm_host.hpp"

namespace scarf {

// Constants of one FIM realization (src/scarflib_fim.f90:754-837 and
// compute_spectrum :2440-2482), at `real` precision.
template <class Real>
struct FimSetup {
    int ndim = 3;
    int npts[3] = {1, 1, 1};     // internal grid (even along every used axis; 1 along z in 2D)
    Real off_axis[3] = {};       // absolute position of internal point 1 (negated z in 2D)
    Real dh = 0, ds = 0;
    Real dk[3] = {};             // 2 pi / (npts dh)
    Real kc = 0;                 // pi / ds, the filter corner
    bool filter = false;         // ds != dh: amplitudes above kc are zeroed
    int acf = 0;
    Real cl[3] = {};
    Real nu = 0, cnst = 0;       // PSD exponent and constant
    Real scaling = 0;            // 1 / sqrt(npts1 npts2 npts3 dh^n)
    int info1 = 0, info2 = 0;    // spectrum coverage flags
    Real sigma = 0;
    std::vector<Real> pt;        // points of interest on the internal grid (3 per point)
    long internal_points() const { return long(npts[0]) * npts[1] * npts[2]; }
};

template <class Real>
int fim_setup(const SrmConfig &cfg, FimSetup<Real> &out, std::string &err);

// Device timings of the last FIM run, seconds: spectrum, symmetry, transforms,
// mapping (the reference's info(5:8) under TIMING).
struct FimTimings {
    double spectrum = 0, symmetry = 0, fft = 0, mapping = 0, total = 0;
};

class FimDevice {
public:
    FimDevice();
    ~FimDevice();
    FimDevice(const FimDevice &) = delete;
    FimDevice &operator=(const FimDevice &) = delete;

    // One realization into d_field (device, `Real`): the caller's structured
    // block (g.structured) or point list (points: FP64 device coordinates,
    // 3 x npts, z zero in 2D). stats[0..3] = {info3, info4} as the
    // reference reports them after an optional rescale, then FP64 mean and
    // variance of the internal field. Runs on `stream`. reserved_elsewhere:
    // device bytes the library holds outside FIM, counted against
    // SCARF_DEVICE_MEMORY_LIMIT=<bytes> when that is set (a buffer that would
    // take the total above it fails as a failed hipMalloc does).
    template <class Real>
    bool run(const FimSetup<Real> &s, int rescale, double mute, double taper, int seed, int fs[3], int fe[3],
             bool structured, const double *d_points, long npoints, Real *d_field, double stats[4],
             hipStream_t stream, size_t reserved_elsewhere, std::string &err);

    const FimTimings &timings() const { return timings_; }
    // device bytes held by the FIM buffers
    size_t reserved_bytes() const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
    FimTimings timings_;
};

}  // namespace scarf
