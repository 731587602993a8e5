// Device half of the SRM: buffers kept across executes and the launch
// sequence of one realization. All O(points x harmonics) work happens here.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "srm_host.hpp"

namespace scarf {

// Field kernels. Values are part of the public extension API (scarf3d_srm.h).
enum Kernel : int {
    kKernelAuto = 0,
    kKernelDirectF32 = 1,   // FP32 cos on FP64-reduced per-axis phases
    kKernelDirectF64 = 2,   // FP64 throughout (cospi), the exact path
    kKernelMfmaBf16x3 = 3,  // per-slab GEMM, 3-term bf16 split, 6 MFMA products
    kKernelMfmaF16x2 = 4,   // per-slab GEMM, scaled 2-term fp16 split, 3 products
    kKernelMfmaBf16x2 = 5,  // per-slab GEMM, 2-term bf16 split, 3 products (below 1e-5 sigma)
};
const char *kernel_name(int k);

// Geometry of one realization as the kernels see it.
struct Geometry {
    bool structured = true;
    int nx = 1, ny = 1, nz = 1;       // structured extents (local block)
    int ox = 0, oy = 0, oz = 0;       // global 0-based offset of the block (fs - 1)
    double ds = 0;                    // grid step
    long npts = 0;                    // unstructured point count
    long points() const { return structured ? long(nx) * ny * nz : npts; }
};

// Epilogue parameters (rescale + taper, src/scarflib_srm.f90:897-925).
struct Epilogue {
    int rescale = 0;
    double sigma = 0;
    double mute = -999, taper = -999;
    std::vector<double> poi3;  // npoi x 3 (missing coordinates zero)
    int ndim = 3;
    bool active() const;
};

// Per-phase device timings of the last run, milliseconds.
struct Timings {
    double tables = 0, field = 0, stats = 0, epilogue = 0, total = 0;
};

class TileTuner;

class Device {
public:
    Device() = default;
    ~Device();
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;

    // Selects the current HIP device and creates the stream on first use.
    bool ensure(std::string &err);

    // Uploads the harmonic table (FP64 SoA) and, for point lists, the
    // coordinates (widened to FP64).
    bool upload_table(const HarmonicTable &t, std::string &err);
    template <class Real>
    bool upload_points(const Real *x, const Real *y, const Real *z, long n, std::string &err);

    // Runs tables + field + statistics (+ epilogue) into d_field (device,
    // `Real` elements). stats_out[0..3] = {count, mean, variance, sigma_d}
    // in FP64 and real_stats[0..1] = {info3, info4} rounded to Real.
    // on_slab, if given, is called with (first point, points) for each part
    // of the field as soon as it is final: in z order while the rest is still
    // computed (MFMA grids with no rescale or taper), else once for the whole
    // field at the end.
    using SlabHook = std::function<bool(long, long, std::string &)>;
    template <class Real>
    bool run(const Geometry &g, int kernel, double scaling, const Epilogue &ep, Real *d_field,
             double *stats_out, std::string &err, const SlabHook *on_slab = nullptr);

    hipStream_t stream() const { return stream_; }
    const Timings &timings() const { return timings_; }

    // Device buffer for scarf_execute's host-field path (grow-only).
    void *field_buffer(size_t bytes, std::string &err);

    int nharm_resident() const { return nharm_; }

    // tile autotuner of the MFMA kernels (shared with ensembles)
    TileTuner &tuner();
    // x-tile width (32 * nbx) and column tile tk of the last MFMA run
    int last_nbx() const { return last_nbx_; }
    int last_tk() const { return last_tk_; }

private:
    bool reserve(void **ptr, size_t &cap, size_t bytes, std::string &err);

    int device_ = -1;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_[6] = {};
    static constexpr int kSlabs = 8;  // field parts handed out while the rest is computed
    hipEvent_t slab_ev_[kSlabs] = {};
    int nharm_ = 0;
    double amp_max_ = 0;  // largest amplitude of the resident table
    std::unique_ptr<TileTuner> tuner_;
    int last_nbx_ = 0, last_tk_ = 0;
    // harmonic record: wx, wy, wz (rad/m), phase (rad), amp; plus revolution copies
    double *d_harm_ = nullptr;
    size_t harm_cap_ = 0;
    // per-axis tables (FP32 or FP64 revolutions, operand tables of the GEMM)
    void *d_tab_ = nullptr;
    size_t tab_cap_ = 0;
    // point coordinates (FP64 x, y, z)
    double *d_pts_ = nullptr;
    size_t pts_cap_ = 0;
    long npts_ = 0;
    // statistics partials and results
    double *d_part_ = nullptr;
    size_t part_cap_ = 0;
    double *d_stat_ = nullptr;
    double *h_stat_ = nullptr;  // pinned
    void *d_field_ = nullptr;
    size_t field_cap_ = 0;
    double *d_poi_ = nullptr;
    size_t poi_cap_ = 0;
    Timings timings_;
};

}  // namespace scarf
