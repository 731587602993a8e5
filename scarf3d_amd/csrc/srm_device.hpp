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

    // Source of a realization's harmonics in draw order: feed(kept, t, last)
    // appends to t the kept harmonics that bring their total to `kept`
    // (fewer only when the draw has ended; last = true once every harmonic
    // has been drawn).
    using HarmonicFeed = std::function<bool(int kept, HarmonicTable &t, bool &last, std::string &err)>;
    // The same realization as run() on an MFMA kernel of a structured grid,
    // with the harmonics taken from `feed` in chunks while the device works:
    // each chunk's record upload, operand tables and field pass are queued as
    // soon as it is drawn, and the next chunk is drawn while they run. The
    // field is the one run() gives on the same harmonics, bit for bit. At
    // most nharm harmonics are kept. Returns false with *restart = true, and
    // nothing written, when a chunk holds an amplitude that the scaled-fp16
    // scheme can only take with a rescaled table: the caller then draws the
    // rest and uses run().
    // chunk > 0 (a multiple of 128) fixes the kept harmonics per chunk.
    template <class Real>
    bool run_streamed(const Geometry &g, int kernel, int nharm, double scaling, const Epilogue &ep, Real *d_field,
                      double *stats_out, std::string &err, const HarmonicFeed &feed, bool *restart,
                      const SlabHook *on_slab = nullptr, int chunk = 0);
    // kept-harmonic totals at which run_streamed cuts the chunks of a draw of
    // nharm harmonics (the last entry is nharm: draw to the end)
    static std::vector<int> chunk_targets(int nharm, int chunk = 0);

    // run() on the resident table with no device allocation at all: the
    // operand tables, statistics partials, statistics and points of
    // interest live in a static device array of kNoAllocScratch bytes
    // (per GPU, for the life of the process), the tables one chunk of
    // harmonics at a time (chunks at multiples of 128 harmonics, running
    // sums carried through the field: the field equals run()'s bit for
    // bit). MFMA kernels on structured grids only; no tile tuning (a tile
    // chosen earlier for the grid, else the least-padding one). Refuses, with
    // an error, grids whose tables for 128 harmonics (or whose partials and
    // points of interest) do not fit. Every call on one GPU uses the same
    // array: calls that overlap on one GPU race.
    static constexpr size_t kNoAllocScratch = size_t(64) << 20;
    template <class Real>
    bool run_noalloc(const Geometry &g, int kernel, double scaling, const Epilogue &ep, Real *d_field,
                     double *stats_out, std::string &err);

    hipStream_t stream() const { return stream_; }
    const Timings &timings() const { return timings_; }

    // Device buffer for scarf_execute's host-field path (grow-only).
    void *field_buffer(size_t bytes, std::string &err);

    int nharm_resident() const { return nharm_; }

    // FP64 coordinates of the last uploaded point list (x, y, z blocks of
    // points_count() values each)
    const double *points() const { return d_pts_; }
    long points_count() const { return npts_; }

    // tile autotuner of the MFMA kernels (shared with ensembles)
    TileTuner &tuner();
    // x-tile width (32 * nbx) and column tile tk of the last MFMA run
    int last_nbx() const { return last_nbx_; }
    int last_tk() const { return last_tk_; }

    // device bytes held by the library's grow-only buffers
    size_t reserved_bytes() const { return reserved_; }
    // device bytes the library holds in other buffers (FIM's), counted with
    // these against SCARF_DEVICE_MEMORY_LIMIT
    void set_external_reserved(size_t bytes) { external_ = bytes; }

private:
    // Grows *ptr to `bytes` (keeps it when large enough). With
    // SCARF_DEVICE_MEMORY_LIMIT=<bytes> set, a call that would leave the
    // buffers holding more than that fails with an error instead, as a
    // failed hipMalloc does: the allocation-failure paths can be exercised.
    bool reserve(void **ptr, size_t &cap, size_t bytes, std::string &err);
    // the same on a typed member: the pointer is written back whatever the
    // outcome (after a failed hipMalloc it is null, its capacity 0)
    template <class T>
    bool reserve(T *&ptr, size_t &cap, size_t bytes, std::string &err)
    {
        void *p = ptr;
        const bool ok = reserve(&p, cap, bytes, err);
        ptr = static_cast<T *>(p);
        return ok;
    }
    size_t reserved_ = 0;
    size_t external_ = 0;
    // statistics, epilogue, field hand-over and timings after the field
    // kernels; slab_first/slab_points name the parts whose events are
    // slab_ev_. part/stat/poi: device room for the partials, the statistics
    // and the points of interest (null: the library's own buffers)
    template <class Real>
    bool finish(const Geometry &g, const Epilogue &ep, Real *d_field, int nparts, double *stats_out,
                const std::vector<long> &slab_first, const std::vector<long> &slab_points, std::string &err,
                const SlabHook *on_slab, double *part = nullptr, double *stat = nullptr, double *poi = nullptr);

    int device_ = -1;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_[6] = {};
    static constexpr int kSlabs = 8;  // field parts handed out while the rest is computed
    hipEvent_t slab_ev_[kSlabs] = {};
    void *scratch_ = nullptr;  // this device's copy of the no-alloc scratch array
    int nharm_ = 0;
    double amp_max_ = 0;  // largest amplitude of the resident table
    std::unique_ptr<TileTuner> tuner_;
    int last_nbx_ = 0, last_tk_ = 0;
    // harmonic record: wx, wy, wz (rad/m), phase (rad), amp; plus revolution copies
    double *d_harm_ = nullptr;
    size_t harm_cap_ = 0;
    double *h_rec_ = nullptr;  // pinned staging of a streamed record
    size_t h_rec_cap_ = 0;
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
