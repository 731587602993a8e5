// C ABI of the library and the object layer behind it.
//
// One process-global object, as in the reference (`obj`, src/scarflib.f90:87,
// and the binding's `dims`/`structured`, src/f_binding.f90:44-47): initialize
// stores the parameters with their defaults, execute runs one realization,
// finalize releases the state. The library is not re-entrant, like the
// reference. Differences, all deliberate:
//   - points of interest are copied at initialize (the reference keeps a
//     pointer, and its unstructured path assigns into a zero-size pointer);
//   - errors are reported (stderr + scarf_last_error) instead of crashing:
//     initialize twice, execute before initialize, acf = 2, von Karman with
//     hurst <= 0;
//   - nharm is a runtime setting (default 10000);
//   - FIM (method 0) runs as the reference does on one process holding the
//     whole internal grid (fim.hpp); its rotation angles are ignored, as in
//     the reference's default build (DIP off).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/scarf3d.h"
#include "../../include/scarf3d_srm.h"
#include "fim.hpp"
#include "srm_device.hpp"
#include "srm_host.hpp"
#include "srm_mfma.hpp"
#include "worker_pool.hpp"

#include <climits>
#include <thread>

namespace {

using scarf::SrmConfig;

struct State {
    bool initialized = false;
    SrmConfig cfg;
    const real *x = nullptr, *y = nullptr, *z = nullptr;  // borrowed point coordinates
    int kernel = -1;                                       // -1: not chosen yet (environment / auto)
    int nharm = 10000;
    int draw_threads = 0;  // host draw workers (0: default)
    long draw_block = 0;   // pairs per draw block (0: default)
    bool stream_draw = true;  // MFMA grids: device work starts on the first drawn harmonics
    int draw_chunk = 0;       // scarf_draw_harmonics: kept harmonics per cut of the walk (0: as streamed)
    int stream_chunk = 0;     // kept harmonics per streamed chunk, a multiple of 128 (0: growing chunks)
    std::string error;
    std::unique_ptr<scarf::Device> device;
    std::unique_ptr<scarf::FimDevice> fim;
    std::unique_ptr<scarf::Ensemble> ensemble;
    // resident realization (scarf_srm_draw -> scarf_srm_run)
    bool table_ready = false;
    scarf::SrmSetup setup;
    double host_seconds = 0;
    double ensemble_ms = 0;
};

// Never destroyed: its device buffers, streams and graphs must not be
// released by static destructors after the HIP runtime (or a profiler hooked
// into it) has begun its own teardown; the process exit returns them.
State &state()
{
    static State *s = new State;
    return *s;
}

int fail(const std::string &msg)
{
    State &s = state();
    s.error = msg;
    std::fprintf(stderr, "scarf3d: error: %s\n", msg.c_str());
    return -1;
}

constexpr bool kDouble = sizeof(real) == sizeof(double);

// The field back into the caller's array (pageable memory): chunks go by DMA
// into two pinned staging buffers and from there, on up to 8 pool threads,
// into the array, the DMA of chunk i+1 running while chunk i is copied out.
// One hipMemcpy into pageable memory moved 537 MB in 52-63 ms (~9 GB/s at
// any size). Copies of a few MiB and more are staged, in at least four
// chunks of 4 to 32 MiB so that DMA and drain overlap. The staging buffers
// are allocated on the first staged copy and kept for the process; if pinned
// memory is refused, every copy is one hipMemcpy.
constexpr size_t kStageBytes = size_t(32) << 20;  // staging buffer size, largest chunk
constexpr size_t kStageMin = size_t(4) << 20;     // smallest chunk and smallest staged copy

scarf::WorkerPool &drain_pool()
{
    static scarf::WorkerPool *p = new scarf::WorkerPool(
        static_cast<int>(std::max(1u, std::min(8u, std::thread::hardware_concurrency()))));  // lives until exit
    return *p;
}

// ensemble members' host draws: one member per thread at a time
scarf::WorkerPool &member_pool()
{
    static scarf::WorkerPool *p = new scarf::WorkerPool(
        static_cast<int>(std::max(1u, std::min(16u, std::thread::hardware_concurrency()))));  // lives until exit
    return *p;
}

bool copy_to_host(void *dst, const void *src, size_t bytes, hipStream_t stream, std::string &err)
{
    auto ok = [&](hipError_t e, const char *what) {
        if (e == hipSuccess) return true;
        err = std::string("copying the field to the host: ") + what + ": " + hipGetErrorString(e);
        return false;
    };
    static char *stage = nullptr;
    static hipEvent_t done[2] = {nullptr, nullptr};
    static bool no_stage = false;  // pinned memory was refused once: one copy from then on
    if (!stage && !no_stage && bytes >= kStageMin) {
        void *p = nullptr;
        if (hipHostMalloc(&p, 2 * kStageBytes, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            no_stage = true;
        } else {
            for (auto &e : done)
                if (!ok(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate")) return false;
            stage = static_cast<char *>(p);
        }
    }
    if (bytes < kStageMin || !stage) return ok(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    const size_t chunk = std::min(kStageBytes, std::max(kStageMin, (bytes / 4 + 4095) / 4096 * 4096));
    const size_t nchunks = (bytes + chunk - 1) / chunk;
    auto len = [&](size_t i) { return std::min(chunk, bytes - i * chunk); };
    auto fetch = [&](size_t i) {
        return ok(hipMemcpyAsync(stage + (i % 2) * kStageBytes, static_cast<const char *>(src) + i * chunk, len(i),
                                 hipMemcpyDeviceToHost, stream),
                  "hipMemcpyAsync") &&
               ok(hipEventRecord(done[i % 2], stream), "hipEventRecord");
    };
    scarf::WorkerPool &pool = drain_pool();
    auto drain = [&](size_t i) {  // chunk i: staging buffer i % 2 -> the caller's array
        if (!ok(hipEventSynchronize(done[i % 2]), "hipEventSynchronize")) return false;
        char *out = static_cast<char *>(dst) + i * chunk;
        const char *in = stage + (i % 2) * kStageBytes;
        const size_t n = len(i), nt = size_t(pool.size()), part = (n + nt - 1) / nt;
        const std::function<void(int)> job = [&](int t) {
            const size_t lo = size_t(t) * part;
            if (lo < n) std::memcpy(out + lo, in + lo, std::min(part, n - lo));
        };
        pool.run(job);
        return true;
    };
    // buffer i % 2 is refilled only after chunk i - 2 has been drained from it
    if (!fetch(0)) return false;
    for (size_t i = 1; i < nchunks; ++i)
        if (!fetch(i) || !drain(i - 1)) return false;
    return drain(nchunks - 1);
}

int kernel_from_env()
{
    const char *e = std::getenv("SCARF_KERNEL");
    if (!e || !*e) return scarf::kKernelAuto;
    for (int k = 0; k <= scarf::kKernelMfmaBf16x2; ++k)
        if (std::strcmp(e, scarf::kernel_name(k)) == 0) return k;
    std::fprintf(stderr, "scarf3d: warning: unknown SCARF_KERNEL=%s, using auto\n", e);
    return scarf::kKernelAuto;
}

int current_kernel()
{
    State &s = state();
    if (s.kernel < 0) s.kernel = kernel_from_env();
    return s.kernel;
}

// kernel actually used for the current configuration
int resolve_kernel(const SrmConfig &cfg)
{
    const int k = current_kernel();
    if (k != scarf::kKernelAuto) return k;
    if (kDouble) return scarf::kKernelDirectF64;
    if (!cfg.structured) return scarf::kKernelDirectF32;
    const long nx = cfg.fe[0] - cfg.fs[0] + 1, ny = cfg.fe[1] - cfg.fs[1] + 1, nz = cfg.fe[2] - cfg.fs[2] + 1;
    // the GEMM needs enough 256 x 128 tiles to fill 256 compute units
    const long tiles = ((nx + 127) / 128) * ((ny * nz + 255) / 256);
    return tiles >= 512 ? scarf::kKernelMfmaBf16x3 : scarf::kKernelDirectF32;
}

scarf::Geometry geometry(const SrmConfig &cfg)
{
    scarf::Geometry g;
    g.structured = cfg.structured;
    if (cfg.structured) {
        g.nx = cfg.fe[0] - cfg.fs[0] + 1;
        g.ny = cfg.fe[1] - cfg.fs[1] + 1;
        g.nz = cfg.fe[2] - cfg.fs[2] + 1;
        g.ox = cfg.fs[0] - 1;
        g.oy = cfg.fs[1] - 1;
        g.oz = cfg.fs[2] - 1;
        g.ds = static_cast<real>(cfg.ds);
    } else {
        g.npts = cfg.npts;
    }
    return g;
}

scarf::Epilogue epilogue(const SrmConfig &cfg, int ndim)
{
    scarf::Epilogue ep;
    ep.rescale = cfg.rescale;
    ep.sigma = cfg.sigma;
    ep.mute = cfg.mute;
    ep.taper = cfg.taper;
    ep.ndim = ndim;
    // pt(1:ndim) = poi(1:ndim, i), pt(3) = 0 in 2D (src/scarflib_srm.f90:920-925)
    for (int q = 0; q < cfg.npoi; ++q) {
        double p3[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < ndim && d < cfg.nd; ++d) p3[d] = cfg.poi[size_t(q) * cfg.nd + d];
        ep.poi3.insert(ep.poi3.end(), p3, p3 + 3);
    }
    return ep;
}

bool ensure_device()
{
    State &s = state();
    if (!s.device) s.device = std::make_unique<scarf::Device>();
    std::string err;
    if (!s.device->ensure(err)) {
        fail(err);
        return false;
    }
    s.device->set_external_reserved(s.fim ? s.fim->reserved_bytes() : 0);
    return true;
}

void clear_stats(real stats[])
{
    if (stats)
        for (int i = 0; i < 8; ++i) stats[i] = 0;
}

// host half of a realization: setup, draw, upload
int prepare(int seed)
{
    State &s = state();
    s.table_ready = false;
    if (!s.initialized) return fail("execute called before initialize");
    if (s.cfg.method != 1) return fail("the SRM harmonic table needs method = 1 (SRM); method 0 (FIM) runs through execute");
    s.cfg.nharm = s.nharm;
    std::string err;
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = kDouble ? scarf::srm_setup<double>(s.cfg, s.setup, err) : scarf::srm_setup<float>(s.cfg, s.setup, err);
    if (rc != 0) return fail(err);
    if (!ensure_device()) return -1;
    const scarf::HarmonicDraw d = kDouble ? scarf::srm_draw<double>(s.cfg, s.setup, seed, s.draw_threads, s.draw_block)
                                          : scarf::srm_draw<float>(s.cfg, s.setup, seed, s.draw_threads, s.draw_block);
    const scarf::HarmonicTable t = scarf::build_table(s.setup, d);
    s.host_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!s.device->upload_table(t, err)) return fail(err);
    if (!s.cfg.structured) {
        if (!s.device->upload_points<real>(s.x, s.y, s.z, s.cfg.npts, err)) return fail(err);
    }
    s.table_ready = true;
    s.error.clear();
    return d.n_used;
}

// stats(1:8) of a finished realization (src/scarflib.f90:363,377)
void report_stats(real stats[], const double st[6])
{
    State &s = state();
    if (stats) {
        clear_stats(stats);
        stats[0] = static_cast<real>(s.setup.info1);
        stats[1] = static_cast<real>(s.setup.info2);
        // rescale reports sigma and zero mean (src/scarflib_srm.f90:909-910)
        stats[2] = s.cfg.rescale == 1 ? static_cast<real>(s.cfg.sigma) : static_cast<real>(st[3]);
        stats[3] = s.cfg.rescale == 1 ? real(0) : static_cast<real>(st[4]);
        stats[4] = static_cast<real>(s.host_seconds);
        stats[5] = static_cast<real>(s.device->timings().total * 1e-3);
    }
}

// device half: tables, field, statistics, epilogue into d_field
int run(real *d_field, real stats[], const scarf::Device::SlabHook *on_slab = nullptr)
{
    State &s = state();
    if (!s.table_ready) return fail("no harmonic table drawn (call scarf_srm_draw first)");
    const scarf::Geometry g = geometry(s.cfg);
    const scarf::Epilogue ep = epilogue(s.cfg, s.setup.ndim);
    const int kernel = resolve_kernel(s.cfg);
    double st[6] = {};
    std::string err;
    if (!s.device->run<real>(g, kernel, s.setup.scaling, ep, d_field, st, err, on_slab)) return fail(err);
    report_stats(stats, st);
    s.error.clear();
    return 0;
}

long field_points(const SrmConfig &cfg)
{
    if (!cfg.structured) return cfg.npts;
    return long(cfg.fe[0] - cfg.fs[0] + 1) * (cfg.fe[1] - cfg.fs[1] + 1) * (cfg.fe[2] - cfg.fs[2] + 1);
}

// One FIM realization into d_field: stats(1:2) the coverage flags, (3:4)
// standard deviation and mean of the internal field (sigma and 0 after a
// rescale), (5:8) device seconds of the spectrum, the symmetry pass, the
// transforms and the mapping (src/scarflib_fim.f90:614-1104, info(1:8)).
// The field is handed to on_slab in one piece at the end.
int execute_fim(int seed, real *d_field, real stats[], const scarf::Device::SlabHook *on_slab)
{
    State &s = state();
    scarf::FimSetup<real> fs;
    std::string err;
    if (scarf::fim_setup<real>(s.cfg, fs, err) != 0) return fail(err);
    if (!ensure_device()) return -1;
    if (!s.fim) s.fim = std::make_unique<scarf::FimDevice>();
    const double *pts = nullptr;
    if (!s.cfg.structured) {
        if (!s.device->upload_points<real>(s.x, s.y, s.z, s.cfg.npts, err)) return fail(err);
        pts = s.device->points();
    }
    int fsi[3], fei[3];
    for (int i = 0; i < 3; ++i) {
        fsi[i] = s.cfg.fs[i];
        fei[i] = s.cfg.fe[i];
    }
    double st[4] = {};
    if (!s.fim->run<real>(fs, s.cfg.rescale, s.cfg.mute, s.cfg.taper, seed, fsi, fei, s.cfg.structured, pts,
                          s.cfg.npts, d_field, st, s.device->stream(), s.device->reserved_bytes(), err))
        return fail(err);
    if (stats) {
        clear_stats(stats);
        stats[0] = static_cast<real>(fs.info1);
        stats[1] = static_cast<real>(fs.info2);
        stats[2] = static_cast<real>(st[0]);
        stats[3] = static_cast<real>(st[1]);
        const scarf::FimTimings &t = s.fim->timings();
        stats[4] = static_cast<real>(t.spectrum);
        stats[5] = static_cast<real>(t.symmetry);
        stats[6] = static_cast<real>(t.fft);
        stats[7] = static_cast<real>(t.mapping);
    }
    if (on_slab && !(*on_slab)(0, field_points(s.cfg), err)) return fail(err);
    s.error.clear();
    return 0;
}

// One realization into d_field. On the MFMA kernels of a structured grid
// the draw is streamed: the device starts on the first chunk of harmonics
// while the host draws the next (the reference draws each harmonic before
// its device loop, src/scarflib_srm.f90:725-833, so draw and device work
// never overlap there). Elsewhere the whole table is drawn first.
int execute_realization(int seed, real *d_field, real stats[], const scarf::Device::SlabHook *on_slab)
{
    State &s = state();
    s.table_ready = false;
    if (!s.initialized) return fail("execute called before initialize");
    if (s.cfg.method == 0) return execute_fim(seed, d_field, stats, on_slab);
    s.cfg.nharm = s.nharm;
    const int kernel = resolve_kernel(s.cfg);
    const bool mfma = kernel == scarf::kKernelMfmaBf16x3 || kernel == scarf::kKernelMfmaF16x2 ||
                      kernel == scarf::kKernelMfmaBf16x2;
    if (!s.cfg.structured || !mfma || !s.stream_draw) {
        if (prepare(seed) < 0) return -1;
        return run(d_field, stats, on_slab);
    }
    std::string err;
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    const int rc = kDouble ? scarf::srm_setup<double>(s.cfg, s.setup, err) : scarf::srm_setup<float>(s.cfg, s.setup, err);
    if (rc != 0) return fail(err);
    double host = std::chrono::duration<double>(clock::now() - t0).count();
    if (!ensure_device()) return -1;
    std::unique_ptr<scarf::HarmonicSampler> sampler =
        kDouble ? scarf::HarmonicSampler::create<double>(s.cfg, s.setup, seed, s.draw_threads, s.draw_block)
                : scarf::HarmonicSampler::create<float>(s.cfg, s.setup, seed, s.draw_threads, s.draw_block);
    const scarf::Device::HarmonicFeed feed = [&](int kept, scarf::HarmonicTable &t, bool &last, std::string &) {
        const auto a = clock::now();
        const int l0 = sampler->position();
        sampler->draw_until(kept);
        scarf::append_table(s.setup, sampler->draw(), l0, sampler->position(), t);
        last = sampler->finished();
        host += std::chrono::duration<double>(clock::now() - a).count();
        return true;
    };
    const scarf::Geometry g = geometry(s.cfg);
    const scarf::Epilogue ep = epilogue(s.cfg, s.setup.ndim);
    double st[6] = {};
    bool restart = false;
    if (!s.device->run_streamed<real>(g, kernel, s.cfg.nharm, s.setup.scaling, ep, d_field, st, err, feed, &restart,
                                      on_slab, s.stream_chunk)) {
        if (!restart) return fail(err);
        // an amplitude the scaled-fp16 table cannot hold unscaled: the rest of
        // the draw, then the whole table with its power-of-two scale
        const auto a = clock::now();
        sampler->draw_until(INT_MAX);
        const scarf::HarmonicTable t = scarf::build_table(s.setup, sampler->draw());
        s.host_seconds = host + std::chrono::duration<double>(clock::now() - a).count();
        if (!s.device->upload_table(t, err)) return fail(err);
        s.table_ready = true;
        return run(d_field, stats, on_slab);
    }
    s.host_seconds = host;
    report_stats(stats, st);
    s.error.clear();
    return 0;
}

// Runs f; an exception from it is reported as the call's error instead of
// leaving through a C entry point.
template <class F>
void guarded(F &&f)
{
    try {
        f();
    } catch (const std::exception &e) {
        fail(std::string("internal error: ") + e.what());
    } catch (...) {
        fail("internal error");
    }
}

}  // namespace


extern "C" {

void scarf_opt_init(struct scarf_opt *var)
{
    if (!var) return;
    var->method = 0;
    var->hurst = 0;
    var->dh = 0;
    var->ds = 0;
    var->poi = nullptr;
    var->npoi = 0;
    var->mute = 0;
    var->taper = 0;
    var->rescale = 0;
    var->pad = 0;
    var->nc = nullptr;
    var->fc = nullptr;
    var->alpha = 0;
    var->beta = 0;
    var->gamma = 0;
}

// initialize_structured (src/scarflib.f90:97-189) behind struct_initialize
// (src/f_binding.f90:57-113) and the option rules of src/c_binding.c:89-103
static void scarf_cart_initialize_body(const int nd, const int fs[], const int fe[], const real ds, const int acf,
                           const real cl[], const real sigma, struct scarf_opt *var)
{
    State &s = state();
    if (s.initialized) {
        fail("initialize called twice without finalize");
        return;
    }
    if ((nd != 2 && nd != 3) || !fs || !fe || !cl) {
        fail("scarf_cart_initialize: nd must be 2 or 3 and fs, fe, cl non-NULL");
        return;
    }
    SrmConfig c;
    c.nd = nd;
    c.structured = true;
    for (int i = 0; i < 3; ++i) {
        c.fs[i] = 1;
        c.fe[i] = 1;
        c.cl[i] = 0;
    }
    for (int i = 0; i < nd; ++i) {
        c.fs[i] = fs[i];
        c.fe[i] = fe[i];
        c.cl[i] = cl[i];
        if (fe[i] < fs[i]) {
            fail("scarf_cart_initialize: fe < fs");
            return;
        }
    }
    c.ds = ds;
    c.dh = ds;
    c.acf = acf;
    c.sigma = sigma;
    c.method = 0;
    c.hurst = 0;
    c.mute = -999;
    c.taper = -999;
    c.rescale = 0;
    c.pad = 0;
    for (int i = 0; i < 3; ++i) {
        c.nc[i] = static_cast<real>(static_cast<real>(c.fs[i] - 1) * ds);
        c.fc[i] = static_cast<real>(static_cast<real>(c.fe[i] - 1) * ds);
    }
    if (var) {
        if (var->method == 1) c.method = 1;
        if (var->hurst > 0) c.hurst = var->hurst;
        if (var->dh > 0) c.dh = var->dh;
        if (var->mute > 0) c.mute = var->mute;
        if (var->taper > 0) c.taper = var->taper;
        if (var->rescale == 1) c.rescale = 1;
        if (var->pad == 1) c.pad = 1;
        if (var->nc && var->fc)
            for (int i = 0; i < nd; ++i) {
                c.nc[i] = var->nc[i];
                c.fc[i] = var->fc[i];
            }
        if (var->alpha != 0) c.alpha = var->alpha;
        if (var->beta != 0) c.beta = var->beta;
        if (var->gamma != 0) c.gamma = var->gamma;
        if (var->poi && var->npoi > 0) {
            c.npoi = var->npoi;
            c.poi.assign(var->poi, var->poi + size_t(nd) * var->npoi);
        }
    }
    s.cfg = c;
    s.x = s.y = s.z = nullptr;
    s.initialized = true;
    s.table_ready = false;
    s.error.clear();
}

// initialize_unstructured (src/scarflib.f90:195-296) behind unstruct_initialize
// (src/f_binding.f90:119-175) and src/c_binding.c:116-130
static void scarf_nocart_initialize_body(const int nd, const int npts, const real *x, const real *y, const real *z,
                             const real dh, const int acf, const real cl[], const real sigma, struct scarf_opt *var)
{
    State &s = state();
    if (s.initialized) {
        fail("initialize called twice without finalize");
        return;
    }
    if ((nd != 2 && nd != 3) || npts < 1 || !x || !y || !cl || (nd == 3 && !z)) {
        fail("scarf_nocart_initialize: nd must be 2 or 3, npts > 0, x/y (and z in 3D) and cl non-NULL");
        return;
    }
    SrmConfig c;
    c.nd = nd;
    c.structured = false;
    c.npts = npts;
    for (int i = 0; i < 3; ++i) c.cl[i] = 0;
    for (int i = 0; i < nd; ++i) c.cl[i] = cl[i];
    c.dh = dh;
    c.ds = dh;
    c.acf = acf;
    c.sigma = sigma;
    c.method = 0;
    c.hurst = 0;
    c.mute = -999;
    c.taper = -999;
    c.rescale = 0;
    c.pad = 0;
    const real *axes[3] = {x, y, nd == 3 ? z : nullptr};
    for (int a = 0; a < 3; ++a) {
        if (!axes[a]) {
            c.nc[a] = 0;
            c.fc[a] = 0;
            continue;
        }
        real lo = axes[a][0], hi = axes[a][0];
        for (int i = 1; i < npts; ++i) {
            lo = std::min(lo, axes[a][i]);
            hi = std::max(hi, axes[a][i]);
        }
        c.nc[a] = lo;
        c.fc[a] = hi;
    }
    if (var) {
        if (var->method == 1) c.method = 1;
        if (var->hurst > 0) c.hurst = var->hurst;
        if (var->ds > 0) c.ds = var->ds;
        if (var->mute > 0) c.mute = var->mute;
        if (var->taper > 0) c.taper = var->taper;
        if (var->rescale == 1) c.rescale = 1;
        if (var->pad == 1) c.pad = 1;
        if (var->nc && var->fc)
            for (int i = 0; i < nd; ++i) {
                c.nc[i] = var->nc[i];
                c.fc[i] = var->fc[i];
            }
        if (var->alpha != 0) c.alpha = var->alpha;
        if (var->beta != 0) c.beta = var->beta;
        if (var->gamma != 0) c.gamma = var->gamma;
        if (var->poi && var->npoi > 0) {
            c.npoi = var->npoi;
            c.poi.assign(var->poi, var->poi + size_t(nd) * var->npoi);
        }
    }
    s.cfg = c;
    s.x = x;
    s.y = y;
    s.z = nd == 3 ? z : nullptr;
    s.initialized = true;
    s.table_ready = false;
    s.error.clear();
}

static void scarf_execute_body(const int seed, real *field, real stats[])
{
    State &s = state();
    clear_stats(stats);
    if (!field) {
        fail("scarf_execute: field is NULL");
        return;
    }
    if (!s.initialized) {
        fail("execute called before initialize");
        return;
    }
    {  // parameter errors are reported before any device work
        s.cfg.nharm = s.nharm;
        std::string e;
        int rc;
        if (s.cfg.method == 0) {
            scarf::FimSetup<real> fs;
            rc = scarf::fim_setup<real>(s.cfg, fs, e);
        } else {
            rc = kDouble ? scarf::srm_setup<double>(s.cfg, s.setup, e) : scarf::srm_setup<float>(s.cfg, s.setup, e);
        }
        if (rc != 0) {
            fail(e);
            return;
        }
    }
    if (!ensure_device()) return;
    const long n = field_points(s.cfg);
    std::string err;
    void *d = s.device->field_buffer(size_t(n) * sizeof(real), err);
    if (!d) {
        fail(err);
        return;
    }
    // Fields above 64 MiB come back in z slabs, each copied (on a stream of
    // its own) while the slabs after it are computed; the device's run hands
    // the slabs over when they are final, or the whole field at the end.
    static hipStream_t copy_stream = nullptr;
    const scarf::Device::SlabHook to_host = [&](long first, long count, std::string &e) {
        if (!copy_stream) {
            const hipError_t rc = hipStreamCreate(&copy_stream);  // default priority
            if (rc != hipSuccess) {
                e = std::string("hipStreamCreate: ") + hipGetErrorString(rc);
                return false;
            }
        }
        return copy_to_host(field + first, static_cast<const real *>(d) + first, size_t(count) * sizeof(real),
                            copy_stream, e);
    };
    const bool slabs = size_t(n) * sizeof(real) > 2 * kStageBytes;
    if (execute_realization(seed, static_cast<real *>(d), stats, slabs ? &to_host : nullptr) != 0) return;
    if (!slabs && !copy_to_host(field, d, size_t(n) * sizeof(real), s.device->stream(), err)) fail(err);
}

void scarf_finalize(void)
{
    State &s = state();
    s.initialized = false;
    s.table_ready = false;
    s.cfg = SrmConfig{};
    s.x = s.y = s.z = nullptr;
    s.error.clear();
}

// Single-writer versions of write_one / write_slice (src/scarflib_common.f90:245-418,
// :527-646): raw native-endian `real`, column-major (x fastest).
static void scarf_io_one_body(const int nd, const int *npts, const real *field, const char fname[], const int *nwriters)
{
    (void)nwriters;
    if (!npts || !field || !fname || (nd != 2 && nd != 3)) {
        fail("scarf_io_one: bad arguments");
        return;
    }
    size_t n = 1;
    for (int i = 0; i < nd; ++i) {
        if (npts[i] < 1) return (void)fail("scarf_io_one: npts must be positive");
        n *= size_t(npts[i]);
    }
    FILE *f = std::fopen(fname, "wb");
    if (!f || std::fwrite(field, sizeof(real), n, f) != n) fail(std::string("scarf_io_one: cannot write ") + fname);
    if (f) std::fclose(f);
}

static void scarf_io_slice_body(const int *npts, const char axis[], const int plane, const real *field, const char fname[])
{
    if (!npts || !axis || !field || !fname) {
        fail("scarf_io_slice: bad arguments");
        return;
    }
    const long nx = npts[0], ny = npts[1], nz = npts[2];
    if (nx < 1 || ny < 1 || nz < 1) return (void)fail("scarf_io_slice: npts must be positive");
    std::vector<real> out;
    const long p = plane - 1;  // plane is 1-based
    if (axis[0] == 'x') {
        if (p < 0 || p >= nx) return (void)fail("scarf_io_slice: plane out of range");
        for (long k = 0; k < nz; ++k)
            for (long j = 0; j < ny; ++j) out.push_back(field[p + nx * (j + ny * k)]);
    } else if (axis[0] == 'y') {
        if (p < 0 || p >= ny) return (void)fail("scarf_io_slice: plane out of range");
        for (long k = 0; k < nz; ++k)
            for (long i = 0; i < nx; ++i) out.push_back(field[i + nx * (p + ny * k)]);
    } else if (axis[0] == 'z') {
        if (p < 0 || p >= nz) return (void)fail("scarf_io_slice: plane out of range");
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i) out.push_back(field[i + nx * (j + ny * p)]);
    } else {
        fail("scarf_io_slice: axis must be x, y or z");
        return;
    }
    FILE *f = std::fopen(fname, "wb");
    if (!f || std::fwrite(out.data(), sizeof(real), out.size(), f) != out.size())
        fail(std::string("scarf_io_slice: cannot write ") + fname);
    if (f) std::fclose(f);
}

// ---- extensions (scarf3d_srm.h) ----
int scarf_set_nharm(int nharm)
{
    if (nharm < 1) return fail("nharm must be positive");
    state().nharm = nharm;
    state().table_ready = false;
    return 0;
}

int scarf_get_nharm(void) { return state().nharm; }

int scarf_set_kernel(int kernel)
{
    if (kernel < scarf::kKernelAuto || kernel > scarf::kKernelMfmaBf16x2) return fail("unknown kernel");
    state().kernel = kernel;
    return 0;
}

int scarf_get_kernel(void)
{
    State &s = state();
    return s.initialized ? resolve_kernel(s.cfg) : current_kernel();
}

const char *scarf_last_error(void) { return state().error.c_str(); }

static int scarf_execute_device_body(const int seed, real *d_field, real stats[])
{
    clear_stats(stats);
    if (!d_field) return fail("scarf_execute_device: d_field is NULL");
    return execute_realization(seed, d_field, stats, nullptr);
}

static int scarf_srm_draw_body(const int seed, double *host_seconds)
{
    const int n = prepare(seed);
    if (host_seconds) *host_seconds = state().host_seconds;
    return n;
}

// scarf_srm_run with no device allocation (Device::run_noalloc): the
// operand tables, partials, statistics and points of interest go to the
// library's static device scratch array
static int scarf_srm_run_noalloc_body(real *d_field, real stats[])
{
    State &s = state();
    if (!d_field) return fail("scarf_srm_run_noalloc: d_field is NULL");
    if (!s.table_ready || !s.device) return fail("no harmonic table drawn (call scarf_srm_draw first)");
    const scarf::Geometry g = geometry(s.cfg);
    const scarf::Epilogue ep = epilogue(s.cfg, s.setup.ndim);
    const int kernel = resolve_kernel(s.cfg);
    double st[6] = {};
    std::string err;
    if (!s.device->run_noalloc<real>(g, kernel, s.setup.scaling, ep, d_field, st, err)) return fail(err);
    report_stats(stats, st);
    s.error.clear();
    return 0;
}

static int scarf_srm_run_body(real *d_field, real stats[])
{
    if (!d_field) return fail("scarf_srm_run: d_field is NULL");
    return run(d_field, stats);
}

static int scarf_ensemble_execute_body(int n, const struct scarf_member *members, real *const *d_fields, real *stats,
                           int kernel)
{
    State &s = state();
    if (n < 1 || !members || !d_fields) return fail("scarf_ensemble_execute: empty ensemble or NULL arguments");
    if (kernel == scarf::kKernelAuto) kernel = scarf::kKernelMfmaBf16x3;
    if (kernel != scarf::kKernelMfmaBf16x3 && kernel != scarf::kKernelMfmaF16x2 && kernel != scarf::kKernelMfmaBf16x2)
        return fail("scarf_ensemble_execute: ensembles run on the MFMA kernels");
    // member configurations with the defaults of initialize_structured
    std::vector<SrmConfig> cfgs(n);
    for (int m = 0; m < n; ++m) {
        const scarf_member &a = members[m];
        if ((a.nd != 2 && a.nd != 3) || !d_fields[m]) return fail("scarf_ensemble_execute: bad member " + std::to_string(m));
        SrmConfig &c = cfgs[m];
        c.nd = a.nd;
        c.structured = true;
        c.method = 1;
        for (int i = 0; i < 3; ++i) {
            c.fs[i] = i < a.nd ? a.fs[i] : 1;
            c.fe[i] = i < a.nd ? a.fe[i] : 1;
            c.cl[i] = i < a.nd ? a.cl[i] : 0;
            if (c.fe[i] < c.fs[i]) return fail("scarf_ensemble_execute: fe < fs in member " + std::to_string(m));
            c.nc[i] = static_cast<real>(static_cast<real>(c.fs[i] - 1) * a.ds);
            c.fc[i] = static_cast<real>(static_cast<real>(c.fe[i] - 1) * a.ds);
        }
        c.ds = a.ds;
        c.dh = a.dh > 0 ? a.dh : a.ds;
        c.acf = a.acf;
        c.sigma = a.sigma;
        c.hurst = a.hurst;
        c.alpha = a.alpha;
        c.beta = a.beta;
        c.gamma = a.gamma;
        c.rescale = a.rescale;
        c.nharm = a.nharm > 0 ? a.nharm : s.nharm;
    }
    // Host draws. Every member's setup first (one per thread), then its
    // expected draw cost, nharm over the rejection test's acceptance
    // (cnst * intg(kmin, kmax) / (kmax - kmin)). The members that would hold
    // up the rest (cost above a share of the total) are drawn one after the
    // other with every thread of the sampler, the others one per thread,
    // costliest first. The draws are the exact streams either way.
    std::vector<scarf::SrmSetup> setups(n);
    std::vector<std::vector<double>> records(n);
    std::vector<int> nps(n, 0);
    std::vector<double> host_s(n, 0.0), cost(n, 0.0);
    std::vector<std::string> errs(n);
    auto guard = [&](int m, auto &&f) {
        try {
            f();
        } catch (const std::exception &ex) {
            errs[m] = std::string("internal error: ") + ex.what();
        } catch (...) {
            errs[m] = "internal error";
        }
    };
    auto on_members = [&](const std::vector<int> &list, auto &&f) {
        std::atomic<int> next{0};
        const std::function<void(int)> worker = [&](int) {
            for (int q = next++; q < static_cast<int>(list.size()); q = next++) guard(list[q], [&] { f(list[q]); });
        };
        member_pool().run(worker);
    };
    std::vector<int> all(n);
    for (int m = 0; m < n; ++m) all[m] = m;
    on_members(all, [&](int m) {
        const auto t0 = std::chrono::steady_clock::now();
        std::string e;
        const int rc = kDouble ? scarf::srm_setup<double>(cfgs[m], setups[m], e)
                               : scarf::srm_setup<float>(cfgs[m], setups[m], e);
        if (rc != 0) errs[m] = e;
        const scarf::SrmSetup &su = setups[m];
        const double p = su.kmax > su.kmin ? su.cnst * su.band / (su.kmax - su.kmin) : 1.0;
        cost[m] = cfgs[m].nharm / std::min(1.0, std::max(p, 1e-6));
        host_s[m] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    });
    for (int m = 0; m < n; ++m)
        if (!errs[m].empty()) return fail("member " + std::to_string(m) + ": " + errs[m]);
    auto draw_member = [&](int m, int threads) {
        const auto t0 = std::chrono::steady_clock::now();
        const scarf::HarmonicDraw d =
            kDouble ? scarf::srm_draw<double>(cfgs[m], setups[m], members[m].seed, threads)
                    : scarf::srm_draw<float>(cfgs[m], setups[m], members[m].seed, threads);
        const scarf::HarmonicTable t = scarf::build_table(setups[m], d);
        const int np_ = std::max(64, (t.count() + 63) / 64 * 64);
        std::vector<double> &r = records[m];
        r.assign(size_t(5) * np_, 0.0);
        for (int l = 0; l < t.count(); ++l) {
            r[l] = t.wx[l];
            r[np_ + l] = t.wy[l];
            r[2 * size_t(np_) + l] = t.wz[l];
            r[3 * size_t(np_) + l] = t.phase[l];
            r[4 * size_t(np_) + l] = t.amp[l];
        }
        nps[m] = np_;
        host_s[m] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    std::vector<int> order = all;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    const int threads = member_pool().size();
    double total = 0.0;
    for (int m = 0; m < n; ++m) total += cost[m];
    std::vector<int> light;
    for (int m : order) {
        if (threads > 1 && cost[m] > total / (2.0 * threads)) {
            guard(m, [&] { draw_member(m, threads); });  // the sampler's own workers
        } else {
            light.push_back(m);
        }
    }
    on_members(light, [&](int m) { draw_member(m, 1); });
    for (int m = 0; m < n; ++m)
        if (!errs[m].empty()) return fail("member " + std::to_string(m) + ": " + errs[m]);
    if (!ensure_device()) return -1;
    if (!s.ensemble) s.ensemble = std::make_unique<scarf::Ensemble>();
    std::vector<scarf::EnsembleMember> em(n);
    for (int m = 0; m < n; ++m) {
        em[m].g = geometry(cfgs[m]);
        em[m].np_ = nps[m];
        em[m].record = records[m].data();
        for (int l = 0; l < nps[m]; ++l) em[m].amp_max = std::max(em[m].amp_max, records[m][4 * size_t(nps[m]) + l]);
        em[m].scaling = setups[m].scaling;
        em[m].rescale = cfgs[m].rescale;
        em[m].sigma = cfgs[m].sigma;
        em[m].d_field = d_fields[m];
    }
    // tiles per member: in one full launch a member's tile is worth its
    // padded work over the tile's saturated throughput (timed once)
    if (scarf::TileTuner::enabled()) {
        std::string terr;
        for (int m = 0; m < n; ++m) {
            scarf::TileChoice tc;
            if (!s.device->tuner().choose_saturated<real>(em[m].g, kernel, s.device->stream(), tc, terr))
                return fail("tile tuner: " + terr);
            em[m].tk = tc.tk;
        }
    }
    std::vector<double> st;
    double ms = 0;
    std::string err;
    if (!s.ensemble->run<real>(s.device->stream(), kernel, em, st, ms, err)) return fail(err);
    s.ensemble_ms = ms;
    if (stats) {
        for (int m = 0; m < n; ++m) {
            real *o = stats + 8 * size_t(m);
            for (int i = 0; i < 8; ++i) o[i] = 0;
            o[0] = static_cast<real>(setups[m].info1);
            o[1] = static_cast<real>(setups[m].info2);
            o[2] = cfgs[m].rescale == 1 ? static_cast<real>(cfgs[m].sigma) : static_cast<real>(st[6 * m + 3]);
            o[3] = cfgs[m].rescale == 1 ? real(0) : static_cast<real>(st[6 * m + 4]);
            o[4] = static_cast<real>(host_s[m]);
            o[5] = static_cast<real>(ms * 1e-3);
        }
    }
    s.error.clear();
    return 0;
}

double scarf_ensemble_last_ms(void) { return state().ensemble_ms; }

int scarf_last_tile(int out[3])
{
    State &s = state();
    if (!out) return fail("scarf_last_tile: out is NULL");
    out[0] = out[1] = out[2] = 0;
    if (!s.device || s.device->last_nbx() == 0) return 0;
    out[0] = 32 * s.device->last_nbx();
    out[2] = s.device->last_tk();
    out[1] = 256 / out[2];
    return 3;
}

static int scarf_srm_run_fp64_body(double *d_field, double stats_out[4])
{
    State &s = state();
    if (!d_field) return fail("scarf_srm_run_fp64: d_field is NULL");
    if (!s.table_ready) return fail("no harmonic table drawn (call scarf_srm_draw first)");
    const scarf::Geometry g = geometry(s.cfg);
    const scarf::Epilogue ep = epilogue(s.cfg, s.setup.ndim);
    double st[6] = {};
    std::string err;
    if (!s.device->run<double>(g, scarf::kKernelDirectF64, s.setup.scaling, ep, d_field, st, err)) return fail(err);
    if (stats_out)
        for (int i = 0; i < 4; ++i) stats_out[i] = st[i];
    s.error.clear();
    return 0;
}

int scarf_last_timings(double *ms, int n)
{
    State &s = state();
    if (!s.device || !ms) return 0;
    const scarf::Timings &t = s.device->timings();
    const double v[5] = {t.tables, t.field, t.stats, t.epilogue, t.total};
    const int m = std::min(n, 5);
    for (int i = 0; i < m; ++i) ms[i] = v[i];
    return m;
}

void *scarf_stream(void)
{
    if (!ensure_device()) return nullptr;
    return reinterpret_cast<void *>(state().device->stream());
}

static int scarf_setup_values_body(double *out, int n)
{
    State &s = state();
    if (!s.initialized) return fail("not initialized");
    if (!out) return fail("scarf_setup_values: out is NULL");
    s.cfg.nharm = s.nharm;
    scarf::SrmSetup su;
    std::string err;
    const int rc = kDouble ? scarf::srm_setup<double>(s.cfg, su, err) : scarf::srm_setup<float>(s.cfg, su, err);
    if (rc != 0) return fail(err);
    double v[21] = {double(su.ndim), su.kmin, su.kmax, su.kc, su.info1, su.info2, su.scaling, su.band, su.all};
    for (int i = 0; i < 9; ++i) v[9 + i] = su.matrix[i];
    for (int i = 0; i < 3; ++i) v[18 + i] = su.bar[i];
    for (int i = 0; i < std::min(n, 21); ++i) out[i] = v[i];
    return 0;
}

static int scarf_set_option_body(const char *name, double value)
{
    State &s = state();
    if (!s.initialized) return fail("scarf_set_option: not initialized");
    if (!name) return fail("scarf_set_option: name is NULL");
    SrmConfig &c = s.cfg;
    const std::string n(name);
    const double v = static_cast<real>(value);
    // host draw execution (never changes the harmonics drawn)
    if (n == "draw_threads" || n == "draw_block" || n == "stream_draw" || n == "draw_chunk" || n == "stream_chunk") {
        if (value < 0) return fail("scarf_set_option: " + n + " must be >= 0");
        if (n == "stream_chunk" && static_cast<long>(value) % scarf::kFieldChunkHarmonics != 0)
            return fail("scarf_set_option: stream_chunk must be a multiple of 128");
        if (n == "stream_chunk") s.stream_chunk = static_cast<int>(value);
        else if (n == "draw_threads") s.draw_threads = static_cast<int>(value);
        else if (n == "draw_block") s.draw_block = static_cast<long>(value);
        else if (n == "draw_chunk") s.draw_chunk = static_cast<int>(value);
        else s.stream_draw = value != 0;
        s.error.clear();
        return 0;
    }
    if (n == "method") c.method = static_cast<int>(value);
    else if (n == "hurst") c.hurst = v;
    else if (n == "dh") c.dh = v;
    else if (n == "ds") c.ds = v;
    else if (n == "mute") c.mute = v;
    else if (n == "taper") c.taper = v;
    else if (n == "rescale") c.rescale = static_cast<int>(value);
    else if (n == "pad") c.pad = static_cast<int>(value);
    else if (n == "alpha") c.alpha = v;
    else if (n == "beta") c.beta = v;
    else if (n == "gamma") c.gamma = v;
    else if (n.size() == 3 && (n.rfind("nc", 0) == 0 || n.rfind("fc", 0) == 0) && n[2] >= '1' && n[2] <= '3')
        (n[0] == 'n' ? c.nc : c.fc)[n[2] - '1'] = v;
    else
        return fail("scarf_set_option: unknown option " + n);
    s.table_ready = false;
    s.error.clear();
    return 0;
}

int scarf_get_config(double *out, int n)
{
    State &s = state();
    if (!s.initialized) return fail("not initialized");
    if (!out) return fail("scarf_get_config: out is NULL");
    const SrmConfig &c = s.cfg;
    const double v[32] = {double(c.nd), double(c.structured), double(c.fs[0]), double(c.fs[1]), double(c.fs[2]),
                          double(c.fe[0]), double(c.fe[1]), double(c.fe[2]), c.ds, c.dh, double(c.acf), c.cl[0],
                          c.cl[1], c.cl[2], c.sigma, c.hurst, c.nc[0], c.nc[1], c.nc[2], c.fc[0], c.fc[1], c.fc[2],
                          c.alpha, c.beta, c.gamma, c.mute, c.taper, double(c.rescale), double(c.pad),
                          double(c.method), double(c.npoi), double(c.npts)};
    const int m = std::min(n, 32);
    for (int i = 0; i < m; ++i) out[i] = v[i];
    return m;
}

static long scarf_draw_harmonics_body(const int seed, double *harm, unsigned char *used)
{
    State &s = state();
    if (!s.initialized) return fail("not initialized");
    if (!harm || !used) return fail("scarf_draw_harmonics: harm and used must be non-NULL");
    s.cfg.nharm = s.nharm;
    scarf::SrmSetup su;
    std::string err;
    const int rc = kDouble ? scarf::srm_setup<double>(s.cfg, su, err) : scarf::srm_setup<float>(s.cfg, su, err);
    if (rc != 0) return fail(err);
    // the walk cut as the streamed execute cuts it (or every draw_chunk
    // kept harmonics): the harmonics must not depend on the cuts
    std::unique_ptr<scarf::HarmonicSampler> w =
        kDouble ? scarf::HarmonicSampler::create<double>(s.cfg, su, seed, s.draw_threads, s.draw_block)
                : scarf::HarmonicSampler::create<float>(s.cfg, su, seed, s.draw_threads, s.draw_block);
    if (s.draw_chunk > 0) {
        for (int kept = s.draw_chunk; !w->finished(); kept += s.draw_chunk) w->draw_until(kept);
    } else {
        for (int kept : scarf::Device::chunk_targets(s.cfg.nharm, s.stream_chunk)) w->draw_until(kept);
    }
    const scarf::HarmonicDraw &d = w->draw();
    for (int l = 0; l < s.cfg.nharm; ++l) {
        harm[5 * l] = d.v1[l];
        harm[5 * l + 1] = d.v2[l];
        harm[5 * l + 2] = d.v3[l];
        harm[5 * l + 3] = d.a[l];
        harm[5 * l + 4] = d.b[l];
        used[l] = d.used[l];
    }
    return d.uniforms;
}

// ---- the exported entry points that allocate or start threads: an
// exception (std::bad_alloc, std::system_error) never crosses the C ABI; it
// becomes the entry point's error.
void scarf_cart_initialize(const int nd, const int fs[], const int fe[], const real ds, const int acf,
                           const real cl[], const real sigma, struct scarf_opt *var)
{
    guarded([&] { scarf_cart_initialize_body(nd, fs, fe, ds, acf, cl, sigma, var); });
}

void scarf_nocart_initialize(const int nd, const int npts, const real *x, const real *y, const real *z,
                             const real dh, const int acf, const real cl[], const real sigma, struct scarf_opt *var)
{
    guarded([&] { scarf_nocart_initialize_body(nd, npts, x, y, z, dh, acf, cl, sigma, var); });
}

void scarf_execute(const int seed, real *field, real stats[])
{
    guarded([&] { scarf_execute_body(seed, field, stats); });
}

void scarf_io_one(const int nd, const int *npts, const real *field, const char fname[], const int *nwriters)
{
    guarded([&] { scarf_io_one_body(nd, npts, field, fname, nwriters); });
}

void scarf_io_slice(const int *npts, const char axis[], const int plane, const real *field, const char fname[])
{
    guarded([&] { scarf_io_slice_body(npts, axis, plane, field, fname); });
}

int scarf_execute_device(const int seed, real *d_field, real stats[])
{
    int rc = -1;
    guarded([&] { rc = scarf_execute_device_body(seed, d_field, stats); });
    return rc;
}

int scarf_srm_draw(const int seed, double *host_seconds)
{
    int rc = -1;
    guarded([&] { rc = scarf_srm_draw_body(seed, host_seconds); });
    return rc;
}

int scarf_srm_run(real *d_field, real stats[])
{
    int rc = -1;
    guarded([&] { rc = scarf_srm_run_body(d_field, stats); });
    return rc;
}

int scarf_ensemble_execute(int n, const struct scarf_member *members, real *const *d_fields, real *stats,
                           int kernel)
{
    int rc = -1;
    guarded([&] { rc = scarf_ensemble_execute_body(n, members, d_fields, stats, kernel); });
    return rc;
}

int scarf_setup_values(double *out, int n)
{
    int rc = -1;
    guarded([&] { rc = scarf_setup_values_body(out, n); });
    return rc;
}

int scarf_set_option(const char *name, double value)
{
    int rc = -1;
    guarded([&] { rc = scarf_set_option_body(name, value); });
    return rc;
}

long scarf_draw_harmonics(const int seed, double *harm, unsigned char *used)
{
    long rc = -1;
    guarded([&] { rc = scarf_draw_harmonics_body(seed, harm, used); });
    return rc;
}

int scarf_srm_run_noalloc(real *d_field, real stats[])
{
    int rc = -1;
    guarded([&] { rc = scarf_srm_run_noalloc_body(d_field, stats); });
    return rc;
}

int scarf_srm_run_fp64(double *d_field, double stats_out[4])
{
    int rc = -1;
    guarded([&] { rc = scarf_srm_run_fp64_body(d_field, stats_out); });
    return rc;
}

}  // extern "C"
