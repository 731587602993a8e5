// gfx950 kernels of the SRM direct sum and its epilogue, and the launch
// sequence that ties them together (Device::run).
//
// Reference semantics (src/scarflib_srm.f90:794-851, :397-436):
//   field(p) = scaling * sum_n [a_n sin(arg_n(p)) + b_n cos(arg_n(p))],
//   arg_n(p) = v_n . (M p + bar).
// The host rewrites each harmonic as amp_n cos(w_n . p + phase_n) (srm_host).
// On a Cartesian grid the phase splits per axis,
//   w.p + phase = wx*x_i + wy*y_j + (wz*z_k + phase),
// so three small tables of phases, reduced modulo one turn in FP64, replace
// the per-point FP32 products that lose ~1e-4 rad at |arg| ~ 3e3 rad. The
// FP32 kernel then spends one add, one v_cos_f32 (argument in turns) and one
// FMA per point-harmonic; partial sums are promoted to FP64 every 128
// harmonics so that FP32 rounding does not grow with the harmonic count.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "srm_device.hpp"
#include "srm_mfma.hpp"
#include "srm_stats.hpp"

namespace scarf {

namespace {

constexpr double kInvTwoPi = 0.15915494309189533577;  // 1 / (2 pi)
constexpr int kHarmPad = 64;                          // harmonic count padding

__host__ __device__ inline long round_up(long v, long m) { return (v + m - 1) / m * m; }

// Final merge of the per-block partials in a fixed order, then the `real`
// roundings of the reference: info(3) = sqrt(real(var)), info(4) = real(mean),
// rescale factor sigma / info(3) (src/scarflib_srm.f90:880-899).
template <class Real>
__global__ __launch_bounds__(256) void stats_finalize(const double *__restrict__ part, int nparts,
                                                      double sigma, double *__restrict__ stat)
{
    __shared__ double sm[3][256];
    Moments m{0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nparts; b += 256) merge_moments(m, Moments{part[3 * b], part[3 * b + 1], part[3 * b + 2]});
    sm[0][threadIdx.x] = m.n;
    sm[1][threadIdx.x] = m.mean;
    sm[2][threadIdx.x] = m.m2;
    __syncthreads();
    for (int stride = 128; stride >= 1; stride >>= 1) {
        if (threadIdx.x < stride) {
            Moments a{sm[0][threadIdx.x], sm[1][threadIdx.x], sm[2][threadIdx.x]};
            merge_moments(a, Moments{sm[0][threadIdx.x + stride], sm[1][threadIdx.x + stride], sm[2][threadIdx.x + stride]});
            sm[0][threadIdx.x] = a.n;
            sm[1][threadIdx.x] = a.mean;
            sm[2][threadIdx.x] = a.m2;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double n = sm[0][0], mean = sm[1][0], m2 = sm[2][0];
        // n - 1 divisor as the reference's variance (scarflib_common.f90:684):
        // a one-point field gives 0/0 = NaN there and here
        const double var = m2 / (n - 1.0);
        const Real info3 = sqrt(static_cast<Real>(var));
        const Real info4 = static_cast<Real>(mean);
        const Real sc = static_cast<Real>(sigma) / info3;
        stat[0] = n;
        stat[1] = mean;
        stat[2] = var;
        stat[3] = info3;
        stat[4] = info4;
        stat[5] = sc;
    }
}

// ---------------------------------------------------------------------------
// per-axis phase tables, in turns, reduced to [-1/2, 1/2] in FP64
//   tx[n][i] = frac(wx_n * x_i / 2pi), ty[n][j] likewise,
//   tz[n][k] = frac((wz_n * z_k + phase_n) / 2pi)
template <class T>
__global__ __launch_bounds__(256) void axis_tables(const double *__restrict__ harm, int np_, int nx, int ny,
                                                   int nz, int ox, int oy, int oz, double ds,
                                                   T *__restrict__ tx, T *__restrict__ ty, T *__restrict__ tz,
                                                   float *__restrict__ amp32)
{
    const long total = long(np_) * (nx + ny + nz);
    for (long e = blockIdx.x * long(blockDim.x) + threadIdx.x; e < total; e += long(gridDim.x) * blockDim.x) {
        const int n = int(e / (nx + ny + nz));
        const int r = int(e % (nx + ny + nz));
        double t;
        if (r < nx) {
            t = (harm[n] * (double(r + ox) * ds)) * kInvTwoPi;
            tx[long(n) * nx + r] = static_cast<T>(t - rint(t));
        } else if (r < nx + ny) {
            const int j = r - nx;
            t = (harm[np_ + n] * (double(j + oy) * ds)) * kInvTwoPi;
            ty[long(n) * ny + j] = static_cast<T>(t - rint(t));
        } else {
            const int k = r - nx - ny;
            t = (harm[2 * np_ + n] * (double(k + oz) * ds) + harm[3 * np_ + n]) * kInvTwoPi;
            tz[long(n) * nz + k] = static_cast<T>(t - rint(t));
        }
        if (amp32 && r == 0) amp32[n] = static_cast<float>(harm[4 * np_ + n]);
    }
}

// ---------------------------------------------------------------------------
// FP32 direct sum on a Cartesian grid.
// Block tile: 64 x-points by 64 columns (column c = j + ny*k), 256 threads,
// each thread 4 consecutive x-points by 4 columns. Harmonics stream through
// LDS 32 at a time: the tile's x-phases, the tile's (y+z)-phases (summed and
// re-reduced once per column-harmonic) and the amplitudes.
constexpr int kDirH = 32;
constexpr int kDirB = 64;

template <class Out>
__global__ __launch_bounds__(256) void direct_grid_f32(const float *__restrict__ tx, const float *__restrict__ ty,
                                                       const float *__restrict__ tz,
                                                       const float *__restrict__ amp, int np_, int nx, int ny,
                                                       int nz, int tiles_i, double scaling,
                                                       Out *__restrict__ field, double *__restrict__ part)
{
    __shared__ float4 sx[kDirH][kDirB / 4];
    __shared__ float4 syz[kDirH][kDirB / 4];
    __shared__ float sa[kDirH];
    const int tid = threadIdx.x;
    const int ti = tid & 15, tc = tid >> 4;
    const int i0 = (blockIdx.x % tiles_i) * kDirB;
    const long c0 = long(blockIdx.x / tiles_i) * kDirB;
    const long ncols = long(ny) * nz;

    // loader role: one x-point / one column, harmonic rows lh, lh+8, ...
    const int lcol = tid & 63, lh = tid >> 6;
    const long lc = c0 + lcol;
    const bool lc_ok = lc < ncols;
    const int lj = lc_ok ? int(lc % ny) : 0, lk = lc_ok ? int(lc / ny) : 0;
    const int li = i0 + lcol;
    const bool li_ok = li < nx;
    float *sxf = reinterpret_cast<float *>(sx);
    float *syzf = reinterpret_cast<float *>(syz);

    float acc[4][4];
    double acc64[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            acc[a][b] = 0.f;
            acc64[a][b] = 0.0;
        }

    int tile = 0;
    for (int n0 = 0; n0 < np_; n0 += kDirH, ++tile) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kDirH / 4; ++q) {
            const int h = lh + 4 * q;
            const long n = n0 + h;
            sxf[h * kDirB + lcol] = li_ok ? tx[n * nx + li] : 0.f;
            float t = 0.f;
            if (lc_ok) {
                t = ty[n * ny + lj] + tz[n * nz + lk];
                t -= rintf(t);
            }
            syzf[h * kDirB + lcol] = t;
        }
        if (tid < kDirH) sa[tid] = amp[n0 + tid];
        __syncthreads();
#pragma unroll 4
        for (int h = 0; h < kDirH; ++h) {
            const float4 xv = sx[h][ti];
            const float4 yv = syz[h][tc];
            const float a = sa[h];
            const float xs[4] = {xv.x, xv.y, xv.z, xv.w};
            const float ys[4] = {yv.x, yv.y, yv.z, yv.w};
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[p][c] = fmaf(a, __builtin_amdgcn_cosf(xs[p] + ys[c]), acc[p][c]);
        }
        if ((tile & 3) == 3) {
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    acc64[p][c] += static_cast<double>(acc[p][c]);
                    acc[p][c] = 0.f;
                }
        }
    }

    double v[16];
    bool ok[16];
    const int ib = i0 + 4 * ti;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const long col = c0 + 4 * tc + c;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const Out val = static_cast<Out>((acc64[p][c] + static_cast<double>(acc[p][c])) * scaling);
            ok[4 * c + p] = (col < ncols) && (ib + p < nx);
            v[4 * c + p] = static_cast<double>(val);
            if (ok[4 * c + p]) field[col * nx + ib + p] = val;
        }
    }
    block_moments(local_moments(v, ok), part + 3 * blockIdx.x);
}

// FP64 direct sum on a Cartesian grid (exact path): same tiling, FP64 phase
// tables, cos(2 pi t) = cospi(2 t) in FP64, FP64 accumulation.
template <class Out>
__global__ __launch_bounds__(256) void direct_grid_f64(const double *__restrict__ tx, const double *__restrict__ ty,
                                                       const double *__restrict__ tz,
                                                       const double *__restrict__ amp, int np_, int nx, int ny,
                                                       int nz, int tiles_i, double scaling,
                                                       Out *__restrict__ field, double *__restrict__ part)
{
    constexpr int H = 16;
    __shared__ double sx[H][kDirB];
    __shared__ double syz[H][kDirB];
    __shared__ double sa[H];
    const int tid = threadIdx.x;
    const int ti = tid & 15, tc = tid >> 4;
    const int i0 = (blockIdx.x % tiles_i) * kDirB;
    const long c0 = long(blockIdx.x / tiles_i) * kDirB;
    const long ncols = long(ny) * nz;
    const int lcol = tid & 63, lh = tid >> 6;
    const long lc = c0 + lcol;
    const bool lc_ok = lc < ncols;
    const int lj = lc_ok ? int(lc % ny) : 0, lk = lc_ok ? int(lc / ny) : 0;
    const int li = i0 + lcol;
    const bool li_ok = li < nx;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int n0 = 0; n0 < np_; n0 += H) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < H / 4; ++q) {
            const int h = lh + 4 * q;
            const long n = n0 + h;
            sx[h][lcol] = li_ok ? tx[n * nx + li] : 0.0;
            double t = 0.0;
            if (lc_ok) {
                t = ty[n * ny + lj] + tz[n * nz + lk];
                t -= rint(t);
            }
            syz[h][lcol] = t;
        }
        if (tid < H) sa[tid] = amp[n0 + tid];
        __syncthreads();
        for (int h = 0; h < H; ++h) {
            const double a = sa[h];
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    acc[p][c] = fma(a, cospi(2.0 * (sx[h][4 * ti + p] + syz[h][4 * tc + c])), acc[p][c]);
        }
    }
    double v[16];
    bool ok[16];
    const int ib = i0 + 4 * ti;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const long col = c0 + 4 * tc + c;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const Out val = static_cast<Out>(acc[p][c] * scaling);
            ok[4 * c + p] = (col < ncols) && (ib + p < nx);
            v[4 * c + p] = static_cast<double>(val);
            if (ok[4 * c + p]) field[col * nx + ib + p] = val;
        }
    }
    block_moments(local_moments(v, ok), part + 3 * blockIdx.x);
}

// ---------------------------------------------------------------------------
// Direct sum on a point list: the phase of every point-harmonic is formed
// and reduced modulo one turn in FP64 (|w.p| reaches ~1e3 rad, where FP32
// keeps only ~1e-4 rad); cos and the accumulation run in FP32 (or FP64 for
// the exact path). 4 points per thread, harmonics staged 64 at a time.
constexpr int kPtsH = 64;
constexpr int kPtsP = 4;

template <class Out, bool F64>
__global__ __launch_bounds__(256) void direct_points(const double *__restrict__ px, const double *__restrict__ py,
                                                     const double *__restrict__ pz, long npts,
                                                     const double *__restrict__ harm, int np_, double scaling,
                                                     Out *__restrict__ field, double *__restrict__ part)
{
    __shared__ double sw[kPtsH][4];
    __shared__ double sa[kPtsH];
    const int tid = threadIdx.x;
    const long base = long(blockIdx.x) * 256 * kPtsP;
    double x[kPtsP], y[kPtsP], z[kPtsP];
    bool ok[kPtsP];
#pragma unroll
    for (int q = 0; q < kPtsP; ++q) {
        const long p = base + tid + 256 * q;
        ok[q] = p < npts;
        x[q] = ok[q] ? px[p] : 0.0;
        y[q] = ok[q] ? py[p] : 0.0;
        z[q] = ok[q] ? pz[p] : 0.0;
    }
    float acc[kPtsP] = {};
    double acc64[kPtsP] = {};
    int tile = 0;
    for (int n0 = 0; n0 < np_; n0 += kPtsH, ++tile) {
        __syncthreads();
        if (tid < kPtsH) {
            const int n = n0 + tid;
            sw[tid][0] = harm[n] * kInvTwoPi;
            sw[tid][1] = harm[np_ + n] * kInvTwoPi;
            sw[tid][2] = harm[2 * np_ + n] * kInvTwoPi;
            sw[tid][3] = harm[3 * np_ + n] * kInvTwoPi;
            sa[tid] = harm[4 * np_ + n];
        }
        __syncthreads();
        for (int h = 0; h < kPtsH; ++h) {
            const double wx = sw[h][0], wy = sw[h][1], wz = sw[h][2], ph = sw[h][3];
#pragma unroll
            for (int q = 0; q < kPtsP; ++q) {
                // phase in turns, reduced to [0, 1) by one v_fract_f64 (exact:
                // t - floor(t)); cos(2 pi t) is the same for t - rint(t)
                const double t = __builtin_amdgcn_fract(fma(wx, x[q], fma(wy, y[q], fma(wz, z[q], ph))));
                if constexpr (F64) {
                    acc64[q] = fma(sa[h], cospi(2.0 * t), acc64[q]);
                } else {
                    acc[q] = fmaf(static_cast<float>(sa[h]), __builtin_amdgcn_cosf(static_cast<float>(t)), acc[q]);
                }
            }
        }
        if constexpr (!F64) {
            if ((tile & 1) == 1) {
#pragma unroll
                for (int q = 0; q < kPtsP; ++q) {
                    acc64[q] += static_cast<double>(acc[q]);
                    acc[q] = 0.f;
                }
            }
        }
    }
    double v[kPtsP];
#pragma unroll
    for (int q = 0; q < kPtsP; ++q) {
        const Out val = static_cast<Out>((acc64[q] + static_cast<double>(acc[q])) * scaling);
        v[q] = static_cast<double>(val);
        if (ok[q]) field[base + tid + 256 * q] = val;
    }
    block_moments(local_moments(v, ok), part + 3 * blockIdx.x);
}

// ---------------------------------------------------------------------------
// Epilogue: rescale (field*s - mean, src/scarflib_srm.f90:897-912) and the
// mute/taper of every point of interest in order (taper_factor), at `real`
// precision without contraction, as the reference computes them on the host.
template <class Real>
__global__ __launch_bounds__(256) void epilogue_grid(Real *__restrict__ field, int nx, int ny, int nz, int ox,
                                                     int oy, int oz, double ds_, const double *__restrict__ stat,
                                                     int rescale, const double *__restrict__ poi, int npoi,
                                                     double mute_, double taper_)
{
#pragma clang fp contract(off)
    const long n = long(nx) * ny * nz;
    const Real ds = static_cast<Real>(ds_), mute = static_cast<Real>(mute_), taper = static_cast<Real>(taper_);
    const Real sc = static_cast<Real>(stat[5]), mu = static_cast<Real>(stat[4]);
    for (long e = blockIdx.x * long(blockDim.x) + threadIdx.x; e < n; e += long(gridDim.x) * blockDim.x) {
        Real v = field[e];
        if (rescale) v = v * sc - mu;
        if (npoi > 0) {
            const int i = int(e % nx), j = int((e / nx) % ny), k = int(e / (long(nx) * ny));
            for (int q = 0; q < npoi; ++q) {
                Real dx = static_cast<Real>(i + ox) * ds - static_cast<Real>(poi[3 * q]);
                Real dy = static_cast<Real>(j + oy) * ds - static_cast<Real>(poi[3 * q + 1]);
                Real dz = static_cast<Real>(k + oz) * ds - static_cast<Real>(poi[3 * q + 2]);
                v = taper_factor(v, dx * dx, dy * dy, dz * dz, mute, taper);
            }
        }
        field[e] = v;
    }
}

template <class Real>
__global__ __launch_bounds__(256) void epilogue_points(Real *__restrict__ field, long n, const double *__restrict__ px,
                                                       const double *__restrict__ py, const double *__restrict__ pz,
                                                       const double *__restrict__ stat, int rescale,
                                                       const double *__restrict__ poi, int npoi, double mute_,
                                                       double taper_)
{
#pragma clang fp contract(off)
    const Real mute = static_cast<Real>(mute_), taper = static_cast<Real>(taper_);
    const Real sc = static_cast<Real>(stat[5]), mu = static_cast<Real>(stat[4]);
    for (long e = blockIdx.x * long(blockDim.x) + threadIdx.x; e < n; e += long(gridDim.x) * blockDim.x) {
        Real v = field[e];
        if (rescale) v = v * sc - mu;
        for (int q = 0; q < npoi; ++q) {
            const Real ax = static_cast<Real>(px[e]) - static_cast<Real>(poi[3 * q]);
            const Real ay = static_cast<Real>(py[e]) - static_cast<Real>(poi[3 * q + 1]);
            const Real az = static_cast<Real>(pz[e]) - static_cast<Real>(poi[3 * q + 2]);
            v = taper_factor(v, ax * ax, ay * ay, az * az, mute, taper);
        }
        field[e] = v;
    }
}

}  // namespace

// ===========================================================================
const char *kernel_name(int k)
{
    switch (k) {
    case kKernelAuto: return "auto";
    case kKernelDirectF32: return "direct_f32";
    case kKernelDirectF64: return "direct_f64";
    case kKernelMfmaBf16x3: return "mfma_bf16x3";
    case kKernelMfmaF16x2: return "mfma_f16x2";
    case kKernelMfmaBf16x2: return "mfma_bf16x2";
    default: return "unknown";
    }
}

bool Epilogue::active() const
{
    if (rescale == 1) return true;
    // with mute + taper <= 0 no distance satisfies either branch (defaults -999)
    return !poi3.empty() && (mute + taper > 0.0 || mute >= 0.0);
}

// Scratch of Device::run_noalloc: operand tables, statistics partials,
// statistics and points of interest, for callers that cannot allocate device
// memory. Each GPU the library runs on holds its own copy for the life of
// the process. In host code the array's name is not a device address: host
// code uses the address hipGetSymbolAddress gives (Device::scratch_), taken
// once per device in Device::ensure.
static __device__ __align__(256) unsigned char scratch[Device::kNoAllocScratch];

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        const hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                              \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);                         \
            return false;                                                                    \
        }                                                                                    \
    } while (0)

TileTuner &Device::tuner()
{
    if (!tuner_) tuner_.reset(new TileTuner);
    return *tuner_;
}

Device::~Device()
{
    if (device_ < 0) return;
    hipSetDevice(device_);
    if (stream_) hipStreamSynchronize(stream_);
    void *bufs[] = {d_harm_, d_tab_, d_pts_, d_part_, d_stat_, d_field_, d_poi_};
    for (void *b : bufs)
        if (b) hipFree(b);
    if (h_stat_) hipHostFree(h_stat_);
    if (h_rec_) hipHostFree(h_rec_);
    for (auto &e : ev_)
        if (e) hipEventDestroy(e);
    for (auto &e : slab_ev_)
        if (e) hipEventDestroy(e);
    if (stream_) hipStreamDestroy(stream_);
}

bool Device::ensure(std::string &err)
{
    if (device_ >= 0) {
        HIP_TRY(hipSetDevice(device_));
        return true;
    }
    int count = 0;
    const hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count < 1) {
        err = "no HIP device available (SRM and FIM run only on the GPU; there is no CPU fallback)";
        return false;
    }
    HIP_TRY(hipGetDevice(&device_));
    HIP_TRY(hipSetDevice(device_));
    HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    for (auto &ev : ev_) HIP_TRY(hipEventCreate(&ev));
    HIP_TRY(hipMalloc(&d_stat_, 16 * sizeof(double)));
    HIP_TRY(hipHostMalloc(&h_stat_, 16 * sizeof(double)));
    // this device's copy of the no-alloc scratch array (with the device current)
    HIP_TRY(hipGetSymbolAddress(&scratch_, HIP_SYMBOL(scratch)));
    return true;
}

bool Device::reserve(void **ptr, size_t &cap, size_t bytes, std::string &err)
{
    const bool grow = !(bytes <= cap && *ptr);
    if (const char *e = std::getenv("SCARF_DEVICE_MEMORY_LIMIT")) {
        const double limit = std::atof(e);
        const size_t after = reserved_ + external_ + (grow ? bytes - std::min(bytes, cap) : 0);
        if (limit > 0 && double(after) > limit) {
            err = "device memory limit SCARF_DEVICE_MEMORY_LIMIT=" + std::string(e) + " bytes: the library holds " +
                  std::to_string(reserved_ + external_) + " and would need " + std::to_string(after);
            return false;
        }
    }
    if (!grow) return true;
    if (*ptr) {
        HIP_TRY(hipStreamSynchronize(stream_));
        HIP_TRY(hipFree(*ptr));
        *ptr = nullptr;
        reserved_ -= cap;
        cap = 0;
    }
    HIP_TRY(hipMalloc(ptr, std::max<size_t>(bytes, 256)));
    cap = bytes;
    reserved_ += cap;
    return true;
}

void *Device::field_buffer(size_t bytes, std::string &err)
{
    if (!reserve(d_field_, field_cap_, bytes, err)) return nullptr;
    return d_field_;
}

bool Device::upload_table(const HarmonicTable &t, std::string &err)
{
    const int n = t.count();
    const int np_ = static_cast<int>(round_up(std::max(n, 1), kHarmPad));
    std::vector<double> rec(size_t(5) * np_, 0.0);
    for (int l = 0; l < n; ++l) {
        rec[l] = t.wx[l];
        rec[np_ + l] = t.wy[l];
        rec[2 * size_t(np_) + l] = t.wz[l];
        rec[3 * size_t(np_) + l] = t.phase[l];
        rec[4 * size_t(np_) + l] = t.amp[l];
    }
    if (!reserve(d_harm_, harm_cap_, rec.size() * sizeof(double), err)) return false;
    HIP_TRY(hipMemcpyAsync(d_harm_, rec.data(), rec.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    nharm_ = np_;
    amp_max_ = 0;
    for (int l = 0; l < n; ++l) amp_max_ = std::max(amp_max_, t.amp[l]);
    return true;
}

template <class Real>
bool Device::upload_points(const Real *x, const Real *y, const Real *z, long n, std::string &err)
{
    std::vector<double> pts(size_t(3) * n);
    for (long i = 0; i < n; ++i) {
        pts[i] = x[i];
        pts[n + i] = y[i];
        pts[2 * n + i] = z ? static_cast<double>(z[i]) : 0.0;
    }
    if (!reserve(d_pts_, pts_cap_, pts.size() * sizeof(double), err)) return false;
    HIP_TRY(hipMemcpyAsync(d_pts_, pts.data(), pts.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    npts_ = n;
    return true;
}

template <class Real>
bool Device::run(const Geometry &g, int kernel, double scaling, const Epilogue &ep, Real *d_field,
                 double *stats_out, std::string &err, const SlabHook *on_slab)
{
    HIP_TRY(hipSetDevice(device_));
    if (!d_harm_ || nharm_ <= 0) {
        err = "no harmonic table resident on the device";
        return false;
    }
    const int np_ = nharm_;
    const long npts = g.points();
    if (npts < 1) {
        err = "empty field";
        return false;
    }
    const bool mfma = g.structured && (kernel == kKernelMfmaBf16x3 || kernel == kKernelMfmaF16x2 ||
                                       kernel == kKernelMfmaBf16x2);
    TileChoice tile;
    if (mfma && !TileTuner::forced(tile) && TileTuner::enabled() &&
        !tuner().choose<Real>(g, kernel, d_field, stream_, tile, err))
        return false;
    timings_ = Timings{};
    HIP_TRY(hipEventRecord(ev_[0], stream_));

    int nparts = 0;
    // field parts (first point, points, event) handed to on_slab while the
    // rest is computed: z slabs of whole k-tile rows, MFMA grids only, and
    // only when no rescale or taper changes the field afterwards
    std::vector<long> slab_first, slab_points;
    if (g.structured) {
        if (mfma) {
            MfmaPlan plan;
            if (!mfma_plan(g, np_, kernel, plan, err, tile.nbx, tile.tk)) return false;
            last_nbx_ = plan.nbx;
            last_tk_ = plan.tk;
            plan.zscale = mfma_zscale(kernel, amp_max_);
            if (!reserve(d_tab_, tab_cap_, plan.table_bytes, err)) return false;
            nparts = plan.nparts;
            if (!reserve(d_part_, part_cap_, size_t(3) * nparts * sizeof(double), err)) return false;
            if (!mfma_tables(plan, g, d_harm_, np_, d_tab_, stream_, err)) return false;
            HIP_TRY(hipEventRecord(ev_[1], stream_));
            if (on_slab && !ep.active() && plan.tiles_k > 1) {
                const long row = long(plan.tiles_x) * plan.tiles_j;  // tiles of one k-tile row
                const int nslab = std::min(kSlabs, plan.tiles_k);
                const int per = (plan.tiles_k + nslab - 1) / nslab;
                const long plane = long(g.nx) * g.ny;
                for (int k0 = 0; k0 < plan.tiles_k; k0 += per) {
                    const int k1 = std::min(k0 + per, plan.tiles_k);
                    const int c = int(slab_first.size());
                    if (!slab_ev_[c]) HIP_TRY(hipEventCreateWithFlags(&slab_ev_[c], hipEventDisableTiming));
                    if (!mfma_field(plan, g, d_tab_, scaling, d_field, d_part_, stream_, err, (k1 - k0) * row,
                                    k0 * row))
                        return false;
                    HIP_TRY(hipEventRecord(slab_ev_[c], stream_));
                    const long z0 = long(k0) * plan.tk, z1 = std::min<long>(long(k1) * plan.tk, g.nz);
                    slab_first.push_back(z0 * plane);
                    slab_points.push_back((z1 - z0) * plane);
                }
            } else if (!mfma_field(plan, g, d_tab_, scaling, d_field, d_part_, stream_, err)) {
                return false;
            }
        } else {
            const bool f64 = kernel == kKernelDirectF64;
            const size_t elt = f64 ? sizeof(double) : sizeof(float);
            const size_t tab_bytes = size_t(np_) * (size_t(g.nx) + g.ny + g.nz) * elt + size_t(np_) * sizeof(float);
            if (!reserve(d_tab_, tab_cap_, tab_bytes, err)) return false;
            const int tiles_i = int((g.nx + kDirB - 1) / kDirB);
            const long tiles_c = (long(g.ny) * g.nz + kDirB - 1) / kDirB;
            nparts = int(tiles_i * tiles_c);
            if (!reserve(d_part_, part_cap_, size_t(3) * nparts * sizeof(double), err)) return false;
            const long tab_elems = long(np_) * (g.nx + g.ny + g.nz);
            const int tblocks = int(std::min<long>((tab_elems + 255) / 256, 65536));
            if (f64) {
                double *tx = static_cast<double *>(d_tab_);
                double *ty = tx + size_t(np_) * g.nx;
                double *tz = ty + size_t(np_) * g.ny;
                axis_tables<double><<<tblocks, 256, 0, stream_>>>(d_harm_, np_, g.nx, g.ny, g.nz, g.ox, g.oy, g.oz,
                                                                  g.ds, tx, ty, tz, nullptr);
                HIP_TRY(hipGetLastError());
                HIP_TRY(hipEventRecord(ev_[1], stream_));
                direct_grid_f64<Real><<<nparts, 256, 0, stream_>>>(tx, ty, tz, d_harm_ + 4 * size_t(np_), np_, g.nx,
                                                                   g.ny, g.nz, tiles_i, scaling, d_field, d_part_);
            } else {
                float *tx = static_cast<float *>(d_tab_);
                float *ty = tx + size_t(np_) * g.nx;
                float *tz = ty + size_t(np_) * g.ny;
                float *amp = tz + size_t(np_) * g.nz;
                axis_tables<float><<<tblocks, 256, 0, stream_>>>(d_harm_, np_, g.nx, g.ny, g.nz, g.ox, g.oy, g.oz,
                                                                 g.ds, tx, ty, tz, amp);
                HIP_TRY(hipGetLastError());
                HIP_TRY(hipEventRecord(ev_[1], stream_));
                direct_grid_f32<Real><<<nparts, 256, 0, stream_>>>(tx, ty, tz, amp, np_, g.nx, g.ny, g.nz, tiles_i,
                                                                   scaling, d_field, d_part_);
            }
        }
        HIP_TRY(hipGetLastError());
    } else {
        if (!d_pts_ || npts_ != npts) {
            err = "point coordinates are not resident on the device";
            return false;
        }
        if (kernel != kKernelDirectF32 && kernel != kKernelDirectF64) {
            err = "point lists run on the direct kernels only (direct_f32 or direct_f64)";
            return false;
        }
        HIP_TRY(hipEventRecord(ev_[1], stream_));
        nparts = int((npts + 256 * kPtsP - 1) / (256 * kPtsP));
        if (!reserve(d_part_, part_cap_, size_t(3) * nparts * sizeof(double), err)) return false;
        if (kernel == kKernelDirectF64)
            direct_points<Real, true><<<nparts, 256, 0, stream_>>>(d_pts_, d_pts_ + npts, d_pts_ + 2 * npts, npts,
                                                                   d_harm_, np_, scaling, d_field, d_part_);
        else
            direct_points<Real, false><<<nparts, 256, 0, stream_>>>(d_pts_, d_pts_ + npts, d_pts_ + 2 * npts, npts,
                                                                    d_harm_, np_, scaling, d_field, d_part_);
        HIP_TRY(hipGetLastError());
    }
    return finish<Real>(g, ep, d_field, nparts, stats_out, slab_first, slab_points, err, on_slab);
}

template <class Real>
bool Device::finish(const Geometry &g, const Epilogue &ep, Real *d_field, int nparts, double *stats_out,
                    const std::vector<long> &slab_first, const std::vector<long> &slab_points, std::string &err,
                    const SlabHook *on_slab, double *part, double *stat, double *poi)
{
    const long npts = g.points();
    if (!part) part = d_part_;
    if (!stat) stat = d_stat_;
    HIP_TRY(hipEventRecord(ev_[2], stream_));
    stats_finalize<Real><<<1, 256, 0, stream_>>>(part, nparts, ep.sigma, stat);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev_[3], stream_));

    if (ep.active()) {
        const int npoi = static_cast<int>(ep.poi3.size() / 3);
        if (npoi > 0) {
            if (!poi) {
                if (!reserve(d_poi_, poi_cap_, ep.poi3.size() * sizeof(double), err)) return false;
                poi = d_poi_;
            }
            HIP_TRY(hipMemcpyAsync(poi, ep.poi3.data(), ep.poi3.size() * sizeof(double), hipMemcpyHostToDevice,
                                   stream_));
        }
        const int blocks = int(std::min<long>((npts + 255) / 256, 16384));
        if (g.structured)
            epilogue_grid<Real><<<blocks, 256, 0, stream_>>>(d_field, g.nx, g.ny, g.nz, g.ox, g.oy, g.oz, g.ds,
                                                             stat, ep.rescale == 1, poi, npoi, ep.mute, ep.taper);
        else
            epilogue_points<Real><<<blocks, 256, 0, stream_>>>(d_field, npts, d_pts_, d_pts_ + npts,
                                                               d_pts_ + 2 * npts, stat, ep.rescale == 1, poi, npoi,
                                                               ep.mute, ep.taper);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(ev_[4], stream_));
    HIP_TRY(hipMemcpyAsync(h_stat_, stat, 6 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    for (size_t c = 0; c < slab_first.size(); ++c) {
        HIP_TRY(hipEventSynchronize(slab_ev_[c]));
        if (!(*on_slab)(slab_first[c], slab_points[c], err)) return false;
    }
    HIP_TRY(hipStreamSynchronize(stream_));
    if (on_slab && slab_first.empty() && !(*on_slab)(0, npts, err)) return false;
    float ms = 0.f;
    hipEventElapsedTime(&ms, ev_[0], ev_[1]);
    timings_.tables = ms;
    hipEventElapsedTime(&ms, ev_[1], ev_[2]);
    timings_.field = ms;
    hipEventElapsedTime(&ms, ev_[2], ev_[3]);
    timings_.stats = ms;
    hipEventElapsedTime(&ms, ev_[3], ev_[4]);
    timings_.epilogue = ms;
    hipEventElapsedTime(&ms, ev_[0], ev_[4]);
    timings_.total = ms;
    for (int i = 0; i < 6; ++i) stats_out[i] = h_stat_[i];
    return true;
}

template <class Real>
bool Device::run_noalloc(const Geometry &g, int kernel, double scaling, const Epilogue &ep, Real *d_field,
                         double *stats_out, std::string &err)
{
    HIP_TRY(hipSetDevice(device_));
    const bool mfma = kernel == kKernelMfmaBf16x3 || kernel == kKernelMfmaF16x2 || kernel == kKernelMfmaBf16x2;
    if (!g.structured || !mfma) {
        err = "the allocation-free run takes the MFMA kernels of structured grids";
        return false;
    }
    if (!d_harm_ || nharm_ <= 0) {
        err = "no harmonic table resident on the device";
        return false;
    }
    if (!scratch_) {
        err = "the no-alloc scratch array has no device address";
        return false;
    }
    const int np_ = nharm_;
    // a tile chosen earlier for this grid, else the least-padding one: the
    // tuner's timing runs would allocate
    TileChoice tile;
    if (!TileTuner::forced(tile) && TileTuner::enabled()) tuner().cached(g, kernel, sizeof(Real), tile);
    MfmaPlan full;
    if (!mfma_plan(g, np_, kernel, full, err, tile.nbx, tile.tk)) return false;
    // scratch layout: statistics (16 doubles), points of interest, partials,
    // then the operand tables of one chunk
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t stat_off = 0, poi_off = 256;
    const size_t part_off = poi_off + up(ep.poi3.size() * sizeof(double));
    const size_t tab_off = part_off + up(size_t(3) * full.nparts * sizeof(double));
    if (tab_off >= kNoAllocScratch) {
        err = "the grid's statistics partials and points of interest exceed the no-alloc scratch (" +
              std::to_string(kNoAllocScratch >> 20) + " MiB)";
        return false;
    }
    const size_t room = kNoAllocScratch - tab_off;
    // the largest chunk (a multiple of 128 harmonics, at most the table) whose tables fit
    int cap = static_cast<int>(round_up(np_, kFieldChunkHarmonics));
    MfmaPlan plan;
    for (;;) {
        if (!mfma_plan(g, cap, kernel, plan, err, full.nbx, full.tk)) return false;
        if (plan.table_bytes <= room) break;
        const double per = double(plan.table_bytes) / cap;
        const int next = static_cast<int>(room / per) / kFieldChunkHarmonics * kFieldChunkHarmonics;
        cap = std::min(cap - kFieldChunkHarmonics, next);
        if (cap < kFieldChunkHarmonics) {
            err = "the grid's operand tables for 128 harmonics exceed the no-alloc scratch (" +
                  std::to_string(kNoAllocScratch >> 20) + " MiB)";
            return false;
        }
    }
    last_nbx_ = plan.nbx;
    last_tk_ = plan.tk;
    plan.zscale = mfma_zscale(kernel, amp_max_);
    char *base = static_cast<char *>(scratch_);
    double *stat = reinterpret_cast<double *>(base + stat_off);
    double *poi = reinterpret_cast<double *>(base + poi_off);
    double *part = reinterpret_cast<double *>(base + part_off);
    void *tab = base + tab_off;

    timings_ = Timings{};
    HIP_TRY(hipEventRecord(ev_[0], stream_));
    // chunk by chunk: each chunk's tables overwrite the last one's (stream
    // order: its field pass has read them) and the field carries the sums
    for (int h0 = 0; h0 < np_; h0 += cap) {
        const int nh = std::min(cap, np_ - h0);
        plan.tab_s0 = h0 / 16;
        if (!mfma_tables(plan, g, d_harm_, np_, tab, stream_, err, h0, nh)) return false;
        if (h0 == 0) HIP_TRY(hipEventRecord(ev_[1], stream_));
        const int flags = (h0 > 0 ? kFieldCarryIn : 0) | (h0 + nh == np_ ? kFieldFinal : 0);
        if (!mfma_field(plan, g, tab, scaling, d_field, part, stream_, err, 0, 0, h0 / 16, nh / 16, flags))
            return false;
    }
    const std::vector<long> none;
    return finish<Real>(g, ep, d_field, plan.nparts, stats_out, none, none, err, nullptr, part, stat, poi);
}

std::vector<int> Device::chunk_targets(int nharm, int chunk)
{
    std::vector<int> t;
    if (chunk > 0) {
        for (long kept = chunk; kept < nharm; kept += chunk) t.push_back(static_cast<int>(kept));
        t.push_back(nharm);
        return t;
    }
    // The first chunk is small, so that the device starts after a short part
    // of the draw; each later one doubles, since the draw runs several times
    // faster than the device consumes it. A tail shorter than the chunk
    // before it joins that chunk. Every cut but the last is a multiple of the
    // promotion period of the field kernel.
    const int q = kFieldChunkHarmonics;
    long c = std::max<long>(8 * q, round_up(nharm / 16, q));
    long kept = 0;
    while (kept + 2 * c < nharm) {  // what is left after this chunk is at least a chunk
        kept += c;
        t.push_back(static_cast<int>(kept));
        c *= 2;
    }
    t.push_back(nharm);
    return t;
}

template <class Real>
bool Device::run_streamed(const Geometry &g, int kernel, int nharm, double scaling, const Epilogue &ep,
                          Real *d_field, double *stats_out, std::string &err, const HarmonicFeed &feed,
                          bool *restart, const SlabHook *on_slab, int chunk)
{
    if (restart) *restart = false;
    if (chunk < 0 || chunk % kFieldChunkHarmonics != 0) {
        err = "the streamed chunk must be a multiple of 128 harmonics";
        return false;
    }
    HIP_TRY(hipSetDevice(device_));
    const bool mfma = kernel == kKernelMfmaBf16x3 || kernel == kKernelMfmaF16x2 || kernel == kKernelMfmaBf16x2;
    if (!g.structured || !mfma || nharm < 1) {
        err = "streamed realizations run on the MFMA kernels of structured grids";
        return false;
    }
    TileChoice tile;
    if (!TileTuner::forced(tile) && TileTuner::enabled() &&
        !tuner().choose<Real>(g, kernel, d_field, stream_, tile, err))
        return false;
    // record and tables sized for every harmonic plus one spare padding block:
    // a draw whose last harmonics are all dropped ends with an empty chunk,
    // which then runs one block of zero-amplitude harmonics (each adds 0)
    const int cap = static_cast<int>(round_up(nharm, kHarmPad)) + kHarmPad;
    MfmaPlan plan;
    if (!mfma_plan(g, cap, kernel, plan, err, tile.nbx, tile.tk)) return false;
    last_nbx_ = plan.nbx;
    last_tk_ = plan.tk;
    plan.zscale = 1.0;  // a chunk that would need another scale restarts on run()
    const size_t rec_bytes = size_t(5) * cap * sizeof(double);
    if (!reserve(d_harm_, harm_cap_, rec_bytes, err)) return false;
    nharm_ = 0;  // the streamed record is not a resident table for run()
    if (h_rec_cap_ < rec_bytes) {
        HIP_TRY(hipStreamSynchronize(stream_));
        if (h_rec_) HIP_TRY(hipHostFree(h_rec_));
        h_rec_ = nullptr;
        h_rec_cap_ = 0;
        HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&h_rec_), rec_bytes, hipHostMallocDefault));
        h_rec_cap_ = rec_bytes;
    }
    if (!reserve(d_tab_, tab_cap_, plan.table_bytes, err)) return false;
    const int nparts = plan.nparts;
    if (!reserve(d_part_, part_cap_, size_t(3) * nparts * sizeof(double), err)) return false;

    timings_ = Timings{};
    HIP_TRY(hipEventRecord(ev_[0], stream_));
    HIP_TRY(hipStreamSynchronize(stream_));  // the staging buffer is free
    const std::vector<int> targets = chunk_targets(nharm, chunk);
    std::vector<long> slab_first, slab_points;
    HarmonicTable t;
    int kept = 0;  // harmonics queued so far (padding included in the last chunk)
    bool last = false;
    for (size_t c = 0; !last; ++c) {
        t = HarmonicTable{};
        // the host draws this chunk while the device runs the chunks before it
        if (!feed(c < targets.size() ? targets[c] : nharm, t, last, err)) return false;
        const int n = t.count();
        double amp_max = 0;
        for (int l = 0; l < n; ++l) amp_max = std::max(amp_max, t.amp[l]);
        if (mfma_zscale(kernel, amp_max) != 1.0) {
            HIP_TRY(hipStreamSynchronize(stream_));
            if (restart) *restart = true;
            err = "amplitude beyond the scaled-fp16 range: the realization needs a rescaled table";
            return false;
        }
        if (kept + n > cap - kHarmPad || (!last && n % kFieldChunkHarmonics != 0)) {
            err = "harmonic feed returned more harmonics than asked for";
            return false;
        }
        int nh = last ? static_cast<int>(round_up(kept + n, kHarmPad)) - kept : n;
        if (nh == 0) nh = kHarmPad;
        // the chunk's part of each record field, zero-padded, staged in
        // pinned memory (each chunk has its own range: nothing is rewritten
        // while its copy may still run)
        const double *src[5] = {t.wx.data(), t.wy.data(), t.wz.data(), t.phase.data(), t.amp.data()};
        for (int f = 0; f < 5; ++f) {
            double *dst = h_rec_ + size_t(f) * cap + kept;
            std::copy(src[f], src[f] + n, dst);
            std::fill(dst + n, dst + nh, 0.0);
            HIP_TRY(hipMemcpyAsync(d_harm_ + size_t(f) * cap + kept, dst, size_t(nh) * sizeof(double),
                                   hipMemcpyHostToDevice, stream_));
        }
        if (!mfma_tables(plan, g, d_harm_, cap, d_tab_, stream_, err, kept, nh)) return false;
        if (c == 0) HIP_TRY(hipEventRecord(ev_[1], stream_));
        const int s0 = kept / 16, ns = nh / 16;
        const int flags = (kept > 0 ? kFieldCarryIn : 0) | (last ? kFieldFinal : 0);
        if (last && on_slab && !ep.active() && plan.tiles_k > 1) {
            // the final chunk in z slabs, each handed over as soon as it is final
            const long row = long(plan.tiles_x) * plan.tiles_j;
            const int nslab = std::min(kSlabs, plan.tiles_k);
            const int per = (plan.tiles_k + nslab - 1) / nslab;
            const long plane = long(g.nx) * g.ny;
            for (int k0 = 0; k0 < plan.tiles_k; k0 += per) {
                const int k1 = std::min(k0 + per, plan.tiles_k);
                const int sl = int(slab_first.size());
                if (!slab_ev_[sl]) HIP_TRY(hipEventCreateWithFlags(&slab_ev_[sl], hipEventDisableTiming));
                if (!mfma_field(plan, g, d_tab_, scaling, d_field, d_part_, stream_, err, (k1 - k0) * row, k0 * row,
                                s0, ns, flags))
                    return false;
                HIP_TRY(hipEventRecord(slab_ev_[sl], stream_));
                const long z0 = long(k0) * plan.tk, z1 = std::min<long>(long(k1) * plan.tk, g.nz);
                slab_first.push_back(z0 * plane);
                slab_points.push_back((z1 - z0) * plane);
            }
        } else if (!mfma_field(plan, g, d_tab_, scaling, d_field, d_part_, stream_, err, 0, 0, s0, ns, flags)) {
            return false;
        }
        kept += nh;
    }
    return finish<Real>(g, ep, d_field, nparts, stats_out, slab_first, slab_points, err, on_slab);
}

template bool Device::upload_points<float>(const float *, const float *, const float *, long, std::string &);
template bool Device::upload_points<double>(const double *, const double *, const double *, long, std::string &);
template bool Device::run<float>(const Geometry &, int, double, const Epilogue &, float *, double *, std::string &,
                                 const SlabHook *);
template bool Device::run<double>(const Geometry &, int, double, const Epilogue &, double *, double *,
                                  std::string &, const SlabHook *);
template bool Device::run_noalloc<float>(const Geometry &, int, double, const Epilogue &, float *, double *,
                                         std::string &);
template bool Device::run_noalloc<double>(const Geometry &, int, double, const Epilogue &, double *, double *,
                                          std::string &);
template bool Device::run_streamed<float>(const Geometry &, int, int, double, const Epilogue &, float *, double *,
                                          std::string &, const HarmonicFeed &, bool *, const SlabHook *,
                                          int);
template bool Device::run_streamed<double>(const Geometry &, int, int, double, const Epilogue &, double *, double *,
                                           std::string &, const HarmonicFeed &, bool *, const SlabHook *,
                                          int);

}  // namespace scarf
