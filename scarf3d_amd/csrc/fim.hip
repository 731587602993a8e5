// FIM device passes and their launch sequence (see fim.hpp).
//
// Layout in HBM, one realization:
//   S  half spectrum, complex `real`, H x N2 x N3 with H = N1/2 + 1, index
//      i + H (j + N2 k) (kx fastest): the only part of the reference's full
//      spectrum that its c2r transform along x reads
//   R  internal field, `real`, N1 x N2 x N3, index i + N1 (j + N2 k)
//   partials (3 doubles per block), statistics (6 doubles), points of
//   interest; the generator tables (fim_stream.hpp: 18 KB of powers of the
//   primitive root, the companion-matrix powers), uploaded once per process.
// Work per internal point (bytes of HBM): spectrum 8 (written), symmetry
// reads/writes 2 planes, the c2r transform ~3 passes over S plus R, statistics 4
// (R read), epilogue 8 if active, mapping 4 per output point plus the
// gathered R values.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>

#include "fim.hpp"
#include "srm_stats.hpp"
#include "fim_stream.hpp"

namespace scarf {
namespace {

using fim_stream::kJumpBits;
using fim_stream::kPowWords;
using GenTables = fim_stream::Tables;

template <class Real>
struct Cplx {
    Real re, im;
};

// Spectrum parameters (compute_spectrum, src/scarflib_fim.f90:2440-2510)
template <class Real>
struct SpecArgs {
    int n1, n2, n3, h;      // internal grid, h = n1/2 + 1
    long lines;             // n2 * n3 x-lines
    int lines_per_item;
    uint32_t s1[2];         // generator state after its first draw (state of stream position 0)
    uint32_t line_step[4];  // A^n1
    Real dk[3], cl[3], kc, nu, cnst;
    int filter, acf;
};

// Half spectrum with the random phases of the global stream: internal
// point (i, j, k) (0-based) takes the uniform at stream position
// i + n1 (j + n2 k), the Fortran-order index that the reference's prng gives
// every point of the grid whatever the process split
// (src/prng.cpp:66-130: seed, then jumps over the points other processes own).
// One wave per item of lines_per_item consecutive x-lines: the wave jumps to
// its first line's start (companion-matrix powers A^(2^b)), lane l steps to
// position l, then every lane advances by A^64 per point and every line by
// A^n1. Stores are coalesced (lane l writes point i = l + 64 t).
template <class Real>
__global__ __launch_bounds__(256) void fim_spectrum(Cplx<Real> *__restrict__ S, SpecArgs<Real> a,
                                                    const GenTables *__restrict__ gt, long items)
{
#pragma clang fp contract(off)
    __shared__ uint32_t pw[kPowWords];
    for (int q = threadIdx.x; q < kPowWords; q += blockDim.x) pw[q] = gt->pw[q];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const long waves = long(gridDim.x) * (blockDim.x >> 6);
    const Real pi = static_cast<Real>(3.141592653589793);
    for (long item = long(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6); item < items; item += waves) {
        const long l0 = item * a.lines_per_item;
        const long l1 = min(l0 + a.lines_per_item, a.lines);
        fim_stream::walk_item(pw, gt, a.s1, a.line_step, a.n1, a.h, l0, l1, lane, [&](long line, int i, uint32_t raw) {
#pragma clang fp contract(off)
            const int j = int(line % a.n2), k = int(line / a.n2);
            const Real ky = static_cast<Real>(j <= a.n2 / 2 ? j : a.n2 - j) * a.dk[1];
            const Real kz = static_cast<Real>(k <= a.n3 / 2 ? k : a.n3 - k) * a.dk[2];
            const Real u = static_cast<Real>(static_cast<double>(raw) / 2147483647.0);
            const Real kx = static_cast<Real>(i) * a.dk[0];
            Real butter = Real(1);
            if (a.filter && sqrt(kx * kx + ky * ky + kz * kz) > a.kc) butter = Real(0);
            const Real x1 = kx * a.cl[0], x2 = ky * a.cl[1], x3 = kz * a.cl[2];
            const Real kl2 = x1 * x1 + x2 * x2 + x3 * x3;  // |k cl|^2
            Real amp, c, sn;
            if constexpr (sizeof(Real) == sizeof(float)) {
                // the hardware's log2, exp2 and square root, and sin and cos
                // of the phase in turns (2 pi u is u turns): ~1e-6 relative,
                // within the float library's parity bar
                const float fn = a.acf == 0 ? __builtin_amdgcn_exp2f(-a.nu * __builtin_amdgcn_logf(1.f + kl2))
                                            : __builtin_amdgcn_exp2f(-0.25f * 1.44269504088896341f * kl2);
                amp = __builtin_amdgcn_sqrtf(a.cnst * fn) * butter;
                sn = __builtin_amdgcn_sinf(u);
                c = __builtin_amdgcn_cosf(u);
            } else {
                const Real fn = a.acf == 0 ? Real(1) / pow(Real(1) + kl2, a.nu) : exp(Real(-0.25) * kl2);
                amp = sqrt(a.cnst * fn) * butter;
                sincos((u * Real(2)) * pi, &sn, &c);
            }
            S[size_t(i) + size_t(a.h) * size_t(line)] = Cplx<Real>{c * amp, sn * amp};
        });
    }
}

// Hermitian conditions on the kx = 0 and kx = Nyquist planes as one process
// holding the whole grid applies them (enforce_symmetry,
// src/scarflib_fim.f90:2660-3008), 0-based with jn = n2/2, kn = n3/2:
//   point conditions, in order: S(0,0,0) = 0 (plane 0); S(N,0,0) real (plane
//   N = n1/2); S(p,jn,kn), S(p,0,kn), S(p,jn,0) real;
//   then, from the values after the point conditions (B):
//   k > kn, every j:        S(p,j,k)  = conj(B(p, -j mod n2, n3 - k))
//   k = kn, j > jn:         S(p,j,kn) = conj(B(p, n2 - j, kn))
//   k = 0,  j > jn:         S(p,j,0)  = conj(B(p, n2 - j, 0)) in 2D, and
//                           conj(B(p, j, 0)) in 3D: with one process the
//                           reference's exchange of that line (its "beta"
//                           region) is replaced by the exchange of the j = 0
//                           column, whose subarray types overwrite it
//                           (:2866-2960), so the line is conjugated in place.
// After these rules each plane is Hermitian (S(-j,-k) = conj S(j,k)) except
// that 3D line, and the reference's c2r along x reads only the real parts of
// the planes' values after the (y, z) transforms. The real part of a 2D
// transform is the transform of the plane's Hermitian part
// (S(j,k) + conj S(-j,-k)) / 2, so the line's pairs (j, n2 - j), 0 < j < jn,
// are replaced by v = (B(j,0) + B(n2-j,0)) / 2 and conj(v): the planes are
// then exactly Hermitian, one 3D c2r transform reproduces the reference's
// z, y, x sequence, and no transform depends on how it treats imaginary parts
// that should be zero.
// Every position is written by one thread at most, and each thread reads only
// positions that no other thread writes, so one pass in place is exact.
template <class Real>
__global__ __launch_bounds__(256) void fim_symmetry(Cplx<Real> *__restrict__ S, int n1, int n2, int n3, int h)
{
    const long per_plane = long(n2) * n3;
    const long e = blockIdx.x * long(blockDim.x) + threadIdx.x;
    if (e >= 2 * per_plane) return;
    const int p = e < per_plane ? 0 : n1 / 2;
    const long r = e % per_plane;
    const int j = int(r % n2), k = int(r / n2);
    const int jn = n2 / 2, kn = n3 / 2;
    auto at = [&](int jj, int kk) -> Cplx<Real> & { return S[size_t(p) + size_t(h) * (size_t(jj) + size_t(n2) * kk)]; };
    Cplx<Real> &v = at(j, k);
    const bool point = (j == 0 && k == 0) || (j == jn && k == kn) || (j == 0 && k == kn) || (j == jn && k == 0);
    if (point) {
        if (p == 0 && j == 0 && k == 0) v = Cplx<Real>{Real(0), Real(0)};
        if (p == n1 / 2 && j == 0 && k == 0) v.im = Real(0);
        if (j == jn && k == kn) v.im = Real(0);
        if (j == 0 && k == kn) v.im = Real(0);
        if (j == jn && k == 0) v.im = Real(0);
        return;
    }
    if (k > kn) {
        const Cplx<Real> b = at(j == 0 ? 0 : n2 - j, n3 - k);
        v = Cplx<Real>{b.re, -b.im};
    } else if (j > jn && k == kn) {
        const Cplx<Real> b = at(n2 - j, kn);
        v = Cplx<Real>{b.re, -b.im};
    } else if (k == 0 && n3 == 1 && j > jn) {
        const Cplx<Real> b = at(n2 - j, 0);
        v = Cplx<Real>{b.re, -b.im};
    } else if (k == 0 && n3 > 1 && j > 0 && j < jn) {
        // the pair (j, n2 - j) of the in-place line, by its Hermitian part
        Cplx<Real> &w = at(n2 - j, 0);
        const Cplx<Real> hp{(v.re + w.re) / Real(2), (v.im + w.im) / Real(2)};
        v = hp;
        w = Cplx<Real>{hp.re, -hp.im};
    }
}

// Moments of buffer = REAL(spec) * scaling (src/scarflib_fim.f90:955-961):
// each value is the `real` product the reference stores, accumulated in
// FP64 as a sum and a sum of squares per thread (the internal field has zero
// mean up to rounding: its kx = ky = kz = 0 coefficient is zero, so the sum
// of squares carries the variance without cancellation), then merged per
// block as (count, mean, M2). R itself is left unscaled: the pass that next
// writes or reads each value (epilogue or mapping) applies the same product.
template <class Real>
__global__ __launch_bounds__(256) void fim_moments(const Real *__restrict__ R, long n, Real scaling,
                                                   double *__restrict__ part)
{
    double s1[4] = {0.0, 0.0, 0.0, 0.0}, s2[4] = {0.0, 0.0, 0.0, 0.0};
    const long stride = long(gridDim.x) * blockDim.x;
    long e = blockIdx.x * long(blockDim.x) + threadIdx.x;
    for (; e + 3 * stride < n; e += 4 * stride) {  // four loads in flight per thread
        Real x[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) x[q] = R[e + q * stride];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double v = static_cast<double>(x[q] * scaling);
            s1[q] += v;
            s2[q] += v * v;
        }
    }
    for (; e < n; e += stride) {
        const double v = static_cast<double>(R[e] * scaling);
        s1[0] += v;
        s2[0] += v * v;
    }
    const long mine = n > blockIdx.x * long(blockDim.x) + threadIdx.x
                          ? (n - 1 - (blockIdx.x * long(blockDim.x) + threadIdx.x)) / stride + 1
                          : 0;
    const double c = static_cast<double>(mine);
    const double t1 = (s1[0] + s1[1]) + (s1[2] + s1[3]), t2 = (s2[0] + s2[1]) + (s2[2] + s2[3]);
    Moments m{c, 0.0, 0.0};
    if (c > 0.0) {
        m.mean = t1 / c;
        m.m2 = t2 - t1 * m.mean;
    }
    block_moments(m, part + 3 * blockIdx.x);
}

// mean, variance, info(3) = sqrt(real(var)), info(4) = real(mean) and the
// rescale factor sigma / info(3) (src/scarflib_fim.f90:1001-1031; one
// process: parallel_variance returns its single set unchanged)
template <class Real>
__global__ __launch_bounds__(256) void fim_finalize(const double *__restrict__ part, int nparts, double sigma,
                                                    double *__restrict__ stat)
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
            Moments x{sm[0][threadIdx.x], sm[1][threadIdx.x], sm[2][threadIdx.x]};
            merge_moments(x, Moments{sm[0][threadIdx.x + stride], sm[1][threadIdx.x + stride], sm[2][threadIdx.x + stride]});
            sm[0][threadIdx.x] = x.n;
            sm[1][threadIdx.x] = x.mean;
            sm[2][threadIdx.x] = x.m2;
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
        stat[0] = n;
        stat[1] = mean;
        stat[2] = var;
        stat[3] = info3;
        stat[4] = info4;
        stat[5] = static_cast<Real>(sigma) / info3;
    }
}

// scaling, rescale (buffer * (sigma / info(3)) - info(4)) and the mute/taper
// of every point of interest on the internal grid, positions (i - 1) dh
// (src/scarflib_fim.f90:955-961, :1014-1059, tapering_structured with
// ls = 1). Blocks walk the x-lines with a grid stride.
template <class Real>
__global__ __launch_bounds__(256) void fim_epilogue(Real *__restrict__ R, int n1, int n2, int lines, Real scaling,
                                                    Real dh, const double *__restrict__ stat, int rescale,
                                                    const Real *__restrict__ pt, int npoi, Real mute, Real taper)
{
#pragma clang fp contract(off)
    const Real sc = static_cast<Real>(stat[5]), mu = static_cast<Real>(stat[4]);
    for (int line = blockIdx.x; line < lines; line += gridDim.x) {
        const int j = line % n2, k = line / n2;
        Real *row = R + size_t(line) * n1;
        for (int i = threadIdx.x; i < n1; i += blockDim.x) {
            Real v = row[i] * scaling;
            if (rescale) v = v * sc - mu;
            for (int q = 0; q < npoi; ++q) {
                const Real dx = static_cast<Real>(i) * dh - pt[3 * q];
                const Real dy = static_cast<Real>(j) * dh - pt[3 * q + 1];
                const Real dz = static_cast<Real>(k) * dh - pt[3 * q + 2];
                v = taper_factor(v, dx * dx, dy * dy, dz * dz, mute, taper);
            }
            row[i] = v;
        }
    }
}

// 1-based internal index along one axis for coordinate x, 0 for the halo,
// -1 outside
template <class Real>
__device__ inline int axis_index(Real x, Real dh, int np)
{
#pragma clang fp contract(off)
    const Real c = (x / dh + Real(3)) - Real(1);
    if (!(c >= Real(1) && c < static_cast<Real>(np + 1))) return -1;
    return static_cast<int>(lround(c)) - 1;
}
// value at 1-based axis indices (x, y halo 0: zero; z halo 0: the last plane)
// times `scale` (the buffer's scaling when the internal field is unscaled)
template <class Real>
__device__ inline Real gather(const Real *__restrict__ R, int n1, int n2, int n3, int ix, int iy, int iz, Real scale)
{
    if (ix <= 0 || iy <= 0 || iz < 0) return Real(0);
    if (iz == 0) iz = n3;
    return R[size_t(ix - 1) + size_t(n1) * (size_t(iy - 1) + size_t(n2) * size_t(iz - 1))] * scale;
}

// Per-axis internal indices of a structured block fs..fe of step ds:
// external point (fs - 1 + i0) ds, axis_index of it (grid_mapping_struct /
// order_points_struct, src/scarflib_fim.f90:1303-1512, :1598-1700). The
// nearest node along each axis depends on that axis' coordinate only, so
// the block's nx + ny + nz indices are computed once; tab holds x, then y,
// then z.
template <class Real>
__global__ __launch_bounds__(256) void fim_axis_tables(int *__restrict__ tab, int nx, int ny, int nz, int fx, int fy,
                                                       int fz, Real ds, Real dh, Real ox, Real oy, Real oz, int n1,
                                                       int n2, int n3)
{
#pragma clang fp contract(off)
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nx) {
        tab[e] = axis_index(static_cast<Real>(fx - 1 + e) * ds - ox, dh, n1);
    } else if (e < nx + ny) {
        const int j = e - nx;
        tab[e] = axis_index(static_cast<Real>(fy - 1 + j) * ds - oy, dh, n2);
    } else if (e < nx + ny + nz) {
        const int k = e - nx - ny;
        tab[e] = axis_index(static_cast<Real>(fz - 1 + k) * ds - oz, dh, n3);
    }
}

// The block's values: each block walks groups of 4 output x-lines
// (row = j + ny k) with a grid stride, its threads along x; the four rows'
// gathers are independent loads in flight together
template <class Real>
__global__ __launch_bounds__(256) void fim_map_grid(const Real *__restrict__ R, int n1, int n2, int n3, Real scale,
                                                    Real *__restrict__ out, const int *__restrict__ tab, int nx,
                                                    int ny, int rows)
{
    constexpr int G = 4;
    for (int l0 = blockIdx.x * G; l0 < rows; l0 += gridDim.x * G) {
        int iy[G], iz[G];
#pragma unroll
        for (int r = 0; r < G; ++r) {
            const int line = min(l0 + r, rows - 1);
            iy[r] = tab[nx + line % ny];
            iz[r] = tab[nx + ny + line / ny];
        }
        for (int i = threadIdx.x; i < nx; i += blockDim.x) {
            const int ix = tab[i];
            Real v[G];
#pragma unroll
            for (int r = 0; r < G; ++r) v[r] = gather(R, n1, n2, n3, ix, iy[r], iz[r], scale);
#pragma unroll
            for (int r = 0; r < G; ++r)
                if (l0 + r < rows) out[size_t(l0 + r) * nx + i] = v[r];
        }
    }
}

// point list (FP64 copies of the caller's `real` coordinates)
// (grid_mapping_unstruct / order_points_unstruct, src/scarflib_fim.f90:1110-1300, :1514-1596)
template <class Real>
__global__ __launch_bounds__(256) void fim_map_points(const Real *__restrict__ R, int n1, int n2, int n3, Real dh,
                                                      Real scale, Real *__restrict__ out,
                                                      const double *__restrict__ px, const double *__restrict__ py,
                                                      const double *__restrict__ pz, long n, Real ox, Real oy,
                                                      Real oz)
{
#pragma clang fp contract(off)
    for (long e = blockIdx.x * long(blockDim.x) + threadIdx.x; e < n; e += long(gridDim.x) * blockDim.x) {
        const int ix = axis_index(static_cast<Real>(px[e]) - ox, dh, n1);
        const int iy = axis_index(static_cast<Real>(py[e]) - oy, dh, n2);
        const int iz = axis_index(static_cast<Real>(pz[e]) - oz, dh, n3);
        out[e] = (ix < 0 || iy < 0) ? Real(0) : gather(R, n1, n2, n3, ix, iy, iz, scale);
    }
}

template <class Real>
struct FftTypes;
template <>
struct FftTypes<float> {
    static constexpr hipfftType c2r = HIPFFT_C2R;
    static hipfftResult exec_c2r(hipfftHandle p, void *in, void *out) { return hipfftExecC2R(p, (hipfftComplex *)in, (hipfftReal *)out); }
};
template <>
struct FftTypes<double> {
    static constexpr hipfftType c2r = HIPFFT_Z2D;
    static hipfftResult exec_c2r(hipfftHandle p, void *in, void *out) { return hipfftExecZ2D(p, (hipfftDoubleComplex *)in, (hipfftDoubleReal *)out); }
};

constexpr int kScaleBlocks = 2048;  // 8 blocks of 256 per compute unit
constexpr long kRowBlocks = 8192;   // row-walking passes: 32 blocks per compute unit

}  // namespace

struct FimDevice::Impl {
    GenTables *d_gen = nullptr;
    void *d_spec = nullptr;
    size_t spec_cap = 0;
    void *d_real = nullptr;
    size_t real_cap = 0;
    double *d_part = nullptr, *d_stat = nullptr, *h_stat = nullptr;
    void *d_pt = nullptr;
    size_t pt_cap = 0;
    int *d_tab = nullptr;  // per-axis nearest indices of a structured block
    size_t tab_cap = 0;
    hipfftHandle plan = 0;  // c2r over the whole internal grid
    bool plans = false;
    int plan_n[3] = {0, 0, 0};
    int plan_real = 0;  // sizeof(Real) of the plans
    hipEvent_t ev[6] = {};
    bool ready = false;  // the fixed buffers and events exist
    size_t reserved = 0;

    // grows *p to `bytes` (kept when large enough); the stream is drained
    // before a buffer it may still use is freed
    bool grow(void **p, size_t &cap, size_t bytes, hipStream_t stream, size_t elsewhere, std::string &err)
    {
        if (bytes <= cap && *p) return true;
        if (const char *e = std::getenv("SCARF_DEVICE_MEMORY_LIMIT")) {
            const double limit = std::atof(e);
            const size_t after = elsewhere + reserved + bytes - std::min(bytes, cap);
            if (limit > 0 && double(after) > limit) {
                err = "device memory limit SCARF_DEVICE_MEMORY_LIMIT=" + std::string(e) + " bytes: the library holds " +
                      std::to_string(elsewhere + reserved) + " and would need " + std::to_string(after);
                return false;
            }
        }
        if (*p) {
            (void)hipStreamSynchronize(stream);
            (void)hipFree(*p);
            reserved -= cap;
            *p = nullptr;
            cap = 0;
        }
        const hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();  // a failed allocation leaves no error for the next call
            *p = nullptr;
            err = std::string("FIM: hipMalloc of ") + std::to_string(bytes) + " bytes: " + hipGetErrorString(e);
            return false;
        }
        cap = bytes;
        reserved += bytes;
        return true;
    }
    void drop_plans()
    {
        if (!plans) return;
        (void)hipfftDestroy(plan);
        plans = false;
    }
};

FimDevice::FimDevice() : impl_(new Impl) {}

FimDevice::~FimDevice()
{
    // like the SRM device buffers, released by the process exit (the HIP
    // runtime may already be tearing down when static objects are destroyed)
    impl_.release();
}

size_t FimDevice::reserved_bytes() const { return impl_ ? impl_->reserved : 0; }

template <class Real>
bool FimDevice::run(const FimSetup<Real> &s, int rescale, double mute, double taper, int seed, int fs[3], int fe[3],
                    bool structured, const double *d_points, long npoints, Real *d_field, double stats[4],
                    hipStream_t stream, size_t reserved_elsewhere, std::string &err)
{
    Impl &m = *impl_;
    auto ok = [&](hipError_t e, const char *what) {
        if (e == hipSuccess) return true;
        (void)hipGetLastError();  // leave no error behind for the launch checks of a later call
        err = std::string("FIM: ") + what + ": " + hipGetErrorString(e);
        return false;
    };
    auto fft_ok = [&](hipfftResult r, const char *what) {
        if (r == HIPFFT_SUCCESS) return true;
        err = std::string("FIM: ") + what + " failed (hipfftResult " + std::to_string(int(r)) + ")";
        return false;
    };
    const int n1 = s.npts[0], n2 = s.npts[1], n3 = s.npts[2], h = n1 / 2 + 1;
    const long lines = long(n2) * n3;
    const long total = long(n1) * lines;
    if (total >= (long(1) << kJumpBits) || lines >= (long(1) << 31)) {
        err = "FIM: internal grid too large";
        return false;
    }
    // fixed buffers, made once (each on its own, so that a failure leaves the
    // rest to the next call)
    if (!m.ready) {
        if (!m.d_gen) {
            GenTables t;
            fim_stream::make_tables(t);
            void *g = nullptr;
            if (!ok(hipMalloc(&g, sizeof(GenTables)), "hipMalloc")) return false;
            if (!ok(hipMemcpy(g, &t, sizeof t, hipMemcpyHostToDevice), "hipMemcpy")) {
                (void)hipFree(g);
                return false;
            }
            m.d_gen = static_cast<GenTables *>(g);
            m.reserved += sizeof(GenTables);
        }
        if (!m.d_part) {
            if (!ok(hipMalloc(&m.d_part, sizeof(double) * 3 * kScaleBlocks), "hipMalloc")) return false;
            m.reserved += sizeof(double) * 3 * kScaleBlocks;
        }
        if (!m.d_stat) {
            if (!ok(hipMalloc(&m.d_stat, sizeof(double) * 8), "hipMalloc")) return false;
            m.reserved += sizeof(double) * 8;
        }
        if (!m.h_stat && !ok(hipHostMalloc(reinterpret_cast<void **>(&m.h_stat), sizeof(double) * 8,
                                           hipHostMallocDefault),
                             "hipHostMalloc"))
            return false;
        for (auto &e : m.ev)
            if (!e && !ok(hipEventCreate(&e), "hipEventCreate")) return false;
        m.ready = true;
    }
    if (!m.grow(&m.d_spec, m.spec_cap, sizeof(Cplx<Real>) * size_t(h) * size_t(lines), stream, reserved_elsewhere,
                err))
        return false;
    if (!m.grow(&m.d_real, m.real_cap, sizeof(Real) * size_t(total), stream, reserved_elsewhere, err)) return false;
    const int npoi = static_cast<int>(s.pt.size() / 3);
    if (npoi > 0) {
        if (!m.grow(&m.d_pt, m.pt_cap, sizeof(Real) * s.pt.size(), stream, reserved_elsewhere, err)) return false;
        if (!ok(hipMemcpyAsync(m.d_pt, s.pt.data(), sizeof(Real) * s.pt.size(), hipMemcpyHostToDevice, stream),
                "hipMemcpyAsync"))
            return false;
    }
    // transform: one c2r over the internal grid (backward, unnormalized; the
    // kx = 0 and Nyquist planes are Hermitian after fim_symmetry)
    if (!m.plans || m.plan_n[0] != n1 || m.plan_n[1] != n2 || m.plan_n[2] != n3 || m.plan_real != int(sizeof(Real))) {
        m.drop_plans();
        const hipfftResult r = n3 > 1 ? hipfftPlan3d(&m.plan, n3, n2, n1, FftTypes<Real>::c2r)
                                      : hipfftPlan2d(&m.plan, n2, n1, FftTypes<Real>::c2r);
        if (!fft_ok(r, "hipfftPlan (c2r)")) return false;
        m.plans = true;
        m.plan_n[0] = n1;
        m.plan_n[1] = n2;
        m.plan_n[2] = n3;
        m.plan_real = int(sizeof(Real));
    }
    if (!fft_ok(hipfftSetStream(m.plan, stream), "hipfftSetStream")) return false;

    Cplx<Real> *S = static_cast<Cplx<Real> *>(m.d_spec);
    Real *R = static_cast<Real *>(m.d_real);

    // spectrum
    SpecArgs<Real> a{};
    a.n1 = n1;
    a.n2 = n2;
    a.n3 = n3;
    a.h = h;
    a.lines = lines;
    a.lines_per_item = 8;
    fim_stream::position0(seed, a.s1);
    fim_stream::power(uint64_t(n1), a.line_step);
    for (int d = 0; d < 3; ++d) {
        a.dk[d] = s.dk[d];
        a.cl[d] = s.cl[d];
    }
    a.kc = s.kc;
    a.nu = s.nu;
    a.cnst = s.cnst;
    a.filter = s.filter ? 1 : 0;
    a.acf = s.acf;
    const long items = (lines + a.lines_per_item - 1) / a.lines_per_item;
    const int sblocks = static_cast<int>(std::min<long>((items + 3) / 4, 2048));
    if (!ok(hipEventRecord(m.ev[0], stream), "hipEventRecord")) return false;
    fim_spectrum<Real><<<sblocks, 256, 0, stream>>>(S, a, m.d_gen, items);
    if (!ok(hipGetLastError(), "fim_spectrum launch")) return false;
    if (!ok(hipEventRecord(m.ev[1], stream), "hipEventRecord")) return false;
    const long sym = 2 * lines;
    fim_symmetry<Real><<<static_cast<int>((sym + 255) / 256), 256, 0, stream>>>(S, n1, n2, n3, h);
    if (!ok(hipGetLastError(), "fim_symmetry launch")) return false;
    if (!ok(hipEventRecord(m.ev[2], stream), "hipEventRecord")) return false;
    if (!fft_ok(FftTypes<Real>::exec_c2r(m.plan, S, R), "hipfftExecC2R")) return false;
    if (!ok(hipEventRecord(m.ev[3], stream), "hipEventRecord")) return false;
    // statistics of the scaled internal field; scaling, rescale and taper
    fim_moments<Real><<<kScaleBlocks, 256, 0, stream>>>(R, total, s.scaling, m.d_part);
    if (!ok(hipGetLastError(), "fim_moments launch")) return false;
    fim_finalize<Real><<<1, 256, 0, stream>>>(m.d_part, kScaleBlocks, static_cast<double>(s.sigma), m.d_stat);
    if (!ok(hipGetLastError(), "fim_finalize launch")) return false;
    Real scale = s.scaling;  // still to apply to R when the mapping reads it
    if (rescale == 1 || npoi > 0) {
        const int blocks = static_cast<int>(std::min<long>(lines, kRowBlocks));
        fim_epilogue<Real><<<blocks, 256, 0, stream>>>(R, n1, n2, static_cast<int>(lines), s.scaling, s.dh, m.d_stat,
                                                       rescale == 1 ? 1 : 0, static_cast<const Real *>(m.d_pt),
                                                       npoi, static_cast<Real>(mute), static_cast<Real>(taper));
        if (!ok(hipGetLastError(), "fim_epilogue launch")) return false;
        scale = Real(1);
    }
    if (!ok(hipEventRecord(m.ev[4], stream), "hipEventRecord")) return false;
    // nearest-neighbour values at the caller's points
    if (structured) {
        const int nx = fe[0] - fs[0] + 1, ny = fe[1] - fs[1] + 1, nz = fe[2] - fs[2] + 1;
        const long rows = long(ny) * nz;
        if (rows >= (long(1) << 31)) {
            err = "FIM: output grid too large for the mapping launch";
            return false;
        }
        const int blocks = static_cast<int>(std::min<long>(rows, kRowBlocks));
        if (!m.grow(reinterpret_cast<void **>(&m.d_tab), m.tab_cap, sizeof(int) * size_t(nx + ny + nz), stream,
                    reserved_elsewhere, err))
            return false;
        fim_axis_tables<Real><<<(nx + ny + nz + 255) / 256, 256, 0, stream>>>(
            m.d_tab, nx, ny, nz, fs[0], fs[1], fs[2], s.ds, s.dh, s.off_axis[0], s.off_axis[1], s.off_axis[2], n1, n2,
            n3);
        if (!ok(hipGetLastError(), "fim_axis_tables launch")) return false;
        fim_map_grid<Real><<<blocks, 256, 0, stream>>>(R, n1, n2, n3, scale, d_field, m.d_tab, nx, ny,
                                                       static_cast<int>(rows));
    } else {
        const int mblocks = static_cast<int>(std::max<long>(1, std::min<long>((npoints + 255) / 256, 8192)));
        fim_map_points<Real><<<mblocks, 256, 0, stream>>>(R, n1, n2, n3, s.dh, scale, d_field, d_points,
                                                          d_points + npoints, d_points + 2 * npoints, npoints,
                                                          s.off_axis[0], s.off_axis[1], s.off_axis[2]);
    }
    if (!ok(hipGetLastError(), "fim_map launch")) return false;
    if (!ok(hipEventRecord(m.ev[5], stream), "hipEventRecord")) return false;
    if (!ok(hipMemcpyAsync(m.h_stat, m.d_stat, sizeof(double) * 6, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync"))
        return false;
    if (!ok(hipStreamSynchronize(stream), "FIM run")) return false;
    float ms[5] = {};
    for (int q = 0; q < 5; ++q)
        if (!ok(hipEventElapsedTime(&ms[q], m.ev[q], m.ev[q + 1]), "hipEventElapsedTime")) return false;
    timings_.spectrum = ms[0] * 1e-3;
    timings_.symmetry = ms[1] * 1e-3;
    timings_.fft = ms[2] * 1e-3;
    timings_.mapping = ms[4] * 1e-3;
    timings_.total = (ms[0] + ms[1] + ms[2] + ms[3] + ms[4]) * 1e-3;
    const double *st = m.h_stat;
    if (rescale == 1) {
        stats[0] = static_cast<double>(s.sigma);
        stats[1] = 0.0;
    } else {
        stats[0] = st[3];
        stats[1] = st[4];
    }
    stats[2] = st[1];
    stats[3] = st[2];
    return true;
}

template bool FimDevice::run<float>(const FimSetup<float> &, int, double, double, int, int[3], int[3], bool,
                                    const double *, long, float *, double[4], hipStream_t, size_t, std::string &);
template bool FimDevice::run<double>(const FimSetup<double> &, int, double, double, int, int[3], int[3], bool,
                                     const double *, long, double *, double[4], hipStream_t, size_t, std::string &);

}  // namespace scarf
