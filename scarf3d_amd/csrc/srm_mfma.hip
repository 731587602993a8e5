// MFMA formulation of the structured-grid SRM (see srm_mfma.hpp).
//
// Block: 512 threads = 8 waves, output tile 256 grid columns x 128 x-points.
// Wave w owns 32 columns and all 128 x: four v_mfma_f32_32x32x16 blocks.
// Per stage of 16 harmonics (K = 32):
//   - the x-operand tile (128 x 32 K per part, pre-split at table time) is
//     copied global -> LDS once per block and read by all 8 waves;
//   - the generated operand G[c, K] = Re/Im(Z[n,k] Y[n,j]) is formed by the
//     loader threads as FP32 complex products in LDS (one per column and
//     harmonic), and each lane then splits its own 8 K values into 16-bit
//     parts directly in MFMA operand layout (row c = lane % 32,
//     K = 8 (lane / 32) .. +7);
//   - next-stage global loads are issued before the MFMAs of the current
//     stage and written to the other LDS buffer after them (one barrier per
//     stage).
// Accumulation: FP32 MFMA accumulators, promoted into a second FP32
// accumulator every 128 harmonics (8 stages), so the rounding of the long
// sum stays ~1e-7 of the field's standard deviation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "srm_mfma.hpp"
#include "srm_stats.hpp"

namespace scarf {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kStageH = 16;     // harmonics per stage
constexpr int kStageK = 32;     // K values per stage (cos, -sin per harmonic)
constexpr int kTileX = 128;     // x-points per block
constexpr int kTileC = 256;     // grid columns per block
constexpr int kThreads = 512;
constexpr int kPromote = 8;     // stages between promotions
constexpr double kInvTwoPi = 0.15915494309189533577;

enum Scheme { kBf16x3 = 0, kBf16x2 = 1, kF16x2s = 2 };

template <int S>
struct SchemeOf;
template <>
struct SchemeOf<kBf16x3> {
    static constexpr int parts = 3;
    static constexpr int accs = 1;
    using vec = bf16x8;
    using elt = __bf16;
};
template <>
struct SchemeOf<kBf16x2> {
    static constexpr int parts = 2;
    static constexpr int accs = 1;
    using vec = bf16x8;
    using elt = __bf16;
};
template <>
struct SchemeOf<kF16x2s> {
    static constexpr int parts = 2;
    static constexpr int accs = 2;
    using vec = f16x8;
    using elt = _Float16;
};

constexpr float kLoScale = 4096.f;  // fp16 low parts are stored times 2^12

// split x into `parts` 16-bit values whose sum reproduces x to 2^-24 (bf16
// x3), 2^-16 (bf16 x2) or 2^-22 (fp16 x2, low part scaled by 2^12)
template <int S>
__device__ inline void split(float x, typename SchemeOf<S>::elt (&out)[SchemeOf<S>::parts])
{
    using E = typename SchemeOf<S>::elt;
    const E h = static_cast<E>(x);
    const float r = x - static_cast<float>(h);
    out[0] = h;
    if constexpr (S == kBf16x3) {
        const E m = static_cast<E>(r);
        out[1] = m;
        out[2] = static_cast<E>(r - static_cast<float>(m));
    } else if constexpr (S == kBf16x2) {
        out[1] = static_cast<E>(r);
    } else {
        out[1] = static_cast<E>(r * kLoScale);
    }
}

__device__ inline f32x16 mma(bf16x8 a, bf16x8 b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }
__device__ inline f32x16 mma(f16x8 a, f16x8 b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0); }

// LDS position of 16-byte chunk q (K = 8q..8q+7) of row r: the chunk index
// is XOR-swizzled with row bits so that 8 consecutive rows reading the same
// chunk hit 8 distinct 16-byte bank groups.
__device__ inline int chunk_slot(int r, int q) { return r * 4 + (q ^ ((r >> 1) & 3)); }

// ---------------------------------------------------------------------------
// Operand tables, built once per realization:
//   X: [stage][part][nx_pad][32] 16-bit, K = 2h: cos(2 pi tx), 2h+1: -sin(2 pi tx)
//   Y: [np][ny] float2 = (cos, sin)(2 pi ty)
//   Z: [np][nz] float2 = amp * (cos, sin)(2 pi tz)
// Phases are formed and reduced in FP64; sin/cos in FP64 (sincospi).
template <int S>
__global__ __launch_bounds__(256) void mfma_operand_tables(const double *__restrict__ harm, int np_, int nx,
                                                           int nx_pad, int ny, int nz, int ox, int oy, int oz,
                                                           double ds, typename SchemeOf<S>::elt *__restrict__ xt,
                                                           float2 *__restrict__ yt, float2 *__restrict__ zt)
{
    using E = typename SchemeOf<S>::elt;
    constexpr int P = SchemeOf<S>::parts;
    const int width = nx_pad + ny + nz;
    const long total = long(np_) * width;
    for (long e = blockIdx.x * long(blockDim.x) + threadIdx.x; e < total; e += long(gridDim.x) * blockDim.x) {
        const int n = int(e / width);
        const int r = int(e % width);
        if (r < nx_pad) {
            float c = 0.f, s = 0.f;
            if (r < nx) {
                double t = (harm[n] * (double(r + ox) * ds)) * kInvTwoPi;
                t -= rint(t);
                double sd, cd;
                sincospi(2.0 * t, &sd, &cd);
                c = static_cast<float>(cd);
                s = static_cast<float>(-sd);
            }
            E pc[P], ps[P];
            split<S>(c, pc);
            split<S>(s, ps);
            const int stage = n / kStageH, h = n % kStageH;
#pragma unroll
            for (int p = 0; p < P; ++p) {
                E *row = xt + ((size_t(stage) * P + p) * nx_pad + r) * kStageK;
                row[2 * h] = pc[p];
                row[2 * h + 1] = ps[p];
            }
        } else if (r < nx_pad + ny) {
            const int j = r - nx_pad;
            double t = (harm[np_ + n] * (double(j + oy) * ds)) * kInvTwoPi;
            t -= rint(t);
            double sd, cd;
            sincospi(2.0 * t, &sd, &cd);
            yt[size_t(n) * ny + j] = make_float2(static_cast<float>(cd), static_cast<float>(sd));
        } else {
            const int k = r - nx_pad - ny;
            double t = (harm[2 * np_ + n] * (double(k + oz) * ds) + harm[3 * np_ + n]) * kInvTwoPi;
            t -= rint(t);
            double sd, cd;
            sincospi(2.0 * t, &sd, &cd);
            const double a = harm[4 * np_ + n];
            zt[size_t(n) * nz + k] = make_float2(static_cast<float>(a * cd), static_cast<float>(a * sd));
        }
    }
}

// ---------------------------------------------------------------------------
template <int S, class Out>
__global__ __launch_bounds__(kThreads, 1) void mfma_grid(const typename SchemeOf<S>::elt *__restrict__ xt,
                                                         const float2 *__restrict__ yt,
                                                         const float2 *__restrict__ zt, int nx, int ny, int nz,
                                                         int nx_pad, int stages, int tiles_x, int tj_shift, int tk,
                                                         int tiles_j, double scaling, Out *__restrict__ field,
                                                         double *__restrict__ part)
{
    using Sch = SchemeOf<S>;
    using E = typename Sch::elt;
    using V = typename Sch::vec;
    constexpr int P = Sch::parts;
    constexpr int A = Sch::accs;

    __shared__ uint4 sx[2][P][kTileX * 4];         // 16-byte chunks, swizzled
    __shared__ float2 syz[2][kStageH][kTileC];     // generated complex operand

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long b = blockIdx.x;
    const int xt_idx = int(b % tiles_x);
    const long ct = b / tiles_x;
    const int j0 = int(ct % tiles_j) << tj_shift;
    const int k0 = int(ct / tiles_j) * tk;
    const int x0 = xt_idx * kTileX;
    const int tj_mask = (1 << tj_shift) - 1;

    // loader roles
    const int lcol = tid & (kTileC - 1), lh0 = tid >> 8;  // column, first harmonic row (0 or 1)
    const int lj = j0 + (lcol & tj_mask), lk = k0 + (lcol >> tj_shift);
    const bool lok = lj < ny && lk < nz;
    const int xrow = tid >> 2, xq = tid & 3;  // x-table chunk copied by this thread
    const uint4 *xsrc = reinterpret_cast<const uint4 *>(xt);

    uint4 xreg[P];
    float2 yreg[kStageH / 2], zreg[kStageH / 2];
    auto issue_loads = [&](int s) {
#pragma unroll
        for (int p = 0; p < P; ++p) xreg[p] = xsrc[((size_t(s) * P + p) * nx_pad + x0) * 4 + tid];
#pragma unroll
        for (int q = 0; q < kStageH / 2; ++q) {
            const size_t n = size_t(s) * kStageH + lh0 + 2 * q;
            yreg[q] = lok ? yt[n * ny + lj] : make_float2(0.f, 0.f);
            zreg[q] = lok ? zt[n * nz + lk] : make_float2(0.f, 0.f);
        }
    };
    auto commit = [&](int buf) {
#pragma unroll
        for (int p = 0; p < P; ++p) sx[buf][p][chunk_slot(xrow, xq)] = xreg[p];
#pragma unroll
        for (int q = 0; q < kStageH / 2; ++q) {
            const float2 y = yreg[q], z = zreg[q];
            syz[buf][lh0 + 2 * q][lcol] = make_float2(z.x * y.x - z.y * y.y, z.x * y.y + z.y * y.x);
        }
    };

    f32x16 acc[A][4], main_acc[4];
#pragma unroll
    for (int bx = 0; bx < 4; ++bx) {
#pragma unroll
        for (int a = 0; a < A; ++a) acc[a][bx] = f32x16{};
        main_acc[bx] = f32x16{};
    }

    issue_loads(0);
    commit(0);
    __syncthreads();

    const int crow = wave * 32 + (lane & 31);  // generated-operand row of this lane
    const int khalf = lane >> 5;
    for (int s = 0; s < stages; ++s) {
        const int buf = s & 1;
        if (s + 1 < stages) issue_loads(s + 1);
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            // generated operand: 4 harmonics x (re, im) of this lane's column
            E gp[P][8];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float2 v = syz[buf][kk * 8 + khalf * 4 + t][crow];
                E pr[P], pi[P];
                split<S>(v.x, pr);
                split<S>(v.y, pi);
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    gp[p][2 * t] = pr[p];
                    gp[p][2 * t + 1] = pi[p];
                }
            }
            V g[P];
#pragma unroll
            for (int p = 0; p < P; ++p)
#pragma unroll
                for (int e = 0; e < 8; ++e) g[p][e] = gp[p][e];
#pragma unroll
            for (int bx = 0; bx < 4; ++bx) {
                const int xr = bx * 32 + (lane & 31);
                const int slot = chunk_slot(xr, kk * 2 + khalf);
                V xv[P];
#pragma unroll
                for (int p = 0; p < P; ++p) xv[p] = __builtin_bit_cast(V, sx[buf][p][slot]);
                if constexpr (S == kBf16x3) {
                    acc[0][bx] = mma(g[1], xv[1], acc[0][bx]);
                    acc[0][bx] = mma(g[0], xv[2], acc[0][bx]);
                    acc[0][bx] = mma(g[2], xv[0], acc[0][bx]);
                    acc[0][bx] = mma(g[0], xv[1], acc[0][bx]);
                    acc[0][bx] = mma(g[1], xv[0], acc[0][bx]);
                    acc[0][bx] = mma(g[0], xv[0], acc[0][bx]);
                } else if constexpr (S == kBf16x2) {
                    acc[0][bx] = mma(g[0], xv[1], acc[0][bx]);
                    acc[0][bx] = mma(g[1], xv[0], acc[0][bx]);
                    acc[0][bx] = mma(g[0], xv[0], acc[0][bx]);
                } else {
                    acc[1][bx] = mma(g[0], xv[1], acc[1][bx]);
                    acc[1][bx] = mma(g[1], xv[0], acc[1][bx]);
                    acc[0][bx] = mma(g[0], xv[0], acc[0][bx]);
                }
            }
        }
        if (s + 1 < stages) commit(buf ^ 1);
        if ((s % kPromote) == kPromote - 1 || s + 1 == stages) {
#pragma unroll
            for (int bx = 0; bx < 4; ++bx) {
                if constexpr (A == 2) {
                    main_acc[bx] += acc[0][bx] + acc[1][bx] * (1.f / kLoScale);
                } else {
                    main_acc[bx] += acc[0][bx];
                }
#pragma unroll
                for (int a = 0; a < A; ++a) acc[a][bx] = f32x16{};
            }
        }
        __syncthreads();
    }

    // epilogue: scale, round to the field type, store, statistics
    double vals[64];
    bool ok[64];
#pragma unroll
    for (int bx = 0; bx < 4; ++bx) {
        const int x = x0 + bx * 32 + (lane & 31);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = 8 * (r >> 2) + 4 * khalf + (r & 3);
            const int cl = wave * 32 + row;
            const int j = j0 + (cl & tj_mask), k = k0 + (cl >> tj_shift);
            const bool inside = x < nx && j < ny && k < nz;
            const Out val = static_cast<Out>(static_cast<double>(main_acc[bx][r]) * scaling);
            vals[bx * 16 + r] = static_cast<double>(val);
            ok[bx * 16 + r] = inside;
            if (inside) field[size_t(x) + size_t(nx) * (size_t(j) + size_t(ny) * k)] = val;
        }
    }
    block_moments(local_moments(vals, ok), part + 3 * b);
}

template <int S>
size_t x_table_bytes(int nx_pad, int np_)
{
    return size_t(nx_pad) * np_ * 2 * SchemeOf<S>::parts * sizeof(typename SchemeOf<S>::elt);
}

int scheme_of(int kernel)
{
    switch (kernel) {
    case kKernelMfmaBf16x3: return kBf16x3;
    case kKernelMfmaBf16x2: return kBf16x2;
    case kKernelMfmaF16x2: return kF16x2s;
    default: return -1;
    }
}

}  // namespace

bool mfma_plan(const Geometry &g, int np_, int kernel, MfmaPlan &plan, std::string &err)
{
    const int s = scheme_of(kernel);
    if (s < 0) {
        err = "not an MFMA kernel";
        return false;
    }
    if (np_ % kStageH != 0) {
        err = "harmonic table not padded to the MFMA stage";
        return false;
    }
    plan = MfmaPlan{};
    plan.kernel = kernel;
    plan.parts = (s == kBf16x3) ? 3 : 2;
    plan.np_ = np_;
    plan.stages = np_ / kStageH;
    plan.tiles_x = (g.nx + kTileX - 1) / kTileX;
    plan.nx_pad = plan.tiles_x * kTileX;
    // column tile tj (j) x tk (k) = 256 columns: the shape wasting the fewest
    // padded columns, ties going to more k (fewer distinct Y rows per stage)
    long best_waste = -1;
    for (int tk = 8; tk >= 1; tk /= 2) {
        const int tj = kTileC / tk;
        const long waste = long((g.ny + tj - 1) / tj) * tj * long((g.nz + tk - 1) / tk) * tk - long(g.ny) * g.nz;
        if (best_waste < 0 || waste < best_waste) {
            best_waste = waste;
            plan.tj = tj;
            plan.tk = tk;
        }
    }
    plan.tiles_j = (g.ny + plan.tj - 1) / plan.tj;
    plan.tiles_k = (g.nz + plan.tk - 1) / plan.tk;
    plan.blocks = long(plan.tiles_x) * plan.tiles_j * plan.tiles_k;
    if (plan.blocks > (1L << 30)) {
        err = "grid too large for one launch";
        return false;
    }
    plan.nparts = int(plan.blocks);
    plan.x_bytes = (s == kBf16x3) ? x_table_bytes<kBf16x3>(plan.nx_pad, np_)
                                  : x_table_bytes<kBf16x2>(plan.nx_pad, np_);
    plan.y_bytes = size_t(np_) * g.ny * sizeof(float2);
    plan.z_bytes = size_t(np_) * g.nz * sizeof(float2);
    plan.table_bytes = plan.x_bytes + plan.y_bytes + plan.z_bytes + 256;
    return true;
}

bool mfma_tables(const MfmaPlan &plan, const Geometry &g, const double *d_harm, int np_, void *d_tab,
                 hipStream_t stream, std::string &err)
{
    char *base = static_cast<char *>(d_tab);
    float2 *yt = reinterpret_cast<float2 *>(base + plan.x_bytes);
    float2 *zt = reinterpret_cast<float2 *>(base + plan.x_bytes + plan.y_bytes);
    const long total = long(np_) * (plan.nx_pad + g.ny + g.nz);
    const int blocks = int(std::min<long>((total + 255) / 256, 65536));
    switch (scheme_of(plan.kernel)) {
    case kBf16x3:
        mfma_operand_tables<kBf16x3><<<blocks, 256, 0, stream>>>(d_harm, np_, g.nx, plan.nx_pad, g.ny, g.nz, g.ox,
                                                                  g.oy, g.oz, g.ds, reinterpret_cast<__bf16 *>(base),
                                                                  yt, zt);
        break;
    case kBf16x2:
        mfma_operand_tables<kBf16x2><<<blocks, 256, 0, stream>>>(d_harm, np_, g.nx, plan.nx_pad, g.ny, g.nz, g.ox,
                                                                  g.oy, g.oz, g.ds, reinterpret_cast<__bf16 *>(base),
                                                                  yt, zt);
        break;
    default:
        mfma_operand_tables<kF16x2s><<<blocks, 256, 0, stream>>>(d_harm, np_, g.nx, plan.nx_pad, g.ny, g.nz, g.ox,
                                                                  g.oy, g.oz, g.ds,
                                                                  reinterpret_cast<_Float16 *>(base), yt, zt);
        break;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        err = std::string("mfma_operand_tables: ") + hipGetErrorString(e);
        return false;
    }
    return true;
}

template <class Real>
bool mfma_field(const MfmaPlan &plan, const Geometry &g, const void *d_tab, double scaling, Real *d_field,
                double *d_part, hipStream_t stream, std::string &err)
{
    const char *base = static_cast<const char *>(d_tab);
    const float2 *yt = reinterpret_cast<const float2 *>(base + plan.x_bytes);
    const float2 *zt = reinterpret_cast<const float2 *>(base + plan.x_bytes + plan.y_bytes);
    int tj_shift = 0;
    while ((1 << tj_shift) < plan.tj) ++tj_shift;
    const dim3 grid(static_cast<unsigned>(plan.blocks));
    switch (scheme_of(plan.kernel)) {
    case kBf16x3:
        mfma_grid<kBf16x3, Real><<<grid, kThreads, 0, stream>>>(reinterpret_cast<const __bf16 *>(base), yt, zt, g.nx,
                                                                g.ny, g.nz, plan.nx_pad, plan.stages, plan.tiles_x,
                                                                tj_shift, plan.tk, plan.tiles_j, scaling, d_field,
                                                                d_part);
        break;
    case kBf16x2:
        mfma_grid<kBf16x2, Real><<<grid, kThreads, 0, stream>>>(reinterpret_cast<const __bf16 *>(base), yt, zt, g.nx,
                                                                g.ny, g.nz, plan.nx_pad, plan.stages, plan.tiles_x,
                                                                tj_shift, plan.tk, plan.tiles_j, scaling, d_field,
                                                                d_part);
        break;
    default:
        mfma_grid<kF16x2s, Real><<<grid, kThreads, 0, stream>>>(reinterpret_cast<const _Float16 *>(base), yt, zt,
                                                                g.nx, g.ny, g.nz, plan.nx_pad, plan.stages,
                                                                plan.tiles_x, tj_shift, plan.tk, plan.tiles_j,
                                                                scaling, d_field, d_part);
        break;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        err = std::string("mfma_grid: ") + hipGetErrorString(e);
        return false;
    }
    return true;
}

template bool mfma_field<float>(const MfmaPlan &, const Geometry &, const void *, double, float *, double *,
                                hipStream_t, std::string &);
template bool mfma_field<double>(const MfmaPlan &, const Geometry &, const void *, double, double *, double *,
                                 hipStream_t, std::string &);

}  // namespace scarf
