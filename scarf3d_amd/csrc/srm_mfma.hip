// MFMA formulation of the structured-grid SRM (see srm_mfma.hpp).
//
// Block: 512 threads = 8 waves, output tile 256 grid columns x TX x-points
// (TX = 32 * NBX = 64, 96 or 128). Wave w owns 32 columns and all TX x:
// NBX v_mfma_f32_32x32x16 blocks. Per stage of 16 harmonics (K = 32):
//   - the x-operand tile (TX x 32 K per part, pre-split at table time) and
//     the raw complex factors Y (16 x tj) and Z (16 x tk) move global -> LDS
//     by LDS DMA (global_load_lds), issued one stage ahead;
//   - each lane forms the generated operand G[c, K] = Re/Im(Z[n,k] Y[n,j])
//     of its own column and 8 K values and splits it into 16-bit parts
//     directly in MFMA operand layout (row c = lane % 32,
//     K = 8 (lane / 32) .. +7);
//   - each wave reads all x chunks of a k half and issues the products in
//     product-major order, so that the NBX accumulators' dependent chains
//     interleave on the matrix core; the second half's LDS reads and operand
//     formation are scheduled between the first half's products;
//   - one vmcnt wait and one barrier end each stage.
// Accumulation: FP32 MFMA accumulators, promoted into a second FP32
// accumulator every 128 harmonics (8 stages), so the rounding of the long
// sum stays ~1e-7 of the field's standard deviation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <queue>
#include <utility>
#include <vector>

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
static_assert(kFieldChunkHarmonics == kPromote * kStageH, "chunks start at promotion boundaries");
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
    static constexpr int accs = 1;
    using vec = f16x8;
    using elt = _Float16;
};

// fp16 low parts are stored times 2^12 (out of the subnormal range). The
// high part of the generated operand is multiplied by 2^12 too before its
// product with the x high part, so all three products carry the same 2^12 and
// share one accumulator; |G| < 8 (mfma_zscale) keeps 2^12 G inside fp16.
constexpr float kLoScale = 4096.f;

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

// The same splits on a (re, im) pair, one 32-bit word per part: the low half
// holds re's part, the high half im's, which is how a harmonic's two K
// values sit in an MFMA operand register. Each rounding is one packed
// conversion; the rounded values come back to FP32 exactly by a shift or a
// conversion.
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2v __attribute__((ext_vector_type(2)));

template <int S>
__device__ inline uint32_t pack2(float a, float b)
{
    if constexpr (S == kF16x2s) {
        return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, f16x2v));
    } else {
        return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2v));
    }
}
template <int S>
__device__ inline void unpack2(uint32_t w, float &a, float &b)
{
    if constexpr (S == kF16x2s) {
        const f16x2v h = __builtin_bit_cast(f16x2v, w);
        a = static_cast<float>(h[0]);
        b = static_cast<float>(h[1]);
    } else {
        a = __builtin_bit_cast(float, w << 16);
        b = __builtin_bit_cast(float, w & 0xffff0000u);
    }
}
template <int S>
__device__ inline void split2(float re, float im, uint32_t (&out)[SchemeOf<S>::parts])
{
    out[0] = pack2<S>(re, im);
    float hr, hi;
    unpack2<S>(out[0], hr, hi);
    const float rr = re - hr, ri = im - hi;
    if constexpr (S == kBf16x3) {
        out[1] = pack2<S>(rr, ri);
        float mr, mi;
        unpack2<S>(out[1], mr, mi);
        out[2] = pack2<S>(rr - mr, ri - mi);
    } else if constexpr (S == kBf16x2) {
        out[1] = pack2<S>(rr, ri);
    } else {
        out[1] = pack2<S>(rr * kLoScale, ri * kLoScale);
    }
}

__device__ inline f32x16 mma(bf16x8 a, bf16x8 b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }
__device__ inline f32x16 mma(f16x8 a, f16x8 b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0); }

// Instruction-scheduler pattern for the block in front of it: NMF MFMAs, each
// preceded by its share of NDS LDS reads and followed by its share of NVA
// VALU instructions, spread evenly. Work that is independent of the MFMAs
// (the next operand's formation) then issues between them, while each MFMA
// keeps the matrix core busy for 32 cycles. The pattern does not fix which
// MFMA fills which slot: a block whose MFMAs must keep their order is fenced
// into groups of MFMAs that may run in any order.
template <int NMF, int NDS, int NVA, int I = 0>
__device__ __forceinline__ void interleave_mfma()
{
    if constexpr (I < NMF) {
        constexpr int ds = (I + 1) * NDS / NMF - I * NDS / NMF;
        constexpr int va = (I + 1) * NVA / NMF - I * NVA / NMF;
        if constexpr (ds > 0) __builtin_amdgcn_sched_group_barrier(0x100, ds, 0);  // LDS reads
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                         // one MFMA
        if constexpr (va > 0) __builtin_amdgcn_sched_group_barrier(0x002, va, 0);  // VALU
        interleave_mfma<NMF, NDS, NVA, I + 1>();
    }
}

// LDS position of 16-byte chunk q (K = 8q..8q+7) of row r (64 bytes per
// row and part): the chunk index is XOR-permuted by row bits so that the
// lanes one ds_read_b128 serves together, consecutive rows reading the same
// chunk, hit distinct 16-byte bank groups: any 8 rows 8m..8m+7 fill 128
// bytes of banks and any 16 rows 16m..16m+15 fill 256 bytes. (Permuting by
// (r >> 1) & 3 alone covers the 8-row case only; with 16 lanes served per
// cycle it left rows r and r + 8 on the same banks, a 2-way conflict on every
// x-operand read.)
__device__ inline int chunk_perm(int r) { return ((r >> 1) & 3) ^ ((r >> 3) & 1); }
__device__ inline int chunk_slot(int r, int q) { return r * 4 + (q ^ chunk_perm(r)); }

// ---------------------------------------------------------------------------
// Operand tables, built once per realization:
//   X: [stage][part][nx_pad][32] 16-bit, K = 2h: cos(2 pi tx), 2h+1: -sin(2 pi tx)
//   Y: [np][ny] float2 = (cos, sin)(2 pi ty)
//   Z: [np][nz] float2 = amp * (cos, sin)(2 pi tz)
// Phases are formed and reduced in FP64; sin/cos in FP64 (sincospi).
// Work items: one x-point of one stage (its 16 harmonics, so that each part's
// 64-byte row is written whole: consecutive threads write consecutive rows),
// then one (harmonic, j) of Y, then one (harmonic, k) of Z.
template <int S>
__device__ inline uint32_t half_bits(typename SchemeOf<S>::elt v)
{
    return __builtin_bit_cast(uint16_t, v);
}
// one realization's table parameters (nx_pad, and the table pointers, from its plan)
struct TableArgs {
    const double *harm;  // SoA record, np_ values per field
    void *xt;
    float2 *yt, *zt;
    int np_, nx, nx_pad, ny, nz, ox, oy, oz;
    int h0, nh;          // harmonics [h0, h0 + nh) (multiples of the stage)
    int t0;              // first harmonic the tables hold (a multiple of the stage)
    double ds, zscale;
};

// the work items first, first + step, ... of one realization's tables
template <int S>
__device__ __forceinline__ void table_items(const TableArgs &a, long first, long step)
{
    using E = typename SchemeOf<S>::elt;
    constexpr int P = SchemeOf<S>::parts;
    const double *__restrict__ harm = a.harm;
    E *__restrict__ xt = static_cast<E *>(a.xt);
    const int np_ = a.np_, nx = a.nx, nx_pad = a.nx_pad, ny = a.ny, nz = a.nz;
    const double ds = a.ds;
    const int h0 = a.h0, nh = a.nh, t0 = a.t0;
    const long nxi = long(nh / kStageH) * nx_pad, nyi = long(nh) * ny;
    const long total = nxi + nyi + long(nh) * nz;
    for (long e = first; e < total; e += step) {
        if (e < nxi) {
            const int stage = h0 / kStageH + int(e / nx_pad), r = int(e % nx_pad);
            const int slot = stage - t0 / kStageH;  // the stage's place in the tables
            uint32_t words[P][kStageH];  // harmonic h: (cos, -sin) parts as one little-endian word
#pragma unroll
            for (int h = 0; h < kStageH; ++h) {
                const int n = stage * kStageH + h;
                float c = 0.f, s = 0.f;
                if (r < nx) {
                    double t = (harm[n] * (double(r + a.ox) * ds)) * kInvTwoPi;
                    t -= rint(t);
                    double sd, cd;
                    sincospi(2.0 * t, &sd, &cd);
                    c = static_cast<float>(cd);
                    s = static_cast<float>(-sd);
                }
                E pc[P], ps[P];
                split<S>(c, pc);
                split<S>(s, ps);
#pragma unroll
                for (int p = 0; p < P; ++p) words[p][h] = half_bits<S>(pc[p]) | (half_bits<S>(ps[p]) << 16);
            }
#pragma unroll
            for (int p = 0; p < P; ++p) {
                uint4 *row = reinterpret_cast<uint4 *>(xt + ((size_t(slot) * P + p) * nx_pad + r) * kStageK);
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    row[q] = uint4{words[p][4 * q], words[p][4 * q + 1], words[p][4 * q + 2], words[p][4 * q + 3]};
            }
        } else if (e < nxi + nyi) {
            const int n = h0 + int((e - nxi) / ny), j = int((e - nxi) % ny);
            double t = (harm[np_ + n] * (double(j + a.oy) * ds)) * kInvTwoPi;
            t -= rint(t);
            double sd, cd;
            sincospi(2.0 * t, &sd, &cd);
            a.yt[size_t(n - t0) * ny + j] = make_float2(static_cast<float>(cd), static_cast<float>(sd));
        } else {
            const int n = h0 + int((e - nxi - nyi) / nz), k = int((e - nxi - nyi) % nz);
            double t = (harm[2 * np_ + n] * (double(k + a.oz) * ds) + harm[3 * np_ + n]) * kInvTwoPi;
            t -= rint(t);
            double sd, cd;
            sincospi(2.0 * t, &sd, &cd);
            const double amp = harm[4 * np_ + n] * a.zscale;
            a.zt[size_t(n - t0) * nz + k] = make_float2(static_cast<float>(amp * cd), static_cast<float>(amp * sd));
        }
    }
}

template <int S>
__global__ __launch_bounds__(256) void mfma_operand_tables(TableArgs a)
{
    table_items<S>(a, blockIdx.x * long(blockDim.x) + threadIdx.x, long(gridDim.x) * blockDim.x);
}

// the tables of every member of an ensemble in one launch: blockIdx.y is the
// member, the blocks along x share its work items
template <int S>
__global__ __launch_bounds__(256) void ensemble_tables(const TableArgs *__restrict__ args)
{
    const TableArgs a = args[blockIdx.y];
    table_items<S>(a, blockIdx.x * long(blockDim.x) + threadIdx.x, long(gridDim.x) * blockDim.x);
}

// ---------------------------------------------------------------------------
// One stage buffer in LDS; a block double-buffers two of them, declared as
// two separate __shared__ variables so that the compiler can tell a DMA into
// one from reads of the other (and does not wait for the DMA before them).
//   x: P parts x (up to 128 rows x 4 16-byte chunks), chunks XOR-swizzled
//   y: 16 harmonics x tj values of Y (complex), row stride tj
//   z: 16 harmonics x tk values of Z (complex), row stride tk
template <int P>
struct StageLds {
    uint4 x[P][kTileX * 4];
    float2 y[kStageH * kTileC];
    float2 z[kStageH * 8];
};

// LDS DMA of 16 or 4 bytes per lane from base + off (base the same for the
// whole wave, off a 32-bit per-lane byte offset) to lds_wave_base + lane * size
typedef const __attribute__((address_space(1))) char *global_bytes;
// base is uniform; taking it through readfirstlane keeps it in scalar
// registers, so that the load uses the scalar-base + 32-bit lane offset form
// instead of a 64-bit per-lane address
__device__ inline global_bytes uniform_base(const char *base)
{
    const uint64_t b = reinterpret_cast<uint64_t>(base);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(b));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(b >> 32));
    return (global_bytes)((static_cast<uint64_t>(hi) << 32) | lo);
}
__device__ inline void lds_dma16(const char *base, uint32_t off, void *lds_wave_base)
{
    __builtin_amdgcn_global_load_lds(uniform_base(base) + off,
                                     (__attribute__((address_space(3))) void *)lds_wave_base, 16, 0, 0);
}
__device__ inline void lds_dma4(const char *base, uint32_t off, void *lds_wave_base)
{
    __builtin_amdgcn_global_load_lds(uniform_base(base) + off,
                                     (__attribute__((address_space(3))) void *)lds_wave_base, 4, 0, 0);
}

// Flags of a field launch over a range of stages. A realization whose
// harmonics arrive in chunks runs one launch per chunk: every launch but the
// last stores its tile's unscaled FP32 running sums into the field, and
// every launch but the first starts from them. Chunks start at multiples of
// the promotion period, so each value sums the same products in the same
// order as one launch over all stages.
constexpr int kCarryIn = kFieldCarryIn;  // start from the running sums stored in the field
constexpr int kFinal = kFieldFinal;      // scale, round, store and take the statistics

// One output tile: 256 grid columns (tj values of j by tk values of k,
// starting at j0, k0) by TX = 32*NBX x-points starting at x0, over
// `stages` stages of 16 harmonics. Every thread of the block calls this with
// the same arguments.
//
// Stage data moves global -> LDS with the LDS-DMA loads of gfx950
// (global_load_lds: each lane of a wave writes 4 or 16 bytes at
// M0 + lane * size), issued for stage s+1 before the MFMAs of stage s and
// waited for once, right before the stage barrier: no registers hold
// prefetched data. Reads past the ragged grid edges are clamped to valid
// rows; those columns are never stored.
template <int S, int NBX, class Out>
__device__ __forceinline__ void gemm_tile(const typename SchemeOf<S>::elt *__restrict__ xt,
                                          const float2 *__restrict__ yt, const float2 *__restrict__ zt, int nx,
                                          int ny, int nz, int nx_pad, int stages, int tj_shift, int tk, int x0,
                                          int j0, int k0, double scaling, Out *__restrict__ field,
                                          double *__restrict__ part_out, StageLds<SchemeOf<S>::parts> &lds0,
                                          StageLds<SchemeOf<S>::parts> &lds1, int flags = kFinal)
{
    using Sch = SchemeOf<S>;
    using E = typename Sch::elt;
    using V = typename Sch::vec;
    constexpr int P = Sch::parts;
    constexpr int A = Sch::accs;
    constexpr int TX = 32 * NBX;

    // the wave index is uniform: in a scalar register, the LDS targets of the
    // DMAs (M0) are scalar arithmetic
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int tj = 1 << tj_shift;
    const int tj_mask = tj - 1;

    // DMA sources: a per-lane byte offset fixed for the tile plus a base that
    // is the same for the whole block (stage, part, y row group), so that the
    // per-stage address work is scalar.
    //   x: LDS slot tid holds the chunk the permutation puts there
    const bool xload = tid < TX * 4;
    const int xr = tid >> 2;
    const int xq = (tid & 3) ^ chunk_perm(xr);
    const uint32_t xoff = uint32_t((x0 + xr) * 4 + xq) * 16u;
    const size_t xplane = size_t(nx_pad) * 64;  // bytes of one part of one stage
    //   y: float i = q * 512 + tid of the stage tile = row h, element e of
    //      the tile row (2 * tj floats); 2 * tj divides 512, so h advances by
    //      512 / (2 tj) rows per q and e is the same for every q
    const int ny_floats = 32 * tj, nz_floats = 32 * tk;
    const int ye = tid & (2 * tj - 1), yh = tid >> (tj_shift + 1);
    const uint32_t yoff = uint32_t((yh * ny + min(j0 + (ye >> 1), ny - 1)) * 2 + (ye & 1)) * 4u;
    const size_t yq_step = size_t(kThreads / (2 * tj)) * ny * 8;  // bytes per q
    const size_t ystage = size_t(kStageH) * ny * 8;
    //   z: float tid of the stage tile (fewer than 512)
    const int ze = tid % (2 * tk), zh = tid / (2 * tk);
    const uint32_t zoff = uint32_t((zh * nz + min(k0 + (ze >> 1), nz - 1)) * 2 + (ze & 1)) * 4u;
    const size_t zstage = size_t(kStageH) * nz * 8;
    const char *xb = reinterpret_cast<const char *>(xt);
    const char *yb = reinterpret_cast<const char *>(yt);
    const char *zb = reinterpret_cast<const char *>(zt);

    auto issue_loads = [&](int s, StageLds<P> &dst) {
        if (xload) {
#pragma unroll
            for (int p = 0; p < P; ++p)
                lds_dma16(xb + (size_t(s) * P + p) * xplane, xoff, &dst.x[p][wave * 64]);
        }
        const char *ys = yb + size_t(s) * ystage;
        for (int q = 0; q * kThreads < ny_floats; ++q)
            lds_dma4(ys + q * yq_step, yoff, reinterpret_cast<float *>(dst.y) + q * kThreads + wave * 64);
        if (tid < nz_floats) lds_dma4(zb + size_t(s) * zstage, zoff, reinterpret_cast<float *>(dst.z) + wave * 64);
    };

    f32x16 acc[A][NBX], main_acc[NBX];
#pragma unroll
    for (int bx = 0; bx < NBX; ++bx) {
#pragma unroll
        for (int a = 0; a < A; ++a) acc[a][bx] = f32x16{};
        main_acc[bx] = f32x16{};
    }
    // output value r of x block bx: grid position of this lane's entry
    auto out_index = [&](int bx, int r, bool &inside) __attribute__((always_inline)) {
        const int x = x0 + bx * 32 + (lane & 31);
        const int row = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int cl = wave * 32 + row;
        const int j = j0 + (cl & tj_mask), k = k0 + (cl >> tj_shift);
        inside = x < nx && j < ny && k < nz;
        return size_t(x) + size_t(nx) * (size_t(j) + size_t(ny) * k);
    };
    if (flags & kCarryIn) {
#pragma unroll
        for (int bx = 0; bx < NBX; ++bx)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                bool inside;
                const size_t idx = out_index(bx, r, inside);
                main_acc[bx][r] = inside ? static_cast<float>(field[idx]) : 0.f;
            }
    }

    __syncthreads();  // the LDS buffers may still be read by the previous tile
    issue_loads(0, lds0);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();

    const int crow = wave * 32 + (lane & 31);  // generated-operand row (grid column) of this lane
    const int cj = crow & tj_mask, ck = crow >> tj_shift;
    const int khalf = lane >> 5;
    // generated operand of k half kk: Z*Y for 4 harmonics x (re, im) of this
    // lane's column (read_yz), split into P parts in MFMA operand layout (form_g)
    auto read_yz = [&](const StageLds<P> &buf, int kk, float2 (&y)[4], float2 (&z)[4]) __attribute__((always_inline)) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int h = kk * 8 + khalf * 4 + t;
            y[t] = buf.y[h * tj + cj];
            z[t] = buf.z[h * tk + ck];
        }
    };
    // one harmonic t of that operand: the words of part p go to gw[p][t]
    auto form_one = [&](const float2 &y, const float2 &z, uint32_t (&gw)[P][4], uint32_t (&gsw)[4], int t)
        __attribute__((always_inline)) {
        const float re = z.x * y.x - z.y * y.y;
        const float im = z.x * y.y + z.y * y.x;
        if constexpr (S == kF16x2s) {
            // high part h, its 2^12 multiple (exact in fp16: |G| < 8), and the low
            // part 2^12 (x - h) = fma(x, 2^12, -2^12 h): the difference is exact, so the
            // fma rounds nothing and the low part equals the split2 one
            const uint32_t h = pack2<S>(re, im);
            const f16x2v hs = __builtin_bit_cast(f16x2v, h) * f16x2v{static_cast<E>(kLoScale), static_cast<E>(kLoScale)};
            gw[0][t] = h;
            gw[1][t] = pack2<S>(fmaf(re, kLoScale, -static_cast<float>(hs[0])),
                                fmaf(im, kLoScale, -static_cast<float>(hs[1])));
            gsw[t] = __builtin_bit_cast(uint32_t, hs);
        } else {
            uint32_t w[P];
            split2<S>(re, im, w);
#pragma unroll
            for (int p = 0; p < P; ++p) gw[p][t] = w[p];
        }
    };
    auto assemble_g = [&](const uint32_t (&gw)[P][4], const uint32_t (&gsw)[4], V (&g)[P], V &g0s)
        __attribute__((always_inline)) {
#pragma unroll
        for (int p = 0; p < P; ++p) g[p] = __builtin_bit_cast(V, uint4{gw[p][0], gw[p][1], gw[p][2], gw[p][3]});
        if constexpr (S == kF16x2s) g0s = __builtin_bit_cast(V, uint4{gsw[0], gsw[1], gsw[2], gsw[3]});
    };
    auto form_g = [&](const float2 (&y)[4], const float2 (&z)[4], V (&g)[P], V &g0s) __attribute__((always_inline)) {
        uint32_t gw[P][4], gsw[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) form_one(y[t], z[t], gw, gsw, t);
        assemble_g(gw, gsw, g, g0s);
    };
    // all x chunks of k half kk
    auto read_x = [&](const StageLds<P> &buf, int kk, V (&xv)[NBX][P]) __attribute__((always_inline)) {
#pragma unroll
        for (int bx = 0; bx < NBX; ++bx) {
            const int slot = chunk_slot(bx * 32 + (lane & 31), kk * 2 + khalf);
#pragma unroll
            for (int p = 0; p < P; ++p) xv[bx][p] = __builtin_bit_cast(V, buf.x[p][slot]);
        }
    };
    // product q of one k half for every x block (different accumulators);
    // a k half's products run q = 0 .. NP-1, so each accumulator receives
    // them in that order
    constexpr int NP = S == kBf16x3 ? 6 : 3;
    auto product_q = [&](int q, const V (&g)[P], const V &g0s, const V (&xv)[NBX][P]) __attribute__((always_inline)) {
#pragma unroll
        for (int bx = 0; bx < NBX; ++bx) {
            f32x16 &a = acc[0][bx];
            if constexpr (S == kBf16x3) {
                constexpr int gi[6] = {1, 0, 2, 0, 1, 0}, xi[6] = {1, 2, 0, 1, 0, 0};
                a = mma(g[gi[q]], xv[bx][xi[q]], a);
            } else if constexpr (S == kBf16x2) {
                constexpr int gi[3] = {0, 1, 0}, xi[3] = {1, 0, 0};
                a = mma(g[gi[q]], xv[bx][xi[q]], a);
            } else {
                if (q == 0) a = mma(g[0], xv[bx][1], a);
                else if (q == 1) a = mma(g[1], xv[bx][0], a);
                else a = mma(g0s, xv[bx][0], a);
            }
        }
    };
    auto products = [&](const V (&g)[P], const V &g0s, const V (&xv)[NBX][P]) __attribute__((always_inline)) {
#pragma unroll
        for (int q = 0; q < NP; ++q) product_q(q, g, g0s, xv);
    };
    // VALU instructions of one operand formation (4 harmonics: complex product
    // and split), an estimate for the interleave pattern
    constexpr int kFormValu = S == kBf16x3 ? 60 : 48;

    // stage s from `cur` while stage s+1 streams into `nxt`. Three phases:
    //   1. first-half operand G0 and its x chunks;
    //   2. first-half products, with the second half's LDS reads and operand
    //      formation issued between them (their own work for the matrix core
    //      while the VALU forms G1), in fenced groups of one product each:
    //      left to itself the scheduler issued runs of dependent MFMAs on one
    //      accumulator, and this order is 4% faster (bf16 x3, 512^3);
    //   3. second-half products.
    auto stage = [&](int s, StageLds<P> &cur, StageLds<P> &nxt) __attribute__((always_inline)) {
        if (s + 1 < stages) issue_loads(s + 1, nxt);
        V g0[P], g0s0{}, g1[P], g0s1{};
        V xv0[NBX][P], xv1[NBX][P];
        float2 y0[4], z0[4], y1[4], z1[4];
        read_yz(cur, 0, y0, z0);
        read_yz(cur, 1, y1, z1);
        form_g(y0, z0, g0, g0s0);
        read_x(cur, 0, xv0);
        __builtin_amdgcn_sched_barrier(0);
        {
            // phase 2 as NP fenced groups: group q holds product q of every
            // x block, then its share of G1's formation (harmonics t with
            // t % NP == q) and of the second half's x reads, spread between
            // its MFMAs
            uint32_t gw1[P][4], gsw1[4];
            constexpr int NR = NBX * P;  // x reads of one k half
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                product_q(q, g0, g0s0, xv0);
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    if (t % NP == q) form_one(y1[t], z1[t], gw1, gsw1, t);
#pragma unroll
                for (int r = q * NR / NP; r < (q + 1) * NR / NP; ++r) {
                    const int bx = r / P, p = r % P;
                    xv1[bx][p] = __builtin_bit_cast(V, cur.x[p][chunk_slot(bx * 32 + (lane & 31), 2 + khalf)]);
                }
                interleave_mfma<NBX, NR / NP, (kFormValu / 4) * ((4 + NP - 1) / NP)>();
                __builtin_amdgcn_sched_barrier(0);
            }
            assemble_g(gw1, gsw1, g1, g0s1);
        }
        products(g1, g0s1, xv1);
        if ((s % kPromote) == kPromote - 1 || s + 1 == stages) {
#pragma unroll
            for (int bx = 0; bx < NBX; ++bx) {
                if constexpr (S == kF16x2s) {
                    main_acc[bx] += acc[0][bx] * (1.f / kLoScale);
                } else {
                    main_acc[bx] += acc[0][bx];
                }
#pragma unroll
                for (int a = 0; a < A; ++a) acc[a][bx] = f32x16{};
            }
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // stage s+1 has landed in LDS
        __syncthreads();
    };
    for (int s = 0; s < stages; s += 2) {
        stage(s, lds0, lds1);
        if (s + 1 == stages) break;
        stage(s + 1, lds1, lds0);
    }

    if (!(flags & kFinal)) {  // running sums for the next chunk
#pragma unroll
        for (int bx = 0; bx < NBX; ++bx)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                bool inside;
                const size_t idx = out_index(bx, r, inside);
                if (inside) field[idx] = static_cast<Out>(main_acc[bx][r]);
            }
        return;
    }
    // epilogue: scale, round to the field type, store, statistics
    double vals[16 * NBX];
    bool ok[16 * NBX];
#pragma unroll
    for (int bx = 0; bx < NBX; ++bx) {
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
    block_moments(local_moments(vals, ok), part_out);
}

// CHUNK = false: one launch over every stage (flags is kFinal); true: one
// chunk of a realization (flags as passed), a separate instantiation so that
// the single-launch kernel keeps its registers.
template <int S, int NBX, class Out, bool CHUNK>
__global__ __launch_bounds__(kThreads, 1) void mfma_grid(const typename SchemeOf<S>::elt *__restrict__ xt,
                                                         const float2 *__restrict__ yt,
                                                         const float2 *__restrict__ zt, int nx, int ny, int nz,
                                                         int nx_pad, int stages, int tiles_x, int tj_shift, int tk,
                                                         int tiles_j, double scaling, Out *__restrict__ field,
                                                         double *__restrict__ part, long first, int flags)
{
    __shared__ StageLds<SchemeOf<S>::parts> lds0, lds1;
    const long b = first + blockIdx.x;
    const int x0 = int(b % tiles_x) * 32 * NBX;
    const long ct = b / tiles_x;
    const int j0 = int(ct % tiles_j) << tj_shift;
    const int k0 = int(ct / tiles_j) * tk;
    gemm_tile<S, NBX, Out>(xt, yt, zt, nx, ny, nz, nx_pad, stages, tj_shift, tk, x0, j0, k0, scaling, field,
                           part + 3 * b, lds0, lds1, CHUNK ? flags : kFinal);
}

// Ensemble: one persistent launch over the tiles of many realizations.
struct MemberDesc {
    const void *xt;
    const float2 *yt, *zt;
    void *field;
    int nx, ny, nz, nx_pad, stages, tiles_x, tiles_j, tiles_k, tj_shift, tk;
    int x_full, x_tail0, x_tail1;  // mixed x tiles: x_full of 128 points, then tails of 32 * x_tail points
    int part_base;  // first statistics triple of this member
    double scaling, sigma;
};

// per-member merge of the tile triples, then the `real` roundings of
// src/scarflib_srm.f90:880-899 (one block per member)
template <class Real>
__global__ __launch_bounds__(256) void ensemble_finalize(const MemberDesc *__restrict__ members,
                                                         const double *__restrict__ part, double *__restrict__ stat)
{
    __shared__ double sm[3][256];
    const MemberDesc m = members[blockIdx.x];
    const int ntiles = m.tiles_x * m.tiles_j * m.tiles_k;
    Moments acc{0.0, 0.0, 0.0};
    for (int t = threadIdx.x; t < ntiles; t += 256) {
        const double *q = part + 3 * (size_t(m.part_base) + t);
        merge_moments(acc, Moments{q[0], q[1], q[2]});
    }
    sm[0][threadIdx.x] = acc.n;
    sm[1][threadIdx.x] = acc.mean;
    sm[2][threadIdx.x] = acc.m2;
    __syncthreads();
    for (int stride = 128; stride >= 1; stride >>= 1) {
        if (threadIdx.x < stride) {
            Moments a{sm[0][threadIdx.x], sm[1][threadIdx.x], sm[2][threadIdx.x]};
            merge_moments(a, Moments{sm[0][threadIdx.x + stride], sm[1][threadIdx.x + stride],
                                     sm[2][threadIdx.x + stride]});
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
        double *o = stat + 6 * blockIdx.x;
        o[0] = n;
        o[1] = mean;
        o[2] = var;
        o[3] = info3;
        o[4] = static_cast<Real>(mean);
        o[5] = static_cast<Real>(m.sigma) / info3;
    }
}

// rescale of one member: field * (sigma / sigma_d) - mean (Q2), no contraction
template <class Real>
__global__ __launch_bounds__(256) void ensemble_rescale(Real *__restrict__ field, long n,
                                                        const double *__restrict__ stat)
{
#pragma clang fp contract(off)
    const Real sc = static_cast<Real>(stat[5]), mu = static_cast<Real>(stat[4]);
    for (long e = blockIdx.x * long(blockDim.x) + threadIdx.x; e < n; e += long(gridDim.x) * blockDim.x)
        field[e] = field[e] * sc - mu;
}

template <int S, class Out>
__global__ __launch_bounds__(kThreads, 1) void mfma_ensemble(const MemberDesc *__restrict__ members,
                                                             const int2 *__restrict__ items,
                                                             const int *__restrict__ first,
                                                             double *__restrict__ part)
{
    using E = typename SchemeOf<S>::elt;
    __shared__ StageLds<SchemeOf<S>::parts> lds0, lds1;
    // Block b runs items first[b] .. first[b + 1] - 1, a list the host deals
    // out by cost (largest item to the least-loaded block). The bounds depend
    // on blockIdx only, so the loop is uniform by construction. (A work
    // counter claimed by thread 0 and broadcast through LDS was laid out by
    // the compiler as a loop nest in which the other lanes re-entered the item
    // loop and its barriers ahead of thread 0's atomic, and never saw the
    // counter advance.)
    const int end = first[blockIdx.x + 1];
    for (int it = first[blockIdx.x]; it < end; ++it) {
        const int2 item = items[it];  // (member, tile within the member, x fastest)
        const MemberDesc m = members[item.x];
        const int xi = item.y % m.tiles_x;  // x tile: full ones first, then the tails
        int x0 = 128 * xi, nbx = 4;
        if (xi == m.x_full) {
            nbx = m.x_tail0;
        } else if (xi > m.x_full) {
            x0 = 128 * m.x_full + 32 * m.x_tail0;
            nbx = m.x_tail1;
        }
        const int ct = item.y / m.tiles_x;
        const int j0 = (ct % m.tiles_j) << m.tj_shift;
        const int k0 = (ct / m.tiles_j) * m.tk;
        const E *xt = static_cast<const E *>(m.xt);
        Out *field = static_cast<Out *>(m.field);
        double *po = part + 3 * (size_t(m.part_base) + item.y);
        switch (nbx) {
        case 2:
            gemm_tile<S, 2, Out>(xt, m.yt, m.zt, m.nx, m.ny, m.nz, m.nx_pad, m.stages, m.tj_shift, m.tk, x0, j0,
                                 k0, m.scaling, field, po, lds0, lds1);
            break;
        case 3:
            gemm_tile<S, 3, Out>(xt, m.yt, m.zt, m.nx, m.ny, m.nz, m.nx_pad, m.stages, m.tj_shift, m.tk, x0, j0,
                                 k0, m.scaling, field, po, lds0, lds1);
            break;
        default:
            gemm_tile<S, 4, Out>(xt, m.yt, m.zt, m.nx, m.ny, m.nz, m.nx_pad, m.stages, m.tj_shift, m.tk, x0, j0,
                                 k0, m.scaling, field, po, lds0, lds1);
            break;
        }
    }
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

bool mfma_plan(const Geometry &g, int np_, int kernel, MfmaPlan &plan, std::string &err, int nbx, int tk_force)
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
    // x tile of 64, 96 or 128 points: the fewest padded x, ties to the wider tile
    if (nbx < 2 || nbx > 4) {
        nbx = 4;
        long best = -1;
        for (int c = 4; c >= 2; --c) {
            const long waste = long((g.nx + 32 * c - 1) / (32 * c)) * 32 * c - g.nx;
            if (best < 0 || waste < best) {
                best = waste;
                nbx = c;
            }
        }
    }
    plan.nbx = nbx;
    const int tx = 32 * nbx;
    plan.tiles_x = (g.nx + tx - 1) / tx;
    plan.nx_pad = plan.tiles_x * tx;
    // column tile tj (j) x tk (k) = 256 columns: the shape wasting the fewest
    // padded columns, ties going to more k (fewer distinct Y rows per stage)
    long best_waste = -1;
    for (int tk = 8; tk >= 1; tk /= 2) {
        if ((tk_force == 1 || tk_force == 2 || tk_force == 4 || tk_force == 8) && tk != tk_force) continue;
        const int tj = kTileC / tk;
        const long waste = long((g.ny + tj - 1) / tj) * tj * long((g.nz + tk - 1) / tk) * tk - long(g.ny) * g.nz;
        if (best_waste < 0 || waste < best_waste) {
            best_waste = waste;
            plan.tj = tj;
            plan.tk = tk;
        }
    }
    plan.tj_shift = 0;
    while ((1 << plan.tj_shift) < plan.tj) ++plan.tj_shift;
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
    plan.table_bytes = (plan.x_bytes + plan.y_bytes + plan.z_bytes + 255) / 256 * 256;
    return true;
}

bool mfma_plan_mixed(const Geometry &g, int np_, int kernel, MfmaPlan &plan, std::string &err, int tk)
{
    if (!mfma_plan(g, np_, kernel, plan, err, 4, tk)) return false;
    const int chunks = (g.nx + 31) / 32;  // 32-point chunks of x
    int full = chunks / 4, t0 = 0, t1 = 0;
    switch (chunks % 4) {
    case 1:  // 128 + 32 = 96 + 64; a single chunk takes one 64-point tile
        if (full >= 1) {
            --full;
            t0 = 3;
            t1 = 2;
        } else {
            t0 = 2;
        }
        break;
    case 2: t0 = 2; break;
    case 3: t0 = 3; break;
    default: break;
    }
    plan.mixed = true;
    plan.x_full = full;
    plan.x_tail[0] = t0;
    plan.x_tail[1] = t1;
    plan.tiles_x = full + (t0 > 0) + (t1 > 0);
    plan.nx_pad = 32 * (4 * full + t0 + t1);
    plan.blocks = long(plan.tiles_x) * plan.tiles_j * plan.tiles_k;
    if (plan.blocks > (1L << 30)) {
        err = "grid too large for one launch";
        return false;
    }
    plan.nparts = int(plan.blocks);
    plan.x_bytes = (scheme_of(kernel) == kBf16x3) ? x_table_bytes<kBf16x3>(plan.nx_pad, np_)
                                                  : x_table_bytes<kBf16x2>(plan.nx_pad, np_);
    plan.table_bytes = (plan.x_bytes + plan.y_bytes + plan.z_bytes + 255) / 256 * 256;
    return true;
}

double mfma_zscale(int kernel, double amp_max)
{
    double z = 1.0;
    if (scheme_of(kernel) != kF16x2s) return z;
    // SCARF_FP16_AMP_LIMIT lowers the limit (tests reach the rescaled path)
    double limit = 8.0;
    if (const char *e = std::getenv("SCARF_FP16_AMP_LIMIT")) {
        const double v = std::atof(e);
        if (v > 0.0 && v < 8.0) limit = v;
    }
    for (int i = 0; i < 2000 && amp_max * z >= limit; ++i) z *= 0.5;
    return z;
}

namespace {

TableArgs table_args(const MfmaPlan &plan, const Geometry &g, const double *d_harm, int np_, void *d_tab)
{
    char *base = static_cast<char *>(d_tab);
    TableArgs a;
    a.harm = d_harm;
    a.xt = base;
    a.yt = reinterpret_cast<float2 *>(base + plan.x_bytes);
    a.zt = reinterpret_cast<float2 *>(base + plan.x_bytes + plan.y_bytes);
    a.np_ = np_;
    a.nx = g.nx;
    a.nx_pad = plan.nx_pad;
    a.ny = g.ny;
    a.nz = g.nz;
    a.ox = g.ox;
    a.oy = g.oy;
    a.oz = g.oz;
    a.ds = g.ds;
    a.zscale = plan.zscale;
    a.h0 = 0;
    a.nh = np_;
    a.t0 = 0;
    return a;
}

}  // namespace

bool mfma_tables(const MfmaPlan &plan, const Geometry &g, const double *d_harm, int np_, void *d_tab,
                 hipStream_t stream, std::string &err, int h0, int nh)
{
    TableArgs a = table_args(plan, g, d_harm, np_, d_tab);
    if (nh < 0) nh = np_ - h0;
    const int t0 = plan.tab_s0 * kStageH;
    if (h0 < t0 || h0 % kStageH != 0 || nh % kStageH != 0 || h0 + nh > np_ || h0 - t0 + nh > plan.np_) {
        err = "mfma_tables: harmonic range outside the table";
        return false;
    }
    if (nh == 0) return true;
    a.h0 = h0;
    a.nh = nh;
    a.t0 = t0;
    const long total = long(nh / kStageH) * plan.nx_pad + long(nh) * (g.ny + g.nz);  // work items
    const int blocks = int(std::min<long>((total + 255) / 256, 65536));
    switch (scheme_of(plan.kernel)) {
    case kBf16x3: mfma_operand_tables<kBf16x3><<<blocks, 256, 0, stream>>>(a); break;
    case kBf16x2: mfma_operand_tables<kBf16x2><<<blocks, 256, 0, stream>>>(a); break;
    default: mfma_operand_tables<kF16x2s><<<blocks, 256, 0, stream>>>(a); break;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        err = std::string("mfma_operand_tables: ") + hipGetErrorString(e);
        return false;
    }
    return true;
}

namespace {

template <int S, class Real>
void launch_grid(const MfmaPlan &plan, const Geometry &g, const char *base, double scaling, Real *d_field,
                 double *d_part, hipStream_t stream, long blocks, long first, int s0, int nstages, int flags)
{
    using E = typename SchemeOf<S>::elt;
    // operand tables from stage s0 on: X is stage-major, Y and Z harmonic-major
    const size_t ts = size_t(s0 - plan.tab_s0);
    const E *xt = reinterpret_cast<const E *>(base) + ts * SchemeOf<S>::parts * plan.nx_pad * kStageK;
    const float2 *yt = reinterpret_cast<const float2 *>(base + plan.x_bytes) + ts * kStageH * g.ny;
    const float2 *zt = reinterpret_cast<const float2 *>(base + plan.x_bytes + plan.y_bytes) + ts * kStageH * g.nz;
    const dim3 grid(static_cast<unsigned>(blocks));
#define SCARF_GRID_ARGS xt, yt, zt, g.nx, g.ny, g.nz, plan.nx_pad, nstages, plan.tiles_x, plan.tj_shift, plan.tk, \
                        plan.tiles_j, scaling, d_field, d_part, first, flags
    if (flags == kFinal) {
        switch (plan.nbx) {
        case 2: mfma_grid<S, 2, Real, false><<<grid, kThreads, 0, stream>>>(SCARF_GRID_ARGS); break;
        case 3: mfma_grid<S, 3, Real, false><<<grid, kThreads, 0, stream>>>(SCARF_GRID_ARGS); break;
        default: mfma_grid<S, 4, Real, false><<<grid, kThreads, 0, stream>>>(SCARF_GRID_ARGS); break;
        }
    } else {
        switch (plan.nbx) {
        case 2: mfma_grid<S, 2, Real, true><<<grid, kThreads, 0, stream>>>(SCARF_GRID_ARGS); break;
        case 3: mfma_grid<S, 3, Real, true><<<grid, kThreads, 0, stream>>>(SCARF_GRID_ARGS); break;
        default: mfma_grid<S, 4, Real, true><<<grid, kThreads, 0, stream>>>(SCARF_GRID_ARGS); break;
        }
    }
#undef SCARF_GRID_ARGS
}

}  // namespace

template <class Real>
bool mfma_field(const MfmaPlan &plan, const Geometry &g, const void *d_tab, double scaling, Real *d_field,
                double *d_part, hipStream_t stream, std::string &err, long max_blocks, long first_block, int s0,
                int nstages, int flags)
{
    const char *base = static_cast<const char *>(d_tab);
    scaling /= plan.zscale;  // exact: a power of two
    if (first_block < 0 || first_block >= plan.blocks) {
        err = "mfma_field: first tile out of range";
        return false;
    }
    if (nstages < 0) nstages = plan.stages - (s0 - plan.tab_s0);
    // a chunk starts at a promotion boundary, holds at least one stage and
    // lies inside the tables
    if (s0 < plan.tab_s0 || nstages < 1 || s0 - plan.tab_s0 + nstages > plan.stages || s0 % kPromote != 0 ||
        (flags & ~(kFieldCarryIn | kFieldFinal)) != 0 || ((flags & kFieldCarryIn) == 0) != (s0 == 0) ||
        (((flags & kFieldFinal) == 0) && nstages % kPromote != 0)) {
        err = "mfma_field: stage range or flags do not fit the table";
        return false;
    }
    const long left = plan.blocks - first_block;
    const long blocks = max_blocks > 0 ? std::min(left, max_blocks) : left;
    const long f = first_block;
    switch (scheme_of(plan.kernel)) {
    case kBf16x3:
        launch_grid<kBf16x3, Real>(plan, g, base, scaling, d_field, d_part, stream, blocks, f, s0, nstages, flags);
        break;
    case kBf16x2:
        launch_grid<kBf16x2, Real>(plan, g, base, scaling, d_field, d_part, stream, blocks, f, s0, nstages, flags);
        break;
    default:
        launch_grid<kF16x2s, Real>(plan, g, base, scaling, d_field, d_part, stream, blocks, f, s0, nstages, flags);
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
                                 hipStream_t, std::string &, long, long, int, int, int);

// ---------------------------------------------------------------------------
Ensemble::~Ensemble() { release(); }

void Ensemble::release()
{
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
    void *bufs[] = {d_rec_, d_tab_, d_part_, d_desc_, d_items_, d_first_, d_stat_, d_targs_};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    d_rec_ = d_tab_ = d_part_ = d_desc_ = d_items_ = d_first_ = d_stat_ = d_targs_ = nullptr;
    for (auto &e : ev_)
        if (e) (void)hipEventDestroy(e);
    ev_[0] = ev_[1] = nullptr;
    key_ = 0;
}

#define ENS_TRY(expr)                                                                \
    do {                                                                             \
        const hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                      \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);                 \
            return false;                                                            \
        }                                                                            \
    } while (0)

template <class Real>
bool Ensemble::run(hipStream_t stream, int kernel, const std::vector<EnsembleMember> &members,
                   std::vector<double> &stats, double &device_ms, std::string &err)
{
    const int s = scheme_of(kernel);
    if (s < 0) {
        err = "ensembles run on the MFMA kernels (mfma_bf16x3, mfma_f16x2, mfma_bf16x2)";
        return false;
    }
    const int n = static_cast<int>(members.size());
    if (n < 1 || n > 65535) {
        err = n < 1 ? "empty ensemble" : "more than 65535 ensemble members";
        return false;
    }
    // plans, offsets and the shape key
    std::vector<MfmaPlan> plans(n);
    std::vector<size_t> rec_off(n), tab_off(n);
    std::vector<int> part_base(n);
    size_t rec_total = 0, tab_total = 0;
    int parts_total = 0;
    size_t key = 1469598103934665603ull;
    auto mix = [&](size_t v) { key = (key ^ v) * 1099511628211ull; };
    mix(size_t(kernel));
    mix(sizeof(Real));
    for (int m = 0; m < n; ++m) {
        const EnsembleMember &e = members[m];
        if (!e.g.structured) {
            err = "ensemble members must be structured grids";
            return false;
        }
        if (!mfma_plan_mixed(e.g, e.np_, kernel, plans[m], err, e.tk)) return false;
        plans[m].zscale = mfma_zscale(kernel, e.amp_max);
        rec_off[m] = rec_total;
        rec_total += size_t(5) * e.np_;
        tab_off[m] = tab_total;
        tab_total += plans[m].table_bytes;
        part_base[m] = parts_total;
        parts_total += plans[m].nparts;
        const size_t vals[] = {size_t(e.g.nx), size_t(e.g.ny), size_t(e.g.nz), size_t(e.g.ox), size_t(e.g.oy),
                               size_t(e.g.oz), size_t(e.np_), size_t(plans[m].nbx), size_t(plans[m].tk),
                               size_t(e.rescale),
                               reinterpret_cast<size_t>(e.d_field)};
        for (size_t v : vals) mix(v);
        double ds = e.g.ds, sg = e.sigma;
        size_t bits;
        std::memcpy(&bits, &ds, sizeof bits);
        mix(bits);
        std::memcpy(&bits, &sg, sizeof bits);
        mix(bits);
        std::memcpy(&bits, &plans[m].zscale, sizeof bits);  // baked into the captured table kernel
        mix(bits);
    }

    if (key != key_ || !exec_) {
        // work items, most costly first (cost = stages x tile width; ties keep
        // member and tile order)
        std::vector<std::pair<long, int2>> costed;
        for (int m = 0; m < n; ++m) {
            const MfmaPlan &p = plans[m];
            for (int t = 0; t < p.nparts; ++t) {
                const int xi = t % p.tiles_x;
                const int nbx = xi < p.x_full ? 4 : (xi == p.x_full ? p.x_tail[0] : p.x_tail[1]);
                costed.push_back({long(p.stages) * nbx, make_int2(m, t)});
            }
        }
        using Costed = std::pair<long, int2>;
        std::stable_sort(costed.begin(), costed.end(),
                         [](const Costed &a, const Costed &b) { return a.first > b.first; });

        release();
        ENS_TRY(hipEventCreate(&ev_[0]));
        ENS_TRY(hipEventCreate(&ev_[1]));
        ENS_TRY(hipMalloc(&d_rec_, rec_total * sizeof(double)));
        ENS_TRY(hipMalloc(&d_tab_, std::max<size_t>(tab_total, 256)));
        ENS_TRY(hipMalloc(&d_part_, size_t(3) * parts_total * sizeof(double)));
        ENS_TRY(hipMalloc(&d_desc_, size_t(n) * sizeof(MemberDesc)));
        ENS_TRY(hipMalloc(&d_items_, costed.size() * sizeof(int2)));
        ENS_TRY(hipMalloc(&d_stat_, size_t(6) * n * sizeof(double)));
        rec_bytes_ = rec_total * sizeof(double);
        nitems_ = static_cast<int>(costed.size());
        nmembers_ = n;
        // member descriptors and the work list do not change with the seeds
        desc_host_.assign(size_t(n) * sizeof(MemberDesc), 0);
        MemberDesc *desc = reinterpret_cast<MemberDesc *>(desc_host_.data());
        for (int m = 0; m < n; ++m) {
            const MfmaPlan &p = plans[m];
            char *base = static_cast<char *>(d_tab_) + tab_off[m];
            MemberDesc &d = desc[m];
            d.xt = base;
            d.yt = reinterpret_cast<const float2 *>(base + p.x_bytes);
            d.zt = reinterpret_cast<const float2 *>(base + p.x_bytes + p.y_bytes);
            d.field = members[m].d_field;
            d.nx = members[m].g.nx;
            d.ny = members[m].g.ny;
            d.nz = members[m].g.nz;
            d.nx_pad = p.nx_pad;
            d.stages = p.stages;
            d.tiles_x = p.tiles_x;
            d.tiles_j = p.tiles_j;
            d.tiles_k = p.tiles_k;
            d.tj_shift = p.tj_shift;
            d.tk = p.tk;
            d.x_full = p.x_full;
            d.x_tail0 = p.x_tail[0];
            d.x_tail1 = p.x_tail[1];
            d.part_base = part_base[m];
            d.scaling = members[m].scaling / p.zscale;
            d.sigma = members[m].sigma;
        }
        ENS_TRY(hipMemcpy(d_desc_, desc_host_.data(), desc_host_.size(), hipMemcpyHostToDevice));

        int dev = 0, cus = 256;
        ENS_TRY(hipGetDevice(&dev));
        ENS_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const int persistent = std::max(1, std::min(nitems_, cus));
        // Deal the items, largest first, each to the block with the least
        // cost so far (ties to the lower block index), then lay out each
        // block's list contiguously. On the 64-member benchmark ensemble the
        // busiest block carries 0.07% more than the mean, where the dealing
        // b, b + G, b + 2G, ... of the sorted list carried 1.3% more.
        std::vector<std::vector<int2>> lists(persistent);
        {
            using Load = std::pair<long, int>;
            std::priority_queue<Load, std::vector<Load>, std::greater<Load>> least;
            for (int b = 0; b < persistent; ++b) least.push({0L, b});
            for (const auto &c : costed) {
                const Load l = least.top();
                least.pop();
                lists[l.second].push_back(c.second);
                least.push({l.first + c.first, l.second});
            }
        }
        std::vector<int2> items;
        std::vector<int> first(size_t(persistent) + 1, 0);
        items.reserve(costed.size());
        for (int b = 0; b < persistent; ++b) {
            items.insert(items.end(), lists[b].begin(), lists[b].end());
            first[b + 1] = static_cast<int>(items.size());
        }
        ENS_TRY(hipMalloc(&d_first_, first.size() * sizeof(int)));
        ENS_TRY(hipMemcpy(d_items_, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
        ENS_TRY(hipMemcpy(d_first_, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice));

        // capture the launch sequence
        // operand-table parameters of every member (one launch builds all tables)
        std::vector<TableArgs> targs(n);
        for (int m = 0; m < n; ++m)
            targs[m] = table_args(plans[m], members[m].g, static_cast<const double *>(d_rec_) + rec_off[m],
                                  members[m].np_, static_cast<char *>(d_tab_) + tab_off[m]);
        ENS_TRY(hipMalloc(&d_targs_, size_t(n) * sizeof(TableArgs)));
        ENS_TRY(hipMemcpy(d_targs_, targs.data(), size_t(n) * sizeof(TableArgs), hipMemcpyHostToDevice));

        ENS_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        const TableArgs *dt = static_cast<const TableArgs *>(d_targs_);
        const dim3 tgrid(256, static_cast<unsigned>(n));  // 256 blocks share each member's work items
        switch (s) {
        case kBf16x3: ensemble_tables<kBf16x3><<<tgrid, 256, 0, stream>>>(dt); break;
        case kBf16x2: ensemble_tables<kBf16x2><<<tgrid, 256, 0, stream>>>(dt); break;
        default: ensemble_tables<kF16x2s><<<tgrid, 256, 0, stream>>>(dt); break;
        }
        const MemberDesc *dd = static_cast<const MemberDesc *>(d_desc_);
        const int2 *di = static_cast<const int2 *>(d_items_);
        const int *df = static_cast<const int *>(d_first_);
        double *dp = static_cast<double *>(d_part_);
        switch (s) {
        case kBf16x3: mfma_ensemble<kBf16x3, Real><<<persistent, kThreads, 0, stream>>>(dd, di, df, dp); break;
        case kBf16x2: mfma_ensemble<kBf16x2, Real><<<persistent, kThreads, 0, stream>>>(dd, di, df, dp); break;
        default: mfma_ensemble<kF16x2s, Real><<<persistent, kThreads, 0, stream>>>(dd, di, df, dp); break;
        }
        ensemble_finalize<Real><<<n, 256, 0, stream>>>(dd, dp, static_cast<double *>(d_stat_));
        for (int m = 0; m < n; ++m) {
            if (members[m].rescale != 1) continue;
            const long pts = members[m].g.points();
            const int blocks = int(std::min<long>((pts + 255) / 256, 8192));
            ensemble_rescale<Real><<<blocks, 256, 0, stream>>>(static_cast<Real *>(members[m].d_field), pts,
                                                               static_cast<const double *>(d_stat_) + 6 * m);
        }
        ENS_TRY(hipStreamEndCapture(stream, &graph_));
        ENS_TRY(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
        key_ = key;
        reused_ = false;
    } else {
        reused_ = true;
        // scaling can differ between runs of the same shape (sigma, spectrum)
        MemberDesc *desc = reinterpret_cast<MemberDesc *>(desc_host_.data());
        bool changed = false;
        for (int m = 0; m < n; ++m) {
            const double sc = members[m].scaling / plans[m].zscale;
            if (desc[m].scaling != sc) {
                desc[m].scaling = sc;
                changed = true;
            }
        }
        if (changed) ENS_TRY(hipMemcpy(d_desc_, desc_host_.data(), desc_host_.size(), hipMemcpyHostToDevice));
    }
    for (int m = 0; m < n; ++m)
        ENS_TRY(hipMemcpyAsync(static_cast<double *>(d_rec_) + rec_off[m], members[m].record,
                               size_t(5) * members[m].np_ * sizeof(double), hipMemcpyHostToDevice, stream));
    ENS_TRY(hipEventRecord(ev_[0], stream));
    ENS_TRY(hipGraphLaunch(exec_, stream));
    ENS_TRY(hipEventRecord(ev_[1], stream));
    stats.assign(size_t(6) * n, 0.0);
    ENS_TRY(hipMemcpyAsync(stats.data(), d_stat_, size_t(6) * n * sizeof(double), hipMemcpyDeviceToHost, stream));
    ENS_TRY(hipStreamSynchronize(stream));
    float ms = 0.f;
    ENS_TRY(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
    device_ms = ms;
    return true;
}

template bool Ensemble::run<float>(hipStream_t, int, const std::vector<EnsembleMember> &, std::vector<double> &,
                                   double &, std::string &);
template bool Ensemble::run<double>(hipStream_t, int, const std::vector<EnsembleMember> &, std::vector<double> &,
                                    double &, std::string &);
// ---------------------------------------------------------------------------
// Tile autotuner
namespace {

// Single-field candidates run the real grid's first tiles, one block per CU
// (the kernel's steady state: LDS holds one block per CU), for 320 stages
// (5120 harmonics: an x table far larger than the caches, as in real runs).
// Short runs on a small table ranked the tiles differently from full runs.
constexpr int kTuneStages = 320;
constexpr int kTuneStagesSat = 24;  // stages per candidate on the device-filling rate-table grid

// synthetic record: distinct small wavenumbers, unit amplitudes
__global__ void tune_record(double *rec, int np_)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= np_) return;
    rec[n] = 1.0e-3 * (1 + n % 7);
    rec[np_ + n] = 2.0e-3 * (1 + n % 5);
    rec[2 * np_ + n] = 3.0e-3 * (1 + n % 3);
    rec[3 * np_ + n] = 0.1 * n;
    rec[4 * np_ + n] = 1.0;
}

bool tune_reserve(void **p, size_t &cap, size_t bytes, std::string &err)
{
    if (bytes <= cap && *p) return true;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    cap = 0;
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        err = std::string("tile tuner: hipMalloc: ") + hipGetErrorString(e);
        return false;
    }
    cap = bytes;
    return true;
}

}  // namespace

TileTuner::~TileTuner()
{
    if (d_sat_) (void)hipFree(d_sat_);
    if (d_rec_) (void)hipFree(d_rec_);
    if (d_rec_sat_) (void)hipFree(d_rec_sat_);
    if (d_tab_) (void)hipFree(d_tab_);
    if (d_part_) (void)hipFree(d_part_);
    for (hipEvent_t &e : ev_)
        if (e) (void)hipEventDestroy(e);
}

bool TileTuner::prepare(double **rec, int np_, hipStream_t stream, std::string &err)
{
    if (!ev_[0]) {
        ENS_TRY(hipEventCreate(&ev_[0]));
        ENS_TRY(hipEventCreate(&ev_[1]));
    }
    if (!*rec) {
        ENS_TRY(hipMalloc(rec, size_t(5) * np_ * sizeof(double)));
        tune_record<<<(np_ + 255) / 256, 256, 0, stream>>>(*rec, np_);
        ENS_TRY(hipGetLastError());
    }
    return true;
}

bool TileTuner::enabled()
{
    const char *e = std::getenv("SCARF_TUNE");
    return !(e && e[0] == '0');
}

void TileTuner::load_disk()
{
    if (disk_loaded_) return;
    disk_loaded_ = true;
    const char *path = std::getenv("SCARF_TUNE_CACHE");
    if (!path || !*path) return;
    FILE *f = std::fopen(path, "r");
    if (!f) return;
    long k[5];
    int nbx, tk;
    while (std::fscanf(f, "%ld %ld %ld %ld %ld %d %d", &k[0], &k[1], &k[2], &k[3], &k[4], &nbx, &tk) == 7) {
        if (nbx < 2 || nbx > 4 || (tk != 1 && tk != 2 && tk != 4 && tk != 8)) continue;
        TileChoice c;
        c.nbx = nbx;
        c.tk = tk;
        c.tuned = true;
        cache_[Key{k[0], k[1], k[2], k[3], k[4]}] = c;
    }
    std::fclose(f);
}

void TileTuner::save_disk(const Key &k, const TileChoice &c)
{
    const char *path = std::getenv("SCARF_TUNE_CACHE");
    if (!path || !*path) return;
    FILE *f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%ld %ld %ld %ld %ld %d %d\n", k[0], k[1], k[2], k[3], k[4], c.nbx, c.tk);
    std::fclose(f);
}

// SCARF_TILE=x,j,k (x in {64, 96, 128}, j * k = 256, k in {1, 2, 4, 8})
// fixes the block tile of single-field runs, e.g. to profile one tile
bool TileTuner::cached(const Geometry &g, int kernel, size_t real_bytes, TileChoice &out)
{
    const Key key{long(kernel), long(real_bytes), long(g.nx), long(g.ny), long(g.nz)};
    load_disk();
    const auto it = cache_.find(key);
    if (it == cache_.end()) return false;
    out = it->second;
    return true;
}

bool TileTuner::forced(TileChoice &out)
{
    const char *e = std::getenv("SCARF_TILE");
    if (!e || !*e) return false;
    int x = 0, j = 0, k = 0;
    if (std::sscanf(e, "%d,%d,%d", &x, &j, &k) != 3) return false;
    if ((x != 64 && x != 96 && x != 128) || (k != 1 && k != 2 && k != 4 && k != 8) || j * k != kTileC) return false;
    out = TileChoice{};
    out.nbx = x / 32;
    out.tk = k;
    return true;
}

template <class Real>
bool TileTuner::choose(const Geometry &g, int kernel, Real *d_scratch, hipStream_t stream, TileChoice &out,
                       std::string &err)
{
    const Key key{long(kernel), long(sizeof(Real)), long(g.nx), long(g.ny), long(g.nz)};
    load_disk();
    const auto it = cache_.find(key);
    if (it != cache_.end()) {
        out = it->second;
        return true;
    }
    const int np_ = kTuneStages * kStageH;
    MfmaPlan def;
    if (!mfma_plan(g, np_, kernel, def, err)) return false;
    if (!prepare(&d_rec_, np_, stream, err)) return false;
    int dev = 0, cus = 256;
    ENS_TRY(hipGetDevice(&dev));
    ENS_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    auto padded = [&](const MfmaPlan &p) { return double(p.blocks) * kTileC * 32.0 * p.nbx; };
    TileChoice best;
    double best_ms = -1.0, def_ms = -1.0;
    // The column shape is tuned at the least-padding x width: every candidate
    // then does the same work per block, so one wave of blocks compares them
    // fairly. (Across x widths the first wave misranked them: at 1024 x 1024
    // x 512 it preferred the 96-wide tile, 15% slower in full.) When that
    // width leaves CUs without a block, narrower x tiles (more blocks) are
    // tried too, as long as all their blocks still run at once: their timing
    // then covers the whole grid.
    for (int nbx = def.nbx; nbx >= 2; --nbx) {
        if (nbx != def.nbx && def.blocks >= cus) break;
        for (int tk = 8; tk >= 1; tk /= 2) {
            MfmaPlan p;
            if (!mfma_plan(g, np_, kernel, p, err, nbx, tk)) return false;
            if (p.nbx != nbx) continue;
            if (nbx != def.nbx && p.blocks > cus) continue;
            const bool is_def = p.nbx == def.nbx && p.tk == def.tk;
            if (!is_def && padded(p) > 2.0 * padded(def)) continue;  // far more padding: cannot win
            if (!tune_reserve(&d_tab_, tab_cap_, p.table_bytes, err)) return false;
            if (!tune_reserve(&d_part_, part_cap_, size_t(3) * p.nparts * sizeof(double), err)) return false;
            if (!mfma_tables(p, g, d_rec_, np_, d_tab_, stream, err)) return false;
            if (!mfma_field(p, g, d_tab_, 1.0, d_scratch, static_cast<double *>(d_part_), stream, err, cus)) return false;
            float t_min = -1.f;
            for (int rep = 0; rep < 3; ++rep) {
                ENS_TRY(hipEventRecord(ev_[0], stream));
                if (!mfma_field(p, g, d_tab_, 1.0, d_scratch, static_cast<double *>(d_part_), stream, err, cus))
                    return false;
                ENS_TRY(hipEventRecord(ev_[1], stream));
                ENS_TRY(hipEventSynchronize(ev_[1]));
                float ms = 0.f;
                ENS_TRY(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
                if (t_min < 0.f || ms < t_min) t_min = ms;
            }
            // the full grid runs in ceil(blocks / CUs) such waves
            const double est = double(t_min) * double((p.blocks + cus - 1) / cus);
            if (is_def) def_ms = est;
            if (best_ms < 0.0 || est < best_ms) {
                best_ms = est;
                best.nbx = p.nbx;
                best.tk = p.tk;
            }
        }
    }
    // keep the least-padding tile unless another is clearly faster: the
    // candidates' short timings differ by a few percent from run to run
    if (def_ms > 0.0 && best_ms > 0.95 * def_ms) {
        best.nbx = def.nbx;
        best.tk = def.tk;
        best_ms = def_ms;
    }
    best.ms = best_ms;
    best.ms_default = def_ms;
    best.tuned = true;
    cache_[key] = best;
    save_disk(key, best);
    out = best;
    return true;
}

template <class Real>
bool TileTuner::choose_saturated(const Geometry &g, int kernel, hipStream_t stream, TileChoice &out, std::string &err)
{
    const int np_ = kTuneStagesSat * kStageH;
    const long ks = long(sizeof(Real));
    // the throughput table of this kernel and output type, timed once
    if (rate_.find({long(kernel), ks, 4L, 8L}) == rate_.end()) {
        Geometry sg;  // 384 x 512 x 128: no x padding for 64/96/128, no column padding for any tj x tk
        sg.structured = true;
        sg.nx = 384;
        sg.ny = 512;
        sg.nz = 128;
        sg.ds = 10.0;
        if (!tune_reserve(&d_sat_, sat_cap_, size_t(sg.points()) * sizeof(Real), err)) return false;
        if (!prepare(&d_rec_sat_, np_, stream, err)) return false;
        Real *scratch = static_cast<Real *>(d_sat_);
        for (int nbx = 4; nbx >= 2; --nbx) {
            for (int tk = 8; tk >= 1; tk /= 2) {
                MfmaPlan p;
                if (!mfma_plan(sg, np_, kernel, p, err, nbx, tk)) return false;
                if (p.nbx != nbx || p.tk != tk) continue;
                if (!tune_reserve(&d_tab_, tab_cap_, p.table_bytes, err)) return false;
                if (!tune_reserve(&d_part_, part_cap_, size_t(3) * p.nparts * sizeof(double), err)) return false;
                if (!mfma_tables(p, sg, d_rec_sat_, np_, d_tab_, stream, err)) return false;
                if (!mfma_field(p, sg, d_tab_, 1.0, scratch, static_cast<double *>(d_part_), stream, err)) return false;
                float t_min = -1.f;
                for (int rep = 0; rep < 3; ++rep) {
                    ENS_TRY(hipEventRecord(ev_[0], stream));
                    if (!mfma_field(p, sg, d_tab_, 1.0, scratch, static_cast<double *>(d_part_), stream, err))
                        return false;
                    ENS_TRY(hipEventRecord(ev_[1], stream));
                    ENS_TRY(hipEventSynchronize(ev_[1]));
                    float ms = 0.f;
                    ENS_TRY(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
                    if (t_min < 0.f || ms < t_min) t_min = ms;
                }
                rate_[{long(kernel), ks, long(nbx), long(tk)}] = double(sg.points()) * np_ / std::max(t_min, 1e-6f);
            }
        }
    }
    // the member's column tile: over its mixed x tiles, the padded work of each
    // tile width over that width's saturated rate
    double best = -1.0;
    for (int tk = 8; tk >= 1; tk /= 2) {
        MfmaPlan p;
        if (!mfma_plan_mixed(g, np_, kernel, p, err, tk)) return false;
        if (p.tk != tk) continue;
        const long col_tiles = long(p.tiles_j) * p.tiles_k;
        double cost = 0.0;
        bool have = true;
        const int widths[3] = {4, p.x_tail[0], p.x_tail[1]};
        const int counts[3] = {p.x_full, p.x_tail[0] > 0, p.x_tail[1] > 0};
        for (int w = 0; w < 3; ++w) {
            if (counts[w] == 0) continue;
            const auto it = rate_.find({long(kernel), ks, long(widths[w]), long(tk)});
            if (it == rate_.end()) {
                have = false;
                break;
            }
            cost += double(counts[w]) * col_tiles * kTileC * 32.0 * widths[w] / it->second;  // points incl. padding / rate
        }
        if (!have) continue;
        if (best < 0.0 || cost < best) {
            best = cost;
            out.nbx = 0;
            out.tk = tk;
        }
    }
    out.ms = best * np_;
    out.ms_default = 0.0;
    out.tuned = true;
    return best >= 0.0;
}

template bool TileTuner::choose_saturated<float>(const Geometry &, int, hipStream_t, TileChoice &, std::string &);
template bool TileTuner::choose_saturated<double>(const Geometry &, int, hipStream_t, TileChoice &, std::string &);
template bool TileTuner::choose<float>(const Geometry &, int, float *, hipStream_t, TileChoice &, std::string &);
template bool TileTuner::choose<double>(const Geometry &, int, double *, hipStream_t, TileChoice &, std::string &);

template bool mfma_field<double>(const MfmaPlan &, const Geometry &, const void *, double, double *, double *,
                                 hipStream_t, std::string &, long, long, int, int, int);

}  // namespace scarf
