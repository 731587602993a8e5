// Sustained rate of v_mfma_f32_32x32x16_bf16 on this device, the ceiling the
// MFMA field kernels are measured against: 8 waves per block (2 per SIMD),
// 4 blocks per CU, 8 independent accumulators per wave, no memory traffic in
// the timed loop. Two operand modes:
//   constant: every MFMA of a wave reads the same A and B registers;
//   changing: A and B cycle through 4 random register sets, so the operand
//             bits toggle on every instruction, as in a real GEMM.
// The chip's clock falls as the matrix cores' switching power rises, so the
// second figure is the one a GEMM on random data can approach.
//
// Usage: mfma_ceiling [iterations per launch]   (default 8192)
//        mfma_ceiling coexec
// The second form measures how much independent VALU work a wave can issue
// between its MFMAs: one block per CU (100 KB of LDS keeps a second one out),
// W waves per SIMD with NACC accumulators each, NV FP32 FMAs after every MFMA,
// changing operands.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// src holds 8 uint4 per thread: (A, B) for each of the 4 register sets
template <bool kChanging>
__global__ __launch_bounds__(512, 1) void mfma_loop(const uint4 *__restrict__ src, float *__restrict__ out, int iters)
{
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    bf16x8 a[4], b[4];
    for (int i = 0; i < 4; ++i) {
        const int set = kChanging ? i : 0;
        a[i] = __builtin_bit_cast(bf16x8, src[t * 8 + 2 * set]);
        b[i] = __builtin_bit_cast(bf16x8, src[t * 8 + 2 * set + 1]);
    }
    f32x16 acc[8];
    for (int c = 0; c < 8; ++c) acc[c] = f32x16{};
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < 8; ++c)
            acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[c & 3], b[(c + 1) & 3], acc[c], 0, 0, 0);
    }
    float s = 0.f;
    for (int c = 0; c < 8; ++c)
        for (int r = 0; r < 16; ++r) s += acc[c][r];
    out[t] = s;  // keeps the loop live
}

// one MFMA, then NV VALU instructions, NACC times (instruction-scheduler pattern)
template <int I, int NACC, int NV>
__device__ __forceinline__ void coexec_pattern()
{
    if constexpr (I < NACC) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        if constexpr (NV > 0) __builtin_amdgcn_sched_group_barrier(0x002, NV, 0);
        coexec_pattern<I + 1, NACC, NV>();
    }
}

template <int THREADS, int NACC, int NV>
__global__ __launch_bounds__(THREADS, 1) void coexec_loop(const uint4 *__restrict__ src, float *__restrict__ out,
                                                          int iters)
{
    extern __shared__ float lds[];
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    bf16x8 a[4], b[4];
    for (int i = 0; i < 4; ++i) {
        a[i] = __builtin_bit_cast(bf16x8, src[t * 8 + 2 * i]);
        b[i] = __builtin_bit_cast(bf16x8, src[t * 8 + 2 * i + 1]);
    }
    float v[8];
    for (int k = 0; k < 8; ++k) v[k] = 0.001f * float(threadIdx.x + k);
    const float k1 = 0.999f + 1e-7f * float(blockIdx.x), k2 = 1e-6f;
    f32x16 acc[NACC];
    for (int c = 0; c < NACC; ++c) acc[c] = f32x16{};
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < NACC; ++c) {
            acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[c & 3], b[(c + 1) & 3], acc[c], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < NV; ++j) v[(c * NV + j) & 7] = __builtin_fmaf(v[(c * NV + j) & 7], k1, k2);
        }
        coexec_pattern<0, NACC, NV>();
    }
    float s = 0.f;
    for (int c = 0; c < NACC; ++c)
        for (int r = 0; r < 16; ++r) s += acc[c][r];
    for (int k = 0; k < 8; ++k) s += v[k];
    if (threadIdx.x == 0) lds[0] = s;
    out[t] = s;  // keeps the loop live
}

#define CHECK(x)                                                                      \
    do {                                                                              \
        const hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) {                                                       \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));              \
            return 1;                                                                 \
        }                                                                             \
    } while (0)

template <int THREADS, int NACC, int NV>
int coexec_row(int cus, const uint4 *src, float *out, hipEvent_t e0, hipEvent_t e1)
{
    const int iters = 4096;
    const size_t lds = 100 * 1024;
    CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(coexec_loop<THREADS, NACC, NV>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    coexec_loop<THREADS, NACC, NV><<<cus, THREADS, lds>>>(src, out, 256);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    float best = -1.f;
    for (int rep = 0; rep < 3; ++rep) {
        CHECK(hipEventRecord(e0));
        coexec_loop<THREADS, NACC, NV><<<cus, THREADS, lds>>>(src, out, iters);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (best < 0.f || ms < best) best = ms;
    }
    const double flops = double(cus) * (THREADS / 64) * iters * NACC * (32.0 * 32 * 16 * 2);
    std::printf("| %d | %d | %d | %.3f | %.1f |\n", THREADS / 256, NACC, NV, best, flops / (best * 1e-3) / 1e12);
    std::fflush(stdout);
    return 0;
}

int main(int argc, char **argv)
{
    const bool coexec = argc > 1 && std::strcmp(argv[1], "coexec") == 0;
    const int iters = (argc > 1 && !coexec) ? std::atoi(argv[1]) : 8192;
    if (iters < 1 || iters > (1 << 20)) {
        std::fprintf(stderr, "iterations must be in [1, 2^20]\n");
        return 2;
    }
    int dev = 0, cus = 0;
    CHECK(hipGetDevice(&dev));
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int blocks = cus * 4;
    const size_t threads = size_t(blocks) * 512;
    std::vector<uint32_t> host(threads * 8 * 4);
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> u(-1.f, 1.f);
    for (uint32_t &w : host) {  // two random bf16 (truncated floats) per word
        const uint32_t lo = __builtin_bit_cast(uint32_t, u(rng)) >> 16;
        const uint32_t hi = __builtin_bit_cast(uint32_t, u(rng)) >> 16;
        w = lo | (hi << 16);
    }
    uint4 *src = nullptr;
    float *out = nullptr;
    CHECK(hipMalloc(&src, host.size() * sizeof(uint32_t)));
    CHECK(hipMalloc(&out, threads * sizeof(float)));
    CHECK(hipMemcpy(src, host.data(), host.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    if (coexec) {
        std::printf("mfma_ceiling coexec: %d CUs, one block per CU, 4096 iterations, changing operands\n\n", cus);
        std::printf("| waves per SIMD | accumulators per wave | VALU per MFMA | ms | TFLOP/s bf16 dense |\n");
        std::printf("|---|---|---|---|---|\n");
        int rc = 0;
        rc |= coexec_row<512, 4, 0>(cus, src, out, e0, e1);
        rc |= coexec_row<512, 4, 2>(cus, src, out, e0, e1);
        rc |= coexec_row<512, 4, 3>(cus, src, out, e0, e1);
        rc |= coexec_row<512, 4, 5>(cus, src, out, e0, e1);
        rc |= coexec_row<256, 8, 0>(cus, src, out, e0, e1);
        rc |= coexec_row<256, 8, 2>(cus, src, out, e0, e1);
        rc |= coexec_row<256, 8, 3>(cus, src, out, e0, e1);
        rc |= coexec_row<256, 8, 5>(cus, src, out, e0, e1);
        CHECK(hipFree(src));
        CHECK(hipFree(out));
        return rc;
    }
    const double flops = double(threads / 64) * iters * 8 * (32.0 * 32 * 16 * 2);
    std::printf("mfma_ceiling: %d CUs, %d blocks x 512 threads, %d iterations x 8 MFMA per wave\n", cus, blocks, iters);
    for (int rep = 0; rep < 2; ++rep) {
        for (int mode = 0; mode < 2; ++mode) {
            // short warm launch, then the timed one
            if (mode == 0) mfma_loop<false><<<blocks, 512>>>(src, out, 256);
            else mfma_loop<true><<<blocks, 512>>>(src, out, 256);
            CHECK(hipGetLastError());
            CHECK(hipDeviceSynchronize());
            CHECK(hipEventRecord(e0));
            if (mode == 0) mfma_loop<false><<<blocks, 512>>>(src, out, iters);
            else mfma_loop<true><<<blocks, 512>>>(src, out, iters);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            std::printf("%-8s operands: %8.1f TFLOP/s bf16 dense (%.3f ms)\n", mode ? "changing" : "constant",
                        flops / (ms * 1e-3) / 1e12, ms);
        }
    }
    CHECK(hipFree(src));
    CHECK(hipFree(out));
    return 0;
}
