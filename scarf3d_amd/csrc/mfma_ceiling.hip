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
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
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

#define CHECK(x)                                                                      \
    do {                                                                              \
        const hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) {                                                       \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));              \
            return 1;                                                                 \
        }                                                                             \
    } while (0)

int main(int argc, char **argv)
{
    const int iters = argc > 1 ? std::atoi(argv[1]) : 8192;
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
