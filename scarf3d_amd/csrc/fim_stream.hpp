// The random stream of the FIM spectrum on the device: position n of the
// yarn2 stream of a seed (the (n+1)-th value its generator returns) without
// walking the n values before it.
//
// The generator state (r0, r1) advances by the companion matrix
// A = [[a0, a1], [1, 0]] over GF(m), m = 2^31 - 1, and a value is g^r0 mod m
// (0 when r0 = 0). So the state of position n is A^n applied to the state of
// position 0, A^n is a product of the precomputed powers A^(2^b), and g^e is
// a product of three table entries. Everything is exact modular arithmetic:
// the values equal Yarn2's sequential ones (yarn2.hpp).
//
// The functions are host and device code: the spectrum kernel walks the
// stream with them, and a host test replays the kernel's walk with the same
// functions against Yarn2's sequential stream.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "yarn2.hpp"

namespace scarf {
namespace fim_stream {

constexpr uint32_t kM = 2147483647u;  // 2^31 - 1
constexpr int kJumpBits = 40;         // stream positions below 2^40
constexpr int kPowWords = 4608;       // 2048 + 2048 + 512 powers of g

// Matrices are {p00, p01, p10, p11}, applied to the column (r0, r1).
//   pw[0..2047] = g^q, pw[2048..4095] = g^(2^11 q), pw[4096..4607] = g^(2^22 q)
//   lane[l] = A^l (l < 64), step64 = A^64, jump[b] = A^(2^b)
struct Tables {
    uint32_t pw[kPowWords];
    uint32_t lane[64][4];
    uint32_t step64[4];
    uint32_t jump[kJumpBits][4];
};

// x mod m for x < 2^63 (two Mersenne folds and a conditional subtraction)
__host__ __device__ inline uint32_t mod_m(uint64_t x)
{
    x = (x & kM) + (x >> 31);
    x = (x & kM) + (x >> 31);
    return static_cast<uint32_t>(x >= kM ? x - kM : x);
}

// (r0, r1) <- P (r0, r1) mod m
__host__ __device__ inline void matvec(const uint32_t p[4], uint32_t &r0, uint32_t &r1)
{
    const uint64_t t0 = uint64_t(p[0]) * r0 + uint64_t(p[1]) * r1;
    const uint64_t t1 = uint64_t(p[2]) * r0 + uint64_t(p[3]) * r1;
    r0 = mod_m(t0);
    r1 = mod_m(t1);
}

// g^e mod m for 0 < e < m
__host__ __device__ inline uint32_t pow_g(const uint32_t *pw, uint32_t e)
{
    const uint32_t a = mod_m(uint64_t(pw[e & 2047u]) * pw[2048 + ((e >> 11) & 2047u)]);
    return mod_m(uint64_t(a) * pw[4096 + (e >> 22)]);
}

// the value of a state
__host__ __device__ inline uint32_t value(const uint32_t *pw, uint32_t r0) { return r0 == 0 ? 0u : pow_g(pw, r0); }

// (r0, r1) <- A^n (r0, r1) for n < 2^40
__host__ __device__ inline void jump(const uint32_t (*jt)[4], uint64_t n, uint32_t &r0, uint32_t &r1)
{
    for (int b = 0; b < kJumpBits; ++b)
        if ((n >> b) & 1u) matvec(jt[b], r0, r1);
}

// The walk of one work item of the spectrum kernel: x-lines [l0, l1) of n1
// positions each (line L starts at stream position n1 L), lane `lane` of a
// 64-wide wave visiting points i = lane, lane + 64, ... < h of every line.
// f(line, i, value) receives the stream value of position i + n1 line.
// pw: the power table (a copy in LDS on the device); line_step = A^n1;
// s0: state of stream position 0.
template <class F>
__host__ __device__ inline void walk_item(const uint32_t *pw, const Tables *t, const uint32_t s0[2],
                                          const uint32_t line_step[4], int n1, int h, long l0, long l1, int lane,
                                          F &&f)
{
    uint32_t r0 = s0[0], r1 = s0[1];
    jump(t->jump, uint64_t(n1) * uint64_t(l0), r0, r1);  // position n1 l0
    matvec(t->lane[lane], r0, r1);                       // + lane
    for (long line = l0; line < l1; ++line) {
        uint32_t q0 = r0, q1 = r1;
        for (int i = lane; i < h; i += 64) {
            f(line, i, value(pw, q0));
            matvec(t->step64, q0, q1);
        }
        matvec(line_step, r0, r1);
    }
}

inline void matmul(const uint32_t x[4], const uint32_t y[4], uint32_t out[4])  // out = x y (out may alias)
{
    uint32_t r[4];
    r[0] = mod_m(uint64_t(x[0]) * y[0] + uint64_t(x[1]) * y[2]);
    r[1] = mod_m(uint64_t(x[0]) * y[1] + uint64_t(x[1]) * y[3]);
    r[2] = mod_m(uint64_t(x[2]) * y[0] + uint64_t(x[3]) * y[2]);
    r[3] = mod_m(uint64_t(x[2]) * y[1] + uint64_t(x[3]) * y[3]);
    std::memcpy(out, r, sizeof r);
}

inline void companion(uint32_t a[4])
{
    a[0] = static_cast<uint32_t>(Yarn2::kA0);
    a[1] = static_cast<uint32_t>(Yarn2::kA1);
    a[2] = 1u;
    a[3] = 0u;
}

// A^n
inline void power(uint64_t n, uint32_t out[4])
{
    uint32_t p[4] = {1, 0, 0, 1}, b[4];
    companion(b);
    while (n > 0) {
        if (n & 1) matmul(p, b, p);
        matmul(b, b, b);
        n >>= 1;
    }
    std::memcpy(out, p, sizeof p);
}

inline void make_tables(Tables &t)
{
    const uint64_t g = static_cast<uint64_t>(Yarn2::kGenerator);
    uint64_t v = 1;
    for (int q = 0; q < 2048; ++q, v = v * g % kM) t.pw[q] = static_cast<uint32_t>(v);
    const uint64_t g11 = v;  // g^(2^11)
    v = 1;
    for (int q = 0; q < 2048; ++q, v = v * g11 % kM) t.pw[2048 + q] = static_cast<uint32_t>(v);
    const uint64_t g22 = v;  // g^(2^22)
    v = 1;
    for (int q = 0; q < 512; ++q, v = v * g22 % kM) t.pw[4096 + q] = static_cast<uint32_t>(v);
    for (int l = 0; l < 64; ++l) power(uint64_t(l), t.lane[l]);
    power(64, t.step64);
    uint32_t p[4];
    companion(p);
    for (int b = 0; b < kJumpBits; ++b) {
        std::memcpy(t.jump[b], p, sizeof p);
        matmul(p, p, p);
    }
}

// state of stream position 0 of a seed: seeded as set_seed / prng seed the
// generator (src/prng.cpp:84-85, :147), then one step
inline void position0(int seed, uint32_t s[2])
{
    Yarn2 y;
    y.seed(static_cast<unsigned long>(seed));
    y.discard(1);
    y.state(s[0], s[1]);
}

}  // namespace fim_stream
}  // namespace scarf
