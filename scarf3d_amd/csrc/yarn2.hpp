// Host random stream of the SRM harmonic draw.
//
// The reference draws every harmonic from one process-global TRNG4 stream,
// `static trng::yarn2 r` seeded by set_seed and read two uniforms at a time
// through trng::uniform01_dist<> (src/prng.cpp:60, :147, :152-165). TRNG4
// (>= 4.20) is an external library that is not part of the reference tree;
// this is a restatement of its published yarn2 design (Bauke & Mertens,
// Phys. Rev. E 75, 066701, 2007): an order-2 multiple recursive generator
// modulo the prime m = 2^31 - 1 with L'Ecuyer's full-period coefficients,
// whose output is the state raised as an exponent of the primitive root
// g = 123567893 (which breaks the linear structure of the MRG).
#pragma once
#include <cstdint>
#include <vector>

namespace scarf {

class Yarn2 {
public:
    static constexpr int64_t kModulus = 2147483647;  // 2^31 - 1
    static constexpr int64_t kA0 = 1498809829;
    static constexpr int64_t kA1 = 1160990996;
    static constexpr int64_t kGenerator = 123567893;

    // seed(s): state (s mod m, 1), as trng::yarn2::seed(unsigned long)
    void seed(unsigned long s)
    {
        int64_t t = static_cast<int64_t>(s);
        t %= kModulus;
        if (t < 0) t += kModulus;
        r0_ = t;
        r1_ = 1;
    }

    // next value in [0, m-1]
    int64_t operator()()
    {
        const uint64_t t = static_cast<uint64_t>(kA0) * static_cast<uint64_t>(r0_) +
                           static_cast<uint64_t>(kA1) * static_cast<uint64_t>(r1_);
        r1_ = r0_;
        r0_ = static_cast<int64_t>(mod_m(t));
        return r0_ == 0 ? 0 : pow_generator(r0_);
    }

    // x mod m for x < 2^63, with m = 2^31 - 1 a Mersenne prime: since
    // 2^31 = 1 (mod m), x = (x >> 31) * 2^31 + (x & m) = (x >> 31) + (x & m);
    // two folds leave a value <= m + 3, one conditional subtraction finishes.
    // Exact, like %, without a division.
    static uint64_t mod_m(uint64_t x)
    {
        x = (x & kM) + (x >> 31);
        x = (x & kM) + (x >> 31);
        return x >= kM ? x - kM : x;
    }

    // uniform01_dist<double>: one draw scaled into [0, 1)
    double uniform01() { return static_cast<double>((*this)()) / static_cast<double>(kModulus); }
    static double to_uniform01(int64_t r) { return static_cast<double>(r) / static_cast<double>(kModulus); }

    // Advance the state by n steps without producing output: the state pair
    // (r0, r1) evolves by the companion matrix [[a0, a1], [1, 0]] over GF(m),
    // so n steps are its n-th power (square-and-multiply, O(log n)).
    void discard(uint64_t n)
    {
        uint64_t p[4] = {1, 0, 0, 1};  // accumulated power
        uint64_t b[4] = {static_cast<uint64_t>(kA0), static_cast<uint64_t>(kA1), 1, 0};
        while (n > 0) {
            if (n & 1) mat_mul(p, b);
            mat_mul(b, b);
            n >>= 1;
        }
        const uint64_t r0 = static_cast<uint64_t>(r0_), r1 = static_cast<uint64_t>(r1_);
        r0_ = static_cast<int64_t>((p[0] * r0 % kM + p[1] * r1 % kM) % kM);
        r1_ = static_cast<int64_t>((p[2] * r0 % kM + p[3] * r1 % kM) % kM);
    }

    // the state pair (r0 = latest value of the recursion, r1 = the one before)
    void state(uint32_t &r0, uint32_t &r1) const
    {
        r0 = static_cast<uint32_t>(r0_);
        r1 = static_cast<uint32_t>(r1_);
    }

private:
    static constexpr uint64_t kM = static_cast<uint64_t>(kModulus);

    static void mat_mul(uint64_t x[4], const uint64_t y[4])  // x = x * y (mod m), y may alias x
    {
        const uint64_t r0 = (x[0] * y[0] % kM + x[1] * y[2] % kM) % kM;
        const uint64_t r1 = (x[0] * y[1] % kM + x[1] * y[3] % kM) % kM;
        const uint64_t r2 = (x[2] * y[0] % kM + x[3] * y[2] % kM) % kM;
        const uint64_t r3 = (x[2] * y[1] % kM + x[3] * y[3] % kM) % kM;
        x[0] = r0;
        x[1] = r1;
        x[2] = r2;
        x[3] = r3;
    }

    // g^e mod m for 0 < e < m by two table lookups: e = hi * 2^16 + lo,
    // g^e = g^(hi * 2^16) * g^lo. Exact modular arithmetic, so identical to
    // square-and-multiply.
    struct PowTables {
        std::vector<uint32_t> lo, hi;
        PowTables() : lo(1u << 16), hi(1u << 15)
        {
            uint64_t v = 1;
            for (uint32_t i = 0; i < (1u << 16); ++i) {
                lo[i] = static_cast<uint32_t>(v);
                v = v * static_cast<uint64_t>(kGenerator) % kM;
            }
            const uint64_t step = v;  // g^(2^16)
            v = 1;
            for (uint32_t i = 0; i < (1u << 15); ++i) {
                hi[i] = static_cast<uint32_t>(v);
                v = v * step % kM;
            }
        }
    };
    static const PowTables &tables()
    {
        static const PowTables t;
        return t;
    }
    static int64_t pow_generator(int64_t e)
    {
        const PowTables &t = tables();
        const uint64_t u = static_cast<uint64_t>(e);
        return static_cast<int64_t>(mod_m(static_cast<uint64_t>(t.hi[u >> 16]) * t.lo[u & 0xffffu]));
    }

    int64_t r0_ = 0;
    int64_t r1_ = 1;
};

}  // namespace scarf
