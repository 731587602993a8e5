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
        r0_ = static_cast<int64_t>(t % static_cast<uint64_t>(kModulus));
        return r0_ == 0 ? 0 : pow_generator(r0_);
    }

    // uniform01_dist<double>: one draw scaled into [0, 1)
    double uniform01() { return static_cast<double>((*this)()) / static_cast<double>(kModulus); }

private:
    static int64_t pow_generator(int64_t e)
    {
        uint64_t result = 1, base = static_cast<uint64_t>(kGenerator);
        while (e > 0) {
            if (e & 1) result = result * base % static_cast<uint64_t>(kModulus);
            base = base * base % static_cast<uint64_t>(kModulus);
            e >>= 1;
        }
        return static_cast<int64_t>(result);
    }

    int64_t r0_ = 0;
    int64_t r1_ = 1;
};

}  // namespace scarf
