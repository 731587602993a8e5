// Device-side field statistics shared by the field kernels: every block
// reduces the `real` values it wrote to one (count, mean, M2) triple, and a
// final pass merges the triples in a fixed order. Triples are merged with
// Chan et al.'s pairwise formula, the same merge the reference applies to
// per-rank statistics (src/scarflib_common.f90:823-875).
#pragma once
#include <hip/hip_runtime.h>

namespace scarf {

struct Moments {
    double n, mean, m2;
};

__device__ inline void merge_moments(Moments &a, const Moments &b)
{
    if (b.n == 0.0) return;
    if (a.n == 0.0) {
        a = b;
        return;
    }
    const double n = a.n + b.n;
    const double d = b.mean - a.mean;
    a.mean = a.mean + d * (b.n / n);
    a.m2 = a.m2 + b.m2 + d * d * (a.n * b.n / n);
    a.n = n;
}

// Block-wide merge (any multiple of 64 threads up to 2048); thread 0 writes
// the triple to out[0..2]. The butterfly merges the lower lane's triple as
// the left operand on both partners, so both hold bit-identical results.
__device__ inline void block_moments(Moments m, double *__restrict__ out)
{
    __shared__ double sm[3][32];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        Moments o;
        o.n = __shfl_xor(m.n, off);
        o.mean = __shfl_xor(m.mean, off);
        o.m2 = __shfl_xor(m.m2, off);
        if (lane & off) {
            Moments t = o;
            merge_moments(t, m);
            m = t;
        } else {
            merge_moments(m, o);
        }
    }
    __syncthreads();  // sm may still be read by a previous call
    if (lane == 0) {
        sm[0][wave] = m.n;
        sm[1][wave] = m.mean;
        sm[2][wave] = m.m2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Moments t{sm[0][0], sm[1][0], sm[2][0]};
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 1; w < nw; ++w) merge_moments(t, Moments{sm[0][w], sm[1][w], sm[2][w]});
        out[0] = t.n;
        out[1] = t.mean;
        out[2] = t.m2;
    }
}

// Two-pass moments of N register values with a validity mask.
template <int N>
__device__ inline Moments local_moments(const double (&v)[N], const bool (&ok)[N])
{
    double n = 0.0, s = 0.0;
#pragma unroll
    for (int q = 0; q < N; ++q)
        if (ok[q]) {
            n += 1.0;
            s += v[q];
        }
    Moments m{n, 0.0, 0.0};
    if (n == 0.0) return m;
    m.mean = s / n;
#pragma unroll
    for (int q = 0; q < N; ++q)
        if (ok[q]) {
            const double d = v[q] - m.mean;
            m.m2 += d * d;
        }
    return m;
}

// Mute/taper of one value at squared axis distances (dx2, dy2, dz2) from a
// point of interest (src/scarflib_common.f90:881-994): zero within `mute`, a
// raised-cosine ramp out to mute + taper, unchanged beyond. `real` precision
// without contraction, as the reference computes it on the host.
template <class Real>
__device__ inline Real taper_factor(Real val, Real dx2, Real dy2, Real dz2, Real mute, Real taper)
{
#pragma clang fp contract(off)
    const Real d = sqrt(dx2 + dy2 + dz2);
    const Real pi = static_cast<Real>(3.141592653589793);
    if (d <= mute) return Real(0);
    if (d > mute && d <= mute + taper) {
        const Real t = (d - mute) / taper;
        return val * (Real(0.5) - Real(0.5) * cos(t * pi));
    }
    return val;
}

}  // namespace scarf
