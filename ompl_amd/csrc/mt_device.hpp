// Device std::mt19937 (the engine of ompl::RNG, util/RandomNumbers.h:190) and
// libstdc++'s generate_canonical<double, 53> over it (the 0-1 distribution
// of RNG::uniformReal, util/RandomNumbers.h:72-76, :191).
//
// The 624-word state lives in LDS.  A twist is done by the whole block in the
// four dependency phases of the in-place recurrence
//   new[i] = src[i+397] ^ twist(old[i], old[i+1])   (indices mod 624)
// phase A: i in [0,227)   src = old,
// phase B: i in [227,454) src = new (phase A),
// phase C: i in [454,623) src = new (phase B),
// phase D: i = 623        old[624 mod 624] = new[0], src = new[396],
// each phase reading its inputs before any thread writes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ompl_amd {

constexpr int kMtN = 624;
constexpr int kMtM = 397;
constexpr uint32_t kMtMatrix = 0x9908b0dfu;
constexpr uint32_t kMtUpper = 0x80000000u, kMtLower = 0x7fffffffu;

__device__ __forceinline__ uint32_t mt_twist_word(uint32_t cur, uint32_t next, uint32_t src)
{
    const uint32_t y = (cur & kMtUpper) | (next & kMtLower);
    return src ^ (y >> 1) ^ ((y & 1u) ? kMtMatrix : 0u);
}

__device__ __forceinline__ uint32_t mt_temper(uint32_t y)
{
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

// std::mt19937::seed(v) by one thread (the recurrence is serial).
__device__ __forceinline__ void mt_seed_serial(uint32_t *mt, uint32_t v)
{
    mt[0] = v;
    for (int i = 1; i < kMtN; ++i) {
        v = 1812433253u * (v ^ (v >> 30)) + (uint32_t)i;
        mt[i] = v;
    }
}

// One phase: words [lo, hi) from (old[i], old[i+1], mt[i + off]), all read
// before any is written.
__device__ __forceinline__ void mt_twist_phase(uint32_t *mt, int lo, int hi, int off)
{
    for (int i0 = lo; i0 < hi; i0 += blockDim.x) {
        const int i = i0 + threadIdx.x;
        uint32_t v = 0;
        if (i < hi) v = mt_twist_word(mt[i], mt[i + 1], mt[i + off]);
        __syncthreads();
        if (i < hi) mt[i] = v;
        __syncthreads();
    }
}

// Twist of the LDS state by the whole block (any block size; every thread
// must call it).
__device__ __forceinline__ void mt_twist_block(uint32_t *mt)
{
    mt_twist_phase(mt, 0, 227, kMtM);
    mt_twist_phase(mt, 227, 454, -227);
    mt_twist_phase(mt, 454, 623, -227);
    if (threadIdx.x == 0) mt[623] = mt_twist_word(mt[623], mt[0], mt[396]);
    __syncthreads();
}

// generate_canonical<double, 53> from two tempered words (w1 first):
// (w1 + w2 * 2^32) / 2^64; a value rounding to 1 becomes the largest double
// below 1, as libstdc++ does.
__device__ __forceinline__ double canonical53(uint32_t w1, uint32_t w2)
{
    double sum = 0.0;
    sum += (double)w1 * 1.0;
    sum += (double)w2 * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    if (r >= 1.0) r = 0x1.fffffffffffffp-1; // nextafter(1, 0)
    return r;
}

} // namespace ompl_amd
