// Batched uniform state sampling: RealVectorStateSampler::sampleUniform
// (base/spaces/src/RealVectorStateSpace.cpp:45-52) over ompl::RNG::uniformReal
// (util/RandomNumbers.h:72-76): std::mt19937 seeded with the RNG's local seed
// (util/src/RandomNumbers.cpp:224-228), uniform_real_distribution<>(0,1) as
// libstdc++'s generate_canonical<double,53> (two 32-bit draws w1, w2:
// (w1 + w2*2^32) / 2^64, values rounding to 1 replaced by the largest double
// below 1), then (hi - lo) * u + lo per dimension in order.
//
// One wave per generator.  The 624-word state lives in LDS; a twist is done in
// the four dependency phases of the in-place recurrence
//   new[i] = src[i+397] ^ twist(old[i], old[i+1])   (indices mod 624)
// phase A: i in [0,227)   src = old,
// phase B: i in [227,454) src = new (phase A),
// phase C: i in [454,623) src = new (phase B),
// phase D: i = 623        old[624 mod 624] = new[0], src = new[396],
// each phase reading its inputs before any lane writes.
#include <hip/hip_runtime.h>

#include "runtime.hpp"

namespace ompl_amd {

constexpr int kMtN = 624;
constexpr int kMtM = 397;
constexpr uint32_t kMtMatrix = 0x9908b0dfu;
constexpr uint32_t kUpper = 0x80000000u, kLower = 0x7fffffffu;

__device__ __forceinline__ uint32_t mt_twist_word(uint32_t cur, uint32_t next, uint32_t src)
{
    const uint32_t y = (cur & kUpper) | (next & kLower);
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

__device__ void mt_twist(uint32_t *mt)
{
    const int t = threadIdx.x;
    // phase A
    for (int i0 = 0; i0 < 227; i0 += 64) {
        const int i = i0 + t;
        uint32_t v = 0;
        if (i < 227) v = mt_twist_word(mt[i], mt[i + 1], mt[i + kMtM]);
        __syncthreads();
        if (i < 227) mt[i] = v;
        __syncthreads();
    }
    // phase B
    for (int i0 = 227; i0 < 454; i0 += 64) {
        const int i = i0 + t;
        uint32_t v = 0;
        if (i < 454) v = mt_twist_word(mt[i], mt[i + 1], mt[i - 227]);
        __syncthreads();
        if (i < 454) mt[i] = v;
        __syncthreads();
    }
    // phase C
    for (int i0 = 454; i0 < 623; i0 += 64) {
        const int i = i0 + t;
        uint32_t v = 0;
        if (i < 623) v = mt_twist_word(mt[i], mt[i + 1], mt[i - 227]);
        __syncthreads();
        if (i < 623) mt[i] = v;
        __syncthreads();
    }
    // phase D
    if (t == 0) mt[623] = mt_twist_word(mt[623], mt[0], mt[396]);
    __syncthreads();
}

__device__ __forceinline__ double canonical53(uint32_t w1, uint32_t w2)
{
    double sum = 0.0;
    sum += (double)w1 * 1.0;
    sum += (double)w2 * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    if (r >= 1.0) r = 0x1.fffffffffffffp-1; // nextafter(1, 0)
    return r;
}

struct SampleArgs {
    const uint32_t *seeds;
    int n_streams;
    int64_t n_per_stream;
    int dim;
    const double *lo;
    const double *hi;
    double *out;       // [stream][n_per_stream][dim]
    float *out_f32;    // [dim][n_streams * n_per_stream] or null
};

__global__ void __launch_bounds__(64) sample_uniform_kernel(SampleArgs A)
{
    __shared__ uint32_t mt[kMtN];
    __shared__ double lo[16], span[16];
    const int s = blockIdx.x;
    const int t = threadIdx.x;
    if (t < A.dim) {
        lo[t] = A.lo[t];
        span[t] = A.hi[t] - A.lo[t];
    }
    if (t == 0) {
        uint32_t v = A.seeds[s];
        mt[0] = v;
        for (int i = 1; i < kMtN; ++i) {
            v = 1812433253u * (v ^ (v >> 30)) + (uint32_t)i;
            mt[i] = v;
        }
    }
    __syncthreads();
    const int64_t total_pairs = A.n_per_stream * A.dim;
    const int64_t total_states = (int64_t)A.n_streams * A.n_per_stream;
    double *out = A.out + (int64_t)s * total_pairs;
    for (int64_t base = 0; base < total_pairs; base += kMtN / 2) {
        mt_twist(mt);
        for (int p = t; p < kMtN / 2; p += 64) {
            const int64_t z = base + p;
            if (z >= total_pairs) break;
            const uint32_t w1 = mt_temper(mt[2 * p]);
            const uint32_t w2 = mt_temper(mt[2 * p + 1]);
            const int d = (int)(z % A.dim);
            const double v = span[d] * canonical53(w1, w2) + lo[d];
            out[z] = v;
            if (A.out_f32) {
                const int64_t st = (int64_t)s * A.n_per_stream + z / A.dim;
                A.out_f32[(int64_t)d * total_states + st] = (float)v;
            }
        }
        __syncthreads();
    }
}

} // namespace ompl_amd

using namespace ompl_amd;

extern "C" int ompl_amd_sample_uniform_dev(const uint32_t *seeds, int n_streams, int64_t n_per_stream, int dim,
                                           const double *lo, const double *hi, double *out_dev, float *out_f32,
                                           void *stream)
{
    return guarded([&] {
        OA_REQUIRE(n_streams >= 0 && n_per_stream >= 0, OMPL_AMD_ERR_ARG, "negative sizes");
        OA_REQUIRE(dim >= 1 && dim <= 16, OMPL_AMD_ERR_ARG, "dim must be in [1, 16]");
        if (n_streams == 0 || n_per_stream == 0) return;
        OA_REQUIRE(seeds && lo && hi && out_dev, OMPL_AMD_ERR_ARG, "null pointer");
        hipStream_t s = resolve_stream(stream);
        // bounds and seeds are small host arrays: stage them on the device
        char *stage = nullptr;
        const size_t bytes = (size_t)n_streams * sizeof(uint32_t) + 2 * 16 * sizeof(double);
        OA_HIP(hipMallocAsync((void **)&stage, bytes, s));
        double *dlo = (double *)stage, *dhi = dlo + 16;
        uint32_t *dseeds = (uint32_t *)(dhi + 16);
        OA_HIP(hipMemcpyAsync(dlo, lo, dim * sizeof(double), hipMemcpyHostToDevice, s));
        OA_HIP(hipMemcpyAsync(dhi, hi, dim * sizeof(double), hipMemcpyHostToDevice, s));
        OA_HIP(hipMemcpyAsync(dseeds, seeds, n_streams * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        SampleArgs A{dseeds, n_streams, n_per_stream, dim, dlo, dhi, out_dev, out_f32};
        {
            ProfScope ps("sample_uniform", s);
            hipLaunchKernelGGL(sample_uniform_kernel, dim3(n_streams), dim3(64), 0, s, A);
            OA_HIP(hipGetLastError());
        }
        OA_HIP(hipFreeAsync(stage, s));
        OA_HIP(hipStreamSynchronize(s)); // host arrays were staged with async copies
    });
}
