// Batched uniform state sampling: RealVectorStateSampler::sampleUniform
// (base/spaces/src/RealVectorStateSpace.cpp:45-52) over ompl::RNG::uniformReal
// (util/RandomNumbers.h:72-76): std::mt19937 seeded with the RNG's local seed
// (util/src/RandomNumbers.cpp:224-228), uniform_real_distribution<>(0,1) as
// libstdc++'s generate_canonical<double,53> (two 32-bit draws w1, w2:
// (w1 + w2*2^32) / 2^64, values rounding to 1 replaced by the largest double
// below 1), then (hi - lo) * u + lo per dimension in order.
//
// One wave per generator; the 624-word state lives in LDS and is twisted by
// the block in the recurrence's four dependency phases (mt_device.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "mt_device.hpp"
#include "runtime.hpp"

namespace ompl_amd {

constexpr int kDim7 = OMPL_AMD_ARM_DOF;

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
        mt_twist_block(mt);
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

// ---------------------------------------------------------------------------
// UniformValidStateSampler::sample (base/samplers/src/UniformValidStateSampler.cpp:
// 46-57), batched over streams: each stream is one sampler making successive
// sample() calls; a call draws uniform states until one is valid or `attempts`
// draws were made.  The draws of a stream depend on the validity of earlier
// ones, so the batch runs in rounds: every unfinished stream draws C more
// states (its generator resumes where it stopped), all are checked in one
// batched validity launch, and one thread per stream hands them out to its
// calls in order.

struct MtState {
    uint32_t mt[kMtN];
    int32_t idx; // next word; kMtN = twist before the next draw
    int32_t pad[3];
};

struct ValidCursor {
    int64_t call;     // next call of this stream
    int64_t used;     // draws consumed from the generator by the calls so far
    int32_t attempts; // draws made by the current call
    int32_t pad;
};

__global__ void mt_seed_kernel(const uint32_t *__restrict__ seeds, int n, MtState *__restrict__ st)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t v = seeds[s];
    st[s].mt[0] = v;
    for (int i = 1; i < kMtN; ++i) {
        v = 1812433253u * (v ^ (v >> 30)) + (uint32_t)i;
        st[s].mt[i] = v;
    }
    st[s].idx = kMtN;
}

// One wave per listed stream: resumes the generator and writes C states of
// `dim` doubles ([a][C][dim]) and, when out_f32 is given, their FP32 copy
// ([dim][n_active * C]); `active` lists the streams (all of them, or the
// unfinished ones of a valid-sampling round).
__global__ void __launch_bounds__(64) resume_uniform_kernel(MtState *__restrict__ st, const int32_t *__restrict__ active,
                                                            int n_active, int64_t C, int dim,
                                                            const double *__restrict__ lo_g,
                                                            const double *__restrict__ hi_g, double *__restrict__ out,
                                                            float *__restrict__ out_f32)
{
    __shared__ uint32_t mt[kMtN];
    __shared__ double lo[16], span[16];
    const int a = blockIdx.x;
    const int t = threadIdx.x;
    MtState &S = st[active ? active[a] : a];
    for (int i = t; i < kMtN; i += 64) mt[i] = S.mt[i];
    if (t < dim) {
        lo[t] = lo_g[t];
        span[t] = hi_g[t] - lo_g[t];
    }
    int idx = S.idx;
    __syncthreads();
    const int64_t total = C * dim;
    const int64_t n_f32 = (int64_t)n_active * C;
    double *o = out + (int64_t)a * total;
    for (int64_t z = 0; z < total;) {
        if (idx >= kMtN) {
            mt_twist_block(mt);
            idx = 0;
        }
        const int64_t cnt = min((int64_t)((kMtN - idx) / 2), total - z);
        for (int p = t; p < cnt; p += 64) {
            const uint32_t w1 = mt_temper(mt[idx + 2 * p]);
            const uint32_t w2 = mt_temper(mt[idx + 2 * p + 1]);
            const int64_t zz = z + p;
            const int d = (int)(zz % dim);
            const double v = span[d] * canonical53(w1, w2) + lo[d];
            o[zz] = v;
            if (out_f32) out_f32[(int64_t)d * n_f32 + (int64_t)a * C + zz / dim] = (float)v;
        }
        __syncthreads();
        idx += 2 * (int)cnt;
        z += cnt;
    }
    for (int i = t; i < kMtN; i += 64) S.mt[i] = mt[i];
    if (t == 0) S.idx = idx;
}

// One wave per stream: st = saved advanced by the cursor's consumed draws
// (2 words per coordinate), i.e. the generator exactly after the last draw
// the calls used; the draws of the last round past that point are dropped.
__global__ void __launch_bounds__(64) mt_advance_kernel(MtState *__restrict__ st, const MtState *__restrict__ saved,
                                                        const ValidCursor *__restrict__ cur, int dim)
{
    __shared__ uint32_t mt[kMtN];
    const int s = blockIdx.x;
    const int t = threadIdx.x;
    for (int i = t; i < kMtN; i += 64) mt[i] = saved[s].mt[i];
    int idx = saved[s].idx;
    __syncthreads();
    int64_t skip = cur[s].used * 2 * dim;
    while (skip > 0) {
        if (idx >= kMtN) {
            mt_twist_block(mt);
            idx = 0;
        }
        const int take = (int)min((int64_t)(kMtN - idx), skip);
        idx += take;
        skip -= take;
    }
    for (int i = t; i < kMtN; i += 64) st[s].mt[i] = mt[i];
    if (t == 0) st[s].idx = idx;
}

// One thread per active stream: hands the round's draws to its calls in order.
__global__ void assign_valid_kernel(ValidCursor *__restrict__ cur, const int32_t *__restrict__ active, int n_active,
                                    int64_t C, const uint8_t *__restrict__ valid, const double *__restrict__ cand,
                                    int64_t n_calls, int attempts, double *__restrict__ out,
                                    uint8_t *__restrict__ found, int32_t *__restrict__ draws,
                                    unsigned long long *__restrict__ stats)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_active) return;
    const int s = active[a];
    ValidCursor c = cur[s];
    int64_t used = 0, used_valid = 0;
    for (int64_t j = 0; j < C && c.call < n_calls; ++j) {
        const bool v = valid[(int64_t)a * C + j] != 0;
        ++c.attempts;
        ++used;
        used_valid += v;
        if (v || c.attempts >= attempts) {
            const int64_t o = (int64_t)s * n_calls + c.call;
#pragma unroll
            for (int d = 0; d < kDim7; ++d) out[o * kDim7 + d] = cand[((int64_t)a * C + j) * kDim7 + d];
            found[o] = v ? 1 : 0;
            if (draws) draws[o] = c.attempts;
            ++c.call;
            c.attempts = 0;
        }
    }
    c.used += used;
    cur[s] = c;
    atomicAdd(&stats[0], (unsigned long long)used);
    atomicAdd(&stats[1], (unsigned long long)used_valid);
    if (c.call < n_calls) atomicMax(&stats[2], (unsigned long long)(n_calls - c.call));
}

__global__ void next_active_kernel(const ValidCursor *__restrict__ cur, int n_streams, int64_t n_calls,
                                   int32_t *__restrict__ active, unsigned long long *__restrict__ stats)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool open = s < n_streams && cur[s].call < n_calls;
    const uint64_t m = __ballot(open);
    int base = 0;
    const int lane = threadIdx.x & 63;
    if (m) {
        if (lane == __ffsll((unsigned long long)m) - 1) base = (int)atomicAdd(&stats[3], (unsigned long long)__popcll(m));
        base = __shfl(base, __ffsll((unsigned long long)m) - 1);
    }
    if (open) active[base + __popcll(m & (((uint64_t)1 << lane) - 1))] = s;
}

// SE3StateSampler::sampleUniform (a compound sampler: the R^3 component's
// RealVectorStateSampler, then SO3StateSampler::sampleUniform = RNG::quaternion,
// util/src/RandomNumbers.cpp:266-279, Shoemake) from six canonical draws per
// state u[0..5]: x_k = (hi_k - lo_k) u_k + lo_k; x0 = u3, t1 = 2 pi u4,
// t2 = 2 pi u5, q = (sin t1 sqrt(1-x0), cos t1 sqrt(1-x0), sin t2 sqrt(x0),
// cos t2 sqrt(x0)) in FP64 (the device's sin/cos are within an ulp or two of
// the host's).  Output rows (x, y, z, qx, qy, qz, qw) as FP64 and FP32.
__global__ void se3_from_uniform_kernel(const double *__restrict__ u, int64_t n, const double *__restrict__ lo,
                                        const double *__restrict__ hi, double *__restrict__ out,
                                        float *__restrict__ out_f32)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *v = u + i * 6;
    double r[7];
#pragma unroll
    for (int k = 0; k < 3; ++k) r[k] = (hi[k] - lo[k]) * v[k] + lo[k];
    const double pi = 3.141592653589793238462643383279502884;
    const double x0 = v[3];
    const double r1 = sqrt(1.0 - x0), r2 = sqrt(x0);
    const double t1 = 2.0 * pi * v[4], t2 = 2.0 * pi * v[5];
    r[3] = sin(t1) * r1;
    r[4] = cos(t1) * r1;
    r[5] = sin(t2) * r2;
    r[6] = cos(t2) * r2;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        out[i * 7 + k] = r[k];
        if (out_f32) out_f32[i * 7 + k] = (float)r[k];
    }
}

// Device bounds of a sampling call (host arrays staged per call).
struct Bounds {
    double *lo, *hi;
};

static Bounds stage_bounds(const double *lo, const double *hi, int dim, hipStream_t s)
{
    double *d = nullptr;
    OA_HIP(hipMallocAsync((void **)&d, 32 * sizeof(double), s));
    OA_HIP(hipMemcpyAsync(d, lo, dim * sizeof(double), hipMemcpyHostToDevice, s));
    OA_HIP(hipMemcpyAsync(d + 16, hi, dim * sizeof(double), hipMemcpyHostToDevice, s));
    return Bounds{d, d + 16};
}

// UniformValidStateSampler::sample x n_calls on each of n_streams generators
// (gen, device), continuing them; on return every generator stands exactly
// after the last draw its calls consumed.  Blocking.
static void sample_valid_rounds(ompl_amd_env env, MtState *gen, int n_streams, int64_t n_calls, int attempts,
                                const double *lo, const double *hi, double *out_dev, uint8_t *found_dev,
                                int32_t *draws_dev, hipStream_t s)
{
    const size_t head = 256;
    const size_t bytes = head + (size_t)n_streams * (sizeof(MtState) + sizeof(ValidCursor) + sizeof(int32_t));
    char *base = nullptr;
    OA_HIP(hipMallocAsync((void **)&base, bytes, s));
    unsigned long long *stats = (unsigned long long *)base; // [0] used [1] valid [2] max remaining [3] active
    MtState *saved = (MtState *)(base + head);
    ValidCursor *cur = (ValidCursor *)(saved + n_streams);
    int32_t *active = (int32_t *)(cur + n_streams);
    const Bounds bd = stage_bounds(lo, hi, kDim7, s);
    OA_HIP(hipMemcpyAsync(saved, gen, (size_t)n_streams * sizeof(MtState), hipMemcpyDeviceToDevice, s));
    OA_HIP(hipMemsetAsync(cur, 0, (size_t)n_streams * sizeof(ValidCursor), s));
    std::vector<int32_t> all((size_t)n_streams);
    for (int i = 0; i < n_streams; ++i) all[(size_t)i] = i;
    OA_HIP(hipMemcpyAsync(active, all.data(), all.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    int n_active = n_streams;
    int64_t remaining = n_calls;
    double draws_per_call = 1.0; // refined from every round
    double *cand = nullptr;
    float *cand32 = nullptr;
    uint8_t *valid = nullptr;
    size_t cap = 0;
    const int64_t max_round = ((int64_t)64 << 20) / ((int64_t)n_streams * kDim7 * 8) + 16; // ~64 MiB of draws
    int64_t rounds = 0;
    const int64_t round_limit = n_calls * (int64_t)attempts + 2;
    while (n_active > 0) {
        if (++rounds > round_limit) fail(OMPL_AMD_ERR_INTERNAL, "valid sampler did not finish");
        const double want = (double)remaining * std::min((double)attempts, draws_per_call) * 1.25 + 16.0;
        const int64_t C = std::max<int64_t>(
            16, std::min<int64_t>((int64_t)want, std::min<int64_t>(max_round, remaining * (int64_t)attempts)));
        const size_t need = (size_t)n_active * (size_t)C;
        if (need > cap) {
            if (cand) {
                OA_HIP(hipFreeAsync(cand, s));
                OA_HIP(hipFreeAsync(cand32, s));
                OA_HIP(hipFreeAsync(valid, s));
            }
            cap = need;
            OA_HIP(hipMallocAsync((void **)&cand, cap * kDim7 * sizeof(double), s));
            OA_HIP(hipMallocAsync((void **)&cand32, cap * kDim7 * sizeof(float), s));
            OA_HIP(hipMallocAsync((void **)&valid, cap, s));
        }
        {
            ProfScope ps("sample_valid", s);
            hipLaunchKernelGGL(resume_uniform_kernel, dim3(n_active), dim3(64), 0, s, gen, active, n_active, C, kDim7,
                               bd.lo, bd.hi, cand, cand32);
            OA_HIP(hipGetLastError());
        }
        const int rc = ompl_amd_validate_states_dev(env, cand32, (int64_t)need, valid, s);
        if (rc != OMPL_AMD_OK) fail(rc, "%s", last_error());
        OA_HIP(hipMemsetAsync(stats, 0, 4 * sizeof(unsigned long long), s));
        hipLaunchKernelGGL(assign_valid_kernel, dim3((n_active + 63) / 64), dim3(64), 0, s, cur, active, n_active, C,
                           valid, cand, n_calls, attempts, out_dev, found_dev, draws_dev, stats);
        hipLaunchKernelGGL(next_active_kernel, dim3((n_streams + 255) / 256), dim3(256), 0, s, cur, n_streams,
                           n_calls, active, stats);
        OA_HIP(hipGetLastError());
        unsigned long long h[4];
        OA_HIP(hipMemcpyAsync(h, stats, sizeof(h), hipMemcpyDeviceToHost, s));
        OA_HIP(hipStreamSynchronize(s));
        n_active = (int)h[3];
        remaining = (int64_t)h[2];
        if (h[1] > 0)
            draws_per_call = std::max(1.0, (double)h[0] / (double)h[1]);
        else
            draws_per_call = (double)attempts;
    }
    // drop the draws of the last round that no call used
    hipLaunchKernelGGL(mt_advance_kernel, dim3(n_streams), dim3(64), 0, s, gen, saved, cur, kDim7);
    OA_HIP(hipGetLastError());
    if (cand) {
        OA_HIP(hipFreeAsync(cand, s));
        OA_HIP(hipFreeAsync(cand32, s));
        OA_HIP(hipFreeAsync(valid, s));
    }
    OA_HIP(hipFreeAsync(bd.lo, s));
    OA_HIP(hipFreeAsync(base, s));
    OA_HIP(hipStreamSynchronize(s));
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

// Generators kept on the device between calls: the exact continuation of a
// long-lived sampler's RNG.
struct ompl_amd_rng_s {
    MtState *gen = nullptr;
    int n = 0;
    ~ompl_amd_rng_s()
    {
        if (gen) hipFree(gen);
    }
};

static MtState *seed_generators(const uint32_t *seeds, int n, hipStream_t s)
{
    MtState *gen = nullptr;
    uint32_t *dseeds = nullptr;
    OA_HIP(hipMallocAsync((void **)&gen, (size_t)n * sizeof(MtState), s));
    OA_HIP(hipMallocAsync((void **)&dseeds, (size_t)n * sizeof(uint32_t), s));
    OA_HIP(hipMemcpyAsync(dseeds, seeds, (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(mt_seed_kernel, dim3((n + 63) / 64), dim3(64), 0, s, dseeds, n, gen);
    OA_HIP(hipGetLastError());
    OA_HIP(hipFreeAsync(dseeds, s));
    OA_HIP(hipStreamSynchronize(s)); // the seeds are a host array
    return gen;
}

extern "C" {

int ompl_amd_sample_valid_dev(ompl_amd_env env, const uint32_t *seeds, int n_streams, int64_t n_calls, int attempts,
                              const double *lo, const double *hi, double *out_dev, uint8_t *found_dev,
                              int32_t *draws_dev, void *stream)
{
    return guarded([&] {
        OA_REQUIRE(env, OMPL_AMD_ERR_ARG, "null environment");
        OA_REQUIRE(n_streams >= 0 && n_calls >= 0 && attempts >= 1, OMPL_AMD_ERR_ARG, "bad sizes");
        if (n_streams == 0 || n_calls == 0) return;
        OA_REQUIRE(seeds && lo && hi && out_dev && found_dev, OMPL_AMD_ERR_ARG, "null pointer");
        hipStream_t s = resolve_stream(stream);
        MtState *gen = seed_generators(seeds, n_streams, s);
        try {
            sample_valid_rounds(env, gen, n_streams, n_calls, attempts, lo, hi, out_dev, found_dev, draws_dev, s);
        } catch (...) {
            hipStreamSynchronize(s);
            hipFree(gen);
            throw;
        }
        OA_HIP(hipFreeAsync(gen, s));
        OA_HIP(hipStreamSynchronize(s));
    });
}

int ompl_amd_rng_create(const uint32_t *seeds, int n_streams, ompl_amd_rng *out)
{
    return guarded([&] {
        OA_REQUIRE(out && seeds && n_streams >= 1, OMPL_AMD_ERR_ARG, "bad arguments");
        ensure_init();
        auto *r = new ompl_amd_rng_s();
        try {
            r->gen = seed_generators(seeds, n_streams, default_stream());
        } catch (...) {
            delete r;
            throw;
        }
        r->n = n_streams;
        *out = r;
    });
}

int ompl_amd_rng_destroy(ompl_amd_rng rng)
{
    return guarded([&] {
        if (rng) OA_HIP(hipDeviceSynchronize());
        delete rng;
    });
}

int ompl_amd_rng_sample_uniform_dev(ompl_amd_rng rng, int64_t n_per_stream, int dim, const double *lo,
                                    const double *hi, double *out_dev, float *out_f32_soa_dev, void *stream)
{
    return guarded([&] {
        OA_REQUIRE(rng, OMPL_AMD_ERR_ARG, "null generator");
        OA_REQUIRE(n_per_stream >= 0 && dim >= 1 && dim <= 16, OMPL_AMD_ERR_ARG, "bad sizes");
        if (n_per_stream == 0) return;
        OA_REQUIRE(lo && hi && out_dev, OMPL_AMD_ERR_ARG, "null pointer");
        hipStream_t s = resolve_stream(stream);
        const Bounds bd = stage_bounds(lo, hi, dim, s);
        {
            ProfScope ps("sample_uniform", s);
            hipLaunchKernelGGL(resume_uniform_kernel, dim3(rng->n), dim3(64), 0, s, rng->gen, (const int32_t *)nullptr,
                               rng->n, n_per_stream, dim, bd.lo, bd.hi, out_dev, out_f32_soa_dev);
            OA_HIP(hipGetLastError());
        }
        OA_HIP(hipFreeAsync(bd.lo, s));
        OA_HIP(hipStreamSynchronize(s)); // host bounds were staged with async copies
    });
}

int ompl_amd_rng_sample_se3_dev(ompl_amd_rng rng, int64_t n_per_stream, const double *lo, const double *hi,
                                double *out_dev, float *out_f32_dev, void *stream)
{
    return guarded([&] {
        OA_REQUIRE(rng, OMPL_AMD_ERR_ARG, "null generator");
        OA_REQUIRE(n_per_stream >= 0, OMPL_AMD_ERR_ARG, "bad sizes");
        if (n_per_stream == 0) return;
        OA_REQUIRE(lo && hi && out_dev, OMPL_AMD_ERR_ARG, "null pointer");
        hipStream_t s = resolve_stream(stream);
        const double zero[6] = {0, 0, 0, 0, 0, 0}, one[6] = {1, 1, 1, 1, 1, 1};
        const Bounds unit = stage_bounds(zero, one, 6, s);
        const Bounds box = stage_bounds(lo, hi, 3, s);
        const int64_t n = (int64_t)rng->n * n_per_stream;
        double *u = nullptr;
        OA_HIP(hipMallocAsync((void **)&u, (size_t)n * 6 * sizeof(double), s));
        {
            ProfScope ps("sample_se3", s);
            hipLaunchKernelGGL(resume_uniform_kernel, dim3(rng->n), dim3(64), 0, s, rng->gen, (const int32_t *)nullptr,
                               rng->n, n_per_stream, 6, unit.lo, unit.hi, u, (float *)nullptr);
            hipLaunchKernelGGL(se3_from_uniform_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, u, n,
                               box.lo, box.hi, out_dev, out_f32_dev);
            OA_HIP(hipGetLastError());
        }
        OA_HIP(hipFreeAsync(u, s));
        OA_HIP(hipFreeAsync(unit.lo, s));
        OA_HIP(hipFreeAsync(box.lo, s));
        OA_HIP(hipStreamSynchronize(s)); // host bounds were staged with async copies
    });
}

int ompl_amd_rng_sample_valid_dev(ompl_amd_rng rng, ompl_amd_env env, int64_t n_calls, int attempts,
                                  const double *lo, const double *hi, double *out_dev, uint8_t *found_dev,
                                  int32_t *draws_dev, void *stream)
{
    return guarded([&] {
        OA_REQUIRE(rng && env, OMPL_AMD_ERR_ARG, "null generator or environment");
        OA_REQUIRE(n_calls >= 0 && attempts >= 1, OMPL_AMD_ERR_ARG, "bad sizes");
        if (n_calls == 0) return;
        OA_REQUIRE(lo && hi && out_dev && found_dev, OMPL_AMD_ERR_ARG, "null pointer");
        sample_valid_rounds(env, rng->gen, rng->n, n_calls, attempts, lo, hi, out_dev, found_dev, draws_dev,
                            resolve_stream(stream));
    });
}

} // extern "C"
