// Batched PRM roadmap construction on the device: the inner loop of
// PRM::growRoadmap / PRM::addMilestone (geometric/planners/prm/src/PRM.cpp:
// 364-385, 570-604) with the KStrategy connection rule
// (geometric/planners/prm/ConnectionStrategy.h:56-81), run for a whole batch:
//
//   1. uniform samples (RealVectorStateSampler::sampleUniform per stream)
//   2. state validity of every sample; valid samples become milestones
//   3. k nearest milestones of every milestone (exact FP32 L2, ties by index)
//   4. candidate edges = the undirected union of those k-NN pairs
//   5. every candidate motion checked (DiscreteMotionValidator semantics,
//      direction: lower milestone index -> higher)
//   6. connected components of the valid edges (union-find: root hooking by
//      atomicCAS + pointer jumping)
//
// The reference connects each new milestone to its k nearest *existing*
// milestones as they are added (incremental); this batch form connects every
// milestone to its k nearest among all milestones, so its roadmap does not
// depend on the order of the milestones.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <queue>
#include <vector>

#include "runtime.hpp"

namespace ompl_amd {

constexpr int kDim = OMPL_AMD_ARM_DOF;

struct PrmRoadmap {
    int64_t samples = 0, nv = 0, ne = 0, valid_edges = 0, components = 0;
    double *samples_f64 = nullptr; // [n][7]
    float *samples_f32 = nullptr;  // [7][n]
    uint8_t *valid = nullptr;      // [n]
    int32_t *milestone = nullptr;  // [nv] sample index of milestone i
    double *nodes = nullptr;       // [nv][7]
    float *nodes_f32 = nullptr;    // [nv][7]
    int32_t *knn = nullptr;        // [nv][k+1]
    float *knn_d = nullptr;
    uint64_t *keys = nullptr;      // candidate (u << 32 | v), u < v
    int32_t *eu = nullptr, *ev = nullptr;
    uint8_t *evalid = nullptr;
    int32_t *label = nullptr;      // component root per milestone
    int64_t *csr_off = nullptr;    // [nv + 1] symmetric adjacency of the valid edges
    int32_t *csr_adj = nullptr;    // [2 * valid_edges], ascending per row
    ompl_amd_prm_stats stats{};
    std::vector<void *> allocs;

    template <class T>
    T *alloc(size_t count)
    {
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, (count ? count : 1) * sizeof(T));
        if (e != hipSuccess) fail(OMPL_AMD_ERR_NOMEM, "roadmap allocation (%zu bytes) failed", count * sizeof(T));
        allocs.push_back(p);
        return (T *)p;
    }
    ~PrmRoadmap()
    {
        for (void *p : allocs) (void)hipFree(p);
    }
};

__global__ void gather_nodes_kernel(const int32_t *__restrict__ milestone, int64_t nv,
                                    const double *__restrict__ samples, double *__restrict__ nodes,
                                    float *__restrict__ nodes_f32)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const int64_t s = milestone[i];
#pragma unroll
    for (int k = 0; k < kDim; ++k) {
        const double v = samples[s * kDim + k];
        nodes[i * kDim + k] = v;
        nodes_f32[i * kDim + k] = (float)v;
    }
}

// Candidate keys: the first k neighbours of milestone i other than i itself.
__global__ void edge_keys_kernel(const int32_t *__restrict__ knn, int64_t nv, int kk, int k,
                                 uint64_t *__restrict__ keys)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    int taken = 0;
    for (int j = 0; j < kk; ++j) {
        const int nb = knn[i * kk + j];
        if (taken < k && nb >= 0 && nb != (int)i) {
            const uint64_t u = (uint64_t)min((int64_t)nb, i), v = (uint64_t)max((int64_t)nb, i);
            keys[i * k + taken] = (u << 32) | v;
            ++taken;
        }
    }
    for (int t = taken; t < k; ++t) keys[i * k + t] = ~(uint64_t)0;
}

__global__ void edge_endpoints_kernel(const uint64_t *__restrict__ keys, int64_t ne, const double *__restrict__ nodes,
                                      int32_t *__restrict__ eu, int32_t *__restrict__ ev, double *__restrict__ a_soa,
                                      double *__restrict__ b_soa)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int32_t u = (int32_t)(keys[e] >> 32), v = (int32_t)(keys[e] & 0xffffffffu);
    eu[e] = u;
    ev[e] = v;
#pragma unroll
    for (int k = 0; k < kDim; ++k) {
        a_soa[(int64_t)k * ne + e] = nodes[(int64_t)u * kDim + k];
        b_soa[(int64_t)k * ne + e] = nodes[(int64_t)v * kDim + k];
    }
}

__global__ void iota_kernel(int32_t *__restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

__device__ __forceinline__ int32_t find_root(const int32_t *parent, int32_t x)
{
    int32_t p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (p != x) {
        x = p;
        p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return x;
}

// Hooks the higher root under the lower one for every valid edge whose ends
// are in different trees; sets *changed when any hook succeeded.
__global__ void hook_kernel(const int32_t *__restrict__ eu, const int32_t *__restrict__ ev,
                            const uint8_t *__restrict__ evalid, int64_t ne, int32_t *parent, int32_t *changed)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne || !evalid[e]) return;
    const int32_t ru = find_root(parent, eu[e]), rv = find_root(parent, ev[e]);
    if (ru == rv) return;
    const int32_t hi = max(ru, rv), lo = min(ru, rv);
    if (atomicCAS(&parent[hi], hi, lo) == hi) *changed = 1;
}

__global__ void jump_kernel(int32_t *parent, int64_t nv)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    parent[i] = find_root(parent, (int32_t)i);
}

// Both directions of every valid edge as (row << 32 | col) keys, at the
// positions given by the exclusive scan of evalid.
__global__ void csr_keys_kernel(const int32_t *__restrict__ eu, const int32_t *__restrict__ ev,
                                const uint8_t *__restrict__ evalid, const int32_t *__restrict__ pos, int64_t ne,
                                uint64_t *__restrict__ keys, int32_t *__restrict__ degree)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne || !evalid[e]) return;
    const uint64_t u = (uint32_t)eu[e], v = (uint32_t)ev[e];
    keys[2 * (int64_t)pos[e]] = (u << 32) | v;
    keys[2 * (int64_t)pos[e] + 1] = (v << 32) | u;
    atomicAdd(&degree[u], 1);
    atomicAdd(&degree[v], 1);
}

__global__ void csr_cols_kernel(const uint64_t *__restrict__ keys, int64_t n, int32_t *__restrict__ adj)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) adj[i] = (int32_t)(keys[i] & 0xffffffffu);
}

__global__ void widen_kernel(const int32_t *__restrict__ in, int64_t n, int64_t *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

__global__ void u8_to_i32_kernel(const uint8_t *__restrict__ in, int64_t n, int32_t *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

__global__ void count_roots_kernel(const int32_t *__restrict__ parent, int64_t nv, unsigned long long *count)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool root = i < nv && parent[i] == (int32_t)i;
    const uint64_t m = __ballot(root);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (unsigned long long)__popcll(m));
}

static unsigned blocks_for(int64_t n, int threads = 256)
{
    return (unsigned)((n + threads - 1) / threads);
}

// Builds the roadmap on stream s.  extra: n_extra states (host, row-major)
// placed before the samples, so that a valid one becomes a milestone with the
// lowest indices (a planning query's start and goal).
void build(PrmRoadmap &R, ompl_amd_env env, uint32_t seed, int n_streams, int64_t n_samples, int k, double lvs,
           unsigned factor, unsigned flags, hipStream_t s, const double *extra, int n_extra)
{
    using Clock = std::chrono::steady_clock;
    auto ms = [](Clock::time_point a) {
        return std::chrono::duration<double, std::milli>(Clock::now() - a).count();
    };
    const auto t_all = Clock::now();
    auto check = [](int rc) {
        if (rc != OMPL_AMD_OK) fail(rc, "%s", last_error());
    };
    // 1. samples
    auto t0 = Clock::now();
    const int64_t per = (n_samples + n_streams - 1) / n_streams;
    const int64_t ns = per * n_streams;
    const int64_t n = ns + n_extra;
    R.samples = n;
    R.samples_f64 = R.alloc<double>((size_t)n * kDim);
    R.samples_f32 = R.alloc<float>((size_t)n * kDim);
    std::vector<uint32_t> seeds((size_t)n_streams);
    for (int i = 0; i < n_streams; ++i) seeds[(size_t)i] = seed + (uint32_t)i;
    const double pi = 3.141592653589793238462643383279502884;
    const double lo[kDim] = {-pi, -pi, -pi, -pi, -pi, -pi, -pi}, hi[kDim] = {pi, pi, pi, pi, pi, pi, pi};
    if (n_extra == 0) {
        check(ompl_amd_sample_uniform_dev(seeds.data(), n_streams, per, kDim, lo, hi, R.samples_f64, R.samples_f32,
                                          s));
    } else {
        double *t64 = R.alloc<double>((size_t)ns * kDim);
        float *t32 = R.alloc<float>((size_t)ns * kDim);
        check(ompl_amd_sample_uniform_dev(seeds.data(), n_streams, per, kDim, lo, hi, t64, t32, s));
        std::vector<float> e32((size_t)n_extra * kDim);
        for (int i = 0; i < n_extra; ++i)
            for (int d = 0; d < kDim; ++d) e32[(size_t)d * n_extra + i] = (float)extra[(size_t)i * kDim + d];
        OA_HIP(hipMemcpyAsync(R.samples_f64, extra, (size_t)n_extra * kDim * sizeof(double), hipMemcpyHostToDevice,
                              s));
        OA_HIP(hipMemcpyAsync(R.samples_f64 + (size_t)n_extra * kDim, t64, (size_t)ns * kDim * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
        for (int d = 0; d < kDim; ++d) {
            OA_HIP(hipMemcpyAsync(R.samples_f32 + (size_t)d * n, e32.data() + (size_t)d * n_extra,
                                  (size_t)n_extra * sizeof(float), hipMemcpyHostToDevice, s));
            OA_HIP(hipMemcpyAsync(R.samples_f32 + (size_t)d * n + n_extra, t32 + (size_t)d * ns,
                                  (size_t)ns * sizeof(float), hipMemcpyDeviceToDevice, s));
        }
        OA_HIP(hipStreamSynchronize(s)); // e32 is a host temporary
    }
    R.stats.ms_sample = ms(t0);
    // 2. validity + milestones
    t0 = Clock::now();
    R.valid = R.alloc<uint8_t>((size_t)n);
    check(ompl_amd_validate_states_dev(env, R.samples_f32, n, R.valid, s));
    int32_t *iota = R.alloc<int32_t>((size_t)n);
    hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(n)), dim3(256), 0, s, iota, n);
    R.milestone = R.alloc<int32_t>((size_t)n);
    int64_t *d_nv = R.alloc<int64_t>(1);
    size_t tmp_bytes = 0;
    OA_HIP(hipcub::DeviceSelect::Flagged(nullptr, tmp_bytes, iota, R.valid, R.milestone, d_nv, (int)n, s));
    void *tmp = R.alloc<char>(tmp_bytes);
    OA_HIP(hipcub::DeviceSelect::Flagged(tmp, tmp_bytes, iota, R.valid, R.milestone, d_nv, (int)n, s));
    OA_HIP(hipMemcpyAsync(&R.nv, d_nv, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    OA_HIP(hipStreamSynchronize(s));
    R.stats.ms_validity = ms(t0);
    const int64_t nv = R.nv;
    R.nodes = R.alloc<double>((size_t)nv * kDim);
    R.nodes_f32 = R.alloc<float>((size_t)nv * kDim);
    if (nv > 0)
        hipLaunchKernelGGL(gather_nodes_kernel, dim3(blocks_for(nv)), dim3(256), 0, s, R.milestone, nv,
                           R.samples_f64, R.nodes, R.nodes_f32);
    R.label = R.alloc<int32_t>((size_t)nv);
    R.stats.k = k;
    if (nv < 2) {
        if (nv) hipLaunchKernelGGL(iota_kernel, dim3(1), dim3(64), 0, s, R.label, nv);
        R.csr_off = R.alloc<int64_t>((size_t)nv + 1);
        OA_HIP(hipMemsetAsync(R.csr_off, 0, ((size_t)nv + 1) * sizeof(int64_t), s));
        OA_HIP(hipStreamSynchronize(s));
        R.components = nv;
        R.stats.ms_total = ms(t_all);
        return;
    }
    if (flags & OMPL_AMD_PRM_KSTAR) {
        // KStarStrategy: kPRMConstant = e + e/d, k = ceil(kPRMConstant * log(n)), n = milestones
        const double e = 2.718281828459045235360287471352662498;
        const double kc = e + e / (double)kDim;
        k = (int)std::min<double>(63.0, std::ceil(kc * std::log((double)nv)));
        if (k < 1) k = 1;
    }
    R.stats.k = k;
    // 3. k nearest milestones (k + 1 to drop the milestone itself)
    t0 = Clock::now();
    const int kk = std::min<int64_t>(k + 1, nv);
    ompl_amd_nn nn = nullptr;
    check(ompl_amd_nn_create(OMPL_AMD_METRIC_L2, kDim, &nn));
    R.knn = R.alloc<int32_t>((size_t)nv * kk);
    R.knn_d = R.alloc<float>((size_t)nv * kk);
    int rc = ompl_amd_nn_add_dev(nn, R.nodes_f32, nv, nullptr, s);
    if (rc == OMPL_AMD_OK) rc = ompl_amd_nn_knn_dev(nn, R.nodes_f32, nv, kk, R.knn, R.knn_d, s);
    if (rc == OMPL_AMD_OK) rc = ompl_amd_synchronize(s);
    ompl_amd_nn_destroy(nn);
    check(rc);
    R.stats.ms_knn = ms(t0);
    // 4. candidate edges: sort + unique of (min, max) keys
    t0 = Clock::now();
    const int64_t nk = nv * k;
    uint64_t *keys = R.alloc<uint64_t>((size_t)nk);
    uint64_t *sorted = R.alloc<uint64_t>((size_t)nk);
    R.keys = R.alloc<uint64_t>((size_t)nk);
    hipLaunchKernelGGL(edge_keys_kernel, dim3(blocks_for(nv)), dim3(256), 0, s, R.knn, nv, kk, k, keys);
    size_t sort_bytes = 0, uniq_bytes = 0;
    OA_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, keys, sorted, (int)nk, 0, 64, s));
    OA_HIP(hipcub::DeviceSelect::Unique(nullptr, uniq_bytes, sorted, R.keys, d_nv, (int)nk, s));
    void *tmp2 = R.alloc<char>(std::max(sort_bytes, uniq_bytes));
    OA_HIP(hipcub::DeviceRadixSort::SortKeys(tmp2, sort_bytes, keys, sorted, (int)nk, 0, 64, s));
    OA_HIP(hipcub::DeviceSelect::Unique(tmp2, uniq_bytes, sorted, R.keys, d_nv, (int)nk, s));
    int64_t nu = 0;
    OA_HIP(hipMemcpyAsync(&nu, d_nv, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    OA_HIP(hipStreamSynchronize(s));
    if (nu > 0) {
        uint64_t last = 0;
        OA_HIP(hipMemcpyAsync(&last, R.keys + nu - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        OA_HIP(hipStreamSynchronize(s));
        if (last == ~(uint64_t)0) --nu; // padding key of milestones with fewer than k neighbours
    }
    R.ne = nu;
    R.eu = R.alloc<int32_t>((size_t)nu);
    R.ev = R.alloc<int32_t>((size_t)nu);
    R.evalid = R.alloc<uint8_t>((size_t)nu);
    double *a_soa = R.alloc<double>((size_t)nu * kDim), *b_soa = R.alloc<double>((size_t)nu * kDim);
    if (nu > 0)
        hipLaunchKernelGGL(edge_endpoints_kernel, dim3(blocks_for(nu)), dim3(256), 0, s, R.keys, nu, R.nodes, R.eu,
                           R.ev, a_soa, b_soa);
    OA_HIP(hipStreamSynchronize(s));
    R.stats.ms_candidates = ms(t0);
    if ((flags & OMPL_AMD_PRM_AUTOTUNE) && nu > 0) {
        // launch shape tuned on (up to) the first 32,768 candidate edges
        const int64_t nt = std::min<int64_t>(nu, 32768);
        double *ta = R.alloc<double>((size_t)nt * kDim), *tb = R.alloc<double>((size_t)nt * kDim);
        int32_t *teu = R.alloc<int32_t>((size_t)nt), *tev = R.alloc<int32_t>((size_t)nt);
        hipLaunchKernelGGL(edge_endpoints_kernel, dim3(blocks_for(nt)), dim3(256), 0, s, R.keys, nt, R.nodes, teu,
                           tev, ta, tb);
        OA_HIP(hipStreamSynchronize(s));
        // the winner becomes the environment's own launch shapes; the
        // library's tuning (other environments') is left as it was
        ompl_amd_tuning saved, best;
        check(ompl_amd_get_tuning(&saved));
        check(ompl_amd_autotune_motions(env, ta, tb, nt, lvs, factor, 3, &best, nullptr));
        check(ompl_amd_set_tuning(&saved));
        check(ompl_amd_env_set_tuning(env, &best));
    }
    // 5. motion checks
    t0 = Clock::now();
    if (nu > 0) check(ompl_amd_check_motions_dev(env, a_soa, b_soa, nu, lvs, factor, R.evalid, nullptr, s));
    OA_HIP(hipStreamSynchronize(s));
    R.stats.ms_motions = ms(t0);
    // 6. components
    t0 = Clock::now();
    hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(nv)), dim3(256), 0, s, R.label, nv);
    // Every round with a successful hook removes at least one root, so at
    // most nv - 1 rounds change anything; the round after the last change
    // sees every valid edge inside one tree and ends the loop.
    int32_t *changed = R.alloc<int32_t>(1);
    for (int64_t round = 0; nu > 0; ++round) {
        if (round > nv) fail(OMPL_AMD_ERR_INTERNAL, "union-find did not converge");
        OA_HIP(hipMemsetAsync(changed, 0, sizeof(int32_t), s));
        hipLaunchKernelGGL(hook_kernel, dim3(blocks_for(nu)), dim3(256), 0, s, R.eu, R.ev, R.evalid, nu, R.label,
                           changed);
        hipLaunchKernelGGL(jump_kernel, dim3(blocks_for(nv)), dim3(256), 0, s, R.label, nv);
        int32_t h = 0;
        OA_HIP(hipMemcpyAsync(&h, changed, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        OA_HIP(hipStreamSynchronize(s));
        if (!h) break;
    }
    unsigned long long *cnt = R.alloc<unsigned long long>(1);
    OA_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(count_roots_kernel, dim3(blocks_for(nv)), dim3(256), 0, s, R.label, nv, cnt);
    // 7. symmetric CSR adjacency of the valid edges (the roadmap graph)
    int32_t *pos = R.alloc<int32_t>((size_t)nu + 1);
    int32_t *degree = R.alloc<int32_t>((size_t)nv + 1);
    OA_HIP(hipMemsetAsync(degree, 0, ((size_t)nv + 1) * sizeof(int32_t), s));
    int32_t *valid32 = R.alloc<int32_t>((size_t)nu);
    size_t scan_bytes = 0;
    OA_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, valid32, pos, (int)nu, s));
    void *tmp3 = R.alloc<char>(scan_bytes);
    int32_t last_pos = 0;
    uint8_t last_valid = 0;
    if (nu) {
        hipLaunchKernelGGL(u8_to_i32_kernel, dim3(blocks_for(nu)), dim3(256), 0, s, R.evalid, nu, valid32);
        OA_HIP(hipcub::DeviceScan::ExclusiveSum(tmp3, scan_bytes, valid32, pos, (int)nu, s));
        OA_HIP(hipMemcpyAsync(&last_pos, pos + nu - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        OA_HIP(hipMemcpyAsync(&last_valid, R.evalid + nu - 1, 1, hipMemcpyDeviceToHost, s));
    }
    unsigned long long roots = 0;
    OA_HIP(hipMemcpyAsync(&roots, cnt, sizeof(roots), hipMemcpyDeviceToHost, s));
    OA_HIP(hipStreamSynchronize(s));
    R.components = (int64_t)roots;
    R.valid_edges = nu ? (int64_t)last_pos + last_valid : 0;
    const int64_t nd = 2 * R.valid_edges;
    uint64_t *dkeys = R.alloc<uint64_t>((size_t)nd), *dsorted = R.alloc<uint64_t>((size_t)nd);
    R.csr_adj = R.alloc<int32_t>((size_t)nd);
    R.csr_off = R.alloc<int64_t>((size_t)nv + 1);
    int32_t *off32 = R.alloc<int32_t>((size_t)nv + 1);
    if (nd) {
        hipLaunchKernelGGL(csr_keys_kernel, dim3(blocks_for(nu)), dim3(256), 0, s, R.eu, R.ev, R.evalid, pos, nu,
                           dkeys, degree);
        size_t b1 = 0;
        OA_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, b1, dkeys, dsorted, (int)nd, 0, 64, s));
        void *tmp4 = R.alloc<char>(b1);
        OA_HIP(hipcub::DeviceRadixSort::SortKeys(tmp4, b1, dkeys, dsorted, (int)nd, 0, 64, s));
        hipLaunchKernelGGL(csr_cols_kernel, dim3(blocks_for(nd)), dim3(256), 0, s, dsorted, nd, R.csr_adj);
    }
    size_t b2 = 0;
    OA_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, degree, off32, (int)(nv + 1), s));
    void *tmp5 = R.alloc<char>(b2);
    OA_HIP(hipcub::DeviceScan::ExclusiveSum(tmp5, b2, degree, off32, (int)(nv + 1), s));
    hipLaunchKernelGGL(widen_kernel, dim3(blocks_for(nv + 1)), dim3(256), 0, s, off32, nv + 1, R.csr_off);
    OA_HIP(hipStreamSynchronize(s));
    OA_HIP(hipGetLastError());
    R.stats.ms_graph = ms(t0);
    R.stats.ms_total = ms(t_all);
}

// Shortest path between milestones 0 and 1 over a symmetric CSR roadmap
// (host): Dijkstra with FP64 RealVectorStateSpace distances as edge weights
// (base/spaces/src/RealVectorStateSpace.cpp:230-242), nodes popped in
// (distance, index) order and relaxed only on a strictly shorter distance.
// The reference searches its roadmap with A* under the same weights
// (PRM::constructSolution, geometric/planners/prm/src/PRM.cpp:696-727), which
// finds a shortest path too.  Returns the node sequence 0 -> 1 (empty when 1 is unreachable).
std::vector<int32_t> roadmap_shortest_path(const std::vector<int64_t> &off, const std::vector<int32_t> &adj,
                                           const std::vector<double> &nodes)
{
    const int64_t nv = (int64_t)off.size() - 1;
    std::vector<double> dist((size_t)nv, INFINITY);
    std::vector<int32_t> prev((size_t)nv, -1);
    using Item = std::pair<double, int32_t>;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
    dist[0] = 0.0;
    pq.push({0.0, 0});
    while (!pq.empty()) {
        const Item it = pq.top();
        pq.pop();
        const int32_t u = it.second;
        if (it.first > dist[(size_t)u]) continue;
        if (u == 1) break;
        for (int64_t e = off[(size_t)u]; e < off[(size_t)u + 1]; ++e) {
            const int32_t v = adj[(size_t)e];
            double d2 = 0.0;
            for (int k = 0; k < kDim; ++k) {
                const double diff = nodes[(size_t)u * kDim + k] - nodes[(size_t)v * kDim + k];
                d2 += diff * diff;
            }
            const double nd = it.first + std::sqrt(d2);
            if (nd < dist[(size_t)v]) {
                dist[(size_t)v] = nd;
                prev[(size_t)v] = u;
                pq.push({nd, v});
            }
        }
    }
    std::vector<int32_t> path;
    if (nv < 2 || !std::isfinite(dist[1])) return path;
    for (int32_t v = 1; v >= 0; v = prev[(size_t)v]) path.push_back(v);
    std::reverse(path.begin(), path.end());
    return path;
}

PrmQuery prm_query(ompl_amd_env env, const double start[kDim], const double goal[kDim], uint32_t seed, int n_streams,
                   int64_t n_samples, int k, double lvs, unsigned factor, hipStream_t s)
{
    PrmQuery q;
    PrmRoadmap R;
    double extra[2 * kDim];
    for (int d = 0; d < kDim; ++d) {
        extra[d] = start[d];
        extra[kDim + d] = goal[d];
    }
    build(R, env, seed, n_streams, n_samples, k, lvs, factor, 0, s, extra, 2);
    q.milestones = R.nv;
    q.candidate_edges = R.ne;
    q.valid_edges = R.valid_edges;
    if (R.nv < 2) return q;
    int32_t first[2], lab[2];
    OA_HIP(hipMemcpyAsync(first, R.milestone, sizeof(first), hipMemcpyDeviceToHost, s));
    OA_HIP(hipMemcpyAsync(lab, R.label, sizeof(lab), hipMemcpyDeviceToHost, s));
    OA_HIP(hipStreamSynchronize(s));
    // start and goal are milestones 0 and 1 only when both are valid
    // (PRM::solve takes valid start and goal states only)
    if (first[0] != 0 || first[1] != 1 || lab[0] != lab[1]) return q;
    std::vector<int64_t> off((size_t)R.nv + 1);
    std::vector<int32_t> adj((size_t)(2 * R.valid_edges));
    std::vector<double> nodes((size_t)R.nv * kDim);
    OA_HIP(hipMemcpyAsync(off.data(), R.csr_off, off.size() * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    if (!adj.empty())
        OA_HIP(hipMemcpyAsync(adj.data(), R.csr_adj, adj.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    OA_HIP(hipMemcpyAsync(nodes.data(), R.nodes, nodes.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    OA_HIP(hipStreamSynchronize(s));
    const std::vector<int32_t> p = roadmap_shortest_path(off, adj, nodes);
    if (p.empty()) return q;
    q.solved = true;
    for (int32_t v : p) q.path.insert(q.path.end(), nodes.begin() + (size_t)v * kDim, nodes.begin() + (size_t)(v + 1) * kDim);
    return q;
}

size_t prm_footprint(int64_t n_samples, int k)
{
    // samples (FP64 rows, FP32 SoA, validity, indices) and at most as many
    // milestones (nodes FP64 + FP32, labels, neighbour lists) and k edges per
    // milestone (keys x3, endpoints, FP64 motion endpoints, flags, CSR)
    const size_t n = (size_t)n_samples + 2;
    return n * (56 + 28 + 1 + 4 + 4 + 8) + n * (56 + 28 + 4 + 8 * (size_t)(k + 1)) +
           n * (size_t)k * (24 + 8 + 112 + 1 + 4 + 16 + 8);
}

} // namespace ompl_amd

struct ompl_amd_prm_s {
    ompl_amd::PrmRoadmap R;
};

using namespace ompl_amd;

extern "C" {

int ompl_amd_prm_build(ompl_amd_env env, uint32_t seed, int n_streams, int64_t n_samples, int k, double lvs,
                       unsigned factor, unsigned flags, ompl_amd_prm *out, ompl_amd_prm_stats *stats)
{
    return guarded([&] {
        OA_REQUIRE(env && out, OMPL_AMD_ERR_ARG, "null argument");
        OA_REQUIRE(n_streams >= 1 && n_samples >= 1 && n_samples < ((int64_t)1 << 31), OMPL_AMD_ERR_ARG,
                   "bad sample counts");
        OA_REQUIRE((flags & ~(OMPL_AMD_PRM_AUTOTUNE | OMPL_AMD_PRM_KSTAR)) == 0, OMPL_AMD_ERR_ARG, "unknown flags");
        const bool kstar = (flags & OMPL_AMD_PRM_KSTAR) != 0;
        OA_REQUIRE(kstar || (k >= 1 && k <= 63), OMPL_AMD_ERR_ARG, "k must be in [1, 63]");
        const int64_t k_max = kstar ? 63 : k;
        OA_REQUIRE(2 * (n_samples + n_streams) * k_max < ((int64_t)1 << 31), OMPL_AMD_ERR_ARG,
                   "n_samples * k too large for one roadmap");
        OA_REQUIRE(lvs > 0.0 && factor >= 1, OMPL_AMD_ERR_ARG, "bad segment settings");
        auto *h = new ompl_amd_prm_s();
        try {
            build(h->R, env, seed, n_streams, n_samples, k, lvs, factor, flags, default_stream(), nullptr, 0);
        } catch (...) {
            delete h;
            throw;
        }
        PrmRoadmap &R = h->R;
        R.stats.samples = R.samples;
        R.stats.milestones = R.nv;
        R.stats.candidate_edges = R.ne;
        R.stats.valid_edges = R.valid_edges;
        R.stats.components = R.components;
        if (stats) *stats = R.stats;
        *out = h;
    });
}

int ompl_amd_prm_destroy(ompl_amd_prm h)
{
    return guarded([&] { delete h; });
}

int ompl_amd_prm_milestones(ompl_amd_prm h, double *nodes, int32_t *sample_index, int64_t cap)
{
    return guarded([&] {
        OA_REQUIRE(h, OMPL_AMD_ERR_ARG, "null roadmap");
        const PrmRoadmap &R = h->R;
        const int64_t n = std::min(cap, R.nv);
        if (nodes && n) OA_HIP(hipMemcpy(nodes, R.nodes, (size_t)n * kDim * sizeof(double), hipMemcpyDeviceToHost));
        if (sample_index && n)
            OA_HIP(hipMemcpy(sample_index, R.milestone, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
    });
}

int ompl_amd_prm_edges(ompl_amd_prm h, int32_t *u, int32_t *v, uint8_t *valid, int64_t cap)
{
    return guarded([&] {
        OA_REQUIRE(h, OMPL_AMD_ERR_ARG, "null roadmap");
        const PrmRoadmap &R = h->R;
        const int64_t n = std::min(cap, R.ne);
        if (!n) return;
        if (u) OA_HIP(hipMemcpy(u, R.eu, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
        if (v) OA_HIP(hipMemcpy(v, R.ev, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
        if (valid) OA_HIP(hipMemcpy(valid, R.evalid, (size_t)n, hipMemcpyDeviceToHost));
    });
}

int ompl_amd_prm_components(ompl_amd_prm h, int32_t *label, int64_t cap)
{
    return guarded([&] {
        OA_REQUIRE(h && label, OMPL_AMD_ERR_ARG, "null argument");
        const PrmRoadmap &R = h->R;
        const int64_t n = std::min(cap, R.nv);
        if (n) OA_HIP(hipMemcpy(label, R.label, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
    });
}

int ompl_amd_prm_graph(ompl_amd_prm h, int64_t *offsets, int32_t *adjacency, int64_t adjacency_cap)
{
    return guarded([&] {
        OA_REQUIRE(h && offsets, OMPL_AMD_ERR_ARG, "null argument");
        const PrmRoadmap &R = h->R;
        OA_REQUIRE(!adjacency || adjacency_cap >= 2 * R.valid_edges, OMPL_AMD_ERR_ARG,
                   "adjacency needs %lld entries", (long long)(2 * R.valid_edges));
        OA_HIP(hipMemcpy(offsets, R.csr_off, ((size_t)R.nv + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
        if (adjacency && R.valid_edges)
            OA_HIP(hipMemcpy(adjacency, R.csr_adj, (size_t)(2 * R.valid_edges) * sizeof(int32_t),
                             hipMemcpyDeviceToHost));
    });
}

} // extern "C"
