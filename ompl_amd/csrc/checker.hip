// Batched state validity and motion validation kernels for gfx950.
//
// Mapping (wave64): a state is checked by a group of LPS lanes (LPS in
// {1,2,4,8}); each lane of the group holds 28/LPS of the arm's sphere centres
// (FK is recomputed per lane: ~0.6k FP32 ops against >= 10k for collision at
// the benchmark sizes).  Obstacles are staged once per block in LDS and read
// as wave-uniform float4 broadcasts; every obstacle iteration ends with a
// ballot so a wave leaves the loop as soon as its exit policy is met.
//
//   validate_kernel       batched StateValidityChecker::isValid
//   motion_endpoint_kernel  check s2 of every edge (DiscreteMotionValidator.cpp:96),
//                         emit survivors (s2 valid and nd >= 2) to a list
//   motion_walk_kernel    one wave per edge: positions 0 (s2) and 1..nd-1 (the
//                         bisection order of DiscreteMotionValidator.cpp:106-134)
//                         in chunks of 64/LPS, stop at the first chunk with a
//                         collision; optional exact first-invalid position
//   motion_single_kernel  the same walk for one edge passed by value (the plain
//                         one-launch-per-call path)
//   motion_last_valid_kernel  the sequential lastValid overload (:48-91)
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "arm_device.hpp"
#include "checker.hpp"
#include "motion_device.hpp"

namespace ompl_amd {

__device__ __forceinline__ int lane_id()
{
    return threadIdx.x & 63;
}

__device__ __forceinline__ int lowest_bit(uint64_t m)
{
    return __builtin_ctzll(m);
}

template <int LPS, typename T>
__global__ void __launch_bounds__(1024) validate_kernel(EnvView env, const T *__restrict__ q_soa, int64_t n,
                                                        uint8_t *__restrict__ out)
{
    extern __shared__ float4 lds[];
    stage_env(env, lds);
    constexpr int G = 64 / LPS;
    const int lane = lane_id();
    const int g = lane % LPS;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t st = wave * G + lane / LPS;
    const bool live = st < n;
    float q[kDof];
#pragma unroll
    for (int j = 0; j < kDof; ++j)
        q[j] = live ? (float)q_soa[(int64_t)j * n + st] : 0.0f;
    LaneSpheres<LPS> S;
    arm_fk<LPS>(q, g, S);
    const bool hit = collide_env<LPS, Exit::kAllGroups>(S, lds, env.ns, env.nb, live);
    const uint64_t H = fold_groups<LPS>(wave_ballot(hit && live));
    if (live && g == 0) out[st] = ((H >> lane) & 1) ? 0 : 1;
}

// Walks one edge from position `start` (0 = s2).  Returns the position of the
// first invalid state found (INT_MAX if none) and the total position count.
template <int LPS, bool EXACT>
__device__ __forceinline__ int walk_edge(const double a[kDof], const double b[kDof], double lvs, unsigned factor,
                                         const float4 *lds, int ns, int nb, int start, int &total)
{
    constexpr int G = 64 / LPS;
    const int lane = lane_id();
    const int g = lane % LPS;
    const int grp = lane / LPS;
    const int nd = valid_segment_count<kDof>(a, b, lvs, factor);
    if (nd > kMaxSegments) {
        total = 0;
        return -1;
    }
    const int m = nd >= 2 ? nd - 1 : 0;
    total = 1 + m;
    for (int base = start; base < total; base += G) {
        const int p = base + grp;
        const bool live = p < total;
        float q[kDof];
        if (p == 0 || !live) {
#pragma unroll
            for (int j = 0; j < kDof; ++j) q[j] = (float)b[j];
        } else {
            const int idx = bisection_index(p - 1, nd);
            const double t = (double)idx / (double)nd;
            interpolate_f32<kDof>(a, b, t, q);
        }
        LaneSpheres<LPS> S;
        arm_fk<LPS>(q, g, S);
        const bool hit = collide_env<LPS, EXACT ? Exit::kLowestGroup : Exit::kAnyGroup>(S, lds, ns, nb, live);
        const uint64_t H = fold_groups<LPS>(wave_ballot(hit && live));
        if (H) return base + lowest_bit(H) / LPS;
    }
    return INT_MAX;
}

template <int LPS>
__global__ void __launch_bounds__(1024) motion_endpoint_kernel(MotionArgs A)
{
    extern __shared__ float4 lds[];
    stage_env(A.env, lds);
    constexpr int G = 64 / LPS;
    const int lane = lane_id();
    const int g = lane % LPS;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t e = wave * G + lane / LPS;
    const bool live = e < A.n;
    double a[kDof], b[kDof];
    float q[kDof];
#pragma unroll
    for (int j = 0; j < kDof; ++j) {
        a[j] = live ? A.a[(int64_t)j * A.n + e] : 0.0;
        b[j] = live ? A.b[(int64_t)j * A.n + e] : 0.0;
        q[j] = (float)b[j];
    }
    LaneSpheres<LPS> S;
    arm_fk<LPS>(q, g, S);
    const bool hit = collide_env<LPS, Exit::kAllGroups>(S, lds, A.env.ns, A.env.nb, live);
    const uint64_t H = fold_groups<LPS>(wave_ballot(hit && live));
    const bool s2_valid = !((H >> lane) & 1);
    bool survivor = false;
    if (live && g == 0) {
        if (!s2_valid) {
            A.out[e] = 0;
            if (A.checked) A.checked[e] = 1;
        } else {
            const int nd = valid_segment_count<kDof>(a, b, A.lvs, A.factor);
            if (nd > kMaxSegments) {
                A.out[e] = 0;
                if (A.checked) A.checked[e] = 0;
            } else if (nd < 2) {
                A.out[e] = 1;
                if (A.checked) A.checked[e] = 1;
            } else {
                survivor = true;
            }
        }
    }
    // wave-aggregated append of the survivors
    const uint64_t sm = wave_ballot(survivor);
    if (sm) {
        int base = 0;
        if (lane == lowest_bit(sm)) base = atomicAdd(A.list_count, __popcll(sm));
        base = __shfl(base, lowest_bit(sm));
        if (survivor) {
            const uint64_t below = sm & (((uint64_t)1 << lane) - 1);
            A.list[base + __popcll(below)] = (int32_t)e;
        }
    }
}

template <int LPS, bool EXACT>
__global__ void __launch_bounds__(1024) motion_walk_kernel(MotionArgs A)
{
    extern __shared__ float4 lds[];
    stage_env(A.env, lds);
    const int lane = lane_id();
    const int waves_total = (gridDim.x * blockDim.x) >> 6;
    const int64_t count = A.list ? (int64_t)(*A.list_count) : A.n;
    // first item by wave id; then a dynamic queue when a counter is given
    // (items waves_total, waves_total+1, ...), otherwise a static wave-stride
    int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    for (;;) {
        if (k >= count) break;
        const int64_t e = A.list ? (int64_t)A.list[k] : k;
        double a[kDof], b[kDof];
#pragma unroll
        for (int j = 0; j < kDof; ++j) {
            a[j] = A.a[(int64_t)j * A.n + e];
            b[j] = A.b[(int64_t)j * A.n + e];
        }
        int total = 0;
        const int bad = walk_edge<LPS, EXACT>(a, b, A.lvs, A.factor, lds, A.env.ns, A.env.nb, A.start, total);
        if (lane == 0) {
            A.out[e] = bad == INT_MAX ? 1 : 0;
            if (A.checked) A.checked[e] = bad == INT_MAX ? total : (bad < 0 ? 0 : bad + 1);
        }
        if (A.queue) {
            int t = 0;
            if (lane == 0) t = atomicAdd(A.queue, 1);
            k = waves_total + (int64_t)__shfl(t, 0);
        } else {
            k += waves_total;
        }
    }
}

// Fused batch: each wave checks s2 of 64/LPS_E edges (LPS_E lanes per edge),
// then walks its own survivors one after another with LPS_W lanes per state.
// One launch, no queue and no counters.
template <int LPS_E, int LPS_W, bool EXACT>
__global__ void __launch_bounds__(1024) motion_fused_kernel(MotionArgs A)
{
    extern __shared__ float4 lds[];
    stage_env(A.env, lds);
    constexpr int G = 64 / LPS_E;
    const int lane = lane_id();
    const int g = lane % LPS_E;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t e = wave * G + lane / LPS_E;
    const bool live = e < A.n;
    double a[kDof], b[kDof];
    float q[kDof];
#pragma unroll
    for (int j = 0; j < kDof; ++j) {
        a[j] = live ? A.a[(int64_t)j * A.n + e] : 0.0;
        b[j] = live ? A.b[(int64_t)j * A.n + e] : 0.0;
        q[j] = (float)b[j];
    }
    LaneSpheres<LPS_E> S;
    arm_fk<LPS_E>(q, g, S);
    const bool hit = collide_env<LPS_E, Exit::kAllGroups>(S, lds, A.env.ns, A.env.nb, live);
    const uint64_t H = fold_groups<LPS_E>(wave_ballot(hit && live));
    const bool s2_valid = !((H >> lane) & 1);
    bool survivor = false;
    if (live && g == 0) {
        const int nd = valid_segment_count<kDof>(a, b, A.lvs, A.factor);
        if (!s2_valid) {
            A.out[e] = 0;
            if (A.checked) A.checked[e] = 1;
        } else if (nd > kMaxSegments) {
            A.out[e] = 0;
            if (A.checked) A.checked[e] = 0;
        } else if (nd < 2) {
            A.out[e] = 1;
            if (A.checked) A.checked[e] = 1;
        } else {
            survivor = true;
        }
    }
    uint64_t sm = wave_ballot(survivor);
    while (sm) {
        const int src = lowest_bit(sm);
        sm &= sm - 1;
        const int64_t es = wave * G + src / LPS_E;
        double ea[kDof], eb[kDof];
#pragma unroll
        for (int j = 0; j < kDof; ++j) {
            ea[j] = A.a[(int64_t)j * A.n + es];
            eb[j] = A.b[(int64_t)j * A.n + es];
        }
        int total = 0;
        const int bad = walk_edge<LPS_W, EXACT>(ea, eb, A.lvs, A.factor, lds, A.env.ns, A.env.nb, 1, total);
        if (lane == 0) {
            A.out[es] = bad == INT_MAX ? 1 : 0;
            if (A.checked) A.checked[es] = bad == INT_MAX ? total : bad + 1;
        }
    }
}

template <int LPS>
__global__ void __launch_bounds__(64) motion_single_kernel(EnvView env, Edge1 E, double lvs, unsigned factor,
                                                           int *result)
{
    extern __shared__ float4 lds[];
    stage_env(env, lds);
    int total = 0;
    const int bad = walk_edge<LPS, false>(E.a, E.b, lvs, factor, lds, env.ns, env.nb, 0, total);
    if (lane_id() == 0) *result = bad == INT_MAX ? 1 : 0; // bad < 0: too many segments, invalid
}

__global__ void __launch_bounds__(64) validate_single_kernel(EnvView env, State1 s, int *result)
{
    extern __shared__ float4 lds[];
    stage_env(env, lds);
    // 8 lanes per state; lane groups 1..7 idle
    const int lane = lane_id();
    const bool live = lane < 8;
    float q[kDof];
#pragma unroll
    for (int j = 0; j < kDof; ++j) q[j] = (float)s.q[j];
    LaneSpheres<8> S;
    arm_fk<8>(q, lane & 7, S);
    const bool hit = collide_env<8, Exit::kAllGroups>(S, lds, env.ns, env.nb, live);
    const uint64_t H = wave_ballot(hit && live);
    if (lane == 0) *result = H ? 0 : 1;
}

// Sequential lastValid overload: positions j = 1..nd-1 in order, then s2.
template <int LPS>
__global__ void __launch_bounds__(1024) motion_last_valid_kernel(MotionArgs A, double *__restrict__ last_t)
{
    extern __shared__ float4 lds[];
    stage_env(A.env, lds);
    constexpr int G = 64 / LPS;
    const int lane = lane_id();
    const int g = lane % LPS;
    const int grp = lane / LPS;
    const int waves_total = (gridDim.x * blockDim.x) >> 6;
    for (int64_t e = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; e < A.n; e += waves_total) {
        double a[kDof], b[kDof];
#pragma unroll
        for (int j = 0; j < kDof; ++j) {
            a[j] = A.a[(int64_t)j * A.n + e];
            b[j] = A.b[(int64_t)j * A.n + e];
        }
        const int nd = valid_segment_count<kDof>(a, b, A.lvs, A.factor);
        if (nd > kMaxSegments) { // reported invalid without a walk
            if (lane == 0) {
                A.out[e] = 0;
                if (A.checked) A.checked[e] = 0;
            }
            continue;
        }
        // positions 1..nd-1 are interpolation indices j, position `last` is s2
        const int last = nd > 1 ? nd : 1;
        int bad = INT_MAX;
        for (int base = 1; base <= last; base += G) {
            const int p = base + grp;
            const bool live = p <= last;
            float q[kDof];
            if (p >= last) {
#pragma unroll
                for (int j = 0; j < kDof; ++j) q[j] = (float)b[j];
            } else {
                interpolate_f32<kDof>(a, b, (double)p / (double)nd, q);
            }
            LaneSpheres<LPS> S;
            arm_fk<LPS>(q, g, S);
            const bool hit = collide_env<LPS, Exit::kLowestGroup>(S, lds, A.env.ns, A.env.nb, live);
            const uint64_t H = fold_groups<LPS>(wave_ballot(hit && live));
            if (H) {
                bad = base + lowest_bit(H) / LPS;
                break;
            }
        }
        if (lane == 0) {
            if (bad == INT_MAX) {
                A.out[e] = 1;
            } else {
                A.out[e] = 0;
                // interior j fails: (j-1)/nd; s2 fails: (nd-1)/nd
                last_t[e] = bad < last ? (double)(bad - 1) / (double)nd : (double)(nd - 1) / (double)nd;
            }
            if (A.checked) A.checked[e] = bad == INT_MAX ? last : bad;
        }
    }
}

__global__ void zero_ints_kernel(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static size_t lds_bytes(const EnvView &e)
{
    return (size_t)(e.ns + 2 * e.nb) * sizeof(float4);
}

template <int LPS, typename T>
static void launch_validate_t(const EnvView &env, const T *q, int64_t n, uint8_t *out, int block, hipStream_t s)
{
    const int64_t states_per_block = (int64_t)(block / 64) * (64 / LPS);
    const int64_t blocks = (n + states_per_block - 1) / states_per_block;
    hipLaunchKernelGGL((validate_kernel<LPS, T>), dim3((unsigned)blocks), dim3(block), lds_bytes(env), s, env, q, n,
                       out);
}

template <typename T>
void launch_validate(const EnvView &env, const T *q, int64_t n, uint8_t *out, int lps, int block, hipStream_t s)
{
    if (n <= 0) return;
    ProfScope ps("validate_states", s);
    switch (lps) {
    case 1: launch_validate_t<1>(env, q, n, out, block, s); break;
    case 2: launch_validate_t<2>(env, q, n, out, block, s); break;
    case 4: launch_validate_t<4>(env, q, n, out, block, s); break;
    default: launch_validate_t<8>(env, q, n, out, block, s); break;
    }
    OA_HIP(hipGetLastError());
}
template void launch_validate<float>(const EnvView &, const float *, int64_t, uint8_t *, int, int, hipStream_t);
template void launch_validate<double>(const EnvView &, const double *, int64_t, uint8_t *, int, int, hipStream_t);

void launch_zero(int *p, int n, hipStream_t s)
{
    hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    OA_HIP(hipGetLastError());
}

template <int LPS>
static void launch_endpoint_t(const MotionArgs &A, int block, hipStream_t s)
{
    const int64_t per_block = (int64_t)(block / 64) * (64 / LPS);
    const int64_t blocks = (A.n + per_block - 1) / per_block;
    hipLaunchKernelGGL((motion_endpoint_kernel<LPS>), dim3((unsigned)blocks), dim3(block), lds_bytes(A.env), s, A);
}

void launch_motion_endpoint(const MotionArgs &A, int lps, int block, hipStream_t s)
{
    if (A.n <= 0) return;
    ProfScope ps("motion_endpoint", s);
    switch (lps) {
    case 1: launch_endpoint_t<1>(A, block, s); break;
    case 2: launch_endpoint_t<2>(A, block, s); break;
    case 4: launch_endpoint_t<4>(A, block, s); break;
    default: launch_endpoint_t<8>(A, block, s); break;
    }
    OA_HIP(hipGetLastError());
}

template <int LPS>
static void launch_walk_t(const MotionArgs &A, bool exact, int block, int64_t waves, hipStream_t s)
{
    const int wpb = block / 64;
    const int64_t blocks = (waves + wpb - 1) / wpb;
    if (exact)
        hipLaunchKernelGGL((motion_walk_kernel<LPS, true>), dim3((unsigned)blocks), dim3(block), lds_bytes(A.env), s,
                           A);
    else
        hipLaunchKernelGGL((motion_walk_kernel<LPS, false>), dim3((unsigned)blocks), dim3(block), lds_bytes(A.env), s,
                           A);
}

void launch_motion_walk(const MotionArgs &A, bool exact, int lps, int block, int64_t waves, hipStream_t s)
{
    if (waves <= 0) return;
    ProfScope ps("motion_walk", s);
    switch (lps) {
    case 1: launch_walk_t<1>(A, exact, block, waves, s); break;
    case 2: launch_walk_t<2>(A, exact, block, waves, s); break;
    case 4: launch_walk_t<4>(A, exact, block, waves, s); break;
    default: launch_walk_t<8>(A, exact, block, waves, s); break;
    }
    OA_HIP(hipGetLastError());
}

template <int LPS_E, int LPS_W>
static void launch_fused_t(const MotionArgs &A, bool exact, int block, hipStream_t s)
{
    const int64_t per_block = (int64_t)(block / 64) * (64 / LPS_E);
    const unsigned blocks = (unsigned)((A.n + per_block - 1) / per_block);
    if (exact)
        hipLaunchKernelGGL((motion_fused_kernel<LPS_E, LPS_W, true>), dim3(blocks), dim3(block), lds_bytes(A.env), s,
                           A);
    else
        hipLaunchKernelGGL((motion_fused_kernel<LPS_E, LPS_W, false>), dim3(blocks), dim3(block), lds_bytes(A.env),
                           s, A);
}

void launch_motion_fused(const MotionArgs &A, bool exact, int lps_end, int lps_walk, int block, hipStream_t s)
{
    if (A.n <= 0) return;
    ProfScope ps("motion_fused", s);
#define OA_FUSED(E, W)                                                                                             \
    if (lps_end == E && lps_walk == W) {                                                                            \
        launch_fused_t<E, W>(A, exact, block, s);                                                                   \
        OA_HIP(hipGetLastError());                                                                                  \
        return;                                                                                                     \
    }
    OA_FUSED(8, 1)
    OA_FUSED(8, 2)
    OA_FUSED(8, 4)
    OA_FUSED(4, 1)
    OA_FUSED(4, 2)
    OA_FUSED(2, 1)
    OA_FUSED(1, 1)
#undef OA_FUSED
    fail(OMPL_AMD_ERR_ARG, "fused motion kernel not built for lanes (%d, %d)", lps_end, lps_walk);
}

void launch_motion_single(const EnvView &env, const Edge1 &E, double lvs, unsigned factor, int *result, hipStream_t s)
{
    hipLaunchKernelGGL((motion_single_kernel<1>), dim3(1), dim3(64), lds_bytes(env), s, env, E, lvs, factor, result);
    OA_HIP(hipGetLastError());
}

void launch_validate_single(const EnvView &env, const State1 &st, int *result, hipStream_t s)
{
    hipLaunchKernelGGL(validate_single_kernel, dim3(1), dim3(64), lds_bytes(env), s, env, st, result);
    OA_HIP(hipGetLastError());
}

void launch_motion_last_valid(const MotionArgs &A, double *last_t, int lps, int block, int64_t waves,
                              hipStream_t s)
{
    if (A.n <= 0) return;
    const int wpb = block / 64;
    const int64_t blocks = (waves + wpb - 1) / wpb;
    switch (lps) {
    case 1:
        hipLaunchKernelGGL((motion_last_valid_kernel<1>), dim3((unsigned)blocks), dim3(block), lds_bytes(A.env), s,
                           A, last_t);
        break;
    case 2:
        hipLaunchKernelGGL((motion_last_valid_kernel<2>), dim3((unsigned)blocks), dim3(block), lds_bytes(A.env), s,
                           A, last_t);
        break;
    case 4:
        hipLaunchKernelGGL((motion_last_valid_kernel<4>), dim3((unsigned)blocks), dim3(block), lds_bytes(A.env), s,
                           A, last_t);
        break;
    default:
        hipLaunchKernelGGL((motion_last_valid_kernel<8>), dim3((unsigned)blocks), dim3(block), lds_bytes(A.env), s,
                           A, last_t);
        break;
    }
    OA_HIP(hipGetLastError());
}

} // namespace ompl_amd

namespace ompl_amd {

__global__ void arm_spheres_kernel(const float *__restrict__ q_soa, int64_t n, float *__restrict__ out)
{
    const int64_t st = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (st >= n) return;
    float q[kDof];
#pragma unroll
    for (int j = 0; j < kDof; ++j) q[j] = q_soa[(int64_t)j * n + st];
    LaneSpheres<1> S;
    arm_fk<1>(q, 0, S);
#pragma unroll
    for (int m = 0; m < kSpheres; ++m) {
        out[(st * kSpheres + m) * 3 + 0] = S.x[m];
        out[(st * kSpheres + m) * 3 + 1] = S.y[m];
        out[(st * kSpheres + m) * 3 + 2] = S.z[m];
    }
}

void launch_arm_spheres(const float *q, int64_t n, float *out, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(arm_spheres_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, q, n, out);
    OA_HIP(hipGetLastError());
}

} // namespace ompl_amd
