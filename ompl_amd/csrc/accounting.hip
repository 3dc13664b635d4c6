// Algorithmic-work accounting for the roofline: counts the work a scalar
// checker does on the same inputs in the reference order -- DiscreteMotionValidator
// checks s2, then the bisection positions up to and including the first invalid
// state (base/src/DiscreteMotionValidator.cpp:93-145); each state's collision test
// visits robot spheres in order and, for each, the sphere obstacles then the
// box obstacles, stopping at the first hit.  One lane per edge / state; run
// outside timed regions.
#include <hip/hip_runtime.h>

#include <functional>

#include "arm_device.hpp"
#include "motion_device.hpp"
#include "runtime.hpp"

namespace ompl_amd {

struct Work {
    unsigned long long states, sphere_pairs, box_pairs;
};

__device__ bool scalar_state_check(const float q[kDof], const float4 *obs, int ns, int nb,
                                   unsigned long long &sp, unsigned long long &bp)
{
    LaneSpheres<1> S;
    arm_fk<1>(q, 0, S);
    for (int s = 0; s < kSpheres; ++s) {
        for (int o = 0; o < ns; ++o) {
            ++sp;
            const float4 ob = obs[o];
            float dx = ob.x - S.x[s], dy = ob.y - S.y[s], dz = ob.z - S.z[s];
            if ((dx * dx + dy * dy) + dz * dz < ob.w) return false;
        }
        for (int o = 0; o < nb; ++o) {
            ++bp;
            const float4 lo = obs[ns + 2 * o], hi = obs[ns + 2 * o + 1];
            float px = fminf(fmaxf(S.x[s], lo.x), hi.x);
            float py = fminf(fmaxf(S.y[s], lo.y), hi.y);
            float pz = fminf(fmaxf(S.z[s], lo.z), hi.z);
            float dx = S.x[s] - px, dy = S.y[s] - py, dz = S.z[s] - pz;
            if ((dx * dx + dy * dy) + dz * dz < kRobotR * kRobotR) return false;
        }
    }
    return true;
}

__global__ void motion_work_kernel(EnvView env, const double *a_soa, const double *b_soa, int64_t n, double lvs,
                                   unsigned factor, Work *acc)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double a[kDof], b[kDof];
    for (int j = 0; j < kDof; ++j) {
        a[j] = a_soa[(int64_t)j * n + e];
        b[j] = b_soa[(int64_t)j * n + e];
    }
    unsigned long long st = 1, sp = 0, bp = 0;
    float q[kDof];
    for (int j = 0; j < kDof; ++j) q[j] = (float)b[j];
    if (scalar_state_check(q, env.obs, env.ns, env.nb, sp, bp)) {
        const int nd = valid_segment_count<kDof>(a, b, lvs, factor);
        if (nd >= 2 && nd <= kMaxSegments) {
            for (int p = 0; p < nd - 1; ++p) {
                const int idx = bisection_index(p, nd);
                interpolate_f32<kDof>(a, b, (double)idx / (double)nd, q);
                ++st;
                if (!scalar_state_check(q, env.obs, env.ns, env.nb, sp, bp)) break;
            }
        }
    }
    atomicAdd(&acc->states, st);
    atomicAdd(&acc->sphere_pairs, sp);
    atomicAdd(&acc->box_pairs, bp);
}

__global__ void validity_work_kernel(EnvView env, const float *q_soa, int64_t n, Work *acc)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float q[kDof];
    for (int j = 0; j < kDof; ++j) q[j] = q_soa[(int64_t)j * n + i];
    unsigned long long sp = 0, bp = 0;
    scalar_state_check(q, env.obs, env.ns, env.nb, sp, bp);
    atomicAdd(&acc->states, 1ull);
    atomicAdd(&acc->sphere_pairs, sp);
    atomicAdd(&acc->box_pairs, bp);
}

static void run_work(const char *what, const std::function<void(Work *, hipStream_t)> &launch, int64_t out[3])
{
    ThreadResources &r = thread_resources();
    Work *acc = nullptr;
    OA_HIP(hipMalloc(&acc, sizeof(Work)));
    OA_HIP(hipMemsetAsync(acc, 0, sizeof(Work), r.stream));
    launch(acc, r.stream);
    OA_HIP(hipGetLastError());
    Work h{};
    OA_HIP(hipMemcpyAsync(&h, acc, sizeof(Work), hipMemcpyDeviceToHost, r.stream));
    OA_HIP(hipStreamSynchronize(r.stream));
    OA_HIP(hipFree(acc));
    out[0] = (int64_t)h.states;
    out[1] = (int64_t)h.sphere_pairs;
    out[2] = (int64_t)h.box_pairs;
    (void)what;
}

} // namespace ompl_amd

using namespace ompl_amd;

extern "C" {

int ompl_amd_motion_work(ompl_amd_env e, const double *a, const double *b, int64_t n, double lvs, unsigned factor,
                         int64_t out[3])
{
    return guarded([&] {
        OA_REQUIRE(e && out && n >= 0, OMPL_AMD_ERR_ARG, "bad arguments");
        OA_REQUIRE(n == 0 || (a && b), OMPL_AMD_ERR_ARG, "null edges");
        OA_REQUIRE(lvs > 0.0 && factor >= 1, OMPL_AMD_ERR_ARG, "bad segment settings");
        const EnvView env = e->env.input_view();
        run_work("motion", [&](Work *acc, hipStream_t s) {
            if (n > 0)
                hipLaunchKernelGGL(motion_work_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, env, a, b, n,
                                   lvs, factor, acc);
        }, out);
    });
}

int ompl_amd_validity_work(ompl_amd_env e, const float *q, int64_t n, int64_t out[3])
{
    return guarded([&] {
        OA_REQUIRE(e && out && n >= 0, OMPL_AMD_ERR_ARG, "bad arguments");
        OA_REQUIRE(n == 0 || q, OMPL_AMD_ERR_ARG, "null states");
        const EnvView env = e->env.input_view();
        run_work("validity", [&](Work *acc, hipStream_t s) {
            if (n > 0)
                hipLaunchKernelGGL(validity_work_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, env, q, n,
                                   acc);
        }, out);
    });
}

} // extern "C"
