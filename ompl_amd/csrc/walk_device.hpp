// Device helpers shared by the batched motion kernels and the planning
// server: wave utilities and the walk of one edge in DiscreteMotionValidator
// order (position 0 = s2, positions p >= 1 = the p-th bisection check,
// base/src/DiscreteMotionValidator.cpp:93-145).
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "arm_device.hpp"
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

// Value of v on lane j (j wave-uniform), as a wave-uniform value.
__device__ __forceinline__ int readlane_i32(int v, int j)
{
    return __builtin_amdgcn_readlane(v, j);
}

__device__ __forceinline__ int64_t readlane_i64(int64_t v, int j)
{
    const unsigned long long u = (unsigned long long)v;
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, j);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), j);
    return (int64_t)(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ double readlane_f64(double v, int j)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, j);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), j);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// State of walk position p (0 = b, else bisection index p-1) as FP32.
__device__ __forceinline__ void walk_state(const double a[kDof], const double b[kDof], int nd, int p, float q[kDof])
{
    if (p == 0) {
#pragma unroll
        for (int j = 0; j < kDof; ++j) q[j] = (float)b[j];
    } else {
        const int idx = bisection_index(p - 1, nd);
        interpolate_f32<kDof>(a, b, (double)idx / (double)nd, q);
    }
}

// Checks walk positions [base, base + STATES) of one edge (one state per lane
// group); returns the lowest position found invalid, or INT_MAX.  With EXACT
// the returned position is the first invalid one of the chunk.
template <int LPS, int OPS, bool EXACT>
__device__ __forceinline__ int walk_chunk(const double a[kDof], const double b[kDof], int nd, int total, int base,
                                          const float4 *lds, int ns, int nb)
{
    using Shape = GroupShape<LPS, OPS>;
    const Shape sh(lane_id());
    const int p = base + sh.group;
    const bool live = p < total;
    float q[kDof];
    walk_state(a, b, nd, live ? p : 0, q);
    LaneSpheres<LPS> S;
    arm_fk<LPS, Shape::GS>(q, sh.slot, S);
    const bool hit =
        collide_env<LPS, OPS, EXACT ? Exit::kLowestGroup : Exit::kAnyGroup>(S, lds, ns, nb, live, sh.part);
    const uint64_t H = fold_groups<Shape::GS>(wave_ballot(hit && live));
    return H ? base + lowest_bit(H) / Shape::GS : INT_MAX;
}

// Shared-FK form of walk_chunk: positions [base, base + 64/LPS) get their FK
// in one pass (LPS lanes each), then are checked OPS passes of STATES
// positions in increasing order; returns at the first pass with an invalid
// position (its lowest one with EXACT), or INT_MAX.
template <int LPS, int OPS, bool EXACT>
__device__ __forceinline__ int walk_chunk_shared(const double a[kDof], const double b[kDof], int nd, int total,
                                                 int base, const float4 *lds, int ns, int nb)
{
    using Shape = GroupShape<LPS, OPS>;
    const int lane = lane_id();
    const int pf = base + lane / LPS;
    float q[kDof];
    walk_state(a, b, nd, pf < total ? pf : 0, q);
    LaneSpheres<LPS> F;
    arm_fk<LPS>(q, lane % LPS, F);
    const Shape sh(lane);
#pragma unroll 1
    for (int t = 0; t < OPS; ++t) {
        const int p0 = base + t * Shape::STATES;
        if (p0 >= total) break;
        LaneSpheres<LPS> S;
        spheres_from_fk<LPS, OPS>(F, t, S);
        const bool live = p0 + sh.group < total;
        const bool hit =
            collide_env<LPS, OPS, EXACT ? Exit::kLowestGroup : Exit::kAnyGroup>(S, lds, ns, nb, live, sh.part);
        const uint64_t H = fold_groups<Shape::GS>(wave_ballot(hit && live));
        if (H) return p0 + lowest_bit(H) / Shape::GS;
    }
    return INT_MAX;
}

// Walks one edge from position `start` (0 = s2) chunk by chunk.  Returns the
// position of the first invalid state found (INT_MAX if none, -1 when the
// edge has more than kMaxSegments segments) and the total position count.
template <int LPS, int OPS, bool EXACT>
__device__ __forceinline__ int walk_edge(const double a[kDof], const double b[kDof], double lvs, unsigned factor,
                                         const float4 *lds, int ns, int nb, int start, int &total)
{
    constexpr int G = GroupShape<LPS, OPS>::STATES;
    const int nd = valid_segment_count<kDof>(a, b, lvs, factor);
    if (nd > kMaxSegments) {
        total = 0;
        return -1;
    }
    const int m = nd >= 2 ? nd - 1 : 0;
    total = 1 + m;
    for (int base = start; base < total; base += G) {
        const int bad = walk_chunk<LPS, OPS, EXACT>(a, b, nd, total, base, lds, ns, nb);
        if (bad != INT_MAX) return bad;
    }
    return INT_MAX;
}

} // namespace ompl_amd
