// Device helpers of DiscreteMotionValidator: FP64 segment count,
// interpolation, and the closed-form position of the p-th check of the
// bisection queue.
//
// Reference: base/src/DiscreteMotionValidator.cpp:93-145 keeps a FIFO of
// intervals [a, b] of interpolation indices, starting with (1, nd-1); each pop
// checks mid = (a+b)/2 and pushes (a, mid-1) if a < mid and (mid+1, b) if
// b > mid.  That is a level-order walk of an implicit binary tree in which a
// node of size s has children of sizes floor((s-1)/2) and floor(s/2).  For
// m = nd-1 positions and L = floor(log2 m): levels 0..L-1 are complete (2^l
// non-empty nodes each) and level L holds the remaining m - (2^L - 1) nodes of
// size 1, so the p-th check is found by an O(log m) descent:
//   p < 2^L - 1: level l = floor(log2(p+1)), node i = p+1-2^l, follow the l
//                bits of i from the top (0 = left, 1 = right);
//   otherwise:   j = p - (2^L - 1); at a node of size s with h levels left,
//                the left child holds floor((s-1)/2) - (2^(h-1) - 1) level-L
//                nodes: go left if j is below that count, else subtract it
//                and go right.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ompl_amd {

// Motions needing more than this many segments are reported invalid without
// being walked (checked count 0): it bounds the work of one wave.
constexpr int kMaxSegments = 1 << 24;

__host__ __device__ __forceinline__ int ilog2_u32(uint32_t v)
{
    return 31 - __builtin_clz(v);
}

// Index (in [1, nd-1]) of the p-th state checked by the bisection, 0 <= p < nd-1.
__host__ __device__ __forceinline__ int bisection_index(int p, int nd)
{
    const int m = nd - 1;
    const int L = ilog2_u32((uint32_t)m);
    int a = 1, b = m;
    const int full = (1 << L) - 1;
    if (p < full) {
        const int l = ilog2_u32((uint32_t)(p + 1));
        const int i = p + 1 - (1 << l);
        for (int t = l - 1; t >= 0; --t) {
            const int mid = (a + b) >> 1;
            if ((i >> t) & 1) a = mid + 1;
            else b = mid - 1;
        }
        return (a + b) >> 1;
    }
    int j = p - full;
    for (int h = L; h > 0; --h) {
        const int mid = (a + b) >> 1;
        const int s = b - a + 1;
        const int left_leaves = ((s - 1) >> 1) - ((1 << (h - 1)) - 1);
        if (j < left_leaves) {
            b = mid - 1;
        } else {
            j -= left_leaves;
            a = mid + 1;
        }
    }
    return (a + b) >> 1;
}

// StateSpace::validSegmentCount (base/src/StateSpace.cpp:851-854) over
// RealVectorStateSpace::distance (base/spaces/src/RealVectorStateSpace.cpp:230-242).
template <int DIM>
__host__ __device__ __forceinline__ int valid_segment_count(const double a[DIM], const double b[DIM], double lvs,
                                                            unsigned factor)
{
    double dist = 0.0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        double diff = a[i] - b[i];
        dist += diff * diff;
    }
    const double d = __builtin_sqrt(dist);
    const double segs = __builtin_ceil(d / lvs);
    // out of range (or NaN): more than kMaxSegments, flagged to the caller
    if (!(segs * (double)factor <= (double)kMaxSegments)) return kMaxSegments + 1;
    return (int)(factor * (unsigned)segs);
}

// RealVectorStateSpace::interpolate then the FP64 -> FP32 cast of
// vamp/Utils.h:26.
template <int DIM>
__device__ __forceinline__ void interpolate_f32(const double a[DIM], const double b[DIM], double t, float out[DIM])
{
#pragma unroll
    for (int i = 0; i < DIM; ++i)
        out[i] = (float)(a[i] + (b[i] - a[i]) * t);
}

} // namespace ompl_amd
