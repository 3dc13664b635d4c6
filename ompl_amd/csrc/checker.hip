// Batched state validity and motion validation kernels for gfx950.
//
// Mapping (wave64): a state is checked by a group of LPS*OPS lanes: LPS lanes
// split the arm's 28 sphere centres (28/LPS each; FK is recomputed per lane)
// and OPS lane sets split the obstacles (partition p visits obstacles p,
// p+OPS, ...).  Obstacles are staged once per block in LDS and read as float4
// broadcasts; every obstacle iteration ends with a ballot so a wave leaves the
// loop as soon as its exit policy is met.  Shapes (LPS, OPS) built:
// (1,1) (2,1) (4,1) (8,1) (8,2) (8,4) (8,8), and the link layout (7,9): one
// state per wave, lane l holds link l % 7 and obstacle partition l / 7, and
// tests its link's bounding sphere before the exact tests of its spheres.
//
//   validate_kernel         batched StateValidityChecker::isValid
//   motion_endpoint_kernel  s2 of every edge (DiscreteMotionValidator.cpp:96);
//                           survivors (s2 valid, nd >= 2) go to a list
//   motion_walk_kernel      one wave per edge: positions 0 (s2) and 1..nd-1 (the
//                           bisection order of DiscreteMotionValidator.cpp:106-134)
//                           in chunks, stop at the first chunk with a collision
//   motion_chunk_kernel     chunk-parallel walk of the survivors in two
//                           launches (early chunks chunk-major, then the
//                           unresolved survivors' remaining chunks); a chunk is
//                           skipped once an earlier position of its edge is
//                           known invalid (atomicMin); motion_chunk_results_kernel
//                           then writes the survivors' results
//   motion_fused_kernel     s2 of the wave's edges, then the wave walks its own
//                           survivors (one launch)
//   motion_single_kernel    one edge passed by value (the plain per-call path)
//   motion_last_valid_kernel  the sequential lastValid overload (:48-91)
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "arm_device.hpp"
#include "checker.hpp"
#include "motion_device.hpp"
#include "walk_device.hpp"

namespace ompl_amd {

// (LPS, OPS) shapes instantiated for every batched kernel.
#define OA_SHAPES(X) X(1, 1) X(2, 1) X(4, 1) X(8, 1) X(8, 2) X(8, 4) X(8, 8) X(7, 1) X(7, 2) X(7, 4) X(7, 9)
// shapes with more than one obstacle partition: the shared-FK kernels
#define OA_SHARED_SHAPES(X) X(8, 2) X(8, 4) X(8, 8)

bool shape_supported(int lps, int ops)
{
#define OA_SHAPE_OK(L, O)                                                                                           \
    if (lps == L && ops == O) return true;
    OA_SHAPES(OA_SHAPE_OK)
#undef OA_SHAPE_OK
    return false;
}

// Configuration i of a dof-major array (n entries per joint) as FP32, zeros
// for a dead lane, which reads entry `fallback` (a valid index) and drops it:
// unconditional reads are issued together and wait for one memory round trip,
// where reads under the lane's condition were issued and awaited a few at a time.
template <typename T>
__device__ __forceinline__ void load_config(const T *soa, int64_t n, int64_t i, bool live, float q[kDof],
                                            int64_t fallback = 0)
{
    const int64_t ic = live ? i : fallback;
    T v[kDof];
#pragma unroll
    for (int j = 0; j < kDof; ++j) v[j] = soa[(int64_t)j * n + ic];
#pragma unroll
    for (int j = 0; j < kDof; ++j) q[j] = live ? (float)v[j] : 0.0f;
}

template <int LPS, int OPS, typename T>
__global__ void __launch_bounds__(1024) validate_kernel(EnvView env, const T *__restrict__ q_soa, int64_t n,
                                                        uint8_t *__restrict__ out)
{
    extern __shared__ float4 lds[];
    using Shape = GroupShape<LPS, OPS>;
    const int lane = lane_id();
    const Shape sh(lane);
    const int64_t wave = (xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) >> 6;
    const int64_t st = wave * Shape::STATES + sh.group;
    const bool live = st < n;
    float q[kDof];
    load_config(q_soa, n, st, live, q); // before the staging: the two reads' latencies overlap
    stage_env<LPS == 7, link_pass(OPS) * OPS>(env, lds);
    LaneSpheres<LPS> S;
    arm_fk<LPS, Shape::GS>(q, sh.slot, S);
    const bool hit = collide_env<LPS, OPS, Exit::kAllGroups>(S, lds, env.ns, env.nb, live, sh.part);
    const uint64_t H = fold_groups<Shape::GS>(wave_ballot(hit && live));
    if (live && lane % Shape::GS == 0) out[st] = ((H >> lane) & 1) ? 0 : 1;
}

// s2 check of one edge per lane group; returns true for the group leader of a
// survivor (s2 valid, 2 <= nd <= kMaxSegments) and writes the final result of
// every other edge.
// Shared-FK batched validity: a wave takes 64/LPS states, computes their FK
// once (LPS lanes per state), then checks them OPS passes at a time.
template <int LPS, int OPS, typename T>
__global__ void __launch_bounds__(1024) validate_shared_kernel(EnvView env, const T *__restrict__ q_soa, int64_t n,
                                                               uint8_t *__restrict__ out)
{
    extern __shared__ float4 lds[];
    stage_env<LPS == 7, link_pass(OPS) * OPS>(env, lds);
    using Shape = GroupShape<LPS, OPS>;
    constexpr int FK = 64 / LPS;
    const int lane = lane_id();
    const Shape sh(lane);
    const int64_t s0 = ((xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) >> 6) * FK;
    const int64_t sf = s0 + lane / LPS;
    float q[kDof];
    load_config(q_soa, n, sf, sf < n, q);
    LaneSpheres<LPS> F;
    arm_fk<LPS>(q, lane % LPS, F);
#pragma unroll 1
    for (int t = 0; t < OPS; ++t) {
        const int64_t st = s0 + t * Shape::STATES + sh.group;
        const bool live = st < n;
        LaneSpheres<LPS> S;
        spheres_from_fk<LPS, OPS>(F, t, S);
        const bool hit = collide_env<LPS, OPS, Exit::kAllGroups>(S, lds, env.ns, env.nb, live, sh.part);
        const uint64_t H = fold_groups<Shape::GS>(wave_ballot(hit && live));
        if (live && lane % Shape::GS == 0) out[st] = ((H >> lane) & 1) ? 0 : 1;
    }
}

// Result of an edge's s2 check (H: folded hit mask of the wave) on the group
// leader: writes the final result of every edge that needs no walk and
// returns true for a survivor (s2 valid, 2 <= nd <= kMaxSegments), m = nd - 1.
template <int GS>
__device__ __forceinline__ bool endpoint_finish(const MotionArgs &A, int64_t e, bool live, uint64_t H, int &m,
                                                double a[kDof], double b[kDof])
{
    const int lane = lane_id();
    m = 0;
    if (!live || lane % GS != 0) return false;
    const bool s2_valid = !((H >> lane) & 1);
    if (!s2_valid) {
        A.out[e] = 0;
        if (A.checked) A.checked[e] = 1;
        return false;
    }
#pragma unroll
    for (int j = 0; j < kDof; ++j) {
        a[j] = A.a[(int64_t)j * A.n + e];
        b[j] = A.b[(int64_t)j * A.n + e];
    }
    const int nd = valid_segment_count<kDof>(a, b, A.lvs, A.factor);
    if (nd > kMaxSegments) {
        A.out[e] = 0;
        if (A.checked) A.checked[e] = 0;
        return false;
    }
    if (nd < 2) {
        A.out[e] = 1;
        if (A.checked) A.checked[e] = 1;
        return false;
    }
    m = nd - 1;
    return true;
}

// s2 (FP32) of edge e, or zeros for a dead lane.
__device__ __forceinline__ void load_s2(const MotionArgs &A, int64_t e, bool live, float q[kDof])
{
    load_config(A.b, A.n, e, live, q);
}

// s2 check of one edge per lane group (q: its s2 from load_s2); returns true
// for the group leader of a survivor and writes the final result of every
// other edge.
template <int LPS, int OPS>
__device__ __forceinline__ bool endpoint_check(const MotionArgs &A, const float4 *lds, int64_t e, bool live,
                                               const float q[kDof], int &m, double a[kDof], double b[kDof])
{
    using Shape = GroupShape<LPS, OPS>;
    const Shape sh(lane_id());
    LaneSpheres<LPS> S;
    arm_fk<LPS, Shape::GS>(q, sh.slot, S);
    const bool hit = collide_env<LPS, OPS, Exit::kAllGroups>(S, lds, A.env.ns, A.env.nb, live, sh.part);
    const uint64_t H = fold_groups<Shape::GS>(wave_ballot(hit && live));
    return endpoint_finish<Shape::GS>(A, e, live, H, m, a, b);
}

// Block-aggregated append of the survivors to the list (with their segment
// count and, when kept, their endpoints contiguous for the chunk walk).
// Every thread of the block calls it (it holds three barriers).  The waves'
// counts and largest segment counts meet in LDS, and one thread per block
// makes the two device-wide atomics: device-scope atomics on one address
// from every CU serialise, and with many survivors (a sparse scene) one
// atomic per survivor cost more than the s2 checks themselves.
__device__ __forceinline__ void append_survivor(const MotionArgs &A, int64_t e, bool survivor, int m,
                                                const double a[kDof], const double b[kDof])
{
    __shared__ int blk_count, blk_max, blk_base;
    const int lane = lane_id();
    if (threadIdx.x == 0) {
        blk_count = 0;
        blk_max = 0;
    }
    __syncthreads();
    const uint64_t sm = wave_ballot(survivor);
    int woff = 0;
    if (sm) {
        int mm = survivor ? m : 0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) mm = max(mm, __shfl_xor(mm, d));
        if (lane == lowest_bit(sm)) {
            woff = atomicAdd(&blk_count, __popcll(sm));
            atomicMax(&blk_max, mm);
        }
        woff = __shfl(woff, lowest_bit(sm));
    }
    __syncthreads();
    if (threadIdx.x == 0 && blk_count > 0) {
        blk_base = atomicAdd(A.list_count, blk_count);
        if (A.max_m) atomicMax(A.max_m, blk_max);
    }
    __syncthreads();
    if (survivor) {
        const int slot = blk_base + woff + __popcll(sm & (((uint64_t)1 << lane) - 1));
        A.out[e] = 1; // until the walk finds an invalid position
        A.list[slot] = (int32_t)e;
        if (A.first_bad) A.first_bad[slot] = INT_MAX;
        if (A.surv_nd) A.surv_nd[slot] = m + 1;
        if (A.surv_ab) {
            double *ab = A.surv_ab + (int64_t)slot * 2 * kDof;
#pragma unroll
            for (int j = 0; j < kDof; ++j) {
                ab[j] = a[j];
                ab[kDof + j] = b[j];
            }
        }
    }
}

template <int LPS, int OPS>
__global__ void __launch_bounds__(1024) motion_endpoint_kernel(MotionArgs A)
{
    extern __shared__ float4 lds[];
    using Shape = GroupShape<LPS, OPS>;
    const int64_t wave = (xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) >> 6;
    const int64_t e = wave * Shape::STATES + lane_id() / Shape::GS;
    // s2 is read before the obstacles are staged, so the two reads' latencies overlap
    float q[kDof];
    load_s2(A, e, e < A.n, q);
    stage_env<LPS == 7, link_pass(OPS) * OPS>(A.env, lds);
    int m = 0;
    double a[kDof], b[kDof];
    const bool survivor = endpoint_check<LPS, OPS>(A, lds, e, e < A.n, q, m, a, b);
    append_survivor(A, e, survivor, m, a, b);
}

// Shared-FK endpoint pass: a wave takes 64/LPS edges, computes the FK of their
// s2 once (LPS lanes per state), then checks them OPS passes at a time.
template <int LPS, int OPS>
__global__ void __launch_bounds__(1024) motion_endpoint_shared_kernel(MotionArgs A)
{
    extern __shared__ float4 lds[];
    stage_env<LPS == 7, link_pass(OPS) * OPS>(A.env, lds);
    using Shape = GroupShape<LPS, OPS>;
    constexpr int FK = 64 / LPS;
    const int lane = lane_id();
    const Shape sh(lane);
    const int64_t e0 = ((xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) >> 6) * FK;
    const int64_t ef = e0 + lane / LPS;
    float q[kDof];
    load_s2(A, ef, ef < A.n, q);
    LaneSpheres<LPS> F;
    arm_fk<LPS>(q, lane % LPS, F);
#pragma unroll 1
    for (int t = 0; t < OPS; ++t) {
        const int64_t e = e0 + t * Shape::STATES + sh.group;
        const bool live = e < A.n;
        LaneSpheres<LPS> S;
        spheres_from_fk<LPS, OPS>(F, t, S);
        const bool hit = collide_env<LPS, OPS, Exit::kAllGroups>(S, lds, A.env.ns, A.env.nb, live, sh.part);
        const uint64_t H = fold_groups<Shape::GS>(wave_ballot(hit && live));
        int m = 0;
        double a[kDof], b[kDof];
        const bool survivor = endpoint_finish<Shape::GS>(A, e, live, H, m, a, b);
        append_survivor(A, e, survivor, m, a, b);
    }
}

template <int LPS, int OPS, bool EXACT>
__global__ void __launch_bounds__(1024) motion_walk_kernel(MotionArgs A)
{
    extern __shared__ float4 lds[];
    stage_env<LPS == 7, link_pass(OPS) * OPS>(A.env, lds);
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
        const int bad = walk_edge<LPS, OPS, EXACT>(a, b, A.lvs, A.factor, lds, A.env.ns, A.env.nb, A.start, total);
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

// Largest survivor count for which the chunk walk numbers the survivors'
// deep chunks edge-major (their chunk-count offsets are scanned into LDS).
constexpr int kFlatSurvivors = 4096;

// Bytes of dynamic LDS the staged environment takes (stage_env<LINKS>).
template <bool LINKS, int PAD>
__device__ __forceinline__ size_t env_lds_bytes(const EnvView &e)
{
    if (!LINKS) return (size_t)(e.ns + 2 * e.nb) * sizeof(float4);
    return (size_t)(pad_to(e.ns, PAD) + 2 * pad_to(e.nb, PAD)) * sizeof(float4) + (size_t)e.ns * sizeof(float);
}

// off[0] = 0, off[i + 1] = sum over survivors t <= i of their chunks past
// the first L0, max(0, ceil((nd_t - 1) / G) - L0), counting only survivors
// with no invalid position in those first L0 chunks (positions 1 .. L0 G, the
// ones phase 1 walked); computed by the whole block (each thread a contiguous
// run of survivors, then a scan of the run totals by the first wave).  Every
// block of phase 2 scans for itself while others already walk and lower
// first_bad, so the test must not see phase 2's own results: those are
// positions past L0 G, which the test ignores, so every block computes the
// same offsets (a chunk behind a phase-2 invalid position is skipped when it
// is walked).
template <int G>
__device__ __forceinline__ void scan_remaining_chunks(const int32_t *__restrict__ nd,
                                                      const int32_t *__restrict__ first_bad, int n, int64_t L0,
                                                      int32_t *off)
{
    __shared__ int32_t part[1024];
    const int T = blockDim.x, t = threadIdx.x;
    const int per = (n + T - 1) / T;
    const int lo = min(n, t * per), hi = min(n, lo + per);
    int sum = 0;
    for (int i = lo; i < hi; ++i) {
        const int64_t chunks = (nd[i] - 1 + G - 1) / G;
        const int len = first_bad[i] > L0 * G ? (int)max((int64_t)0, chunks - L0) : 0;
        off[i + 1] = len;
        sum += len;
    }
    part[t] = sum;
    __syncthreads();
    if (t < 64) {
        int carry = 0;
        for (int c0 = 0; c0 < T; c0 += 64) {
            int v = part[c0 + t];
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int u = __shfl_up(v, d);
                if (t >= d) v += u;
            }
            part[c0 + t] = v + carry;
            carry += __shfl(v, 63);
        }
    }
    __syncthreads();
    int run = t > 0 ? part[t - 1] : 0;
    if (t == 0) off[0] = 0;
    for (int i = lo; i < hi; ++i) {
        run += off[i + 1];
        off[i + 1] = run;
    }
    __syncthreads();
}

// One work item of the chunk walk: positions [bj, bj + G) of survivor sj,
// skipped when bj is past the edge's last position or an earlier position of
// the edge is already known invalid.
template <int LPS, int OPS, bool EXACT, bool SHARED>
__device__ __forceinline__ void walk_item(const MotionArgs &A, const float4 *lds, int sj, int bj)
{
    int ndj;
    // the endpoints, the position count and first_bad are read together (one
    // round trip); the item is dropped past the edge's last position or behind
    // a known invalid one
    const double *ab = A.surv_ab + (int64_t)sj * 2 * kDof;
    double a[kDof], b[kDof];
#pragma unroll
    for (int d = 0; d < kDof; ++d) {
        a[d] = ab[d];
        b[d] = ab[kDof + d];
    }
    ndj = A.surv_nd[sj];
    const int fb = __hip_atomic_load(&A.first_bad[sj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (bj >= ndj || fb < bj) return;
    const int bad = SHARED ? walk_chunk_shared<LPS, OPS, EXACT>(a, b, ndj, ndj, bj, lds, A.env.ns, A.env.nb)
                           : walk_chunk<LPS, OPS, EXACT>(a, b, ndj, ndj, bj, lds, A.env.ns, A.env.nb);
    if (bad != INT_MAX && lane_id() == 0) {
        atomicMin(&A.first_bad[sj], bad);
        A.out[A.list[sj]] = 0; // the same byte from every wave that finds one
    }
}

// Walks the live items of a 64-item window (one per lane) in lane order.
template <int LPS, int OPS, bool EXACT, bool SHARED>
__device__ __forceinline__ void walk_window(const MotionArgs &A, const float4 *lds, int s, int base, bool live)
{
    uint64_t mask = wave_ballot(live);
    while (mask) {
        const int j = lowest_bit(mask);
        mask &= mask - 1;
        walk_item<LPS, OPS, EXACT, SHARED>(A, lds, readlane_i32(s, j), readlane_i32(base, j));
    }
}

// Chunk-parallel walk of the survivors in two launches (PHASE 1, then 2),
// one wave per work item (G positions of one survivor) over a persistent grid
// of W waves, wave w taking items w, w + W, w + 2W, ... of its phase; a wave
// tests the liveness of its next 64 items at once (one per lane) and walks
// the live ones in item order.
//   Phase 1: the first L0 chunks of every survivor in chunk-major order
//     (k = blk * n_surv + s), so that the early bisection positions, where
//     most edges that fail do fail, run first across the whole grid.
//   Phase 2 (survivors <= kFlatSurvivors, with the LDS reserved by the
//     launch): the remaining chunks of the survivors that phase 1 left
//     unresolved, numbered edge-major (survivor s owns numbers off[s] ..
//     off[s+1]-1), so the waves share them evenly; a lane finds its item's
//     survivor by binary search in the offsets.  L0 = W / n_surv (at least
//     1): phase 1 is one item per wave.
//   Without phase 2, phase 1 covers every chunk (L0 = all).
// motion_chunk_results_kernel then writes the survivors' results.
template <int LPS, int OPS, bool EXACT, bool SHARED, int PHASE>
__global__ void __launch_bounds__(1024) motion_chunk_kernel(MotionArgs A)
{
    extern __shared__ float4 lds[];
    // without a results kernel, phase 1 clears the counter set the next batch uses
    if (PHASE == 1 && A.reset_next && blockIdx.x == 0 && threadIdx.x == 0) {
        A.reset_next[0] = 0;
        A.reset_next[1] = 0;
    }
    // both counters are read before the staging, so the reads' latencies
    // overlap; the empty asm keeps the values in vector registers until then
    // (their move to scalar registers would wait for the reads there)
    int n_surv_v = *A.list_count, max_m = *A.max_m;
    stage_env<LPS == 7, link_pass(OPS) * OPS>(A.env, lds);
    asm volatile("" : "+v"(n_surv_v), "+v"(max_m));
    const int64_t n_surv = n_surv_v;
    constexpr int G = SHARED ? 64 / LPS : GroupShape<LPS, OPS>::STATES;
    const int lane = lane_id();
    const int64_t waves_total = (gridDim.x * blockDim.x) >> 6;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (n_surv == 0) return;
    const int64_t nblk = ((int64_t)max_m + G - 1) / G;
    const bool flat = A.flat && n_surv <= kFlatSurvivors && n_surv * nblk < (int64_t)INT_MAX;
    const int64_t L0 = flat ? max((int64_t)1, min(nblk, waves_total / n_surv)) : nblk;
    if (PHASE == 1) {
        const int64_t items1 = n_surv * L0;
        for (int64_t k0 = wave; k0 < items1; k0 += 64 * waves_total) {
            const int64_t k = k0 + lane * waves_total;
            int s = 0, base = 0;
            bool live = false;
            if (k < items1) {
                const int64_t blk = k / n_surv;
                s = (int)(k - blk * n_surv);
                base = 1 + (int)(blk * G);
                // flat: a wave holds at most one item (L0 <= waves / n_surv), or
                // L0 = 1 and none is dead (nd >= 2): walk_item reads the count
                // with its other reads; else drop the many dead items here
                live = flat || base < A.surv_nd[s];
            }
            walk_window<LPS, OPS, EXACT, SHARED>(A, lds, s, base, live);
        }
        return;
    }
    if (flat) {
        int32_t *off = reinterpret_cast<int32_t *>(reinterpret_cast<char *>(lds) + env_lds_bytes<LPS == 7, link_pass(OPS) * OPS>(A.env));
        scan_remaining_chunks<G>(A.surv_nd, A.first_bad, (int)n_surv, L0, off);
        const int64_t R = off[n_surv];
        for (int64_t r0 = wave; r0 < R; r0 += 64 * waves_total) {
            const int64_t r = r0 + lane * waves_total;
            int s = 0, base = 0;
            const bool live = r < R;
            if (live) {
                int lo = 0, hi = (int)n_surv; // off[lo] <= r < off[hi]
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (off[mid] <= r) lo = mid;
                    else hi = mid;
                }
                s = lo;
                base = 1 + (int)((L0 + (r - off[lo])) * G);
            }
            walk_window<LPS, OPS, EXACT, SHARED>(A, lds, s, base, live);
        }
    }
}

// Results of the chunk walk's survivors (one block, after both phases): the
// kernel boundary makes every block's first-invalid updates visible, which
// costs less than an agent-scope fence in each block of the walk.  Clears the
// counters for the next batch.
__global__ void __launch_bounds__(1024) motion_chunk_results_kernel(MotionArgs A)
{
    const int64_t n_surv = *A.list_count;
    for (int64_t sv = threadIdx.x; sv < n_surv; sv += blockDim.x) {
        const int64_t e = A.list[sv];
        const int fb = A.first_bad[sv];
        A.out[e] = fb == INT_MAX ? 1 : 0;
        if (A.checked) A.checked[e] = fb == INT_MAX ? A.surv_nd[sv] : fb + 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *A.list_count = 0;
        *A.max_m = 0;
    }
}

// Fused batch: each wave checks s2 of its edges (endpoint shape LE x OE), then
// walks its own survivors one after another (walk shape LW x OW).
template <int LE, int OE, int LW, int OW, bool EXACT>
__global__ void __launch_bounds__(1024) motion_fused_kernel(MotionArgs A)
{
    extern __shared__ float4 lds[];
    stage_env(A.env, lds);
    using Shape = GroupShape<LE, OE>;
    const int lane = lane_id();
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t e = wave * Shape::STATES + lane / Shape::GS;
    int m = 0;
    double a_unused[kDof], b_unused[kDof];
    float q[kDof];
    load_s2(A, e, e < A.n, q);
    const bool survivor = endpoint_check<LE, OE>(A, lds, e, e < A.n, q, m, a_unused, b_unused);
    uint64_t sm = wave_ballot(survivor);
    while (sm) {
        const int src = lowest_bit(sm);
        sm &= sm - 1;
        const int64_t es = wave * Shape::STATES + src / Shape::GS;
        double ea[kDof], eb[kDof];
#pragma unroll
        for (int j = 0; j < kDof; ++j) {
            ea[j] = A.a[(int64_t)j * A.n + es];
            eb[j] = A.b[(int64_t)j * A.n + es];
        }
        int total = 0;
        const int bad = walk_edge<LW, OW, EXACT>(ea, eb, A.lvs, A.factor, lds, A.env.ns, A.env.nb, 1, total);
        if (lane == 0) {
            A.out[es] = bad == INT_MAX ? 1 : 0;
            if (A.checked) A.checked[es] = bad == INT_MAX ? total : bad + 1;
        }
    }
}

template <int LPS, int OPS>
__global__ void __launch_bounds__(64) motion_single_kernel(EnvView env, Edge1 E, double lvs, unsigned factor,
                                                           int *result)
{
    extern __shared__ float4 lds[];
    stage_env<LPS == 7, link_pass(OPS) * OPS>(env, lds);
    int total = 0;
    const int bad = walk_edge<LPS, OPS, false>(E.a, E.b, lvs, factor, lds, env.ns, env.nb, 0, total);
    if (lane_id() == 0) *result = bad == INT_MAX ? 1 : 0; // bad < 0: too many segments, invalid
}

__global__ void __launch_bounds__(64) validate_single_kernel(EnvView env, State1 s, int *result)
{
    extern __shared__ float4 lds[];
    stage_env(env, lds);
    // one state over the whole wave: 8 sphere slots x 8 obstacle partitions
    using Shape = GroupShape<8, 8>;
    const int lane = lane_id();
    const Shape sh(lane);
    float q[kDof];
#pragma unroll
    for (int j = 0; j < kDof; ++j) q[j] = (float)s.q[j];
    LaneSpheres<8> S;
    arm_fk<8>(q, sh.slot, S);
    const bool hit = collide_env<8, 8, Exit::kAllGroups>(S, lds, env.ns, env.nb, true, sh.part);
    const uint64_t H = wave_ballot(hit);
    if (lane == 0) *result = H ? 0 : 1;
}

// Precomputed state lists (SpaceInformation::checkMotion(states, count[,
// firstInvalidStateIndex]), base/src/SpaceInformation.cpp:288-341), walked in
// chunks of G states in index order with the lowest-invalid exit, so the
// lowest invalid index of the first chunk holding one is the list's first
// invalid state.  Whether a list is valid does not depend on the order the
// reference visits the states in (all must be valid).  Offsets are clamped to
// [0, n_states], so any offsets stay in bounds.
//   state_lists_kernel: one wave per list checks the list's first chunk; a
//     list resolved there (an invalid state, or no states past the chunk)
//     gets its result, the others are appended (one device atomic pair per
//     block) to the unresolved lists, with first_bad = INT_MAX, and the
//     largest chunk count is kept.
//   state_list_chunks_kernel: the unresolved lists' remaining chunks, over a
//     persistent grid, chunk-major (chunk 1 of every list, then chunk 2, ...);
//     a chunk is skipped once an earlier state of its list is known invalid,
//     and an invalid one lowers first_bad (atomicMin).  A long valid list is
//     spread over many waves instead of walked by one.
//   state_list_results_kernel: the unresolved lists' results.
struct StateListArgs {
    EnvView env;
    const double *q; // dof-major [7][n_states]
    int64_t n_states;
    const int64_t *offsets;
    int64_t n_lists;
    uint8_t *out;
    int32_t *first_invalid;
    int32_t *unres;    // unresolved list ids
    int32_t *unres_fb; // their lowest invalid index found so far (INT_MAX: none)
    int32_t *counters; // [0] unresolved lists, [1] largest chunk count (zero at launch)
};

__device__ __forceinline__ void list_bounds(const StateListArgs &A, int64_t list, int64_t &lo, int64_t &hi)
{
    lo = min(max(A.offsets[list], (int64_t)0), A.n_states);
    hi = max(min(A.offsets[list + 1], A.n_states), lo);
}

// Checks states [from, to) (at most G of them, from < to) of one list;
// returns the index of the lowest invalid one, or -1.
template <int LPS, int OPS>
__device__ __forceinline__ int64_t check_list_chunk(const StateListArgs &A, const float4 *lds, int64_t from,
                                                    int64_t to)
{
    using Shape = GroupShape<LPS, OPS>;
    const Shape sh(lane_id());
    const int64_t st = from + sh.group;
    const bool live = st < to;
    float q[kDof];
    load_config(A.q, A.n_states, st, live, q, from);
    LaneSpheres<LPS> S;
    arm_fk<LPS, Shape::GS>(q, sh.slot, S);
    const bool hit = collide_env<LPS, OPS, Exit::kLowestGroup>(S, lds, A.env.ns, A.env.nb, live, sh.part);
    const uint64_t H = fold_groups<Shape::GS>(wave_ballot(hit && live));
    return H ? from + lowest_bit(H) / Shape::GS : -1;
}

template <int LPS, int OPS>
__global__ void __launch_bounds__(1024) state_lists_kernel(StateListArgs A)
{
    extern __shared__ float4 lds[];
    __shared__ int blk_count, blk_max, blk_base;
    if (threadIdx.x == 0) {
        blk_count = 0;
        blk_max = 0;
    }
    stage_env<LPS == 7, link_pass(OPS) * OPS>(A.env, lds); // ends with a barrier
    constexpr int G = GroupShape<LPS, OPS>::STATES;
    const int lane = lane_id();
    const int64_t list = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int64_t lo = 0, hi = 0, bad = -1;
    bool pending = false;
    if (list < A.n_lists) { // wave-uniform
        list_bounds(A, list, lo, hi);
        const int64_t end = min(hi, lo + G);
        if (lo < end) bad = check_list_chunk<LPS, OPS>(A, lds, lo, end);
        if (bad >= 0) bad -= lo;
        pending = bad < 0 && end < hi;
    }
    const int chunks = (int)((hi - lo + G - 1) / G);
    int woff = 0;
    if (pending && lane == 0) {
        woff = atomicAdd(&blk_count, 1);
        atomicMax(&blk_max, chunks);
    }
    __syncthreads();
    if (threadIdx.x == 0 && blk_count > 0) {
        blk_base = atomicAdd(&A.counters[0], blk_count);
        atomicMax(&A.counters[1], blk_max);
    }
    __syncthreads();
    if (lane == 0 && list < A.n_lists) {
        if (pending) {
            A.unres[blk_base + woff] = (int32_t)list;
            A.unres_fb[blk_base + woff] = INT_MAX;
        } else {
            A.out[list] = bad < 0 ? 1 : 0;
            if (A.first_invalid) A.first_invalid[list] = (int32_t)bad;
        }
    }
}

template <int LPS, int OPS>
__global__ void __launch_bounds__(1024) state_list_chunks_kernel(StateListArgs A)
{
    extern __shared__ float4 lds[];
    // the counters are read before the staging, so the reads overlap it
    int n_un_v = A.counters[0], maxc_v = A.counters[1];
    stage_env<LPS == 7, link_pass(OPS) * OPS>(A.env, lds);
    asm volatile("" : "+v"(n_un_v), "+v"(maxc_v));
    const int64_t n_un = n_un_v, maxc = maxc_v;
    if (n_un == 0 || maxc < 2) return;
    constexpr int G = GroupShape<LPS, OPS>::STATES;
    const int64_t waves_total = (gridDim.x * blockDim.x) >> 6;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t items = n_un * (maxc - 1);
    const int lane = lane_id();
    // a wave tests 64 items at once (one per lane: items k0 + lane * W) and
    // walks the live ones; most items of a chunk-major numbering are past
    // their list's last state
    for (int64_t k0 = wave; k0 < items; k0 += 64 * waves_total) {
        const int64_t k = k0 + lane * waves_total;
        int u = 0;
        int64_t lo = 0, hi = 0, base = 0;
        bool live = false;
        if (k < items) {
            const int64_t c = 1 + k / n_un;
            u = (int)(k - (c - 1) * n_un);
            list_bounds(A, A.unres[u], lo, hi);
            base = c * G; // relative to the list's first state
            live = lo + base < hi &&
                   __hip_atomic_load(&A.unres_fb[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= base;
        }
        uint64_t mask = wave_ballot(live);
        while (mask) {
            const int j = lowest_bit(mask);
            mask &= mask - 1;
            const int uj = readlane_i32(u, j);
            const int64_t loj = readlane_i64(lo, j), hij = readlane_i64(hi, j), bj = readlane_i64(base, j);
            if (__hip_atomic_load(&A.unres_fb[uj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < bj) continue;
            const int64_t bad = check_list_chunk<LPS, OPS>(A, lds, loj + bj, min(hij, loj + bj + G));
            if (bad >= 0 && lane == 0) atomicMin(&A.unres_fb[uj], (int)(bad - loj));
        }
    }
}

__global__ void __launch_bounds__(1024) state_list_results_kernel(StateListArgs A)
{
    const int n_un = A.counters[0];
    for (int u = threadIdx.x; u < n_un; u += blockDim.x) {
        const int64_t list = A.unres[u];
        const int fb = A.unres_fb[u];
        A.out[list] = fb == INT_MAX ? 1 : 0;
        if (A.first_invalid) A.first_invalid[list] = fb == INT_MAX ? -1 : fb;
    }
}

// Sequential lastValid overload (DiscreteMotionValidator.cpp:37-93):
// positions j = 1..nd-1 in order, then s2 as position nd; the lowest invalid
// position j gives lastValid (j-1)/nd (s2: (nd-1)/nd).  Walked in chunks of G
// positions like the state lists:
//   motion_last_valid_kernel: one wave per edge checks positions 1..G; an
//     edge resolved there (an invalid position, no positions past the chunk,
//     or too many segments) gets its result, the others are appended (one
//     device atomic pair per workgroup) with their segment count, endpoints
//     and first_bad = INT_MAX, and the largest chunk count is kept;
//   motion_last_valid_chunks_kernel: the unresolved edges' remaining chunks,
//     chunk-major over a persistent grid, skipping a chunk once an earlier
//     position of its edge is known invalid (atomicMin);
//   motion_last_valid_results_kernel: their results.
// Uses list, first_bad, surv_nd, surv_ab, list_count and max_m (the two
// counters zero at launch).

// Positions base .. base+G-1 (those <= last; `last` is s2) of one edge;
// returns the lowest invalid one, or INT_MAX.
template <int LPS, int OPS>
__device__ __forceinline__ int check_positions(const MotionArgs &A, const float4 *lds, const double a[kDof],
                                              const double b[kDof], int nd, int last, int base)
{
    using Shape = GroupShape<LPS, OPS>;
    const Shape sh(lane_id());
    const int p = base + sh.group;
    const bool live = p <= last;
    float q[kDof];
    if (p >= last) {
#pragma unroll
        for (int j = 0; j < kDof; ++j) q[j] = (float)b[j];
    } else {
        interpolate_f32<kDof>(a, b, (double)p / (double)nd, q);
    }
    LaneSpheres<LPS> S;
    arm_fk<LPS, Shape::GS>(q, sh.slot, S);
    const bool hit = collide_env<LPS, OPS, Exit::kLowestGroup>(S, lds, A.env.ns, A.env.nb, live, sh.part);
    const uint64_t H = fold_groups<Shape::GS>(wave_ballot(hit && live));
    return H ? base + lowest_bit(H) / Shape::GS : INT_MAX;
}

__device__ __forceinline__ void last_valid_result(const MotionArgs &A, double *last_t, int64_t e, int bad, int nd,
                                                  int last)
{
    if (bad == INT_MAX) {
        A.out[e] = 1;
    } else {
        A.out[e] = 0;
        // interior j fails: (j-1)/nd; s2 fails: (nd-1)/nd
        last_t[e] = bad < last ? (double)(bad - 1) / (double)nd : (double)(nd - 1) / (double)nd;
    }
    if (A.checked) A.checked[e] = bad == INT_MAX ? last : bad;
}

template <int LPS, int OPS>
__global__ void __launch_bounds__(1024) motion_last_valid_kernel(MotionArgs A, double *__restrict__ last_t)
{
    extern __shared__ float4 lds[];
    __shared__ int blk_count, blk_max, blk_base;
    if (threadIdx.x == 0) {
        blk_count = 0;
        blk_max = 0;
    }
    stage_env<LPS == 7, link_pass(OPS) * OPS>(A.env, lds); // ends with a barrier
    constexpr int G = GroupShape<LPS, OPS>::STATES;
    const int lane = lane_id();
    const int64_t e = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    double a[kDof], b[kDof];
    int nd = 0, last = 0;
    bool pending = false;
    if (e < A.n) { // wave-uniform
#pragma unroll
        for (int j = 0; j < kDof; ++j) {
            a[j] = A.a[(int64_t)j * A.n + e];
            b[j] = A.b[(int64_t)j * A.n + e];
        }
        nd = valid_segment_count<kDof>(a, b, A.lvs, A.factor);
        if (nd > kMaxSegments) { // reported invalid without a walk
            if (lane == 0) {
                A.out[e] = 0;
                if (A.checked) A.checked[e] = 0;
            }
        } else {
            last = nd > 1 ? nd : 1;
            const int bad = check_positions<LPS, OPS>(A, lds, a, b, nd, last, 1);
            pending = bad == INT_MAX && last > G;
            if (!pending && lane == 0) last_valid_result(A, last_t, e, bad, nd, last);
        }
    }
    int woff = 0;
    if (pending && lane == 0) {
        woff = atomicAdd(&blk_count, 1);
        atomicMax(&blk_max, (last - 1) / G + 1);
    }
    __syncthreads();
    if (threadIdx.x == 0 && blk_count > 0) {
        blk_base = atomicAdd(A.list_count, blk_count);
        atomicMax(A.max_m, blk_max);
    }
    __syncthreads();
    if (pending && lane == 0) {
        const int slot = blk_base + woff;
        A.list[slot] = (int32_t)e;
        A.first_bad[slot] = INT_MAX;
        A.surv_nd[slot] = nd;
        double *ab = A.surv_ab + (int64_t)slot * 2 * kDof;
#pragma unroll
        for (int j = 0; j < kDof; ++j) {
            ab[j] = a[j];
            ab[kDof + j] = b[j];
        }
    }
}

template <int LPS, int OPS>
__global__ void __launch_bounds__(1024) motion_last_valid_chunks_kernel(MotionArgs A)
{
    extern __shared__ float4 lds[];
    // the counters are read before the staging, so the reads overlap it
    int n_un_v = *A.list_count, maxc_v = *A.max_m;
    stage_env<LPS == 7, link_pass(OPS) * OPS>(A.env, lds);
    asm volatile("" : "+v"(n_un_v), "+v"(maxc_v));
    const int64_t n_un = n_un_v, maxc = maxc_v;
    if (n_un == 0 || maxc < 2) return;
    constexpr int G = GroupShape<LPS, OPS>::STATES;
    const int64_t waves_total = (gridDim.x * blockDim.x) >> 6;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t items = n_un * (maxc - 1);
    const int lane = lane_id();
    // a wave tests 64 items at once (one per lane: items k0 + lane * W) and
    // walks the live ones; most items of a chunk-major numbering are past
    // their edge's last position
    for (int64_t k0 = wave; k0 < items; k0 += 64 * waves_total) {
        const int64_t k = k0 + lane * waves_total;
        int u = 0, nd = 0, base = 0;
        bool live = false;
        if (k < items) {
            const int64_t c = 1 + k / n_un;
            u = (int)(k - (c - 1) * n_un);
            nd = A.surv_nd[u]; // >= 2: an unresolved edge has positions past its first chunk
            base = 1 + (int)(c * G);
            live = base <= nd &&
                   __hip_atomic_load(&A.first_bad[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= base;
        }
        uint64_t mask = wave_ballot(live);
        while (mask) {
            const int j = lowest_bit(mask);
            mask &= mask - 1;
            const int uj = readlane_i32(u, j), bj = readlane_i32(base, j), nj = readlane_i32(nd, j);
            if (__hip_atomic_load(&A.first_bad[uj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < bj) continue;
            const double *ab = A.surv_ab + (int64_t)uj * 2 * kDof;
            double a[kDof], b[kDof];
#pragma unroll
            for (int d = 0; d < kDof; ++d) {
                a[d] = ab[d];
                b[d] = ab[kDof + d];
            }
            const int bad = check_positions<LPS, OPS>(A, lds, a, b, nj, nj, bj);
            if (bad != INT_MAX && lane == 0) atomicMin(&A.first_bad[uj], bad);
        }
    }
}

__global__ void __launch_bounds__(1024) motion_last_valid_results_kernel(MotionArgs A, double *__restrict__ last_t)
{
    const int n_un = *A.list_count;
    for (int u = threadIdx.x; u < n_un; u += blockDim.x) {
        const int nd = A.surv_nd[u];
        last_valid_result(A, last_t, A.list[u], A.first_bad[u], nd, nd);
    }
}

// The link layout's staging of an environment (stage_env<true>) written to
// HBM, for kernels that read obstacles from HBM (the planning server).
__global__ void stage_links_kernel(EnvView e, float4 *__restrict__ out)
{
    const int nw = env_words(e);
    float *rr = reinterpret_cast<float *>(out + nw);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += gridDim.x * blockDim.x) {
        float4 o = e.obs[i];
        if (i < e.ns) {
            rr[i] = o.w;
            o.w = sphere_bound2(o.w);
        }
        out[i] = o;
    }
}

__global__ void zero_ints_kernel(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

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

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
// Dynamic LDS of a launch: the staged environment; for the link layout
// (lanes_per_state 7) the sphere and box lists padded to whole passes of
// link_pass(ops) * ops entries plus the spheres' exact rr (stage_env<true, ...>).
static size_t lds_bytes(const EnvView &e, int lps = 0, int ops = 1)
{
    if (lps != 7) return (size_t)(e.ns + 2 * e.nb) * sizeof(float4);
    const int pad = link_pass(ops) * ops;
    return (size_t)(pad_to(e.ns, pad) + 2 * pad_to(e.nb, pad)) * sizeof(float4) + (size_t)e.ns * sizeof(float);
}

// Where the link layout's staging exceeds one block's 64 KiB of LDS, the
// (8, 8) shape checks the states instead (same results).
static void fit_link_layout(const EnvView &e, int &lps, int &ops)
{
    if (lps == 7 && lds_bytes(e, 7, ops) > 65536) {
        lps = 8;
        ops = 8;
    }
}

static unsigned grid_for(int64_t units, int64_t units_per_block)
{
    return (unsigned)((units + units_per_block - 1) / units_per_block);
}

template <typename T>
void launch_validate(const EnvView &env, const T *q, int64_t n, uint8_t *out, int lps, int ops, int block,
                     hipStream_t s, bool shared)
{
    if (n <= 0) return;
    fit_link_layout(env, lps, ops);
    ProfScope ps("validate_states", s);
    if (shared && ops > 1) {
#define OA_LAUNCH(L, O)                                                                                             \
    if (lps == L && ops == O) {                                                                                     \
        const int64_t per = (int64_t)(block / 64) * (64 / L);                                                     \
        hipLaunchKernelGGL((validate_shared_kernel<L, O, T>), dim3(grid_for(n, per)), dim3(block), lds_bytes(env, L, O), \
                           s, env, q, n, out);                                                                      \
        OA_HIP(hipGetLastError());                                                                                  \
        return;                                                                                                     \
    }
        OA_SHARED_SHAPES(OA_LAUNCH)
#undef OA_LAUNCH
    }
#define OA_LAUNCH(L, O)                                                                                             \
    if (lps == L && ops == O) {                                                                                     \
        const int64_t per = (int64_t)(block / 64) * GroupShape<L, O>::STATES;                                     \
        hipLaunchKernelGGL((validate_kernel<L, O, T>), dim3(grid_for(n, per)), dim3(block), lds_bytes(env, L, O), s, env, \
                           q, n, out);                                                                              \
        OA_HIP(hipGetLastError());                                                                                  \
        return;                                                                                                     \
    }
    OA_SHAPES(OA_LAUNCH)
#undef OA_LAUNCH
    fail(OMPL_AMD_ERR_ARG, "validity shape (%d, %d) not built", lps, ops);
}
template void launch_validate<float>(const EnvView &, const float *, int64_t, uint8_t *, int, int, int,
                                     hipStream_t, bool);
template void launch_validate<double>(const EnvView &, const double *, int64_t, uint8_t *, int, int, int,
                                      hipStream_t, bool);

size_t links_bytes(const EnvView &env)
{
    return (size_t)(env.ns + 2 * env.nb) * sizeof(float4) + (size_t)env.ns * sizeof(float);
}

void launch_stage_links(const EnvView &env, float4 *out, hipStream_t s)
{
    const int nw = env.ns + 2 * env.nb;
    if (nw == 0) return;
    hipLaunchKernelGGL(stage_links_kernel, dim3((nw + 255) / 256), dim3(256), 0, s, env, out);
    OA_HIP(hipGetLastError());
}

void launch_zero(int *p, int n, hipStream_t s)
{
    hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    OA_HIP(hipGetLastError());
}

void launch_motion_endpoint(const MotionArgs &A, int lps, int ops, int block, hipStream_t s, bool shared)
{
    if (A.n <= 0) return;
    fit_link_layout(A.env, lps, ops);
    ProfScope ps("motion_endpoint", s);
    if (shared && ops > 1) {
#define OA_LAUNCH(L, O)                                                                                             \
    if (lps == L && ops == O) {                                                                                     \
        const int64_t per = (int64_t)(block / 64) * (64 / L);                                                     \
        hipLaunchKernelGGL((motion_endpoint_shared_kernel<L, O>), dim3(grid_for(A.n, per)), dim3(block),          \
                           lds_bytes(A.env, L, O), s, A);                                                                \
        OA_HIP(hipGetLastError());                                                                                  \
        return;                                                                                                     \
    }
        OA_SHARED_SHAPES(OA_LAUNCH)
#undef OA_LAUNCH
    }
#define OA_LAUNCH(L, O)                                                                                             \
    if (lps == L && ops == O) {                                                                                     \
        const int64_t per = (int64_t)(block / 64) * GroupShape<L, O>::STATES;                                     \
        hipLaunchKernelGGL((motion_endpoint_kernel<L, O>), dim3(grid_for(A.n, per)), dim3(block),                 \
                           lds_bytes(A.env, L, O), s, A);                                                                \
        OA_HIP(hipGetLastError());                                                                                  \
        return;                                                                                                     \
    }
    OA_SHAPES(OA_LAUNCH)
#undef OA_LAUNCH
    fail(OMPL_AMD_ERR_ARG, "endpoint shape (%d, %d) not built", lps, ops);
}

void launch_motion_walk(const MotionArgs &A, bool exact, int lps, int ops, int block, int64_t waves, hipStream_t s)
{
    if (waves <= 0) return;
    fit_link_layout(A.env, lps, ops);
    ProfScope ps("motion_walk", s);
    const unsigned g = grid_for(waves, block / 64);
#define OA_LAUNCH(L, O)                                                                                             \
    if (lps == L && ops == O) {                                                                                     \
        if (exact)                                                                                                  \
            hipLaunchKernelGGL((motion_walk_kernel<L, O, true>), dim3(g), dim3(block), lds_bytes(A.env, L, O), s, A);   \
        else                                                                                                        \
            hipLaunchKernelGGL((motion_walk_kernel<L, O, false>), dim3(g), dim3(block), lds_bytes(A.env, L, O), s, A);  \
        OA_HIP(hipGetLastError());                                                                                  \
        return;                                                                                                     \
    }
    OA_SHAPES(OA_LAUNCH)
#undef OA_LAUNCH
    fail(OMPL_AMD_ERR_ARG, "walk shape (%d, %d) not built", lps, ops);
}

void launch_motion_chunks(const MotionArgs &A, bool exact, int lps, int ops, int block, int64_t waves,
                          hipStream_t s, bool shared, bool results)
{
    fit_link_layout(A.env, lps, ops);
    const unsigned g = grid_for(waves, block / 64);
    // LDS after the environment for the survivors' chunk offsets (edge-major
    // phase), where it fits beside the kernel's 4 KiB of static LDS
    MotionArgs B = A;
    const size_t env_b = lds_bytes(A.env, lps, ops), flat_b = (size_t)(kFlatSurvivors + 1) * sizeof(int32_t);
    B.flat = env_b + flat_b + 4096 + 64 <= 65536;
    const size_t dyn = env_b + (B.flat ? flat_b : 0);
    // phase 1 (early chunks), phase 2 (the rest), then (checked counts, or a
    // captured graph) the results; one profiling scope around them (a scope
    // adds its own dispatch latency)
    auto launch = [&](auto early, auto rest) {
        ProfScope ps("motion_chunk", s);
        hipLaunchKernelGGL(early, dim3(g), dim3(block), dyn, s, B);
        hipLaunchKernelGGL(rest, dim3(g), dim3(block), dyn, s, B);
        if (results) hipLaunchKernelGGL(motion_chunk_results_kernel, dim3(1), dim3(1024), 0, s, B);
    };

    if (shared && ops > 1) {
#define OA_LAUNCH(L, O)                                                                                             \
    if (lps == L && ops == O) {                                                                                     \
        if (exact)                                                                                                  \
            launch(motion_chunk_kernel<L, O, true, true, 1>, motion_chunk_kernel<L, O, true, true, 2>);            \
        else                                                                                                        \
            launch(motion_chunk_kernel<L, O, false, true, 1>, motion_chunk_kernel<L, O, false, true, 2>);          \
        OA_HIP(hipGetLastError());                                                                                  \
        return;                                                                                                     \
    }
        OA_SHARED_SHAPES(OA_LAUNCH)
#undef OA_LAUNCH
    }
#define OA_LAUNCH(L, O)                                                                                             \
    if (lps == L && ops == O) {                                                                                     \
        if (exact)                                                                                                  \
            launch(motion_chunk_kernel<L, O, true, false, 1>, motion_chunk_kernel<L, O, true, false, 2>);          \
        else                                                                                                        \
            launch(motion_chunk_kernel<L, O, false, false, 1>, motion_chunk_kernel<L, O, false, false, 2>);        \
        OA_HIP(hipGetLastError());                                                                                  \
        return;                                                                                                     \
    }
    OA_SHAPES(OA_LAUNCH)
#undef OA_LAUNCH
    fail(OMPL_AMD_ERR_ARG, "chunk shape (%d, %d) not built", lps, ops);
}

bool fused_supported(int le, int oe, int lw, int ow)
{
    const bool e_ok = le == 8 && (oe == 1 || oe == 2 || oe == 4);
    const bool w_ok = (lw == 1 && ow == 1) || (lw == 8 && (ow == 1 || ow == 8));
    return e_ok && w_ok;
}

void launch_motion_fused(const MotionArgs &A, bool exact, int le, int oe, int lw, int ow, int block, hipStream_t s)
{
    if (A.n <= 0) return;
    ProfScope ps("motion_fused", s);
#define OA_FUSED_W(LE, OE, LW, OW)                                                                                  \
    if (oe == OE && lw == LW && ow == OW) {                                                                         \
        const int64_t per = (int64_t)(block / 64) * GroupShape<LE, OE>::STATES;                                   \
        if (exact)                                                                                                  \
            hipLaunchKernelGGL((motion_fused_kernel<LE, OE, LW, OW, true>), dim3(grid_for(A.n, per)), dim3(block), \
                               lds_bytes(A.env), s, A);                                                             \
        else                                                                                                        \
            hipLaunchKernelGGL((motion_fused_kernel<LE, OE, LW, OW, false>), dim3(grid_for(A.n, per)),            \
                               dim3(block), lds_bytes(A.env), s, A);                                                \
        OA_HIP(hipGetLastError());                                                                                  \
        return;                                                                                                     \
    }
#define OA_FUSED_E(OE) OA_FUSED_W(8, OE, 1, 1) OA_FUSED_W(8, OE, 8, 1) OA_FUSED_W(8, OE, 8, 8)
    if (le == 8) {
        OA_FUSED_E(1)
        OA_FUSED_E(2)
        OA_FUSED_E(4)
    }
#undef OA_FUSED_E
#undef OA_FUSED_W
    fail(OMPL_AMD_ERR_ARG, "fused shape (%d,%d | %d,%d) not built", le, oe, lw, ow);
}

void launch_motion_single(const EnvView &env, const Edge1 &E, double lvs, unsigned factor, int *result, hipStream_t s)
{
    hipLaunchKernelGGL((motion_single_kernel<1, 1>), dim3(1), dim3(64), lds_bytes(env), s, env, E, lvs, factor,
                       result);
    OA_HIP(hipGetLastError());
}

void launch_validate_single(const EnvView &env, const State1 &st, int *result, hipStream_t s)
{
    hipLaunchKernelGGL(validate_single_kernel, dim3(1), dim3(64), lds_bytes(env), s, env, st, result);
    OA_HIP(hipGetLastError());
}

void launch_motion_last_valid(const MotionArgs &A, double *last_t, int lps, int ops, int64_t waves, hipStream_t s)
{
    if (A.n <= 0) return;
    fit_link_layout(A.env, lps, ops);
    const unsigned g2 = grid_for(waves, 4);
    launch_zero(A.list_count, 2, s); // list_count, then max_m
#define OA_LAUNCH(L, O)                                                                                             \
    if (lps == L && ops == O) {                                                                                     \
        const size_t lds = lds_bytes(A.env, L, O);                                                                  \
        hipLaunchKernelGGL((motion_last_valid_kernel<L, O>), dim3(grid_for(A.n, 16)), dim3(1024), lds, s, A,      \
                           last_t);                                                                                 \
        hipLaunchKernelGGL((motion_last_valid_chunks_kernel<L, O>), dim3(g2), dim3(256), lds, s, A);               \
        hipLaunchKernelGGL(motion_last_valid_results_kernel, dim3(1), dim3(1024), 0, s, A, last_t);                \
        OA_HIP(hipGetLastError());                                                                                  \
        return;                                                                                                     \
    }
    OA_SHAPES(OA_LAUNCH)
#undef OA_LAUNCH
    fail(OMPL_AMD_ERR_ARG, "lastValid shape (%d, %d) not built", lps, ops);
}

void launch_state_lists(const EnvView &env, const double *q_soa, int64_t n_states, const int64_t *offsets,
                        int64_t n_lists, uint8_t *out, int32_t *first_invalid, int lps, int ops, int64_t waves,
                        int32_t *unres, int32_t *unres_fb, int32_t *counters, hipStream_t s)
{
    if (n_lists <= 0) return;
    fit_link_layout(env, lps, ops);
    StateListArgs A{env, q_soa, n_states, offsets, n_lists, out, first_invalid, unres, unres_fb, counters};
    const unsigned g2 = grid_for(waves, 4);
    ProfScope ps("state_lists", s);
    launch_zero(counters, 2, s);
#define OA_LAUNCH(L, O)                                                                                             \
    if (lps == L && ops == O) {                                                                                     \
        const size_t lds = lds_bytes(env, L, O);                                                                    \
        hipLaunchKernelGGL((state_lists_kernel<L, O>), dim3(grid_for(n_lists, 16)), dim3(1024), lds, s, A);        \
        hipLaunchKernelGGL((state_list_chunks_kernel<L, O>), dim3(g2), dim3(256), lds, s, A);                      \
        hipLaunchKernelGGL(state_list_results_kernel, dim3(1), dim3(1024), 0, s, A);                               \
        OA_HIP(hipGetLastError());                                                                                  \
        return;                                                                                                     \
    }
    OA_SHAPES(OA_LAUNCH)
#undef OA_LAUNCH
    fail(OMPL_AMD_ERR_ARG, "state-list shape (%d, %d) not built", lps, ops);
}

void launch_arm_spheres(const float *q, int64_t n, float *out, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(arm_spheres_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, q, n, out);
    OA_HIP(hipGetLastError());
}

} // namespace ompl_amd
