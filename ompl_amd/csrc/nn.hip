// Exact k-nearest and radius neighbour search for ompl::NearestNeighbors<_T>
// (datastructures/NearestNeighbors.h:46-115), brute force over a device-resident
// point set, with the semantics of NearestNeighborsLinear
// (datastructures/NearestNeighborsLinear.h:98-142): results sorted by distance,
// radius inclusive (d <= r), ties broken by lower index (a project rule: the
// reference leaves equal keys in unspecified order).
//
// Distances (FP32, the oracle's operation order, element minus query):
//   L2:  sqrt(sum_i (e_i - q_i)^2) accumulated from 0 in index order
//        (base/spaces/src/RealVectorStateSpace.cpp:230-242)
//   SE3: L2 over (x,y,z) + acos(|q_e . q_q|) (0 when |dot| > 1), weights 1, 1
//        (base/src/StateSpace.cpp:1074-1082, base/spaces/src/SO3StateSpace.cpp:253-262)
//
// Scheme (one wave = 32 queries, one DB tile = 32 points):
//   1. MFMA v_mfma_f32_32x32x2f32 on augmented vectors gives an approximate
//      squared distance for all 32x32 pairs:  x_aug = (x, 1, |x|^2),
//      q_aug = (-2q, |q|^2, 1)  ->  x_aug . q_aug = |q|^2 + |x|^2 - 2 q.x
//      (SE3: the translation part only, a lower bound of the full distance).
//   2. Each lane owns one query (column) and 16 of the tile's points (rows);
//      a value passes if approx <= thr, thr = kth^2 (1 + 8e-6) + M with the
//      rigorous FP32 bound M = 64 u (|q|^2 + max|x|^2) (u = 2^-24) on the
//      cancellation error, so no point of the exact result is ever dropped.
//   3. Passing points get the exact FP32 distance (the oracle's formula) and
//      enter the lane's sorted register top-k; the two lanes of a query merge
//      their lists at the end (wave-level merge), DB splits merge in a second
//      kernel.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <vector>

#include "fp32_math.hpp"
#include "runtime.hpp"

namespace ompl_amd {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr float kPadNorm = 1e30f;
constexpr float kUnit = 5.9604644775390625e-8f; // 2^-24

struct NNIndex {
    int metric = 0;
    int dim = 0;     // coordinates per point
    int tdim = 0;    // coordinates in the MFMA part (dim for L2, 3 for SE3)
    int ksteps = 0;  // MFMA K steps of 2 (augmented length tdim + 2, padded)
    int64_t n = 0, cap = 0;
    float *pts = nullptr;      // [cap][dim]
    float *xt = nullptr;       // [cap/32][ksteps][64] tiled augmented points
    uint8_t *removed = nullptr;
    uint32_t *xn_max = nullptr; // max |x|^2 as FP32 bits (non-negative)
    int64_t n_removed = 0;
    std::vector<uint8_t> removed_host;
    // per-call scratch (grow-only)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
};

} // namespace ompl_amd

struct ompl_amd_nn_s {
    ompl_amd::NNIndex idx;
};

namespace ompl_amd {

__device__ __forceinline__ float l2_exact(const float *__restrict__ e, const float *q, int dim)
{
    float dist = 0.0f;
    for (int i = 0; i < dim; ++i) {
        const float diff = e[i] - q[i];
        dist += diff * diff;
    }
    return __builtin_sqrtf(dist);
}

template <int METRIC, int DIM>
__device__ __forceinline__ float exact_dist(const float *__restrict__ e, const float q[DIM])
{
    if (METRIC == OMPL_AMD_METRIC_L2) {
        float dist = 0.0f;
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            const float diff = e[i] - q[i];
            dist += diff * diff;
        }
        return __builtin_sqrtf(dist);
    } else {
        float t = 0.0f;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float diff = e[i] - q[i];
            t += diff * diff;
        }
        const float dt = __builtin_sqrtf(t);
        const float dq = fabsf(((e[3] * q[3] + e[4] * q[4]) + e[5] * q[5]) + e[6] * q[6]);
        const float dr = dq > 1.0f ? 0.0f : acos_as(dq);
        float dist = 0.0f;
        dist += 1.0f * dt;
        dist += 1.0f * dr;
        return dist;
    }
}

// SE3: true when the point's distance is certainly above `bound` (a finite
// distance), from a lower bound that skips the acos: acos(x) >= sqrt(2 (1 - x))
// on [0, 1] (2 sin(t / 2) <= t), so d >= d_t + sqrt(2 (1 - |q.x|)).  The margin
// (relative 1e-5 plus 1e-5) covers the FP32 rounding of both sides and
// acos_as's error, so a point this skips is one whose exact distance is above
// the bound.  Other metrics: never.
template <int METRIC, int DIM>
__device__ __forceinline__ bool certainly_above(const float *__restrict__ e, const float q[DIM], float bound)
{
    if (METRIC != OMPL_AMD_METRIC_SE3) return false;
    float t = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float diff = e[i] - q[i];
        t += diff * diff;
    }
    const float dq = fabsf(((e[3] * q[3] + e[4] * q[4]) + e[5] * q[5]) + e[6] * q[6]);
    const float lb = __builtin_amdgcn_sqrtf(t) + __builtin_amdgcn_sqrtf(2.0f * (1.0f - dq));
    return lb > bound * (1.0f + 1e-5f) + 1e-5f; // false for NaN (dq > 1) or an infinite bound
}

struct Key {
    float d;
    int i;
};

__device__ __forceinline__ bool key_less(float da, int ia, float db, int ib)
{
    return da < db || (da == db && ia < ib);
}

template <int KMAX>
struct TopK {
    float d[KMAX];
    int i[KMAX];
    __device__ __forceinline__ void init()
    {
#pragma unroll
        for (int s = 0; s < KMAX; ++s) {
            d[s] = INFINITY;
            i[s] = INT_MAX;
        }
    }
    __device__ __forceinline__ void insert(float kd, int ki)
    {
#pragma unroll
        for (int s = KMAX - 1; s >= 0; --s) {
            const bool lt = key_less(kd, ki, d[s], i[s]);
            const bool lt_prev = s > 0 ? key_less(kd, ki, d[s > 0 ? s - 1 : 0], i[s > 0 ? s - 1 : 0]) : false;
            if (lt) {
                d[s] = lt_prev ? d[s > 0 ? s - 1 : 0] : kd;
                i[s] = lt_prev ? i[s > 0 ? s - 1 : 0] : ki;
            }
        }
    }
    __device__ __forceinline__ void kth(int k, float &kd, int &ki) const
    {
        kd = d[0];
        ki = i[0];
#pragma unroll
        for (int s = 1; s < KMAX; ++s)
            if (s == k - 1) {
                kd = d[s];
                ki = i[s];
            }
    }
};

struct KnnArgs {
    const float *pts;
    const float *xt;
    const uint8_t *removed;
    const uint32_t *xn_max;
    int64_t n;
    int64_t n_tiles;
    const float *q; // [nq][dim]
    int64_t nq;
    int k;
    int splits;
    int64_t tiles_per_split;
    float *out_d;    // final [nq][k] or partial [split][nq][k]
    int32_t *out_i;
    float radius;    // radius search
    int64_t *counts; // radius pass 0 / 2: per-query counts; pass 1: per-query cursors
    const int64_t *offsets;
    int slot_cap;    // radius pass 2: keys kept per query in its fixed slot
    // k-NN: splits searched by one workgroup (one wave each, same query
    // tile), and the bound slots its lanes share per query
    int group_waves; // W
    // packed batches: per-query k (clamped to [0, k]; k is then the largest)
    // with results at out_off[q] (one split), per-query radius
    const int32_t *kq;
    const int64_t *out_off;
    const float *rq;
};

// k-NN lanes exchange their bounds after each of a split's first 4 tiles,
// then every 8th tile up to tile 64, then every 64th: bounds tighten fastest
// early, and an exchange costs about as much as filtering a tile.
__device__ __forceinline__ bool share_tile(int64_t tj)
{
    return tj != 0 && (tj < 4 || (tj < 64 ? (tj & 7) == 0 : (tj & 63) == 0));
}

// Words per point row of a wave's tile in LDS: the coordinates padded to a
// multiple of 4, so a candidate's row is read with 16-byte LDS reads.
__host__ __device__ constexpr int knn_row_words(int dim)
{
    return (dim + 3) & ~3;
}

// Dynamic LDS of a k-NN workgroup of W waves: each wave's current tile, then
// the bound slots (M <= 2 W per query, 32 queries).
__host__ __device__ constexpr size_t knn_lds_bytes(int dim, int w)
{
    return (size_t)w * 32 * knn_row_words(dim) * sizeof(float) + (size_t)32 * 2 * w * sizeof(unsigned long long);
}

// Waves of one k-NN workgroup at most: 16 (1,024 threads, 4 waves per SIMD)
// for lists of up to 16 keys, 4 for the longer lists (more registers).
template <int KMAX>
struct KnnGroup {
    static constexpr int kMaxWaves = KMAX <= 16 ? 16 : 4;
};

__device__ __forceinline__ unsigned long long pack_key(float d, int i)
{
    return ((unsigned long long)__float_as_uint(d) << 32) | (uint32_t)i;
}

// Row of the 32x32 MFMA output held in element j of a lane in half h.
__device__ __forceinline__ int mfma_row(int j, int h)
{
    return ((j >> 2) << 3) + (h << 2) + (j & 3);
}

// Shared prologue: the lane's query and its MFMA operand registers.
template <int METRIC, int DIM, int KSTEPS>
struct QueryLane {
    float q[DIM];
    float b[KSTEPS];
    float margin; // M
    bool live;
    int64_t qi;
    __device__ __forceinline__ void load(const KnnArgs &A, int64_t qtile, int lane)
    {
        const int col = lane & 31, h = lane >> 5;
        qi = qtile * 32 + col;
        live = qi < A.nq;
#pragma unroll
        for (int i = 0; i < DIM; ++i) q[i] = live ? A.q[qi * DIM + i] : 0.0f;
        constexpr int TD = METRIC == OMPL_AMD_METRIC_L2 ? DIM : 3;
        float qn = 0.0f;
#pragma unroll
        for (int i = 0; i < TD; ++i) qn += q[i] * q[i];
        // operand row kk of the augmented query (-2 q, |q|^2, 1, 0 ...); the
        // lane's half picks the odd or even row of each K step
        auto row_of = [&](int kk) {
            if (kk < TD) return -2.0f * q[kk];
            if (kk == TD) return qn;
            if (kk == TD + 1) return 1.0f;
            return 0.0f;
        };
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s) b[s] = h ? row_of(2 * s + 1) : row_of(2 * s);
        const float xm = __uint_as_float(*A.xn_max);
        margin = 64.0f * kUnit * (qn + xm);
    }
};

template <int KSTEPS>
__device__ __forceinline__ f32x16 tile_mma(const float *__restrict__ xt, int64_t t, int lane, const float *b)
{
    f32x16 acc = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    const float *base = xt + t * (KSTEPS * 64) + lane;
#pragma unroll
    for (int s = 0; s < KSTEPS; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(base[s * 64], b[s], acc, 0, 0, 0);
    return acc;
}

// The lane's A operands of tile t (one float per K step), and the MFMA chain
// on operands loaded beforehand.
template <int KSTEPS>
__device__ __forceinline__ void tile_load(const float *__restrict__ xt, int64_t t, int lane, float v[KSTEPS])
{
    const float *base = xt + t * (KSTEPS * 64) + lane;
#pragma unroll
    for (int s = 0; s < KSTEPS; ++s) v[s] = base[s * 64];
}

template <int KSTEPS>
__device__ __forceinline__ f32x16 tile_mma_regs(const float v[KSTEPS], const float *b)
{
    f32x16 acc = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < KSTEPS; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(v[s], b[s], acc, 0, 0, 0);
    return acc;
}

// Orders the wave's LDS accesses: its lanes exchange data through the wave's
// own LDS slice, so only the wave itself needs to be ordered.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// A lane's k best keys ((dist bits << 32) | id), ascending, in KMAX slots:
// the first KMAX - k slots hold 0, which no key is below, so the k-th key is
// always the last slot; an empty slot holds ~0 (above every key).
template <int KMAX>
struct TopKey {
    unsigned long long key[KMAX];
    __device__ __forceinline__ void init(int k)
    {
#pragma unroll
        for (int s = 0; s < KMAX; ++s) key[s] = s < KMAX - k ? 0ull : ~0ull;
    }
    __device__ __forceinline__ void insert(unsigned long long nk)
    {
        // the list is ascending, so slot s of the new list is
        // max(old[s - 1], min(old[s], nk)): old[s] below nk stays, the first
        // slot at or above nk takes nk, later slots take their predecessor,
        // and the largest key leaves.  The slots depend only on the old list,
        // not on each other (no carried chain).
#pragma unroll
        for (int s = KMAX - 1; s > 0; --s) {
            const unsigned long long lo = key[s] < nk ? key[s] : nk;
            key[s] = key[s - 1] > lo ? key[s - 1] : lo;
        }
        key[0] = key[0] < nk ? key[0] : nk;
    }
    __device__ __forceinline__ unsigned long long kth() const { return key[KMAX - 1]; }
};

__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int m)
{
    const uint32_t lo = __shfl_xor((uint32_t)v, m), hi = __shfl_xor((uint32_t)(v >> 32), m);
    return ((unsigned long long)hi << 32) | lo;
}

// The filter threshold for k-th key `kth`: its squared distance, widened by
// the relative rounding of the squared distance and the MFMA error bound.
__device__ __forceinline__ float filter_threshold(unsigned long long kth, float margin)
{
    const float kd = __uint_as_float((uint32_t)(kth >> 32));
    return kd * kd * (1.0f + 8e-6f) + margin;
}

// One workgroup = one query tile (32 queries) x W consecutive splits of the
// point tiles, one wave per split.
template <int METRIC, int DIM, int KSTEPS, int KMAX>
__global__ void __launch_bounds__(64 * KnnGroup<KMAX>::kMaxWaves, (KMAX <= 16 ? 4 : (KMAX <= 32 ? 3 : 2)))
knn_kernel(KnnArgs A)
{
    constexpr int MW = KnnGroup<KMAX>::kMaxWaves;
    // the current tile's 32 points, per wave (candidates that pass the MFMA
    // filter read their exact coordinates from here, not from HBM), then the
    // bound slots of the workgroup's 32 queries, M per query
    extern __shared__ __align__(16) unsigned char knn_smem[];
    const int lane = threadIdx.x & 63;
    const int h = lane >> 5;
    // the wave's query tile and split are wave-uniform (scalar registers), so
    // is every tile address below
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr int ROW = knn_row_words(DIM);
    float *tp = reinterpret_cast<float *>(knn_smem) + wid * (32 * ROW);
    unsigned long long *slots =
        reinterpret_cast<unsigned long long *>(knn_smem + (size_t)A.group_waves * 32 * ROW * sizeof(float));
    const int64_t qtiles = (A.nq + 31) / 32;
    const int64_t qtile = blockIdx.x % qtiles;
    const int split = (int)(blockIdx.x / qtiles) * A.group_waves + wid;
    const int SW = 2 * A.group_waves; // slots per query in LDS
    for (int i = threadIdx.x; i < 32 * SW; i += blockDim.x) slots[i] = ~0ull;
    __syncthreads();
    QueryLane<METRIC, DIM, KSTEPS> Q;
    Q.load(A, qtile, lane);
    // the query's k: the batch's, or its own in a packed batch (a query with
    // k = 0 is not searched)
    int kq = A.k;
    if (A.kq && Q.live) {
        const int v = A.kq[Q.qi];
        kq = v < 0 ? 0 : (v > A.k ? A.k : v);
    }
    if (kq == 0) Q.live = false;
    // M = min(k, 2 W) slots per query, each lane publishing its c-th smallest
    // key, c = ceil(k / M)
    const int M = kq < 1 ? 1 : (kq < SW ? kq : SW);
    const int crank = (kq + M - 1) / M;
    TopKey<KMAX> top;
    top.init(kq);
    // bkey: a key that at least k distinct points of the query are at or
    // below, so no point of its k nearest is above it: a point is kept only
    // below it, and the filter threshold follows it.  Sources: the lane's own
    // k-th key, its partner lane's bound (the query's other half of each
    // tile), and the query's slots in the workgroup: its 2 W lanes are dealt
    // round-robin over M slots, each lane publishes its c-th smallest key
    // (c M >= k) into its slot by atomic minimum, and once every slot holds
    // a key their largest is a bound (c keys at or below it in each slot's
    // lane, lanes disjoint).  A key equal to a bound is one that lane holds
    // itself, and no other lane sees that point.
    float thr = INFINITY;
    unsigned long long bkey = ~0ull, published = ~0ull;
    unsigned long long *qslots = slots + (lane & 31) * SW;
    const int my_slot = (2 * wid + h) % M;
    const int64_t t0 = (int64_t)split * A.tiles_per_split;
    const int64_t t1 = t0 + A.tiles_per_split < A.n_tiles ? t0 + A.tiles_per_split : A.n_tiles;
    // A tile's exact coordinates (32 x DIM words, PW per lane) and removal
    // flags are read one tile ahead with its MFMA operands, so a tile with
    // candidates stages them in LDS without waiting on a global read.  The
    // point and flag arrays are allocated in whole tiles, so every read of a
    // tile is in bounds; rows past the last point count as removed.
    constexpr int PW = (32 * DIM + 63) / 64;
    auto pts_load = [&](int64_t t, float v[PW], uint32_t &rm) {
        const float *tile = A.pts + t * (32 * DIM);
#pragma unroll
        for (int w = 0; w < PW; ++w) {
            const int i = lane + 64 * w;
            v[w] = (w + 1 < PW || i < 32 * DIM) ? tile[i < 32 * DIM ? i : 0] : 0.0f;
        }
        rm = A.removed[t * 32 + (lane & 31)];
    };
    // LDS word of each coordinate a lane stages (rows padded to ROW words)
    int soff[PW];
#pragma unroll
    for (int w = 0; w < PW; ++w) {
        const int i = lane + 64 * w;
        soff[w] = (i / DIM) * ROW + i % DIM;
    }
    float cur[KSTEPS], cpts[PW];
    uint32_t crm = 1;
    if (t0 < t1) {
        tile_load<KSTEPS>(A.xt, t0, lane, cur);
        pts_load(t0, cpts, crm);
    }
    for (int64_t t = t0; t < t1; ++t) {
        if (share_tile(t - t0)) { // wave-uniform
            unsigned long long best = bkey;
            const unsigned long long partner = shfl_xor_u64(best, 32);
            if (partner < best) best = partner;
            // publish this lane's c-th smallest key (~0 while it has fewer)
            const int at = KMAX - kq + crank - 1;
            unsigned long long mine = ~0ull;
#pragma unroll
            for (int s = 0; s < KMAX; ++s)
                if (s == at) mine = top.key[s];
            if (Q.live && mine < published) {
                atomicMin(qslots + my_slot, mine);
                published = mine;
            }
            // the slots' largest key, once every slot holds one (slots are
            // only lowered, so any value read is one some lane published)
            unsigned long long b = 0;
#pragma unroll
            for (int s = 0; s < 2 * MW; ++s) {
                if (s < M) { // M of this lane's query
                    const unsigned long long v =
                        __hip_atomic_load(qslots + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    b = v > b ? v : b;
                }
            }
            if (b < best) best = b;
            if (best < bkey) {
                bkey = best;
                thr = filter_threshold(best, Q.margin);
            }
        }
        // the next tile's operands are in flight while this one is filtered
        float nxt[KSTEPS], npts[PW];
        uint32_t nrm;
        const int64_t tn = t + 1 < t1 ? t + 1 : t;
        tile_load<KSTEPS>(A.xt, tn, lane, nxt);
        pts_load(tn, npts, nrm);
        const f32x16 acc = tile_mma_regs<KSTEPS>(cur, Q.b);
#pragma unroll
        for (int s2 = 0; s2 < KSTEPS; ++s2) cur[s2] = nxt[s2];
        // any value of the lane at or below the threshold?
        float lo = fminf(fminf(acc[0], acc[1]), acc[2]);
#pragma unroll
        for (int j = 3; j < 15; j += 2) lo = fminf(fminf(lo, acc[j]), acc[j + 1]);
        lo = fminf(lo, acc[15]);
        if (__ballot(Q.live && lo <= thr) == 0) { // wave-uniform
#pragma unroll
            for (int w = 0; w < PW; ++w) cpts[w] = npts[w];
            crm = nrm;
            continue;
        }
        uint32_t mask = 0;
#pragma unroll
        for (int j = 0; j < 16; ++j) mask |= (acc[j] <= thr ? 1u : 0u) << j;
        if (!Q.live) mask = 0;
        // the tile's removed rows (and rows past the last point) as a 32-bit
        // row mask, then as this lane's 16 elements: element j of half h is
        // row 8 (j >> 2) + 4 h + (j & 3)
        const uint32_t gone = (uint32_t)__ballot(lane < 32 && (crm != 0 || t * 32 + lane >= A.n));
        const uint32_t g = gone >> (4 * h);
        mask &= ~((g & 0xFu) | ((g >> 4) & 0xF0u) | ((g >> 8) & 0xF00u) | ((g >> 12) & 0xF000u));
        wave_lds_sync(); // every lane is done with the previous tile's points
#pragma unroll
        for (int w = 0; w < PW; ++w) {
            const int i = lane + 64 * w;
            if (i < 32 * DIM) tp[soff[w]] = cpts[w];
            cpts[w] = npts[w];
        }
        crm = nrm;
        wave_lds_sync();
        while (mask) {
            const int j = __builtin_ctz(mask);
            mask &= mask - 1;
            const int row = mfma_row(j, h);
            // the row's coordinates in 16-byte reads
            float e[ROW];
#pragma unroll
            for (int c = 0; c < ROW / 4; ++c) {
                const float4 v = reinterpret_cast<const float4 *>(tp + row * ROW)[c];
                e[4 * c] = v.x;
                e[4 * c + 1] = v.y;
                e[4 * c + 2] = v.z;
                e[4 * c + 3] = v.w;
            }
            // SE3: most candidates of the translation filter are far in
            // rotation; skip their acos when they are certainly above bkey
            if (certainly_above<METRIC, DIM>(e, Q.q, __uint_as_float((uint32_t)(bkey >> 32))))
                continue;
            const float d = exact_dist<METRIC, DIM>(e, Q.q);
            const unsigned long long nk = pack_key(d, (int)(t * 32 + row));
            if (KMAX <= 16) {
                if (nk < bkey) top.insert(nk);
            } else {
                // a key at or above the bound is inserted as ~0, which
                // changes nothing: no branch, whose two list versions would
                // both be live in registers
                top.insert(nk < bkey ? nk : ~0ull);
            }
            const unsigned long long kk = top.kth();
            if (kk < bkey) {
                bkey = kk;
                thr = filter_threshold(kk, Q.margin);
            }
        }
    }
    // wave-level merge: lane l (< 32) absorbs the k keys of lane l + 32, which
    // hands over its first slot and shifts its list down, one key a round
    // (one exchange in flight at a time keeps the register count of the
    // search loop)
    const int skip = KMAX - kq;
#pragma unroll 1
    for (int r = 0; r < KMAX; ++r) {
        const unsigned long long ok = shfl_xor_u64(top.key[0], 32);
        if (h == 0) {
            if (r >= skip) top.insert(ok);
        } else {
#pragma unroll
            for (int s = 0; s + 1 < KMAX; ++s) top.key[s] = top.key[s + 1];
            top.key[KMAX - 1] = ~0ull;
        }
    }
    if (h != 0 || !Q.live) return;
    // one split of a packed batch writes its kq results at the query's
    // offset; otherwise each list has k entries (a split's list of a packed
    // query is padded with empty entries for the merge)
    const bool packed_out = A.out_off && A.splits == 1;
    const int64_t base = packed_out ? A.out_off[Q.qi] : ((int64_t)split * A.nq + Q.qi) * A.k;
#pragma unroll
    for (int s = 0; s < KMAX; ++s) {
        if (s >= skip) {
            const unsigned long long kk = top.key[s];
            A.out_d[base + s - skip] = kk == ~0ull ? INFINITY : __uint_as_float((uint32_t)(kk >> 32));
            A.out_i[base + s - skip] = kk == ~0ull ? -1 : (int32_t)(uint32_t)kk;
        }
    }
    if (!packed_out)
        for (int j = kq; j < A.k; ++j) {
            A.out_d[base + j] = INFINITY;
            A.out_i[base + j] = -1;
        }
}

// Merges the per-split partial lists of each query: one wave per query.
// Each lane keeps the k best keys (distance, then index) of its share of the
// splits' entries (lanes read consecutive entries: coalesced), then k rounds
// of a wave-wide minimum take the k best overall in key order, the lane that
// holds each winner dropping it from its list.  Point ids are distinct across
// splits, so every key is held by one lane.
template <int KMAX>
__global__ void __launch_bounds__(256) knn_merge_kernel(const float *__restrict__ pd, const int32_t *__restrict__ pi,
                                                        int64_t nq, int k, int splits, float *__restrict__ out_d,
                                                        int32_t *__restrict__ out_i, const int32_t *__restrict__ kq,
                                                        const int64_t *__restrict__ out_off)
{
    const int lane = threadIdx.x & 63;
    const int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (q >= nq) return; // wave-uniform
    // packed batch: the query's own k (clamped to [0, k]) and output offset
    int kout = k;
    if (kq) {
        const int v = kq[q];
        kout = v < 0 ? 0 : (v > k ? k : v);
    }
    const int64_t obase = out_off ? out_off[q] : q * k;
    TopK<KMAX> top;
    top.init();
    const int total = splits * k;
    for (int c = lane; c < total; c += 64) {
        const int sp = c / k, j = c - sp * k;
        const int64_t at = ((int64_t)sp * nq + q) * k + j;
        const int id = pi[at];
        if (id >= 0) top.insert(pd[at], id);
    }
    for (int r = 0; r < kout; ++r) {
        float bd = top.d[0];
        int bi = top.i[0];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const float od = __shfl_xor(bd, off);
            const int oi = __shfl_xor(bi, off);
            if (key_less(od, oi, bd, bi)) {
                bd = od;
                bi = oi;
            }
        }
        if (bi != INT_MAX && top.i[0] == bi && top.d[0] == bd) {
#pragma unroll
            for (int s = 0; s + 1 < KMAX; ++s) {
                top.d[s] = top.d[s + 1];
                top.i[s] = top.i[s + 1];
            }
            top.d[KMAX - 1] = INFINITY;
            top.i[KMAX - 1] = INT_MAX;
        }
        if (lane == 0) {
            out_d[obase + r] = bd;
            out_i[obase + r] = bi == INT_MAX ? -1 : bi;
        }
    }
}

// Radius search: pass 0 counts; pass 1 writes (idx, dist) keys into CSR
// slots; pass 2 counts and writes the first slot_cap keys of each query into
// a fixed slot per query (one pass when no query has more).
template <int METRIC, int DIM, int KSTEPS, int PASS>
__global__ void __launch_bounds__(256) radius_kernel(KnnArgs A, uint64_t *__restrict__ keys)
{
    const int lane = threadIdx.x & 63;
    const int h = lane >> 5;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t qtiles = (A.nq + 31) / 32;
    const int64_t qtile = wave % qtiles;
    const int split = (int)(wave / qtiles);
    if (split >= A.splits) return;
    QueryLane<METRIC, DIM, KSTEPS> Q;
    Q.load(A, qtile, lane);
    // the batch's radius, or the query's own in a packed batch (a query whose
    // radius is negative or not a number has no results)
    float r = A.radius;
    if (A.rq && Q.live) r = A.rq[Q.qi];
    if (!(r >= 0.0f)) Q.live = false;
    const float thr = r * r * (1.0f + 8e-6f) + Q.margin;
    const int64_t t0 = (int64_t)split * A.tiles_per_split;
    const int64_t t1 = t0 + A.tiles_per_split < A.n_tiles ? t0 + A.tiles_per_split : A.n_tiles;
    int64_t found = 0;
    for (int64_t t = t0; t < t1; ++t) {
        const f32x16 acc = tile_mma<KSTEPS>(A.xt, t, lane, Q.b);
        uint32_t mask = 0;
#pragma unroll
        for (int j = 0; j < 16; ++j) mask |= (acc[j] <= thr ? 1u : 0u) << j;
        if (!Q.live) mask = 0;
        while (mask) {
            const int j = __builtin_ctz(mask);
            mask &= mask - 1;
            const int64_t p = t * 32 + mfma_row(j, h);
            if (p >= A.n || A.removed[p]) continue;
            if (certainly_above<METRIC, DIM>(A.pts + p * DIM, Q.q, r)) continue;
            const float d = exact_dist<METRIC, DIM>(A.pts + p * DIM, Q.q);
            if (d <= r) {
                const uint64_t key = ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)p;
                if (PASS == 0) {
                    ++found;
                } else if (PASS == 1) {
                    const int64_t slot = A.offsets[Q.qi] + atomicAdd((unsigned long long *)&A.counts[Q.qi], 1ull);
                    keys[slot] = key;
                } else {
                    const unsigned long long c = atomicAdd((unsigned long long *)&A.counts[Q.qi], 1ull);
                    if (c < (unsigned long long)A.slot_cap) keys[Q.qi * A.slot_cap + (int64_t)c] = key;
                }
            }
        }
    }
    if (PASS == 0 && Q.live && found) atomicAdd((unsigned long long *)&A.counts[Q.qi], (unsigned long long)found);
}

__global__ void radius_unpack_kernel(const uint64_t *__restrict__ keys, int64_t total, int32_t *__restrict__ idx,
                                     float *__restrict__ dist)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint64_t k = keys[i];
    if (idx) idx[i] = (int32_t)(uint32_t)(k & 0xffffffffu);
    if (dist) dist[i] = __uint_as_float((uint32_t)(k >> 32));
}

// Writes tiled augmented rows for points [from, to) and the padding of the
// last tile; tracks max |x|^2.
__global__ void build_xt_kernel(const float *__restrict__ pts, int dim, int tdim, int ksteps, int64_t from,
                                int64_t to, int64_t n_total, float *__restrict__ xt, uint32_t *__restrict__ xn_max)
{
    const int64_t p = from + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= to) return;
    const bool real = p < n_total;
    float xn = 0.0f;
    if (real)
        for (int i = 0; i < tdim; ++i) xn += pts[p * dim + i] * pts[p * dim + i];
    const int64_t t = p >> 5;
    const int row = (int)(p & 31);
    for (int kk = 0; kk < 2 * ksteps; ++kk) {
        float v;
        if (kk < tdim) v = real ? pts[p * dim + kk] : 0.0f;
        else if (kk == tdim) v = 1.0f;
        else if (kk == tdim + 1) v = real ? xn : kPadNorm;
        else v = 0.0f;
        xt[t * (ksteps * 64) + (kk >> 1) * 64 + (kk & 1) * 32 + row] = v;
    }
    if (real) atomicMax(xn_max, __float_as_uint(xn));
}

// Each query's radius keys (distance bits << 32 | index; unique per query)
// sorted ascending, one wave per query: segments of up to kRadiusSortMax keys
// in the wave's LDS slice by a bitonic network (padding keys ~0 sort last).
// Most queries hold a handful of keys, which a device-wide segmented radix
// sort handles poorly.
constexpr int kRadiusSortMax = 512;

// in_stride: 0 = query keys at [off[q], off[q+1]) of `in`; else at
// [q * in_stride, q * in_stride + count) (the fixed slots of radius pass 2).
__global__ void __launch_bounds__(256) radius_sort_kernel(const uint64_t *__restrict__ in,
                                                          const int64_t *__restrict__ off, int64_t nq,
                                                          int64_t in_stride, uint64_t *__restrict__ out)
{
    __shared__ uint64_t buf[4][kRadiusSortMax];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t qi = (int64_t)blockIdx.x * 4 + w;
    if (qi >= nq) return; // wave-uniform; the kernel uses no block barrier
    const int64_t b = off[qi];
    const int m = (int)(off[qi + 1] - b);
    if (m == 0) return;
    int np2 = 1;
    while (np2 < m) np2 <<= 1;
    uint64_t *sb = buf[w];
    const int64_t src = in_stride ? qi * in_stride : b;
    for (int i = lane; i < np2; i += 64) sb[i] = i < m ? in[src + i] : ~0ull;
    wave_lds_sync();
    for (int k = 2; k <= np2; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = lane; i < np2; i += 64) {
                const int ixj = i ^ j;
                if (ixj > i) { // each pair is handled by one lane
                    const uint64_t a = sb[i], c = sb[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > c) == up) {
                        sb[i] = c;
                        sb[ixj] = a;
                    }
                }
            }
            wave_lds_sync();
        }
    for (int i = lane; i < m; i += 64) out[b + i] = sb[i];
}

__global__ void max_i64_kernel(const int64_t *__restrict__ v, int64_t n, unsigned long long *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicMax(out, (unsigned long long)v[i]);
}

__global__ void fill_i64_kernel(int64_t *p, int64_t n, int64_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static void *scratch(NNIndex &x, size_t bytes)
{
    if (x.scratch_bytes < bytes) {
        if (x.scratch) OA_HIP(hipFree(x.scratch));
        x.scratch = nullptr;
        hipError_t e = hipMalloc(&x.scratch, bytes);
        if (e != hipSuccess) fail(OMPL_AMD_ERR_NOMEM, "nn scratch allocation of %zu bytes failed", bytes);
        x.scratch_bytes = bytes;
    }
    return x.scratch;
}

// Grows the scratch buffer to `bytes`, keeping its first `keep` bytes (the
// copy is finished before the old buffer is freed).
static void *scratch_keep(NNIndex &x, size_t bytes, size_t keep, hipStream_t s)
{
    if (x.scratch_bytes >= bytes) return x.scratch;
    void *fresh = nullptr;
    hipError_t e = hipMalloc(&fresh, bytes);
    if (e != hipSuccess) fail(OMPL_AMD_ERR_NOMEM, "nn scratch allocation of %zu bytes failed", bytes);
    if (x.scratch) {
        if (keep) OA_HIP(hipMemcpyAsync(fresh, x.scratch, keep, hipMemcpyDeviceToDevice, s));
        OA_HIP(hipStreamSynchronize(s));
        OA_HIP(hipFree(x.scratch));
    }
    x.scratch = fresh;
    x.scratch_bytes = bytes;
    return x.scratch;
}

static void reserve(NNIndex &x, int64_t need, hipStream_t s)
{
    if (need <= x.cap) return;
    int64_t cap = x.cap ? x.cap : 1024;
    while (cap < need) cap *= 2;
    cap = (cap + 31) / 32 * 32;
    float *pts = nullptr, *xt = nullptr;
    uint8_t *rem = nullptr;
    OA_HIP(hipMalloc(&pts, (size_t)cap * x.dim * sizeof(float)));
    OA_HIP(hipMalloc(&xt, (size_t)cap * x.ksteps * 2 * sizeof(float)));
    OA_HIP(hipMalloc(&rem, (size_t)cap));
    OA_HIP(hipMemsetAsync(rem, 0, (size_t)cap, s));
    if (x.n) {
        OA_HIP(hipMemcpyAsync(pts, x.pts, (size_t)x.n * x.dim * sizeof(float), hipMemcpyDeviceToDevice, s));
        OA_HIP(hipMemcpyAsync(xt, x.xt, (size_t)((x.n + 31) / 32) * 32 * x.ksteps * 2 * sizeof(float),
                              hipMemcpyDeviceToDevice, s));
        OA_HIP(hipMemcpyAsync(rem, x.removed, (size_t)x.n, hipMemcpyDeviceToDevice, s));
    }
    OA_HIP(hipStreamSynchronize(s));
    if (x.pts) OA_HIP(hipFree(x.pts));
    if (x.xt) OA_HIP(hipFree(x.xt));
    if (x.removed) OA_HIP(hipFree(x.removed));
    x.pts = pts;
    x.xt = xt;
    x.removed = rem;
    x.cap = cap;
    x.removed_host.resize((size_t)cap, 0);
}

static void append(NNIndex &x, const float *src, int64_t n, bool src_dev, hipStream_t s)
{
    if (n <= 0) return;
    reserve(x, x.n + n, s);
    OA_HIP(hipMemcpyAsync(x.pts + x.n * x.dim, src, (size_t)n * x.dim * sizeof(float),
                          src_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    const int64_t from = x.n, to_real = x.n + n;
    const int64_t to_pad = (to_real + 31) / 32 * 32;
    x.n = to_real;
    const int64_t cnt = to_pad - from;
    hipLaunchKernelGGL(build_xt_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, x.pts, x.dim, x.tdim,
                       x.ksteps, from, to_pad, x.n, x.xt, x.xn_max);
    OA_HIP(hipGetLastError());
    if (!src_dev) OA_HIP(hipStreamSynchronize(s)); // host source may be freed after return
}

template <int METRIC, int DIM, int KSTEPS>
static void launch_knn_t(const KnnArgs &A, int kmax, hipStream_t s)
{
    // one workgroup per (query tile, group of W splits)
    const int64_t qtiles = (A.nq + 31) / 32;
    const unsigned blocks = (unsigned)(qtiles * (A.splits / A.group_waves));
    const dim3 block((unsigned)(64 * A.group_waves));
    const size_t lds = knn_lds_bytes(DIM, A.group_waves);
    switch (kmax) {
    case 10: hipLaunchKernelGGL((knn_kernel<METRIC, DIM, KSTEPS, 10>), dim3(blocks), block, lds, s, A); break;
    case 16: hipLaunchKernelGGL((knn_kernel<METRIC, DIM, KSTEPS, 16>), dim3(blocks), block, lds, s, A); break;
    case 32: hipLaunchKernelGGL((knn_kernel<METRIC, DIM, KSTEPS, 32>), dim3(blocks), block, lds, s, A); break;
    default: hipLaunchKernelGGL((knn_kernel<METRIC, DIM, KSTEPS, 64>), dim3(blocks), block, lds, s, A); break;
    }
}

template <int METRIC, int DIM, int KSTEPS, int PASS>
static void launch_radius_t(const KnnArgs &A, uint64_t *keys, int64_t waves, hipStream_t s)
{
    const unsigned blocks = (unsigned)((waves + 3) / 4);
    hipLaunchKernelGGL((radius_kernel<METRIC, DIM, KSTEPS, PASS>), dim3(blocks), dim3(256), 0, s, A, keys);
}

// Dispatch over the supported (metric, dim) shapes: L2 with dim 1..14, SE3.
template <int PASS>
static void dispatch_search(const NNIndex &x, const KnnArgs &A, int kmax, uint64_t *keys, int64_t waves,
                            hipStream_t s, bool radius)
{
#define OA_CASE(D, KS)                                                                                              \
    case D:                                                                                                         \
        if (radius) launch_radius_t<OMPL_AMD_METRIC_L2, D, KS, PASS>(A, keys, waves, s);                          \
        else launch_knn_t<OMPL_AMD_METRIC_L2, D, KS>(A, kmax, s);                                          \
        break;
    if (x.metric == OMPL_AMD_METRIC_SE3) {
        if (radius) launch_radius_t<OMPL_AMD_METRIC_SE3, 7, 3, PASS>(A, keys, waves, s);
        else launch_knn_t<OMPL_AMD_METRIC_SE3, 7, 3>(A, kmax, s);
        return;
    }
    switch (x.dim) {
        OA_CASE(1, 2)
        OA_CASE(2, 2)
        OA_CASE(3, 3)
        OA_CASE(4, 3)
        OA_CASE(5, 4)
        OA_CASE(6, 4)
        OA_CASE(7, 5)
        OA_CASE(8, 5)
        OA_CASE(9, 6)
        OA_CASE(10, 6)
        OA_CASE(11, 7)
        OA_CASE(12, 7)
        OA_CASE(13, 8)
        OA_CASE(14, 8)
    default: fail(OMPL_AMD_ERR_ARG, "L2 dimension %d not supported (1..14)", x.dim);
    }
#undef OA_CASE
}

// Splits of the point tiles for nq queries: about wpc waves per CU, at least
// 8 tiles per split, at most 256 splits.
static int64_t choose_splits(int64_t nq, int64_t n_tiles, int wpc = 16)
{
    const int64_t qtiles = (nq + 31) / 32;
    const int64_t target = (int64_t)num_cus() * wpc;
    int64_t s = (target + qtiles - 1) / qtiles;
    const int64_t max_s = n_tiles / 8 > 1 ? n_tiles / 8 : 1; // >= 8 tiles per split
    if (s > max_s) s = max_s;
    if (s < 1) s = 1;
    if (s > 256) s = 256;
    return s;
}

static KnnArgs base_args(const NNIndex &x, const float *q, int64_t nq)
{
    KnnArgs A{};
    A.pts = x.pts;
    A.xt = x.xt;
    A.removed = x.removed;
    A.xn_max = x.xn_max;
    A.n = x.n;
    A.n_tiles = (x.n + 31) / 32;
    A.q = q;
    A.nq = nq;
    return A;
}

__global__ void packed_counts_kernel(const int32_t *__restrict__ kq, int64_t nq, int kmax, int64_t *__restrict__ cnt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nq) return;
    if (i == nq) {
        cnt[i] = 0;
        return;
    }
    const int v = kq[i];
    cnt[i] = v < 0 ? 0 : (v > kmax ? kmax : v);
}

// k nearest of nq queries.  Uniform k (kq null): results [nq][k].  Packed
// batch (kq non-null, per-query k clamped to [0, k]): results CSR at
// out_off, which this call writes (exclusive scan of the clamped k).
static void knn(NNIndex &x, const float *q, int64_t nq, int k, const int32_t *kq, int64_t *out_off, int32_t *out_i,
                float *out_d, hipStream_t s)
{
    OA_REQUIRE(k >= 1 && k <= 64, OMPL_AMD_ERR_ARG, "k must be in [1, 64] (got %d)", k);
    OA_REQUIRE(nq >= 0, OMPL_AMD_ERR_ARG, "negative query count");
    OA_REQUIRE(!kq || out_off, OMPL_AMD_ERR_ARG, "a packed batch needs an offsets array");
    if (nq == 0) {
        if (out_off) OA_HIP(hipMemsetAsync(out_off, 0, sizeof(int64_t), s));
        return;
    }
    OA_REQUIRE(q && out_i && out_d, OMPL_AMD_ERR_ARG, "null pointer");
    OA_REQUIRE(x.n < INT32_MAX, OMPL_AMD_ERR_ARG, "too many points for int32 ids");
    const int kmax = k <= 16 ? 16 : (k <= 32 ? 32 : 64);
    // the search keeps a 10-slot list for k <= 10 (PRM's default k): fewer
    // slots per insertion; the split merge keeps its 16
    const int kmax_search = k <= 10 ? 10 : kmax;
    KnnArgs A = base_args(x, q, nq);
    A.k = k;
    A.kq = kq;
    A.out_off = kq ? out_off : nullptr;
    // splits of the point tiles, searched W to a workgroup (W a power of two
    // up to the kernel's maximum; the split count rounded up to a multiple)
    int64_t splits = x.n == 0 ? 1 : choose_splits(nq, A.n_tiles);
    const int max_w = kmax_search <= 16 ? 16 : 4;
    int W = 1;
    while (W * 2 <= max_w && W * 2 <= splits) W *= 2;
    splits = (splits + W - 1) / W * W;
    A.splits = (int)splits;
    A.group_waves = W;
    A.tiles_per_split = (A.n_tiles + splits - 1) / splits;
    const int64_t qtiles = (nq + 31) / 32;
    // one scratch allocation: packed counts and scan space, then the splits'
    // partial lists
    size_t cub_bytes = 0;
    if (kq)
        OA_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, cub_bytes, (int64_t *)nullptr, (int64_t *)nullptr,
                                                (int)(nq + 1), s));
    const size_t counts_bytes = kq ? ((size_t)(nq + 1) * sizeof(int64_t) + 255) / 256 * 256 : 0;
    const size_t cub_span = (cub_bytes + 255) / 256 * 256;
    const size_t per = splits > 1 ? (size_t)splits * nq * k : 0;
    char *buf = (char *)scratch(x, counts_bytes + cub_span + per * (sizeof(float) + sizeof(int32_t)) + 256);
    ProfScope ps("nn_knn", s);
    if (kq) {
        int64_t *cnt = (int64_t *)buf;
        hipLaunchKernelGGL(packed_counts_kernel, dim3((unsigned)((nq + 1 + 255) / 256)), dim3(256), 0, s, kq, nq, k,
                           cnt);
        OA_HIP(hipcub::DeviceScan::ExclusiveSum(buf + counts_bytes, cub_bytes, cnt, out_off, (int)(nq + 1), s));
    }
    if (splits == 1) {
        // one split (or an empty structure: every slot is (-1, +inf))
        A.out_d = out_d;
        A.out_i = out_i;
        dispatch_search<0>(x, A, kmax_search, nullptr, qtiles, s, false);
    } else {
        char *lists = buf + counts_bytes + cub_span;
        A.out_d = (float *)lists;
        A.out_i = (int32_t *)(lists + per * sizeof(float));
        dispatch_search<0>(x, A, kmax_search, nullptr, qtiles * splits, s, false);
        const unsigned blocks = (unsigned)((nq + 3) / 4); // one wave per query
        const int64_t *moff = kq ? out_off : nullptr;
        switch (kmax) {
        case 16:
            hipLaunchKernelGGL(knn_merge_kernel<16>, dim3(blocks), dim3(256), 0, s, A.out_d, A.out_i, nq, k,
                               (int)splits, out_d, out_i, kq, moff);
            break;
        case 32:
            hipLaunchKernelGGL(knn_merge_kernel<32>, dim3(blocks), dim3(256), 0, s, A.out_d, A.out_i, nq, k,
                               (int)splits, out_d, out_i, kq, moff);
            break;
        default:
            hipLaunchKernelGGL(knn_merge_kernel<64>, dim3(blocks), dim3(256), 0, s, A.out_d, A.out_i, nq, k,
                               (int)splits, out_d, out_i, kq, moff);
            break;
        }
    }
    OA_HIP(hipGetLastError());
}

// Keys kept per query by the one-pass radius search (pass 2), and the most
// slot memory it may take; a query with more results sends the call to the
// two-pass path (count, then write into CSR slots).
constexpr int kRadiusSlotCap = 64;
constexpr size_t kRadiusSlotBytesMax = (size_t)256 << 20;

// Radius search of nq queries, at radius r or, in a packed batch (rq
// non-null), at each query's own radius.
static void radius(NNIndex &x, const float *q, int64_t nq, float r, const float *rq, int64_t *offsets, int32_t *out_i,
                   float *out_d, int64_t capacity, int64_t *total, hipStream_t s)
{
    OA_REQUIRE(nq >= 0 && offsets && total, OMPL_AMD_ERR_ARG, "bad radius arguments");
    OA_REQUIRE(rq || r >= 0.0f, OMPL_AMD_ERR_ARG, "negative radius");
    *total = 0;
    if (nq == 0) {
        OA_HIP(hipMemsetAsync(offsets, 0, sizeof(int64_t), s));
        return;
    }
    OA_REQUIRE(q, OMPL_AMD_ERR_ARG, "null queries");
    const bool want = out_i || out_d;
    KnnArgs A = base_args(x, q, nq);
    A.radius = r;
    A.rq = rq;
    const int64_t n_tiles = A.n_tiles;
    // more splits than k-NN: a radius wave keeps no list, so the launch gains
    // from more waves to hide its latency (SE3 1M x 65,536: 15.8 ms at 16
    // waves per CU, 13.9 at 32, 13.1 at 64, 12.7 at 128)
    const int64_t splits = n_tiles ? choose_splits(nq, n_tiles, 128) : 1;
    A.splits = (int)splits;
    A.tiles_per_split = n_tiles ? (n_tiles + splits - 1) / splits : 0;
    const int64_t qtiles = (nq + 31) / 32;
    // one pass (count and write the first kRadiusSlotCap keys of each query
    // into its own slot) when results are wanted and the slots fit
    const bool one_pass = want && n_tiles && (size_t)nq * kRadiusSlotCap * sizeof(uint64_t) <= kRadiusSlotBytesMax;
    // scratch: counts [nq + 1] and the largest count (256-byte aligned), cub
    // temp, slots (one pass)
    size_t cub_bytes = 0;
    OA_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, cub_bytes, (int64_t *)nullptr, (int64_t *)nullptr,
                                            (int)(nq + 1), s));
    const size_t counts_bytes = ((size_t)(nq + 2) * sizeof(int64_t) + 255) / 256 * 256;
    const size_t cub_span = (cub_bytes + 255) / 256 * 256;
    const size_t slot_bytes = one_pass ? (size_t)nq * kRadiusSlotCap * sizeof(uint64_t) : 0;
    char *buf = (char *)scratch(x, counts_bytes + cub_span + slot_bytes);
    int64_t *counts = (int64_t *)buf;
    unsigned long long *max_count = (unsigned long long *)(counts + nq + 1);
    void *cub_tmp = buf + counts_bytes;
    uint64_t *slots = one_pass ? (uint64_t *)(buf + counts_bytes + cub_span) : nullptr;
    ProfScope ps("nn_radius", s);
    hipLaunchKernelGGL(fill_i64_kernel, dim3((unsigned)((nq + 2 + 255) / 256)), dim3(256), 0, s, counts, nq + 2,
                       (int64_t)0);
    A.counts = counts;
    if (one_pass) {
        A.slot_cap = kRadiusSlotCap;
        dispatch_search<2>(x, A, 16, slots, qtiles * splits, s, true);
    } else if (n_tiles) {
        dispatch_search<0>(x, A, 16, nullptr, qtiles * splits, s, true);
    }
    hipLaunchKernelGGL(max_i64_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, counts, nq, max_count);
    OA_HIP(hipcub::DeviceScan::ExclusiveSum(cub_tmp, cub_bytes, counts, offsets, (int)(nq + 1), s));
    int64_t tot = 0;
    unsigned long long longest = 0;
    OA_HIP(hipMemcpyAsync(&tot, offsets + nq, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    OA_HIP(hipMemcpyAsync(&longest, max_count, sizeof(longest), hipMemcpyDeviceToHost, s));
    OA_HIP(hipStreamSynchronize(s));
    *total = tot;
    if (!want || tot == 0) return;
    OA_REQUIRE(capacity >= tot, OMPL_AMD_ERR_ARG, "radius output capacity %lld < %lld results", (long long)capacity,
               (long long)tot);
    OA_REQUIRE(tot < INT32_MAX, OMPL_AMD_ERR_ARG, "too many radius results for one call");
    // keys (dist bits << 32 | idx) sorted per query: one wave per query when
    // every query holds at most kRadiusSortMax keys, else a segmented radix sort
    const bool from_slots = one_pass && longest <= (unsigned long long)kRadiusSlotCap;
    const bool wave_sort = longest <= (unsigned long long)kRadiusSortMax;
    size_t sort_bytes = 0;
    if (!wave_sort)
        OA_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, sort_bytes, (uint64_t *)nullptr,
                                                          (uint64_t *)nullptr, (int)tot, (int)nq, (int64_t *)nullptr,
                                                          (int64_t *)nullptr, 0, 64, s));
    const size_t keys_span = ((size_t)tot * sizeof(uint64_t) + 255) / 256 * 256;
    const size_t head = counts_bytes + cub_span + slot_bytes;
    buf = (char *)scratch_keep(x, head + 2 * keys_span + sort_bytes, head, s);
    counts = (int64_t *)buf;
    slots = one_pass ? (uint64_t *)(buf + counts_bytes + cub_span) : nullptr;
    uint64_t *keys = (uint64_t *)(buf + head);
    uint64_t *keys_sorted = (uint64_t *)(buf + head + keys_span);
    void *sort_tmp = buf + head + 2 * keys_span;
    if (from_slots) {
        hipLaunchKernelGGL(radius_sort_kernel, dim3((unsigned)((nq + 3) / 4)), dim3(256), 0, s, slots, offsets, nq,
                           (int64_t)kRadiusSlotCap, keys_sorted);
    } else {
        // a query overflowed its slot (or no slots): write every key into
        // its CSR span, then sort
        hipLaunchKernelGGL(fill_i64_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, counts, nq,
                           (int64_t)0);
        A.counts = counts;
        A.offsets = offsets;
        dispatch_search<1>(x, A, 16, keys, qtiles * splits, s, true);
        if (wave_sort)
            hipLaunchKernelGGL(radius_sort_kernel, dim3((unsigned)((nq + 3) / 4)), dim3(256), 0, s, keys, offsets,
                               nq, (int64_t)0, keys_sorted);
        else
            OA_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(sort_tmp, sort_bytes, keys, keys_sorted, (int)tot,
                                                              (int)nq, offsets, offsets + 1, 0, 64, s));
    }
    hipLaunchKernelGGL(radius_unpack_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, keys_sorted, tot,
                       out_i, out_d);
    OA_HIP(hipGetLastError());
}

static NNIndex &nn_of(ompl_amd_nn h)
{
    OA_REQUIRE(h, OMPL_AMD_ERR_ARG, "null nearest-neighbour handle");
    return h->idx;
}

// Device buffers of one host-convenience call, freed on every exit.
struct DevBufs {
    hipStream_t s;
    std::vector<void *> p;
    explicit DevBufs(hipStream_t st) : s(st) {}
    template <class T> T *alloc(size_t n)
    {
        void *q = nullptr;
        OA_HIP(hipMallocAsync(&q, (n ? n : 1) * sizeof(T), s));
        p.push_back(q);
        return static_cast<T *>(q);
    }
    ~DevBufs()
    {
        for (void *q : p) (void)hipFreeAsync(q, s);
        (void)hipStreamSynchronize(s);
    }
};

} // namespace ompl_amd

using namespace ompl_amd;

extern "C" {

int ompl_amd_nn_create(int metric, int dim, ompl_amd_nn *out)
{
    return guarded([&] {
        OA_REQUIRE(out, OMPL_AMD_ERR_ARG, "null output handle");
        OA_REQUIRE(metric == OMPL_AMD_METRIC_L2 || metric == OMPL_AMD_METRIC_SE3, OMPL_AMD_ERR_ARG, "unknown metric");
        if (metric == OMPL_AMD_METRIC_SE3)
            OA_REQUIRE(dim == 7, OMPL_AMD_ERR_ARG, "SE3 points have 7 coordinates (x,y,z,qx,qy,qz,qw)");
        else
            OA_REQUIRE(dim >= 1 && dim <= 14, OMPL_AMD_ERR_ARG, "L2 dimension must be in [1, 14]");
        ensure_init();
        auto *h = new ompl_amd_nn_s();
        NNIndex &x = h->idx;
        x.metric = metric;
        x.dim = dim;
        x.tdim = metric == OMPL_AMD_METRIC_SE3 ? 3 : dim;
        x.ksteps = (x.tdim + 2 + 1) / 2;
        OA_HIP(hipMalloc(&x.xn_max, sizeof(uint32_t)));
        OA_HIP(hipMemset(x.xn_max, 0, sizeof(uint32_t)));
        *out = h;
    });
}

int ompl_amd_nn_destroy(ompl_amd_nn h)
{
    return guarded([&] {
        if (!h) return;
        NNIndex &x = h->idx;
        if (x.pts) OA_HIP(hipFree(x.pts));
        if (x.xt) OA_HIP(hipFree(x.xt));
        if (x.removed) OA_HIP(hipFree(x.removed));
        if (x.xn_max) OA_HIP(hipFree(x.xn_max));
        if (x.scratch) OA_HIP(hipFree(x.scratch));
        delete h;
    });
}

int ompl_amd_nn_clear(ompl_amd_nn h)
{
    return guarded([&] {
        NNIndex &x = nn_of(h);
        hipStream_t s = default_stream();
        if (x.n && x.removed) OA_HIP(hipMemsetAsync(x.removed, 0, (size_t)x.n, s));
        OA_HIP(hipMemsetAsync(x.xn_max, 0, sizeof(uint32_t), s));
        OA_HIP(hipStreamSynchronize(s));
        std::fill(x.removed_host.begin(), x.removed_host.end(), 0);
        x.n = 0;
        x.n_removed = 0;
    });
}

int ompl_amd_nn_add(ompl_amd_nn h, const float *pts, int64_t n, int64_t *first_id)
{
    return guarded([&] {
        NNIndex &x = nn_of(h);
        OA_REQUIRE(n >= 0 && (n == 0 || pts), OMPL_AMD_ERR_ARG, "bad points");
        if (first_id) *first_id = x.n;
        append(x, pts, n, false, default_stream());
    });
}

int ompl_amd_nn_add_dev(ompl_amd_nn h, const float *pts, int64_t n, int64_t *first_id, void *stream)
{
    return guarded([&] {
        NNIndex &x = nn_of(h);
        OA_REQUIRE(n >= 0 && (n == 0 || pts), OMPL_AMD_ERR_ARG, "bad points");
        if (first_id) *first_id = x.n;
        append(x, pts, n, true, resolve_stream(stream));
    });
}

int ompl_amd_nn_remove(ompl_amd_nn h, int64_t id, int *removed)
{
    return guarded([&] {
        NNIndex &x = nn_of(h);
        int ok = 0;
        if (id >= 0 && id < x.n && !x.removed_host[(size_t)id]) {
            x.removed_host[(size_t)id] = 1;
            const uint8_t one = 1;
            hipStream_t s = default_stream();
            OA_HIP(hipMemcpyAsync(x.removed + id, &one, 1, hipMemcpyHostToDevice, s));
            OA_HIP(hipStreamSynchronize(s));
            ++x.n_removed;
            ok = 1;
        }
        if (removed) *removed = ok;
    });
}

int ompl_amd_nn_size(ompl_amd_nn h, int64_t *size)
{
    return guarded([&] {
        NNIndex &x = nn_of(h);
        OA_REQUIRE(size, OMPL_AMD_ERR_ARG, "null output");
        *size = x.n - x.n_removed;
    });
}

int ompl_amd_nn_build(ompl_amd_nn h, void *stream)
{
    return guarded([&] {
        (void)nn_of(h);
        (void)stream; // tiles are written by add(); nothing is deferred
    });
}

int ompl_amd_nn_knn_dev(ompl_amd_nn h, const float *q, int64_t nq, int k, int32_t *out_i, float *out_d, void *stream)
{
    return guarded([&] { knn(nn_of(h), q, nq, k, nullptr, nullptr, out_i, out_d, resolve_stream(stream)); });
}

int ompl_amd_nn_radius_dev(ompl_amd_nn h, const float *q, int64_t nq, float r, int64_t *offsets, int32_t *out_i,
                           float *out_d, int64_t capacity, int64_t *total, void *stream)
{
    return guarded(
        [&] { radius(nn_of(h), q, nq, r, nullptr, offsets, out_i, out_d, capacity, total, resolve_stream(stream)); });
}

int ompl_amd_nn_knn_packed_dev(ompl_amd_nn h, const float *q, int64_t nq, const int32_t *k, int kmax,
                               int64_t *out_offsets, int32_t *out_i, float *out_d, void *stream)
{
    return guarded([&] {
        OA_REQUIRE(k && out_offsets, OMPL_AMD_ERR_ARG, "null per-query k or offsets");
        knn(nn_of(h), q, nq, kmax, k, out_offsets, out_i, out_d, resolve_stream(stream));
    });
}

int ompl_amd_nn_radius_packed_dev(ompl_amd_nn h, const float *q, int64_t nq, const float *r, int64_t *offsets,
                                  int32_t *out_i, float *out_d, int64_t capacity, int64_t *total, void *stream)
{
    return guarded([&] {
        OA_REQUIRE(r, OMPL_AMD_ERR_ARG, "null per-query radius");
        radius(nn_of(h), q, nq, 0.0f, r, offsets, out_i, out_d, capacity, total, resolve_stream(stream));
    });
}

int ompl_amd_nn_knn(ompl_amd_nn h, const float *q, int64_t nq, int k, int32_t *out_i, float *out_d)
{
    return guarded([&] {
        NNIndex &x = nn_of(h);
        OA_REQUIRE(nq >= 0 && (nq == 0 || (q && out_i && out_d)), OMPL_AMD_ERR_ARG, "bad arguments");
        if (nq == 0) return;
        OA_REQUIRE(k >= 1 && k <= 64, OMPL_AMD_ERR_ARG, "k must be in [1, 64] (got %d)", k);
        hipStream_t s = default_stream();
        DevBufs B(s); // freed (after the copies) on every exit
        float *dq = B.alloc<float>((size_t)nq * x.dim);
        float *dd = B.alloc<float>((size_t)nq * k);
        int32_t *di = B.alloc<int32_t>((size_t)nq * k);
        OA_HIP(hipMemcpyAsync(dq, q, (size_t)nq * x.dim * sizeof(float), hipMemcpyHostToDevice, s));
        knn(x, dq, nq, k, nullptr, nullptr, di, dd, s);
        OA_HIP(hipMemcpyAsync(out_d, dd, (size_t)nq * k * sizeof(float), hipMemcpyDeviceToHost, s));
        OA_HIP(hipMemcpyAsync(out_i, di, (size_t)nq * k * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    });
}

int ompl_amd_nn_radius(ompl_amd_nn h, const float *q, int64_t nq, float r, int64_t *offsets, int32_t *out_i,
                       float *out_d, int64_t capacity, int64_t *total)
{
    return guarded([&] {
        NNIndex &x = nn_of(h);
        OA_REQUIRE(nq >= 0 && offsets && total && (nq == 0 || q), OMPL_AMD_ERR_ARG, "bad arguments");
        hipStream_t s = default_stream();
        DevBufs B(s); // freed (after the copies) on every exit
        const bool want = out_i || out_d;
        float *dq = B.alloc<float>((size_t)nq * x.dim);
        int64_t *doff = B.alloc<int64_t>((size_t)nq + 1);
        int32_t *di = want && capacity > 0 ? B.alloc<int32_t>((size_t)capacity) : nullptr;
        float *dd = want && capacity > 0 ? B.alloc<float>((size_t)capacity) : nullptr;
        if (nq) OA_HIP(hipMemcpyAsync(dq, q, (size_t)nq * x.dim * sizeof(float), hipMemcpyHostToDevice, s));
        int64_t tot = 0;
        *total = 0;
        radius(x, dq, nq, r, nullptr, doff, want ? di : nullptr, want ? dd : nullptr, capacity, &tot, s);
        OA_HIP(hipMemcpyAsync(offsets, doff, (size_t)(nq + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        if (want && tot > 0) {
            if (out_i) OA_HIP(hipMemcpyAsync(out_i, di, (size_t)tot * sizeof(int32_t), hipMemcpyDeviceToHost, s));
            if (out_d) OA_HIP(hipMemcpyAsync(out_d, dd, (size_t)tot * sizeof(float), hipMemcpyDeviceToHost, s));
        }
        OA_HIP(hipStreamSynchronize(s));
        *total = tot;
    });
}


int ompl_amd_nn_knn_packed(ompl_amd_nn h, const float *q, int64_t nq, const int32_t *k, int kmax, int64_t *offsets,
                           int32_t *out_i, float *out_d)
{
    return guarded([&] {
        NNIndex &x = nn_of(h);
        OA_REQUIRE(nq >= 0 && offsets && (nq == 0 || (q && k && out_i && out_d)), OMPL_AMD_ERR_ARG, "bad arguments");
        OA_REQUIRE(kmax >= 1 && kmax <= 64, OMPL_AMD_ERR_ARG, "kmax must be in [1, 64] (got %d)", kmax);
        if (nq == 0) {
            offsets[0] = 0;
            return;
        }
        hipStream_t s = default_stream();
        DevBufs B(s);
        float *dq = B.alloc<float>((size_t)nq * x.dim);
        int32_t *dk = B.alloc<int32_t>((size_t)nq);
        int64_t *doff = B.alloc<int64_t>((size_t)nq + 1);
        int32_t *di = B.alloc<int32_t>((size_t)nq * kmax);
        float *dd = B.alloc<float>((size_t)nq * kmax);
        OA_HIP(hipMemcpyAsync(dq, q, (size_t)nq * x.dim * sizeof(float), hipMemcpyHostToDevice, s));
        OA_HIP(hipMemcpyAsync(dk, k, (size_t)nq * sizeof(int32_t), hipMemcpyHostToDevice, s));
        knn(x, dq, nq, kmax, dk, doff, di, dd, s);
        OA_HIP(hipMemcpyAsync(offsets, doff, (size_t)(nq + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        OA_HIP(hipStreamSynchronize(s));
        const int64_t tot = offsets[nq];
        if (tot > 0) {
            OA_HIP(hipMemcpyAsync(out_i, di, (size_t)tot * sizeof(int32_t), hipMemcpyDeviceToHost, s));
            OA_HIP(hipMemcpyAsync(out_d, dd, (size_t)tot * sizeof(float), hipMemcpyDeviceToHost, s));
        }
    });
}

int ompl_amd_nn_radius_packed(ompl_amd_nn h, const float *q, int64_t nq, const float *r, int64_t *offsets,
                              int32_t *out_i, float *out_d, int64_t capacity, int64_t *total)
{
    return guarded([&] {
        NNIndex &x = nn_of(h);
        OA_REQUIRE(nq >= 0 && offsets && total && (nq == 0 || (q && r)), OMPL_AMD_ERR_ARG, "bad arguments");
        hipStream_t s = default_stream();
        DevBufs B(s);
        const bool want = out_i || out_d;
        float *dq = B.alloc<float>((size_t)nq * x.dim);
        float *dr = B.alloc<float>((size_t)nq);
        int64_t *doff = B.alloc<int64_t>((size_t)nq + 1);
        int32_t *di = want && capacity > 0 ? B.alloc<int32_t>((size_t)capacity) : nullptr;
        float *dd = want && capacity > 0 ? B.alloc<float>((size_t)capacity) : nullptr;
        if (nq) {
            OA_HIP(hipMemcpyAsync(dq, q, (size_t)nq * x.dim * sizeof(float), hipMemcpyHostToDevice, s));
            OA_HIP(hipMemcpyAsync(dr, r, (size_t)nq * sizeof(float), hipMemcpyHostToDevice, s));
        }
        int64_t tot = 0;
        *total = 0;
        radius(x, dq, nq, 0.0f, dr, doff, want ? di : nullptr, want ? dd : nullptr, capacity, &tot, s);
        OA_HIP(hipMemcpyAsync(offsets, doff, (size_t)(nq + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        if (want && tot > 0) {
            if (out_i) OA_HIP(hipMemcpyAsync(out_i, di, (size_t)tot * sizeof(int32_t), hipMemcpyDeviceToHost, s));
            if (out_d) OA_HIP(hipMemcpyAsync(out_d, dd, (size_t)tot * sizeof(float), hipMemcpyDeviceToHost, s));
        }
        OA_HIP(hipStreamSynchronize(s));
        *total = tot;
    });
}

int ompl_amd_nn_nearest(ompl_amd_nn h, const float *q, int64_t *out_id)
{
    return guarded([&] {
        NNIndex &x = nn_of(h);
        OA_REQUIRE(q && out_id, OMPL_AMD_ERR_ARG, "null pointer");
        if (x.n - x.n_removed <= 0) fail(OMPL_AMD_ERR_EMPTY, "No elements found in nearest neighbors data structure");
        int32_t id = -1;
        float d = 0.0f;
        int rc = ompl_amd_nn_knn(h, q, 1, 1, &id, &d);
        if (rc) fail(rc, "%s", last_error());
        *out_id = id;
    });
}

double ompl_amd_rrtstar_radius(int dim, double measure, double rewire_factor, double max_distance, int64_t n)
{
    // unitNBallMeasure (util/src/GeometricEquations.cpp:55-60): pi^(d/2) / Gamma(d/2 + 1)
    const double d = (double)dim;
    const double ball = std::pow(std::sqrt(M_PI), d) / std::tgamma(d / 2.0 + 1.0);
    const double r_rrt = rewire_factor * std::pow(2 * (1.0 + 1.0 / d) * (measure / ball), 1.0 / d);
    const double cardDbl = (double)(n + 1);
    const double r = r_rrt * std::pow(std::log(cardDbl) / cardDbl, 1.0 / d);
    return r < max_distance ? r : max_distance;
}

} // extern "C"
