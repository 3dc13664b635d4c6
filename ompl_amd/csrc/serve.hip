// Planning server: continuous batching of RRT-Connect requests on one GPU,
// each request's planner loop resident on the device.
//
// A request is a single-query RRT-Connect plan (geometric/planners/rrt/src/
// RRTConnect.cpp:116-196 growTree, :245-330 solve) on the 7-DOF arm in one of
// the registered environments, from one or more start states to one or more
// goal states (PlannerInputStates order, base/src/Planner.cpp:228-336).
// Everything the loop touches lives in one region of a device pool sized to
// the HBM budget: both trees (append-only FP64 states + parent indices), the
// start and goal states, the state sampler's mt19937 and the solve loop's
// variables, and the solution path.  The server thread launches slices: one
// block (512 threads, or 1,024) per active request runs its loop
// -- sample, nearest neighbour, steer, motion check, append, connect loop,
// path splice -- for a bounded time, then stores its state; the host then
// retires finished requests and admits queued ones while their region fits
// the pool (continuous batching), and launches the next slice.  A slice
// resumes each loop exactly where the last one stopped, so a request's result
// does not depend on the slices it ran in or the requests it shared them
// with.
//
// Inside a block: the nearest neighbour is a thread-strided FP64 scan of the
// tree (first strict minimum), reduced by (distance, index); a motion check
// takes the block's waves over its positions in rounds (the link layout
// with bounding-sphere culling, the environment staged in LDS once per
// slice), with a block-wide OR after each round as the early exit; the
// sampler's words are tempered from the LDS state and the state is twisted
// by the whole block.
//
// Plain counterpart: ompl_amd_plan_plain runs the same RRT-Connect on the host
// with host nearest neighbours and one GPU launch + synchronisation per
// isValid / checkMotion call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "mt_device.hpp"
#include "runtime.hpp"
#include "walk_device.hpp"

namespace ompl_amd {

enum GrowStatus { kTrapped = 0, kAdvanced = 1, kReached = 2 };

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kEqualEps = 2.220446049250313080847e-16 * 2.0; // StateSpace::equalStates (RealVector)
// Obstacle staging padding shared by the motion check's four partitionings
// (whole passes of 8, 16, 32 and 144 obstacles): their least common multiple.
constexpr int kPlanPad = 288;
// Dynamic LDS of one planner block (the staged environment).
constexpr size_t kPlanMaxEnvBytes = 60 * 1024;
// Start / goal states a request may list (each).
constexpr int kMaxEnds = OMPL_AMD_MAX_ENDPOINTS;

// Persistent state of one request's RRTConnect::solve loop (device).
struct PlanState {
    uint32_t mt[kMtN]; // the state sampler's RNG (std::mt19937)
    int32_t mti;       // next word (kMtN: twist first)
    int32_t status;
    int32_t iteration;
    int32_t n0, n1;     // nodes of the start / goal tree
    int32_t start_tree; // RRTConnect::startTree_
    int32_t tgi_start;  // tgi.start
    int32_t grow_tree;  // tree of the next growTree (0 start, 1 goal)
    int32_t phase;      // 0: sample + grow, 1: first connect step, 2: connect loop
    int32_t added, xmotion;
    int32_t path_len;
    int32_t sampled_goals; // goal states handed out so far (PlannerInputStates)
    int64_t steps, checks;
    double rstate[kDof];
};

__host__ __device__ inline size_t align256(size_t b)
{
    return (b + 255) & ~(size_t)255;
}

// One request's region: [start tree states][goal tree states][start parents]
// [goal parents][path states][path node indices][start and goal states]
// [PlanState].  Tree states
// are dof-major (coordinate k of node i at k * cap + i), so a wave's reads of
// consecutive nodes coalesce; the path is row-major.
struct RegionPtrs {
    double *tree_s[2];
    int32_t *parent[2];
    double *path;
    int32_t *path_idx;
    double *ends; // kMaxEnds start rows, then kMaxEnds goal rows
    PlanState *state;
    size_t bytes;
};

__host__ __device__ inline RegionPtrs region_ptrs(char *base, int cap)
{
    RegionPtrs R;
    size_t off = 0;
    for (int t = 0; t < 2; ++t) {
        R.tree_s[t] = reinterpret_cast<double *>(base + off);
        off += align256((size_t)cap * kDof * sizeof(double));
    }
    for (int t = 0; t < 2; ++t) {
        R.parent[t] = reinterpret_cast<int32_t *>(base + off);
        off += align256((size_t)cap * sizeof(int32_t));
    }
    R.path = reinterpret_cast<double *>(base + off);
    off += align256((size_t)2 * cap * kDof * sizeof(double));
    R.path_idx = reinterpret_cast<int32_t *>(base + off);
    off += align256((size_t)2 * cap * sizeof(int32_t));
    R.ends = reinterpret_cast<double *>(base + off);
    off += align256((size_t)2 * kMaxEnds * kDof * sizeof(double));
    R.state = reinterpret_cast<PlanState *>(base + off);
    off += align256(sizeof(PlanState));
    R.bytes = off;
    return R;
}

static size_t footprint(int max_nodes)
{
    return region_ptrs(nullptr, max_nodes).bytes;
}

// Per active request and slice (device copy of a host array).
struct PlanCmd {
    const float4 *obs; // environment, staging order (spheres, then boxes as lo/hi pairs)
    char *region;
    int32_t ns, nb;
    int32_t cap;       // nodes per tree
    int32_t max_iters;
    double maxd;       // maxDistance_
    double lvs;        // longest valid segment
    uint32_t factor;
    uint32_t seed;
    int32_t init;      // 1: first slice (seed the sampler, plant the roots)
    int32_t ops;       // obstacle partitions of the motion check (0: by motion length)
    int32_t n_starts, n_goals; // rows of the region's start and goal states
    int32_t inter;     // addIntermediateStates
};

// Per active request, written at the end of each slice.  t_*: device
// wall-clock ticks this slice spent sampling, in nearest-neighbour scans, in
// motion checks (steering included) and in whole growTree steps.
struct PlanOut {
    int32_t status, iteration, n0, n1, path_len, pad;
    int64_t steps, checks;
    int64_t t_sample, t_nearest, t_check, t_step;
};

struct SliceArgs {
    const PlanCmd *cmds;
    PlanOut *out;
    int64_t budget; // wall-clock ticks per slice (at least one step is always made)
};

static size_t env_stage_bytes(int ns, int nb)
{
    return (size_t)(pad_to(ns, kPlanPad) + 2 * pad_to(nb, kPlanPad)) * sizeof(float4) + (size_t)ns * sizeof(float);
}

// Obstacle partitions for a motion of npos positions: the widest split whose
// round (NW waves x states per wave) still covers the motion, else one.
template <int NW>
__device__ __forceinline__ int auto_ops(int npos)
{
    return npos <= NW * 1 ? 9 : npos <= NW * 2 ? 4 : npos <= NW * 4 ? 2 : 1;
}

// Nearest node of tree T (n nodes, dof-major with stride cap) to tgt: first
// strict minimum of distance(node, tgt) (NearestNeighborsLinear::nearest
// order).  Every thread returns the same index.
template <int NT>
__device__ __forceinline__ int block_nearest(const double *__restrict__ T, int cap, int n, const double tgt[kDof],
                                             double *s_wd, int *s_wi)
{
    double best = INFINITY;
    int bi = INT_MAX;
    for (int i = threadIdx.x; i < n; i += NT) {
        double dist = 0.0;
#pragma unroll
        for (int k = 0; k < kDof; ++k) {
            const double diff = T[(int64_t)k * cap + i] - tgt[k];
            dist += diff * diff;
        }
        const double d = __builtin_sqrt(dist);
        if (d < best) { // a thread's indices increase: strict < keeps its lowest
            best = d;
            bi = i;
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double od = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (od < best || (od == best && oi < bi)) {
            best = od;
            bi = oi;
        }
    }
    if (lane_id() == 0) {
        s_wd[threadIdx.x >> 6] = best;
        s_wi[threadIdx.x >> 6] = bi;
    }
    __syncthreads();
    best = s_wd[0];
    bi = s_wi[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w)
        if (s_wd[w] < best || (s_wd[w] == best && s_wi[w] < bi)) {
            best = s_wd[w];
            bi = s_wi[w];
        }
    return bi;
}

// Positions [first, npos) of the motion by the whole block, in rounds of NT/64 waves x G
// states: with `extra`, position 0 is isValid(dst) and position p >= 1 the
// DiscreteMotionValidator position p-1 of (a, b) (0 = b, then the bisection
// order); without, position p is the validator's position p.  Returns false
// at the first round holding an invalid position.  Validity of the whole
// motion does not depend on the order positions are visited in.  a, b and
// dst are in LDS (read per round; no registers held across the rounds);
// `stop` is an LDS flag, zero on entry, that lets the other waves of a round
// leave once one has found an invalid position.
template <int NT, int OPS>
__device__ __forceinline__ bool block_check(const double *a, const double *b, const double *dst, int nd, int first,
                                            int npos, bool extra, const float4 *env, int ns, int nb, int *stop)
{
    using Shape = GroupShape<7, OPS>;
    constexpr int G = Shape::STATES;
    const int wave = threadIdx.x >> 6;
    const Shape sh(lane_id());
    for (int base = first; base < npos; base += NT / 64 * G) {
        const int p = base + wave * G + sh.group;
        const bool live = p < npos;
        float q[kDof];
        if (extra && p == 0) {
#pragma unroll
            for (int k = 0; k < kDof; ++k) q[k] = (float)dst[k];
        } else {
            const int mp = extra ? p - 1 : p;
            walk_state(a, b, nd, live ? mp : 0, q);
        }
        LaneSpheres<7> S;
        arm_fk<7, Shape::GS>(q, sh.slot, S);
        const bool hit = collide_links<OPS, Exit::kAnyGroup, true, kPlanPad>(S, env, ns, nb, live, sh.part, stop);
        if (__syncthreads_or(hit && live)) return false;
    }
    return true;
}

// Nearest node of tree T to tgt by one wave (the speculative sampling steps):
// the same first strict minimum as block_nearest; every lane returns it.
__device__ __forceinline__ int wave_nearest(const double *__restrict__ T, int cap, int n, const double tgt[kDof])
{
    double best = INFINITY;
    int bi = INT_MAX;
    for (int i = lane_id(); i < n; i += 64) {
        double dist = 0.0;
#pragma unroll
        for (int k = 0; k < kDof; ++k) {
            const double diff = T[(int64_t)k * cap + i] - tgt[k];
            dist += diff * diff;
        }
        const double d = __builtin_sqrt(dist);
        if (d < best) {
            best = d;
            bi = i;
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double od = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (od < best || (od == best && oi < bi)) {
            best = od;
            bi = oi;
        }
    }
    return bi;
}

// Positions [first, last) of a motion (numbered as in block_check) by one
// wave; false at the first round holding an invalid position.
template <int OPS>
__device__ __forceinline__ bool wave_check(const double *a, const double *b, const double *dst, int nd, int first,
                                           int last, bool extra, const float4 *env, int ns, int nb)
{
    using Shape = GroupShape<7, OPS>;
    constexpr int G = Shape::STATES;
    const Shape sh(lane_id());
    for (int base = first; base < last; base += G) {
        const int p = base + sh.group;
        const bool live = p < last;
        float q[kDof];
        if (extra && p == 0) {
#pragma unroll
            for (int k = 0; k < kDof; ++k) q[k] = (float)dst[k];
        } else {
            const int mp = extra ? p - 1 : p;
            walk_state(a, b, nd, live ? mp : 0, q);
        }
        LaneSpheres<7> S;
        arm_fk<7, Shape::GS>(q, sh.slot, S);
        const bool hit = collide_links<OPS, Exit::kAnyGroup, true, kPlanPad>(S, env, ns, nb, live, sh.part);
        if (wave_ballot(hit && live)) return false;
    }
    return true;
}

// SpaceInformation::satisfiesBounds && isValid of state q (global memory) by
// the whole block (PlannerInputStates, base/src/Planner.cpp:243-245): within
// machine epsilon of [-pi, pi] per joint (RealVectorStateSpace.cpp:200-208),
// then the checker.  q is staged in sx; sa / sb are any LDS states.
template <int NT>
__device__ bool endpoint_usable(const double *q, double *sx, const double *sa, const double *sb, const float4 *env,
                                int ns, int nb, int *stop)
{
    if (threadIdx.x < kDof) sx[threadIdx.x] = q[threadIdx.x];
    if (threadIdx.x == 0) *stop = 0;
    __syncthreads();
    const double eps = 2.220446049250313080847e-16;
    bool in = true;
#pragma unroll
    for (int k = 0; k < kDof; ++k)
        if (sx[k] - eps > kPi || sx[k] + eps < -kPi) in = false;
    if (!in) return false; // block-uniform: every thread read the same state
    return block_check<NT, 9>(sa, sb, sx, 0, 0, 1, true, env, ns, nb, stop);
}

// Appends a valid growTree step to tree (T, P) after node nm: the new state x,
// or with addIntermediateStates (RRTConnect.cpp:148-183) the states of
// getMotionStates(a, b, validSegmentCount(a, b) = nd, endpoints) --
// a + (b - a) j / (nd + 1) -- without the tree's own end: from the start tree
// (a = nearest, b = x) states 1 .. nd + 1, from the goal tree (a = x, b =
// nearest) states nd .. 0, each the parent of the next.  Returns the index of
// the last (the new state), or -1 when the tree cannot hold them.
template <int NT>
__device__ int append_step(double *T, int32_t *P, int cap, int &n, int nm, const double *a, const double *b,
                           const double *x, int nd, bool from_start, bool inter)
{
    const int tid = threadIdx.x;
    if (!inter) {
        if (n + 1 > cap) return -1;
        if (tid < kDof) T[(int64_t)tid * cap + n] = x[tid];
        if (tid == 0) P[n] = nm;
        return n++;
    }
    const int seg = nd + 1;
    if (n + seg > cap) return -1;
    for (int i = tid; i < seg * kDof; i += NT) {
        const int r = i / kDof, k = i - r * kDof; // chain element r (0-based)
        const int j = from_start ? r + 1 : seg - 1 - r;
        double v;
        if (j == 0) v = a[k];
        else if (j == seg) v = b[k];
        else v = a[k] + (b[k] - a[k]) * ((double)j / (double)seg);
        T[(int64_t)k * cap + n + r] = v;
    }
    for (int r = tid; r < seg; r += NT) P[n + r] = r == 0 ? nm : n + r - 1;
    n += seg;
    return n - 1;
}

// Outcome of a speculative sampling step's own check.
enum SpecResult { kSpecTrapped = 0, kSpecValid = 1, kSpecUndecided = 2 };
// Positions a speculative step checks with its own wave: the new state (or
// s2) with the obstacle list split nine ways, then the next eight with one
// state per link group; a motion still valid after them is finished by the
// whole block if its step is the next one to commit.
constexpr int kSpecPositions = 9;

// Runs request blockIdx.x's RRTConnect::solve loop for one slice.  Control
// flow is block-uniform: every thread holds the loop's variables and updates
// them identically; values one thread produces (the sample, the nearest
// node) reach the others through LDS.
template <int NT>
__global__ void __launch_bounds__(NT) plan_slice_kernel(SliceArgs A)
{
    extern __shared__ float4 s_env[];
    __shared__ uint32_t s_mt[kMtN];
    __shared__ double s_rstate[kDof];
    __shared__ double s_wd[NT / 64];
    __shared__ int s_wi[NT / 64];
    __shared__ double s_a[kDof], s_b[kDof], s_x[kDof]; // the step's motion and new state
    __shared__ double s_u[kMtN / 2]; // the 0-1 draws of the generator's current 624-word block
    __shared__ int s_stop, s_hit, s_n1, s_len;
    // speculative sampling steps, one per wave: samples, motions, outcomes
    __shared__ double s_samp[NT / 64 * kDof], s_ga[NT / 64 * kDof], s_gb[NT / 64 * kDof], s_gx[NT / 64 * kDof];
    __shared__ int s_res[NT / 64], s_nm[NT / 64], s_reach[NT / 64], s_chk[NT / 64], s_nd[NT / 64],
        s_npos[NT / 64];
    const int tid = threadIdx.x;
    const int64_t t0 = wall_clock64();
    const PlanCmd &C = A.cmds[blockIdx.x];
    const RegionPtrs R = region_ptrs(C.region, C.cap);
    PlanState *st = R.state;

    int mti, status, iteration, n0, n1, start_tree, tgi_start, grow_tree, phase, added, xmotion, path_len;
    int sampled_goals;
    int64_t steps, checks;
    if (C.init) {
        if (tid == 0) mt_seed_serial(s_mt, C.seed);
        if (tid < kDof) s_rstate[tid] = 0.0;
        mti = kMtN;
        status = OMPL_AMD_PLAN_RUNNING;
        iteration = 0;
        n0 = n1 = 0; // roots: the usable start states below, goal states in the loop
        sampled_goals = 0;
        start_tree = 1;
        tgi_start = 1;
        grow_tree = 0;
        phase = 0;
        added = xmotion = -1;
        path_len = 0;
        steps = checks = 0;
    } else {
        for (int i = tid; i < kMtN; i += NT) s_mt[i] = st->mt[i];
        if (tid < kDof) s_rstate[tid] = st->rstate[tid];
        mti = st->mti;
        status = st->status;
        iteration = st->iteration;
        n0 = st->n0;
        n1 = st->n1;
        start_tree = st->start_tree;
        tgi_start = st->tgi_start;
        grow_tree = st->grow_tree;
        phase = st->phase;
        added = st->added;
        xmotion = st->xmotion;
        path_len = st->path_len;
        sampled_goals = st->sampled_goals;
        steps = st->steps;
        checks = st->checks;
    }
    __syncthreads();
    // the draws of the current block (consecutive word pairs; a sample's seven
    // draws are seven consecutive pairs, so a pair never spans two twists)
    auto refill_draws = [&]() {
        for (int i = tid; i < kMtN / 2; i += NT) s_u[i] = canonical53(mt_temper(s_mt[2 * i]), mt_temper(s_mt[2 * i + 1]));
    };
    if (mti < kMtN) refill_draws();
    int pidx = mti / 2; // next draw (words are consumed in pairs: mti is even)
    stage_env<true, kPlanPad>(EnvView{C.obs, C.ns, C.nb}, s_env); // ends with a barrier
    if (C.init) {
        // PlannerInputStates::nextStart (Planner.cpp:228-256): the start
        // states in order; those in bounds and valid become the start tree's
        // roots
        for (int i = 0; i < C.n_starts; ++i) {
            if (endpoint_usable<NT>(R.ends + (size_t)i * kDof, s_x, s_a, s_b, s_env, C.ns, C.nb, &s_hit) &&
                n0 < C.cap) {
                if (tid < kDof) R.tree_s[0][(int64_t)tid * C.cap + n0] = s_x[tid];
                if (tid == 0) R.parent[0][n0] = -1;
                ++n0;
            }
            __syncthreads(); // s_x is rewritten by the next state
        }
        if (n0 == 0) status = OMPL_AMD_PLAN_INVALID_START;
    }

    int done_here = 0;
    int64_t t_sample = 0, t_nearest = 0, t_check = 0, t_step = 0, c0 = 0, c1 = 0; // thread 0's timers
    while (status == OMPL_AMD_PLAN_RUNNING) {
        if (tid == 0) s_stop = done_here > 0 && wall_clock64() - t0 >= A.budget;
        __syncthreads();
        if (s_stop) break;
        // a step appends at most one node to one tree
        if (n0 >= C.cap || n1 >= C.cap) {
            status = OMPL_AMD_PLAN_TREE_FULL;
            break;
        }
        ++done_here;
        if (tid == 0) c0 = c1 = wall_clock64();
        // ---- speculative sampling steps.  A sampling step (phase 0) that is
        // TRAPPED changes neither tree, and the samples of the next steps are
        // the generator's next draws whatever happens, so the next m steps can
        // run at once, one per wave, against the current trees (step j grows
        // the start tree iff start_tree ^ (j & 1)); they are then taken in order
        // up to the first that is not TRAPPED, which is committed, and the rest
        // are dropped (their draws are not consumed).  Same steps, same results
        // as one at a time.  m stays within the generator's current block of
        // draws and the iterations left.
        // (not while a goal state is due: the goal tree would change between
        // the steps)
        const bool goal_due = n1 == 0 || (sampled_goals < C.n_goals && sampled_goals < n1 / 2);
        const int m_spec =
            phase == 0 && !goal_due ? min(NT / 64, min((kMtN / 2 - pidx) / kDof, C.max_iters - iteration)) : 0;
        if (m_spec >= 2) {
            const int m = m_spec;
            if (tid < kDof * m) s_samp[tid] = (kPi - -kPi) * s_u[pidx + tid] + -kPi;
            __syncthreads();
            if (tid == 0) {
                c1 = wall_clock64();
                t_sample += c1 - c0;
            }
            const int wave = tid >> 6, lane = tid & 63;
            if (wave < m) {
                const int st_j = start_tree ^ (wave & 1);
                const int tree_j = st_j ? 0 : 1;
                const double *Tj = R.tree_s[tree_j];
                double tgt[kDof];
#pragma unroll
                for (int k = 0; k < kDof; ++k) tgt[k] = s_samp[kDof * wave + k];
                const int nm = wave_nearest(Tj, C.cap, tree_j ? n1 : n0, tgt);
                if (tid == 0) {
                    const int64_t c2 = wall_clock64();
                    t_nearest += c2 - c1;
                }
                double nst[kDof], x[kDof];
                double dist = 0.0;
#pragma unroll
                for (int k = 0; k < kDof; ++k) {
                    nst[k] = Tj[(int64_t)k * C.cap + nm];
                    const double diff = nst[k] - tgt[k];
                    dist += diff * diff;
                }
                const double dd = __builtin_sqrt(dist);
                bool reach = true, moved = true;
#pragma unroll
                for (int k = 0; k < kDof; ++k) x[k] = tgt[k];
                if (dd > C.maxd) {
                    const double t = C.maxd / dd;
                    bool equal = true;
#pragma unroll
                    for (int k = 0; k < kDof; ++k) {
                        x[k] = nst[k] + (tgt[k] - nst[k]) * t;
                        if (fabs(nst[k] - x[k]) > kEqualEps) equal = false;
                    }
                    moved = !equal;
                    reach = false;
                }
                int res = kSpecTrapped, nd = 0, npos = 0;
                if (moved) {
                    double *ga = s_ga + kDof * wave, *gb = s_gb + kDof * wave, *gx = s_gx + kDof * wave;
                    double a[kDof], b[kDof];
#pragma unroll
                    for (int k = 0; k < kDof; ++k) {
                        a[k] = st_j ? nst[k] : x[k];
                        b[k] = st_j ? x[k] : nst[k];
                    }
                    nd = valid_segment_count<kDof>(a, b, C.lvs, C.factor);
                    if (lane == 0)
                        for (int k = 0; k < kDof; ++k) {
                            ga[k] = a[k];
                            gb[k] = b[k];
                            gx[k] = x[k];
                        }
                    // lane 0's LDS writes before the wave's reads
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (nd <= kMaxSegments) {
                        const bool extra = !st_j;
                        npos = 1 + (nd >= 2 ? nd - 1 : 0) + (extra ? 1 : 0);
                        bool valid = wave_check<9>(ga, gb, gx, nd, 0, 1, extra, s_env, C.ns, C.nb);
                        if (valid && npos > 1)
                            valid = wave_check<1>(ga, gb, gx, nd, 1, min(npos, kSpecPositions), extra, s_env, C.ns,
                                                  C.nb);
                        res = !valid ? kSpecTrapped : (npos <= kSpecPositions ? kSpecValid : kSpecUndecided);
                    }
                }
                if (lane == 0) {
                    s_res[wave] = res;
                    s_nm[wave] = nm;
                    s_reach[wave] = reach ? 1 : 0;
                    s_chk[wave] = moved ? 1 : 0;
                    s_nd[wave] = nd;
                    s_npos[wave] = npos;
                }
            }
            __syncthreads();
            // take the steps in order
            for (int j = 0; j < m; ++j) {
                const int st_j = start_tree;
                int res = s_res[j];
                ++steps;
                checks += s_chk[j];
                if (res == kSpecUndecided) { // the rest of the motion by the whole block
                    if (tid == 0) s_hit = 0;
                    __syncthreads();
                    const int nd = s_nd[j], npos = s_npos[j];
                    const double *ga = s_ga + kDof * j, *gb = s_gb + kDof * j, *gx = s_gx + kDof * j;
                    const bool extra = !st_j;
                    const int ops = C.ops ? C.ops : auto_ops<NT / 64>(npos - kSpecPositions);
                    bool valid;
                    if (ops == 9)
                        valid = block_check<NT, 9>(ga, gb, gx, nd, kSpecPositions, npos, extra, s_env, C.ns, C.nb,
                                                   &s_hit);
                    else if (ops == 4)
                        valid = block_check<NT, 4>(ga, gb, gx, nd, kSpecPositions, npos, extra, s_env, C.ns, C.nb,
                                                   &s_hit);
                    else if (ops == 2)
                        valid = block_check<NT, 2>(ga, gb, gx, nd, kSpecPositions, npos, extra, s_env, C.ns, C.nb,
                                                   &s_hit);
                    else
                        valid = block_check<NT, 1>(ga, gb, gx, nd, kSpecPositions, npos, extra, s_env, C.ns, C.nb,
                                                   &s_hit);
                    res = valid ? kSpecValid : kSpecTrapped;
                }
                start_tree = !start_tree; // RRTConnect.cpp:249
                pidx += kDof;             // this step's sample is consumed
                if (res == kSpecTrapped) {
                    ++iteration;
                    if (iteration >= C.max_iters) {
                        status = OMPL_AMD_PLAN_FAILED;
                        break;
                    }
                    continue;
                }
                // commit step j: append its new state, then connect (phase 1)
                const int tree_j = st_j ? 0 : 1;
                const int last = append_step<NT>(R.tree_s[tree_j], R.parent[tree_j], C.cap, tree_j ? n1 : n0, s_nm[j],
                                                 s_ga + kDof * j, s_gb + kDof * j, s_gx + kDof * j, s_nd[j], st_j != 0,
                                                 C.inter != 0);
                if (last < 0) {
                    status = OMPL_AMD_PLAN_TREE_FULL;
                    break;
                }
                // rstate: the sample when reached, else the new state
                if (tid < kDof) s_rstate[tid] = s_reach[j] ? s_samp[kDof * j + tid] : s_gx[kDof * j + tid];
                xmotion = last;
                added = last;
                tgi_start = start_tree;
                grow_tree = start_tree ? 0 : 1; // otherTree
                phase = 1;
                break;
            }
            if (tid == 0) {
                const int64_t c3 = wall_clock64();
                t_check += c3 - c1;
                t_step += c3 - c0;
            }
            continue;
        }
        if (phase == 0) {
            // RRTConnect.cpp:247-250: pick the tree, flip startTree_
            grow_tree = start_tree ? 0 : 1;
            tgi_start = start_tree;
            start_tree = !start_tree;
            // RRTConnect.cpp:251-269: goal states (GoalStates, in order) --
            // every remaining one tried while the goal tree is empty
            // (nextGoal(ptc)), else one while fewer than half the goal tree's
            // size have been handed out (nextGoal())
            if (n1 == 0 || sampled_goals < n1 / 2) {
                const bool empty = n1 == 0;
                while (sampled_goals < C.n_goals) {
                    const int g = sampled_goals++;
                    if (endpoint_usable<NT>(R.ends + (size_t)(kMaxEnds + g) * kDof, s_x, s_a, s_b, s_env, C.ns, C.nb,
                                            &s_hit)) {
                        if (n1 >= C.cap) {
                            status = OMPL_AMD_PLAN_TREE_FULL;
                        } else {
                            if (tid < kDof) R.tree_s[1][(int64_t)tid * C.cap + n1] = s_x[tid];
                            if (tid == 0) R.parent[1][n1] = -1;
                            ++n1;
                        }
                        __syncthreads();
                        break;
                    }
                    __syncthreads();
                    if (!empty) break;
                }
                if (status == OMPL_AMD_PLAN_RUNNING && n1 == 0) status = OMPL_AMD_PLAN_INVALID_GOAL;
                if (status != OMPL_AMD_PLAN_RUNNING) break;
            }
            // sampler_->sampleUniform(rstate): RNG::uniformReal(-pi, pi) per joint,
            // from the next seven draws (twisting when the block runs out)
            for (int d0 = 0; d0 < kDof;) {
                if (pidx >= kMtN / 2) {
                    mt_twist_block(s_mt);
                    refill_draws();
                    __syncthreads();
                    pidx = 0;
                }
                const int take = min(kDof - d0, kMtN / 2 - pidx);
                if (tid < take) s_rstate[d0 + tid] = (kPi - -kPi) * s_u[pidx + tid] + -kPi;
                pidx += take;
                d0 += take;
            }
            __syncthreads();
            if (tid == 0) {
                c1 = wall_clock64();
                t_sample += c1 - c0;
            }
        }

        // ---- growTree(tree, tgi, rmotion) (RRTConnect.cpp:116-196)
        double tgt[kDof];
#pragma unroll
        for (int k = 0; k < kDof; ++k) tgt[k] = s_rstate[k];
        const int nt = grow_tree ? n1 : n0;
        double *T = R.tree_s[grow_tree];
        const int nm = block_nearest<NT>(T, C.cap, nt, tgt, s_wd, s_wi);
        if (tid == 0) {
            const int64_t c2 = wall_clock64();
            t_nearest += c2 - c1;
            c1 = c2;
        }
        double nst[kDof], x[kDof];
        double dist = 0.0;
#pragma unroll
        for (int k = 0; k < kDof; ++k) {
            nst[k] = T[(int64_t)k * C.cap + nm];
            const double diff = nst[k] - tgt[k];
            dist += diff * diff;
        }
        const double dd = __builtin_sqrt(dist);
        bool reach = true, moved = true;
#pragma unroll
        for (int k = 0; k < kDof; ++k) x[k] = tgt[k];
        if (dd > C.maxd) {
            const double t = C.maxd / dd;
            bool equal = true;
#pragma unroll
            for (int k = 0; k < kDof; ++k) {
                x[k] = nst[k] + (tgt[k] - nst[k]) * t;
                if (fabs(nst[k] - x[k]) > kEqualEps) equal = false;
            }
            moved = !equal; // no progress: TRAPPED without a check
            reach = false;
        }
        int gs = kTrapped;
        if (moved) {
            // tgi.start ? checkMotion(nearest, dst) : isValid(dst) && checkMotion(dst, nearest)
            const bool from_start = tgi_start != 0;
            double a[kDof], b[kDof];
#pragma unroll
            for (int k = 0; k < kDof; ++k) {
                a[k] = from_start ? nst[k] : x[k];
                b[k] = from_start ? x[k] : nst[k];
            }
            const int nd = valid_segment_count<kDof>(a, b, C.lvs, C.factor);
            if (tid == 0) {
                for (int k = 0; k < kDof; ++k) {
                    s_a[k] = a[k];
                    s_b[k] = b[k];
                    s_x[k] = x[k];
                }
                s_hit = 0;
            }
            __syncthreads();
            ++checks;
            bool valid = false;
            if (nd <= kMaxSegments) { // longer motions are reported invalid
                const int npos = 1 + (nd >= 2 ? nd - 1 : 0) + (from_start ? 0 : 1);
                // First round: one position per wave, the obstacle list split nine
                // ways (a failing motion -- most are -- usually fails at the new
                // state or an early bisection point, found within a pass or two);
                // then the rest in rounds shaped by the configured or auto split.
                const int nfirst = min(npos, NT / 64);
                valid = block_check<NT, 9>(s_a, s_b, s_x, nd, 0, nfirst, !from_start, s_env, C.ns, C.nb, &s_hit);
                if (valid && nfirst < npos) {
                    const int ops = C.ops ? C.ops : auto_ops<NT / 64>(npos - nfirst);
                    if (ops == 9)
                        valid = block_check<NT, 9>(s_a, s_b, s_x, nd, nfirst, npos, !from_start, s_env, C.ns, C.nb,
                                                   &s_hit);
                    else if (ops == 4)
                        valid = block_check<NT, 4>(s_a, s_b, s_x, nd, nfirst, npos, !from_start, s_env, C.ns, C.nb,
                                                   &s_hit);
                    else if (ops == 2)
                        valid = block_check<NT, 2>(s_a, s_b, s_x, nd, nfirst, npos, !from_start, s_env, C.ns, C.nb,
                                                   &s_hit);
                    else
                        valid = block_check<NT, 1>(s_a, s_b, s_x, nd, nfirst, npos, !from_start, s_env, C.ns, C.nb,
                                                   &s_hit);
                }
            }
            if (valid) {
                const int last = append_step<NT>(T, R.parent[grow_tree], C.cap, grow_tree ? n1 : n0, nm, s_a, s_b, s_x,
                                                 nd, from_start, C.inter != 0);
                if (last < 0) {
                    status = OMPL_AMD_PLAN_TREE_FULL;
                    break;
                }
                xmotion = last;
                gs = reach ? kReached : kAdvanced;
            }
        }
        ++steps;
        if (tid == 0) t_check += wall_clock64() - c1;

        // ---- solve loop (RRTConnect.cpp:245-330)
        bool end_iteration = false;
        if (phase == 0) {
            if (gs == kTrapped) {
                end_iteration = true;
            } else {
                added = xmotion;
                // if not reached, grow the other tree towards the added state
                if (gs != kReached && tid == 0)
                    for (int k = 0; k < kDof; ++k) s_rstate[k] = s_x[k];
                tgi_start = start_tree;
                grow_tree = start_tree ? 0 : 1; // otherTree
                phase = 1;
            }
        } else {
            if (phase == 1 && gs == kTrapped) tgi_start = !tgi_start;
            if (gs == kAdvanced) {
                phase = 2;
            } else if (gs == kReached) {
                // RRTConnect.cpp:301-330: step back once on the start side (or
                // the goal side at its root), then start chain reversed + goal chain
                int sm = tgi_start ? xmotion : added, gm = tgi_start ? added : xmotion;
                if (tid == 0) {
                    const int32_t *ps = R.parent[0], *pg = R.parent[1];
                    if (ps[sm] >= 0) sm = ps[sm];
                    else gm = pg[gm];
                    int c = 0;
                    for (int m = sm; m >= 0; m = ps[m]) R.path_idx[c++] = m;
                    s_n1 = c;
                    for (int m = gm; m >= 0; m = pg[m]) R.path_idx[c++] = m;
                    s_len = c;
                }
                __syncthreads();
                const int c1 = s_n1, len = s_len;
                for (int r = tid; r < len * kDof; r += NT) {
                    const int row = r / kDof, k = r - row * kDof;
                    R.path[r] = row < c1 ? R.tree_s[0][(int64_t)k * C.cap + R.path_idx[c1 - 1 - row]]
                                         : R.tree_s[1][(int64_t)k * C.cap + R.path_idx[row]];
                }
                path_len = len;
                status = OMPL_AMD_PLAN_SOLVED;
                ++iteration;
            } else {
                end_iteration = true;
            }
        }
        if (end_iteration) {
            phase = 0;
            ++iteration;
            if (iteration >= C.max_iters) status = OMPL_AMD_PLAN_FAILED;
        }
        if (tid == 0) t_step += wall_clock64() - c0;
    }

    __syncthreads();
    for (int i = tid; i < kMtN; i += NT) st->mt[i] = s_mt[i];
    if (tid < kDof) st->rstate[tid] = s_rstate[tid];
    if (tid == 0) {
        st->mti = 2 * pidx;
        st->status = status;
        st->iteration = iteration;
        st->n0 = n0;
        st->n1 = n1;
        st->start_tree = start_tree;
        st->tgi_start = tgi_start;
        st->grow_tree = grow_tree;
        st->phase = phase;
        st->added = added;
        st->xmotion = xmotion;
        st->path_len = path_len;
        st->sampled_goals = sampled_goals;
        st->steps = steps;
        st->checks = checks;
        PlanOut &o = A.out[blockIdx.x];
        o.status = status;
        o.iteration = iteration;
        o.n0 = n0;
        o.n1 = n1;
        o.path_len = path_len;
        o.pad = 0;
        o.steps = steps;
        o.checks = checks;
        o.t_sample = t_sample;
        o.t_nearest = t_nearest;
        o.t_check = t_check;
        o.t_step = t_step;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
using Clock = std::chrono::steady_clock;

static double ms_since(Clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

struct Planner {
    ompl_amd_plan_request req; // starts / goals point into `ends`
    // kMaxEnds start rows, then kMaxEnds goal rows (the region's layout)
    std::vector<double> ends = std::vector<double>((size_t)2 * kMaxEnds * kDof, 0.0);
    double maxd = 0.0;
    Clock::time_point t_submit, t_start, t_finish;
    // published: written and read under Server::mu only
    int status = OMPL_AMD_PLAN_QUEUED;
    int iterations = 0;
    int64_t nodes = 0;
    int64_t checks = 0;
    std::vector<double> path;
    bool started = false;
    bool cancel = false;
    bool env_held = false; // counted in the environment's server_users
    // server thread only
    size_t off = 0, bytes = 0; // region in the pool
    bool init = true;
    int64_t steps_seen = 0;
};

// Copies a request's start and goal states into the planner (the caller's
// lists need not outlive the submit): the lists, or start / goal when a list
// is empty.
static void take_endpoints(Planner &p)
{
    ompl_amd_plan_request &r = p.req;
    if (r.n_starts > 0) std::copy(r.starts, r.starts + (size_t)r.n_starts * kDof, p.ends.begin());
    else std::copy(r.start, r.start + kDof, p.ends.begin());
    if (r.n_goals > 0) std::copy(r.goals, r.goals + (size_t)r.n_goals * kDof, p.ends.begin() + (size_t)kMaxEnds * kDof);
    else std::copy(r.goal, r.goal + kDof, p.ends.begin() + (size_t)kMaxEnds * kDof);
    r.n_starts = r.n_starts > 0 ? r.n_starts : 1;
    r.n_goals = r.n_goals > 0 ? r.n_goals : 1;
    std::copy(p.ends.begin(), p.ends.begin() + kDof, r.start);
    std::copy(p.ends.begin() + (size_t)kMaxEnds * kDof, p.ends.begin() + (size_t)(kMaxEnds + 1) * kDof, r.goal);
    r.starts = p.ends.data();
    r.goals = p.ends.data() + (size_t)kMaxEnds * kDof;
}

// First-fit allocator over the device pool (offsets; neighbours coalesce).
struct PoolAlloc {
    std::map<size_t, size_t> free; // offset -> size
    void reset(size_t size)
    {
        free.clear();
        if (size) free[0] = size;
    }
    bool alloc(size_t bytes, size_t &off)
    {
        for (auto it = free.begin(); it != free.end(); ++it)
            if (it->second >= bytes) {
                off = it->first;
                const size_t rest = it->second - bytes;
                free.erase(it);
                if (rest) free[off + bytes] = rest;
                return true;
            }
        return false;
    }
    void release(size_t off, size_t bytes)
    {
        auto it = free.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free.erase(nx);
        }
        if (it != free.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free.erase(it);
            }
        }
    }
};

struct Server {
    ompl_amd_server_config cfg;
    hipStream_t stream = nullptr;
    char *pool = nullptr;
    PoolAlloc alloc;
    PlanCmd *cmd_host = nullptr, *cmd_dev = nullptr;
    PlanOut *out_host = nullptr, *out_dev = nullptr;
    int64_t budget_ticks = 0;

    std::mutex mu;
    std::condition_variable cv, done_cv;
    std::deque<int64_t> queue;
    std::unordered_map<int64_t, Planner *> reqs;
    std::vector<int64_t> active;
    int64_t next_id = 1;
    size_t used_bytes = 0;
    bool stop = false;
    std::thread th;
    std::atomic<int64_t> slices{0}, steps{0};
    int64_t phase_ticks[4] = {0, 0, 0, 0}; // sample, nearest, check, step (under mu)
    int64_t clock_khz = 0;
    std::string fault;
    // PRM requests: their own queue, worker thread and stream
    std::deque<int64_t> prm_queue;
    std::condition_variable prm_cv; // PRM arrivals and freed pool memory
    std::thread prm_th;
    hipStream_t prm_stream = nullptr;
    int prm_running = 0;

    void run();
    void run_prm();
    void admit_locked();
    void finish_locked(Planner *p, int status);
};

// Publishes a final status (caller holds mu), returns the region and
// releases the request's hold on its environment.
void Server::finish_locked(Planner *p, int status)
{
    p->status = status;
    p->t_finish = Clock::now();
    if (p->env_held) {
        p->req.env->server_users.fetch_sub(1);
        p->env_held = false;
    }
    if (p->bytes) {
        alloc.release(p->off, p->bytes);
        used_bytes -= p->bytes;
        p->bytes = 0;
        prm_cv.notify_all(); // a PRM request may be waiting for memory
    }
}

static size_t request_footprint(const ompl_amd_plan_request &r)
{
    return r.planner == OMPL_AMD_PLANNER_PRM ? prm_footprint(r.prm_samples, r.prm_k) : footprint(r.max_nodes);
}

// PRM requests, in arrival order, each once its roadmap's footprint is
// reserved in the pool (the roadmap itself is built in memory of its own).
void Server::run_prm()
{
    try {
        OA_HIP(hipSetDevice(device()));
        for (;;) {
            Planner *p = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    if (stop) return;
                    if (!fault.empty()) return;
                    if (!prm_queue.empty()) {
                        Planner *q = reqs[prm_queue.front()];
                        const bool expired = q->req.deadline_ms > 0.0 && ms_since(q->t_submit) > q->req.deadline_ms;
                        if (q->cancel || expired) {
                            prm_queue.pop_front();
                            finish_locked(q, q->cancel ? OMPL_AMD_PLAN_CANCELLED : OMPL_AMD_PLAN_DEADLINE);
                            done_cv.notify_all();
                            continue;
                        }
                        const size_t fp = request_footprint(q->req);
                        size_t off = 0;
                        if (alloc.alloc(fp, off)) {
                            prm_queue.pop_front();
                            q->off = off;
                            q->bytes = fp;
                            used_bytes += fp;
                            q->status = OMPL_AMD_PLAN_RUNNING;
                            q->started = true;
                            q->t_start = Clock::now();
                            ++prm_running;
                            p = q;
                            break;
                        }
                    }
                    prm_cv.wait_for(lk, std::chrono::milliseconds(50));
                }
            }
            const ompl_amd_plan_request &r = p->req;
            const PrmQuery q = prm_query(r.env, r.start, r.goal, r.seed, 64, r.prm_samples, r.prm_k,
                                         r.longest_valid_segment, r.factor, prm_stream);
            std::lock_guard<std::mutex> lk(mu);
            p->iterations = 1;
            p->nodes = q.milestones;
            p->checks = q.candidate_edges;
            p->path = q.path;
            --prm_running;
            finish_locked(p, q.solved ? OMPL_AMD_PLAN_SOLVED : OMPL_AMD_PLAN_FAILED);
            done_cv.notify_all();
        }
    } catch (const std::exception &e) {
        std::lock_guard<std::mutex> lk(mu);
        if (fault.empty()) fault = e.what();
        for (auto &kv : reqs)
            if (kv.second->req.planner == OMPL_AMD_PLANNER_PRM && kv.second->status < OMPL_AMD_PLAN_SOLVED)
                finish_locked(kv.second, OMPL_AMD_PLAN_FAILED);
        prm_queue.clear();
        prm_running = 0;
        done_cv.notify_all();
    }
}

void Server::admit_locked()
{
    while (!queue.empty() && (int)active.size() < cfg.max_active) {
        const int64_t id = queue.front();
        Planner *p = reqs[id];
        const size_t fp = footprint(p->req.max_nodes);
        size_t off = 0;
        if (!alloc.alloc(fp, off)) break; // FIFO: the head waits for memory
        queue.pop_front();
        p->off = off;
        p->bytes = fp;
        used_bytes += fp;
        p->init = true;
        p->status = OMPL_AMD_PLAN_RUNNING;
        p->started = true;
        p->t_start = Clock::now();
        active.push_back(id);
    }
}

void Server::run()
{
    try {
        OA_HIP(hipSetDevice(device()));
        std::vector<Planner *> batch;
        std::vector<int64_t> batch_ids;
        std::vector<std::vector<double>> paths;
        for (;;) {
            batch.clear();
            batch_ids.clear();
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !queue.empty() || !active.empty(); });
                if (stop) return;
                bool notify = false;
                for (auto q = queue.begin(); q != queue.end();) {
                    Planner *p = reqs[*q];
                    if (p->req.deadline_ms > 0.0 && ms_since(p->t_submit) > p->req.deadline_ms) {
                        finish_locked(p, OMPL_AMD_PLAN_DEADLINE);
                        q = queue.erase(q);
                        notify = true;
                    } else {
                        ++q;
                    }
                }
                for (size_t i = 0; i < active.size();) {
                    Planner *p = reqs[active[i]];
                    const bool expired = p->req.deadline_ms > 0.0 && ms_since(p->t_submit) > p->req.deadline_ms;
                    if (p->cancel || expired) {
                        finish_locked(p, p->cancel ? OMPL_AMD_PLAN_CANCELLED : OMPL_AMD_PLAN_DEADLINE);
                        active.erase(active.begin() + (std::ptrdiff_t)i);
                        notify = true;
                    } else {
                        ++i;
                    }
                }
                admit_locked();
                for (int64_t id : active) {
                    batch.push_back(reqs[id]);
                    batch_ids.push_back(id);
                }
                if (notify) done_cv.notify_all();
            }
            if (batch.empty()) continue;
            const int n = (int)batch.size();
            size_t lds = 0;
            for (int i = 0; i < n; ++i) {
                Planner &p = *batch[i];
                const Env &env = p.req.env->env;
                PlanCmd &c = cmd_host[i];
                c.obs = env.obs;
                c.region = pool + p.off;
                c.ns = env.ns;
                c.nb = env.nb;
                c.cap = p.req.max_nodes;
                c.max_iters = p.req.max_iters;
                c.maxd = p.maxd;
                c.lvs = p.req.longest_valid_segment;
                c.factor = p.req.factor;
                c.seed = p.req.seed;
                c.init = p.init ? 1 : 0;
                c.ops = cfg.check_obstacle_parts;
                c.n_starts = p.req.n_starts;
                c.n_goals = p.req.n_goals;
                c.inter = p.req.intermediate_states ? 1 : 0;
                if (p.init) {
                    // the request's start and goal states into its region
                    const RegionPtrs Rp = region_ptrs(pool + p.off, p.req.max_nodes);
                    OA_HIP(hipMemcpyAsync(Rp.ends, p.ends.data(), p.ends.size() * sizeof(double),
                                          hipMemcpyHostToDevice, stream));
                }
                lds = std::max(lds, env_stage_bytes(env.ns, env.nb));
                p.init = false;
            }
            OA_HIP(hipMemcpyAsync(cmd_dev, cmd_host, n * sizeof(PlanCmd), hipMemcpyHostToDevice, stream));
            {
                ProfScope ps("serve_slice", stream);
                const SliceArgs args{cmd_dev, out_dev, budget_ticks};
                if (cfg.block_threads == 1024)
                    hipLaunchKernelGGL(plan_slice_kernel<1024>, dim3(n), dim3(1024), lds, stream, args);
                else
                    hipLaunchKernelGGL(plan_slice_kernel<512>, dim3(n), dim3(512), lds, stream, args);
                OA_HIP(hipGetLastError());
            }
            OA_HIP(hipMemcpyAsync(out_host, out_dev, n * sizeof(PlanOut), hipMemcpyDeviceToHost, stream));
            OA_HIP(hipStreamSynchronize(stream));
            ++slices;
            // solution paths of the requests that solved in this slice
            paths.assign(n, {});
            bool copied = false;
            for (int i = 0; i < n; ++i) {
                const PlanOut &o = out_host[i];
                if (o.status != OMPL_AMD_PLAN_SOLVED || o.path_len <= 0) continue;
                paths[i].resize((size_t)o.path_len * kDof);
                const RegionPtrs Rp = region_ptrs(pool + batch[i]->off, batch[i]->req.max_nodes);
                OA_HIP(hipMemcpyAsync(paths[i].data(), Rp.path, paths[i].size() * sizeof(double),
                                      hipMemcpyDeviceToHost, stream));
                copied = true;
            }
            if (copied) OA_HIP(hipStreamSynchronize(stream));
            int64_t new_steps = 0;
            {
                std::lock_guard<std::mutex> lk(mu);
                bool notify = false;
                for (int i = 0; i < n; ++i) {
                    Planner *p = batch[i];
                    const PlanOut &o = out_host[i];
                    new_steps += o.steps - p->steps_seen;
                    phase_ticks[0] += o.t_sample;
                    phase_ticks[1] += o.t_nearest;
                    phase_ticks[2] += o.t_check;
                    phase_ticks[3] += o.t_step;
                    p->steps_seen = o.steps;
                    p->iterations = o.iteration;
                    p->nodes = (int64_t)o.n0 + o.n1;
                    p->checks = o.checks;
                    if (o.status == OMPL_AMD_PLAN_RUNNING) continue;
                    p->path.swap(paths[i]);
                    finish_locked(p, o.status);
                    active.erase(std::find(active.begin(), active.end(), batch_ids[i]));
                    notify = true;
                }
                if (notify) done_cv.notify_all();
            }
            steps += new_steps;
        }
    } catch (const std::exception &e) {
        std::lock_guard<std::mutex> lk(mu);
        fault = e.what();
        for (int64_t id : active) finish_locked(reqs[id], OMPL_AMD_PLAN_FAILED);
        for (int64_t id : queue) finish_locked(reqs[id], OMPL_AMD_PLAN_FAILED);
        for (int64_t id : prm_queue) finish_locked(reqs[id], OMPL_AMD_PLAN_FAILED);
        active.clear();
        queue.clear();
        prm_queue.clear();
        done_cv.notify_all();
        prm_cv.notify_all();
    }
}

static bool finished_status(int s)
{
    return s >= OMPL_AMD_PLAN_SOLVED;
}

// Caller holds Server::mu.
static void fill_result(const Planner &p, ompl_amd_plan_result *r)
{
    r->status = p.status;
    r->iterations = p.iterations;
    r->path_len = (int)(p.path.size() / kDof);
    r->tree_nodes = p.nodes;
    r->motion_checks = p.checks;
    auto ms = [](Clock::time_point a, Clock::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    const bool done = finished_status(p.status);
    const Clock::time_point end = done ? p.t_finish : Clock::now();
    r->queued_ms = p.started ? ms(p.t_submit, p.t_start) : ms(p.t_submit, end);
    r->run_ms = p.started ? ms(p.t_start, end) : 0.0;
    r->latency_ms = ms(p.t_submit, end);
}

// ---------------------------------------------------------------------------
// plain path: host RRT-Connect, host nearest neighbours, one launch per check
// ---------------------------------------------------------------------------
struct HostTree {
    std::vector<double> s;
    std::vector<int> parent;
    int size() const { return (int)parent.size(); }
};

static int plain_grow(const Env &env, HostTree &t, bool start_tree, const double *target, double maxd, double lvs,
                      unsigned factor, bool inter, int &xmotion, double *xstate, int64_t &checks,
                      ThreadResources &tr)
{
    int nm = 0;
    double dmin = 0.0;
    for (int i = 0; i < t.size(); ++i) {
        double dist = 0.0;
        for (int k = 0; k < kDof; ++k) {
            const double diff = t.s[(size_t)i * kDof + k] - target[k];
            dist += diff * diff;
        }
        const double d = std::sqrt(dist);
        if (i == 0 || dmin > d) {
            nm = i;
            dmin = d;
        }
    }
    double ns[kDof]; // the tree may grow below
    std::copy(t.s.begin() + (size_t)nm * kDof, t.s.begin() + (size_t)(nm + 1) * kDof, ns);
    bool reach = true;
    double x[kDof];
    const double *dst = target;
    if (dmin > maxd) {
        bool equal = true;
        for (int k = 0; k < kDof; ++k) {
            x[k] = ns[k] + (target[k] - ns[k]) * (maxd / dmin);
            if (std::fabs(ns[k] - x[k]) > 2.220446049250313080847e-16 * 2.0) equal = false;
        }
        if (equal) return kTrapped;
        dst = x;
        reach = false;
    }
    auto check = [&](const double *a, const double *b) {
        Edge1 E;
        for (int k = 0; k < kDof; ++k) {
            E.a[k] = a[k];
            E.b[k] = b[k];
        }
        launch_motion_single(env.view(), E, lvs, factor, tr.pinned_dev, tr.stream);
        OA_HIP(hipStreamSynchronize(tr.stream));
        return tr.pinned[0] != 0;
    };
    bool valid;
    ++checks;
    if (start_tree) {
        valid = check(ns, dst);
    } else {
        State1 st;
        for (int k = 0; k < kDof; ++k) st.q[k] = dst[k];
        launch_validate_single(env.view(), st, tr.pinned_dev, tr.stream);
        OA_HIP(hipStreamSynchronize(tr.stream));
        valid = tr.pinned[0] != 0 && check(dst, ns);
    }
    if (!valid) return kTrapped;
    if (inter) {
        // getMotionStates(a, b, validSegmentCount(a, b), endpoints): states
        // 1 .. nd + 1 from the start tree, nd .. 0 from the goal tree
        const double *a = start_tree ? ns : dst, *b = start_tree ? dst : ns;
        const int seg = (int)valid_segment_count<kDof>(a, b, lvs, factor) + 1;
        int parent = nm;
        for (int r = 0; r < seg; ++r) {
            const int j = start_tree ? r + 1 : seg - 1 - r;
            for (int k = 0; k < kDof; ++k)
                t.s.push_back(j == 0 ? a[k] : (j == seg ? b[k] : a[k] + (b[k] - a[k]) * ((double)j / (double)seg)));
            t.parent.push_back(parent);
            parent = t.size() - 1;
        }
    } else {
        t.s.insert(t.s.end(), dst, dst + kDof);
        t.parent.push_back(nm);
    }
    xmotion = t.size() - 1;
    for (int k = 0; k < kDof; ++k) xstate[k] = dst[k];
    return reach ? kReached : kAdvanced;
}

// satisfiesBounds && isValid of one state (PlannerInputStates): one launch.
static bool plain_usable(const Env &env, const double *q, ThreadResources &tr)
{
    const double eps = 2.220446049250313080847e-16;
    for (int k = 0; k < kDof; ++k)
        if (q[k] - eps > kPi || q[k] + eps < -kPi) return false;
    State1 st;
    for (int k = 0; k < kDof; ++k) st.q[k] = q[k];
    launch_validate_single(env.view(), st, tr.pinned_dev, tr.stream);
    OA_HIP(hipStreamSynchronize(tr.stream));
    return tr.pinned[0] != 0;
}

// Plain PRM: the served PRM request's batch roadmap built on the host with one
// GPU launch and synchronisation per isValid / checkMotion call: samples from
// host std::mt19937 generators (the device sampler's streams), k-NN over the
// milestones in FP32 with the device search's distance and (distance, index)
// order, candidate edges checked from the lower milestone to the higher, host
// union-find, and the same shortest path.
static bool plain_prm(const ompl_amd_plan_request &req, const Env &env, ThreadResources &tr, std::vector<double> &out,
                      int64_t &milestones, int64_t &checks)
{
    const int streams = 64;
    const int samples = req.prm_samples > 0 ? req.prm_samples : 4096;
    const int k = req.prm_k > 0 ? req.prm_k : 10;
    const int64_t per = (samples + streams - 1) / streams;
    const int64_t n = 2 + per * streams;
    std::vector<double> smp((size_t)n * kDof);
    for (int d = 0; d < kDof; ++d) {
        smp[(size_t)d] = req.start[d];
        smp[(size_t)kDof + d] = req.goal[d];
    }
    for (int st = 0; st < streams; ++st) {
        std::mt19937 gen(req.seed + (uint32_t)st);
        std::uniform_real_distribution<> uni{0, 1};
        for (int64_t i = 0; i < per; ++i)
            for (int d = 0; d < kDof; ++d)
                smp[(size_t)(2 + st * per + i) * kDof + d] = (kPi - -kPi) * uni(gen) + -kPi;
    }
    std::vector<int32_t> ms;
    for (int64_t r = 0; r < n; ++r) {
        State1 q;
        for (int d = 0; d < kDof; ++d) q.q[d] = smp[(size_t)r * kDof + d];
        launch_validate_single(env.view(), q, tr.pinned_dev, tr.stream);
        OA_HIP(hipStreamSynchronize(tr.stream));
        if (tr.pinned[0] != 0) ms.push_back((int32_t)r);
    }
    const int64_t nv = (int64_t)ms.size();
    milestones = nv;
    checks = 0;
    if (nv < 2 || ms[0] != 0 || ms[1] != 1) return false;
    std::vector<double> nodes((size_t)nv * kDof);
    std::vector<float> f32((size_t)nv * kDof);
    for (int64_t i = 0; i < nv; ++i)
        for (int d = 0; d < kDof; ++d) {
            nodes[(size_t)i * kDof + d] = smp[(size_t)ms[(size_t)i] * kDof + d];
            f32[(size_t)i * kDof + d] = (float)nodes[(size_t)i * kDof + d];
        }
    // k + 1 nearest (the milestone itself among them), first k others kept
    const int64_t kk = std::min<int64_t>(k + 1, nv);
    std::vector<uint64_t> keys;
    std::vector<std::pair<float, int32_t>> cand((size_t)nv);
    for (int64_t i = 0; i < nv; ++i) {
        for (int64_t j = 0; j < nv; ++j) {
            float dist = 0.0f;
            for (int d = 0; d < kDof; ++d) {
                const float diff = f32[(size_t)j * kDof + d] - f32[(size_t)i * kDof + d];
                dist += diff * diff;
            }
            cand[(size_t)j] = {std::sqrt(dist), (int32_t)j};
        }
        std::partial_sort(cand.begin(), cand.begin() + kk, cand.end());
        int taken = 0;
        for (int64_t t = 0; t < kk && taken < k; ++t) {
            const int64_t j = cand[(size_t)t].second;
            if (j == i) continue;
            keys.push_back(((uint64_t)std::min(i, j) << 32) | (uint64_t)std::max(i, j));
            ++taken;
        }
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    checks = (int64_t)keys.size();
    std::vector<std::vector<int32_t>> nbr((size_t)nv);
    for (uint64_t key : keys) {
        const int32_t u = (int32_t)(key >> 32), v = (int32_t)(key & 0xffffffffu);
        Edge1 E;
        for (int d = 0; d < kDof; ++d) {
            E.a[d] = nodes[(size_t)u * kDof + d];
            E.b[d] = nodes[(size_t)v * kDof + d];
        }
        launch_motion_single(env.view(), E, req.longest_valid_segment, req.factor, tr.pinned_dev, tr.stream);
        OA_HIP(hipStreamSynchronize(tr.stream));
        if (tr.pinned[0] != 0) {
            nbr[(size_t)u].push_back(v);
            nbr[(size_t)v].push_back(u);
        }
    }
    std::vector<int64_t> off((size_t)nv + 1, 0);
    std::vector<int32_t> adj;
    for (int64_t i = 0; i < nv; ++i) {
        std::sort(nbr[(size_t)i].begin(), nbr[(size_t)i].end());
        adj.insert(adj.end(), nbr[(size_t)i].begin(), nbr[(size_t)i].end());
        off[(size_t)i + 1] = (int64_t)adj.size();
    }
    const std::vector<int32_t> path = roadmap_shortest_path(off, adj, nodes);
    for (int32_t v : path)
        out.insert(out.end(), nodes.begin() + (size_t)v * kDof, nodes.begin() + (size_t)(v + 1) * kDof);
    return !path.empty();
}

} // namespace ompl_amd

struct ompl_amd_server_s {
    ompl_amd::Server srv;
};

using namespace ompl_amd;

extern "C" {

int ompl_amd_server_create(const ompl_amd_server_config *cfg, ompl_amd_server *out)
{
    return guarded([&] {
        OA_REQUIRE(cfg && out, OMPL_AMD_ERR_ARG, "null argument");
        OA_REQUIRE(cfg->max_active >= 1 && cfg->max_active <= 65536, OMPL_AMD_ERR_ARG, "max_active in [1, 65536]");
        OA_REQUIRE(cfg->max_nodes >= 2, OMPL_AMD_ERR_ARG, "max_nodes >= 2");
        OA_REQUIRE(cfg->check_lanes_per_state == 0 || cfg->check_lanes_per_state == 7, OMPL_AMD_ERR_ARG,
                   "check_lanes_per_state must be 0 or 7 (the link layout)");
        const int co = cfg->check_obstacle_parts;
        OA_REQUIRE(co == 0 || co == 1 || co == 2 || co == 4 || co == 9, OMPL_AMD_ERR_ARG,
                   "check_obstacle_parts must be 0 (by motion length), 1, 2, 4 or 9");
        OA_REQUIRE(cfg->slice_us >= 0, OMPL_AMD_ERR_ARG, "slice_us must be >= 0");
        OA_REQUIRE(cfg->block_threads == 0 || cfg->block_threads == 512 || cfg->block_threads == 1024,
                   OMPL_AMD_ERR_ARG, "block_threads must be 0, 512 or 1024");
        ensure_init();
        auto *h = new ompl_amd_server_s();
        Server &S = h->srv;
        S.cfg = *cfg;
        S.cfg.check_lanes_per_state = 7;
        if (S.cfg.slice_us == 0) S.cfg.slice_us = 200;
        if (S.cfg.block_threads == 0) S.cfg.block_threads = 512; // 1,024 spills registers (measured slower)
        int rate_khz = 0;
        OA_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, device()));
        OA_REQUIRE(rate_khz > 0, OMPL_AMD_ERR_HIP, "device wall clock rate unavailable");
        S.clock_khz = rate_khz;
        S.budget_ticks = (int64_t)S.cfg.slice_us * rate_khz / 1000;
        if (S.budget_ticks < 1) S.budget_ticks = 1;
        OA_HIP(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
        if (cfg->hbm_budget_bytes > 0) {
            hipError_t e = hipMalloc(&S.pool, cfg->hbm_budget_bytes);
            if (e != hipSuccess) {
                (void)hipStreamDestroy(S.stream);
                delete h;
                fail(OMPL_AMD_ERR_NOMEM, "server pool of %zu bytes: %s", cfg->hbm_budget_bytes, hipGetErrorString(e));
            }
        }
        S.alloc.reset(cfg->hbm_budget_bytes);
        OA_HIP(hipHostMalloc((void **)&S.cmd_host, sizeof(PlanCmd) * cfg->max_active, hipHostMallocDefault));
        OA_HIP(hipHostMalloc((void **)&S.out_host, sizeof(PlanOut) * cfg->max_active, hipHostMallocDefault));
        OA_HIP(hipMalloc(&S.cmd_dev, sizeof(PlanCmd) * cfg->max_active));
        OA_HIP(hipMalloc(&S.out_dev, sizeof(PlanOut) * cfg->max_active));
        OA_HIP(hipStreamCreateWithFlags(&S.prm_stream, hipStreamNonBlocking));
        S.th = std::thread([&S] { S.run(); });
        S.prm_th = std::thread([&S] { S.run_prm(); });
        *out = h;
    });
}

int ompl_amd_server_destroy(ompl_amd_server h)
{
    return guarded([&] {
        if (!h) return;
        Server &S = h->srv;
        {
            std::lock_guard<std::mutex> lk(S.mu);
            S.stop = true;
        }
        S.cv.notify_all();
        S.prm_cv.notify_all();
        if (S.th.joinable()) S.th.join();
        if (S.prm_th.joinable()) S.prm_th.join();
        for (auto &kv : S.reqs) {
            if (kv.second->env_held) kv.second->req.env->server_users.fetch_sub(1);
            delete kv.second;
        }
        // teardown: release everything even if one release reports an error
        (void)hipStreamSynchronize(S.stream);
        if (S.pool) (void)hipFree(S.pool);
        (void)hipFree(S.cmd_dev);
        (void)hipFree(S.out_dev);
        (void)hipHostFree(S.cmd_host);
        (void)hipHostFree(S.out_host);
        (void)hipStreamDestroy(S.stream);
        (void)hipStreamSynchronize(S.prm_stream);
        (void)hipStreamDestroy(S.prm_stream);
        delete h;
    });
}

int ompl_amd_server_submit(ompl_amd_server h, const ompl_amd_plan_request *req, int64_t *id)
{
    return guarded([&] {
        OA_REQUIRE(h && req && id, OMPL_AMD_ERR_ARG, "null argument");
        OA_REQUIRE(req->env, OMPL_AMD_ERR_ARG, "request without environment");
        OA_REQUIRE(req->max_iters >= 1 && req->longest_valid_segment > 0.0 && req->factor >= 1, OMPL_AMD_ERR_ARG,
                   "bad request parameters");
        OA_REQUIRE(req->planner == OMPL_AMD_PLANNER_RRT_CONNECT || req->planner == OMPL_AMD_PLANNER_PRM,
                   OMPL_AMD_ERR_ARG, "unknown planner %d", req->planner);
        const Env &env = req->env->env;
        const bool prm = req->planner == OMPL_AMD_PLANNER_PRM;
        OA_REQUIRE(prm || env_stage_bytes(env.ns, env.nb) <= kPlanMaxEnvBytes, OMPL_AMD_ERR_ARG,
                   "environment too large for the server: %d spheres and %d boxes stage %zu bytes (limit %zu)",
                   env.ns, env.nb, env_stage_bytes(env.ns, env.nb), kPlanMaxEnvBytes);
        OA_REQUIRE(req->n_starts >= 0 && req->n_starts <= kMaxEnds && req->n_goals >= 0 && req->n_goals <= kMaxEnds,
                   OMPL_AMD_ERR_ARG, "a request lists at most %d start and %d goal states", kMaxEnds, kMaxEnds);
        OA_REQUIRE((req->n_starts == 0 || req->starts) && (req->n_goals == 0 || req->goals), OMPL_AMD_ERR_ARG,
                   "null start or goal list");
        OA_REQUIRE(!prm || (req->n_starts <= 1 && req->n_goals <= 1 && !req->intermediate_states), OMPL_AMD_ERR_ARG,
                   "PRM requests take one start and one goal state");
        Server &S = h->srv;
        auto *p = new Planner();
        p->req = *req;
        take_endpoints(*p);
        if (prm) {
            if (p->req.prm_samples <= 0) p->req.prm_samples = 4096;
            if (p->req.prm_k <= 0) p->req.prm_k = 10;
            OA_REQUIRE(p->req.prm_k <= 63 && p->req.prm_samples <= (1 << 24), OMPL_AMD_ERR_ARG,
                       "PRM request: prm_k <= 63 and prm_samples <= 2^24");
        }
        if (p->req.max_nodes <= 0 || p->req.max_nodes > S.cfg.max_nodes) p->req.max_nodes = S.cfg.max_nodes;
        if (p->req.max_nodes < 2) p->req.max_nodes = 2;
        if (p->req.max_nodes <= p->req.n_starts) {
            const int mn = p->req.max_nodes, ns = p->req.n_starts;
            delete p;
            fail(OMPL_AMD_ERR_ARG, "max_nodes %d cannot hold the %d start states and a step", mn, ns);
        }
        p->maxd = req->range > 0.0 ? req->range : 0.2 * 2.0 * kPi * std::sqrt(7.0);
        p->nodes = 2;
        p->t_submit = Clock::now();
        std::lock_guard<std::mutex> lk(S.mu);
        if (!S.fault.empty()) {
            delete p;
            fail(OMPL_AMD_ERR_HIP, "server stopped after a fault: %s", S.fault.c_str());
        }
        const int64_t rid = S.next_id++;
        S.reqs[rid] = p;
        if (request_footprint(p->req) > S.cfg.hbm_budget_bytes) {
            p->status = OMPL_AMD_PLAN_REJECTED;
            p->t_finish = p->t_submit;
        } else {
            p->req.env->server_users.fetch_add(1); // the environment must outlive the request
            p->env_held = true;
            if (prm) {
                S.prm_queue.push_back(rid);
                S.prm_cv.notify_all();
            } else {
                S.queue.push_back(rid);
            }
        }
        *id = rid;
        S.cv.notify_all();
    });
}

int ompl_amd_server_poll(ompl_amd_server h, int64_t id, ompl_amd_plan_result *res)
{
    return guarded([&] {
        OA_REQUIRE(h && res, OMPL_AMD_ERR_ARG, "null argument");
        Server &S = h->srv;
        std::lock_guard<std::mutex> lk(S.mu);
        auto it = S.reqs.find(id);
        OA_REQUIRE(it != S.reqs.end(), OMPL_AMD_ERR_ARG, "unknown request %lld", (long long)id);
        fill_result(*it->second, res);
    });
}

int ompl_amd_server_wait(ompl_amd_server h, int64_t id, double timeout_ms, ompl_amd_plan_result *res)
{
    return guarded([&] {
        OA_REQUIRE(h && res, OMPL_AMD_ERR_ARG, "null argument");
        Server &S = h->srv;
        std::unique_lock<std::mutex> lk(S.mu);
        auto it = S.reqs.find(id);
        OA_REQUIRE(it != S.reqs.end(), OMPL_AMD_ERR_ARG, "unknown request %lld", (long long)id);
        Planner *p = it->second;
        auto pred = [&] { return finished_status(p->status); };
        if (timeout_ms < 0) S.done_cv.wait(lk, pred);
        else S.done_cv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), pred);
        fill_result(*p, res);
    });
}

int ompl_amd_server_path(ompl_amd_server h, int64_t id, double *out, int cap, int *n)
{
    return guarded([&] {
        OA_REQUIRE(h && n, OMPL_AMD_ERR_ARG, "null argument");
        Server &S = h->srv;
        std::lock_guard<std::mutex> lk(S.mu);
        auto it = S.reqs.find(id);
        OA_REQUIRE(it != S.reqs.end(), OMPL_AMD_ERR_ARG, "unknown request %lld", (long long)id);
        const auto &path = it->second->path;
        *n = (int)(path.size() / kDof);
        if (out && cap > 0)
            std::copy(path.begin(), path.begin() + std::min<size_t>(path.size(), (size_t)cap * kDof), out);
    });
}

int ompl_amd_server_cancel(ompl_amd_server h, int64_t id)
{
    return guarded([&] {
        OA_REQUIRE(h, OMPL_AMD_ERR_ARG, "null server");
        Server &S = h->srv;
        std::lock_guard<std::mutex> lk(S.mu);
        auto it = S.reqs.find(id);
        OA_REQUIRE(it != S.reqs.end(), OMPL_AMD_ERR_ARG, "unknown request %lld", (long long)id);
        Planner *p = it->second;
        p->cancel = true;
        if (p->status == OMPL_AMD_PLAN_QUEUED) {
            auto q = std::find(S.queue.begin(), S.queue.end(), id);
            if (q != S.queue.end()) S.queue.erase(q);
            auto qp = std::find(S.prm_queue.begin(), S.prm_queue.end(), id);
            if (qp != S.prm_queue.end()) S.prm_queue.erase(qp);
            S.finish_locked(p, OMPL_AMD_PLAN_CANCELLED);
            S.done_cv.notify_all();
        }
        S.cv.notify_all();
    });
}

int ompl_amd_server_release(ompl_amd_server h, int64_t id)
{
    return guarded([&] {
        OA_REQUIRE(h, OMPL_AMD_ERR_ARG, "null server");
        Server &S = h->srv;
        std::lock_guard<std::mutex> lk(S.mu);
        auto it = S.reqs.find(id);
        OA_REQUIRE(it != S.reqs.end(), OMPL_AMD_ERR_ARG, "unknown request %lld", (long long)id);
        OA_REQUIRE(finished_status(it->second->status), OMPL_AMD_ERR_ARG, "request %lld still running",
                   (long long)id);
        delete it->second;
        S.reqs.erase(it);
    });
}

int ompl_amd_server_stats(ompl_amd_server h, int64_t out[6])
{
    return guarded([&] {
        OA_REQUIRE(h && out, OMPL_AMD_ERR_ARG, "null argument");
        Server &S = h->srv;
        std::lock_guard<std::mutex> lk(S.mu);
        out[0] = S.slices.load();
        out[1] = S.steps.load();
        out[2] = (int64_t)S.active.size() + S.prm_running;
        out[3] = (int64_t)(S.queue.size() + S.prm_queue.size());
        out[4] = (int64_t)S.used_bytes;
        out[5] = (int64_t)S.cfg.hbm_budget_bytes;
    });
}

int ompl_amd_server_phase_us(ompl_amd_server h, double out[4])
{
    return guarded([&] {
        OA_REQUIRE(h && out, OMPL_AMD_ERR_ARG, "null argument");
        Server &S = h->srv;
        std::lock_guard<std::mutex> lk(S.mu);
        for (int i = 0; i < 4; ++i) out[i] = 1e3 * (double)S.phase_ticks[i] / (double)S.clock_khz;
    });
}

size_t ompl_amd_server_footprint(int max_nodes)
{
    return max_nodes >= 2 ? footprint(max_nodes) : 0;
}

int ompl_amd_plan_plain(const ompl_amd_plan_request *req, ompl_amd_plan_result *res, double *path, int cap,
                        int *path_len)
{
    return guarded([&] {
        OA_REQUIRE(req && res && req->env, OMPL_AMD_ERR_ARG, "null argument");
        const Clock::time_point t0 = Clock::now();
        const Env &env = req->env->env;
        ThreadResources &tr = thread_resources();
        if (req->planner == OMPL_AMD_PLANNER_PRM) {
            std::vector<double> out;
            int64_t nv = 0, checks = 0;
            const bool ok = plain_prm(*req, env, tr, out, nv, checks);
            const double ms = ms_since(t0);
            res->status = ok ? OMPL_AMD_PLAN_SOLVED : OMPL_AMD_PLAN_FAILED;
            res->iterations = 1;
            res->path_len = (int)(out.size() / kDof);
            res->tree_nodes = nv;
            res->motion_checks = checks;
            res->queued_ms = 0.0;
            res->run_ms = ms;
            res->latency_ms = ms;
            if (path_len) *path_len = res->path_len;
            if (path) std::copy(out.begin(), out.begin() + std::min<size_t>(out.size(), (size_t)cap * kDof), path);
            return;
        }
        OA_REQUIRE(req->planner == OMPL_AMD_PLANNER_RRT_CONNECT, OMPL_AMD_ERR_ARG, "unknown planner %d",
                   req->planner);
        std::mt19937 gen(req->seed);
        std::uniform_real_distribution<> uni{0, 1};
        const double maxd = req->range > 0.0 ? req->range : 0.2 * 2.0 * kPi * std::sqrt(7.0);
        OA_REQUIRE(req->n_starts >= 0 && req->n_starts <= kMaxEnds && req->n_goals >= 0 && req->n_goals <= kMaxEnds &&
                       (req->n_starts == 0 || req->starts) && (req->n_goals == 0 || req->goals),
                   OMPL_AMD_ERR_ARG, "bad start or goal list");
        const double *starts = req->n_starts > 0 ? req->starts : req->start;
        const double *goals = req->n_goals > 0 ? req->goals : req->goal;
        const int n_starts = req->n_starts > 0 ? req->n_starts : 1, n_goals = req->n_goals > 0 ? req->n_goals : 1;
        const bool inter = req->intermediate_states != 0;
        HostTree ts, tg;
        for (int i = 0; i < n_starts; ++i) // nextStart: usable start states become roots
            if (plain_usable(env, starts + (size_t)i * kDof, tr)) {
                ts.s.insert(ts.s.end(), starts + (size_t)i * kDof, starts + (size_t)(i + 1) * kDof);
                ts.parent.push_back(-1);
            }
        bool start_tree = true, solved = false;
        int it = 0, sampled_goals = 0;
        int status = ts.size() == 0 ? OMPL_AMD_PLAN_INVALID_START : OMPL_AMD_PLAN_RUNNING;
        int64_t checks = 0;
        double rstate[kDof], xstate[kDof];
        std::vector<double> out;
        for (it = 0; it < req->max_iters && !solved && status == OMPL_AMD_PLAN_RUNNING; ++it) {
            if (req->deadline_ms > 0.0 && ms_since(t0) > req->deadline_ms) break;
            HostTree &tree = start_tree ? ts : tg;
            bool tgi_start = start_tree;
            start_tree = !start_tree;
            HostTree &other = start_tree ? ts : tg;
            if (tg.size() == 0 || sampled_goals < tg.size() / 2) { // goal states, as the served planner
                const bool empty = tg.size() == 0;
                while (sampled_goals < n_goals) {
                    const double *q = goals + (size_t)(sampled_goals++) * kDof;
                    if (plain_usable(env, q, tr)) {
                        tg.s.insert(tg.s.end(), q, q + kDof);
                        tg.parent.push_back(-1);
                        break;
                    }
                    if (!empty) break;
                }
                if (tg.size() == 0) {
                    status = OMPL_AMD_PLAN_INVALID_GOAL;
                    break;
                }
            }
            for (int k = 0; k < kDof; ++k) rstate[k] = (kPi - -kPi) * uni(gen) + -kPi;
            int xm = -1;
            int gs = plain_grow(env, tree, tgi_start, rstate, maxd, req->longest_valid_segment, req->factor, inter, xm,
                                xstate, checks, tr);
            if (gs == kTrapped) continue;
            const int added = xm;
            if (gs != kReached)
                for (int k = 0; k < kDof; ++k) rstate[k] = xstate[k];
            tgi_start = start_tree;
            int gsc = plain_grow(env, other, tgi_start, rstate, maxd, req->longest_valid_segment, req->factor, inter,
                                 xm, xstate, checks, tr);
            if (gsc == kTrapped) tgi_start = !tgi_start;
            while (gsc == kAdvanced)
                gsc = plain_grow(env, other, tgi_start, rstate, maxd, req->longest_valid_segment, req->factor, inter,
                                 xm, xstate, checks, tr);
            if (gsc != kReached) continue;
            int sm = tgi_start ? xm : added, gm = tgi_start ? added : xm;
            if (ts.parent[sm] >= 0) sm = ts.parent[sm];
            else gm = tg.parent[gm];
            std::vector<int> c1, c2;
            for (int m = sm; m >= 0; m = ts.parent[m]) c1.push_back(m);
            for (int m = gm; m >= 0; m = tg.parent[m]) c2.push_back(m);
            for (int i = (int)c1.size() - 1; i >= 0; --i)
                out.insert(out.end(), ts.s.begin() + (size_t)c1[i] * kDof, ts.s.begin() + (size_t)(c1[i] + 1) * kDof);
            for (int m : c2) out.insert(out.end(), tg.s.begin() + (size_t)m * kDof, tg.s.begin() + (size_t)(m + 1) * kDof);
            solved = true;
        }
        const double ms = ms_since(t0);
        res->status = solved ? OMPL_AMD_PLAN_SOLVED
                      : status != OMPL_AMD_PLAN_RUNNING
                          ? status
                          : ((req->deadline_ms > 0.0 && ms > req->deadline_ms) ? OMPL_AMD_PLAN_DEADLINE
                                                                                  : OMPL_AMD_PLAN_FAILED);
        res->iterations = it;
        res->path_len = (int)(out.size() / kDof);
        res->tree_nodes = ts.size() + tg.size();
        res->motion_checks = checks;
        res->queued_ms = 0.0;
        res->run_ms = ms;
        res->latency_ms = ms;
        if (path_len) *path_len = res->path_len;
        if (path) std::copy(out.begin(), out.begin() + std::min<size_t>(out.size(), (size_t)cap * kDof), path);
    });
}

} // extern "C"
