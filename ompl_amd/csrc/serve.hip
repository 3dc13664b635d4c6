// Planning server: continuous batching of RRT-Connect requests on one GPU.
//
// Each request is a single-query RRT-Connect plan (geometric/planners/rrt/src/
// RRTConnect.cpp:116-196 growTree, :245-330 solve) on the 7-DOF arm in one of
// the registered environments.  Its two trees live in HBM (append-only FP64
// states + parent indices).  A server thread runs ticks: every tick each
// active request contributes exactly one growTree step -- nearest neighbour in
// its tree, steering by the planner range, motion check in the direction of
// growth, append -- executed for all requests by three launches: nearest +
// steer (one 256-thread block per request), the motion checks of every request
// chunk-parallel over the whole GPU (so one long valid motion in a dense
// environment does not hold the tick for its single wave), and commit.  The
// host then advances each request's RRT-Connect state machine.
// New requests are admitted between ticks while their HBM footprint fits the
// budget (continuous batching); a request's result does not depend on which
// other requests share its ticks.
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
#include <mutex>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "runtime.hpp"
#include "walk_device.hpp"

namespace ompl_amd {

enum GrowStatus { kTrapped = 0, kAdvanced = 1, kReached = 2 };

// One growTree step (device).
struct GrowCmd {
    const float4 *obs;
    const float4 *obs_links; // the same obstacles in the link layout's staging
    int32_t ns, nb;
    double *tree_s;        // [cap][7]
    int32_t *tree_parent;  // [cap]
    int32_t tree_n;        // nodes before this step
    int32_t start_tree;    // tgi.start: 1 = motion nearest -> new, 0 = new -> nearest (+ isValid(new))
    double lvs;
    double maxd;
    uint32_t factor;
    int32_t pad;
    double target[kDof];
};

struct GrowRes {
    int32_t status;
    int32_t xmotion; // index of the appended node
    double xstate[kDof];
};

// Per-request scratch of one tick (device).
struct GrowEdge {
    double a[kDof], b[kDof]; // motion checked in the direction of growth
    double dst[kDof];        // state to append
    int32_t nd;              // segments of the motion
    int32_t npos;            // positions to check (0 = none: trapped before the check)
    int32_t extra;           // 1: the last position is isValid(dst) (goal tree)
    int32_t nm;              // nearest node
    int32_t reach;
    int32_t first_bad;       // lowest invalid position found (INT_MAX: none)
};

struct TickArgs {
    const GrowCmd *cmds;
    GrowEdge *edges;
    GrowRes *res;
    int32_t *max_pos; // largest npos of the tick (reset by the commit kernel)
    int ncmd;
};

// Tick kernel 1 (one block of kPrepThreads per request): nearest neighbour in
// the request's tree -- first strict minimum of distance(elem, target),
// thread-strided, then reduced over each wave and across the block's waves
// by (distance, index) -- and steering by the planner range
// (RRTConnect.cpp:119-138).
constexpr int kPrepThreads = 256;

__global__ void __launch_bounds__(kPrepThreads) grow_prepare_kernel(TickArgs T)
{
    const int r = blockIdx.x;
    if (r >= T.ncmd) return;
    const GrowCmd &C = T.cmds[r];
    GrowEdge &E = T.edges[r];
    const int lane = lane_id();
    double best = INFINITY;
    int best_i = INT_MAX;
    for (int i = threadIdx.x; i < C.tree_n; i += kPrepThreads) {
        const double *s = C.tree_s + (int64_t)i * kDof;
        double dist = 0.0;
#pragma unroll
        for (int k = 0; k < kDof; ++k) {
            const double diff = s[k] - C.target[k];
            dist += diff * diff;
        }
        const double d = __builtin_sqrt(dist);
        if (d < best) { // increasing i per lane: strict < keeps the lowest index
            best = d;
            best_i = i;
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double od = __shfl_xor(best, off);
        const int oi = __shfl_xor(best_i, off);
        if (od < best || (od == best && oi < best_i)) {
            best = od;
            best_i = oi;
        }
    }
    __shared__ double wd[kPrepThreads / 64];
    __shared__ int wi[kPrepThreads / 64];
    if (lane == 0) {
        wd[threadIdx.x >> 6] = best;
        wi[threadIdx.x >> 6] = best_i;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w = 1; w < kPrepThreads / 64; ++w)
        if (wd[w] < best || (wd[w] == best && wi[w] < best_i)) {
            best = wd[w];
            best_i = wi[w];
        }
    const int nm = best_i;
    double ns_[kDof], x[kDof];
    for (int k = 0; k < kDof; ++k) ns_[k] = C.tree_s[(int64_t)nm * kDof + k];
    E.nm = nm;
    E.first_bad = INT_MAX;
    const double d = best;
    bool reach = true;
    for (int k = 0; k < kDof; ++k) x[k] = C.target[k];
    if (d > C.maxd) {
        const double t = C.maxd / d;
        bool equal = true;
        for (int k = 0; k < kDof; ++k) {
            x[k] = ns_[k] + (C.target[k] - ns_[k]) * t;
            if (fabs(ns_[k] - x[k]) > 2.220446049250313080847e-16 * 2.0) equal = false;
        }
        if (equal) { // no progress: TRAPPED without a check
            E.npos = 0;
            return;
        }
        reach = false;
    }
    E.reach = reach;
    for (int k = 0; k < kDof; ++k) {
        E.dst[k] = x[k];
        E.a[k] = C.start_tree ? ns_[k] : x[k]; // start tree: checkMotion(nearest, dst)
        E.b[k] = C.start_tree ? x[k] : ns_[k]; // goal tree: isValid(dst) && checkMotion(dst, nearest)
    }
    const int nd = valid_segment_count<kDof>(E.a, E.b, C.lvs, C.factor);
    if (nd > kMaxSegments) { // reported invalid
        E.npos = 1;
        E.first_bad = 0;
        return;
    }
    E.nd = nd;
    E.extra = C.start_tree ? 0 : 1;
    E.npos = 1 + (nd >= 2 ? nd - 1 : 0) + E.extra;
    atomicMax(T.max_pos, E.npos);
}

// Tick kernel 2: every request's positions, chunk-parallel over a persistent
// grid: items (chunk, request) chunk-major, one wave per item (wave w takes
// items w, w + W, ...); a wave tests whether its next 64 items lie inside
// their motions at once, then checks the live ones in item order, skipping a
// request whose motion is already known invalid.  Obstacles are read from
// each request's environment in HBM (wave-uniform loads), so requests in
// different environments share the launch.
template <int LPS, int OPS>
__global__ void __launch_bounds__(256) grow_check_kernel(TickArgs T)
{
    using Shape = GroupShape<LPS, OPS>;
    constexpr int G = Shape::STATES;
    const int lane = lane_id();
    const int64_t waves_total = (gridDim.x * blockDim.x) >> 6;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int n = T.ncmd;
    const int64_t nblk = (*T.max_pos + G - 1) / G;
    const int64_t items = (int64_t)n * nblk;
    const Shape sh(lane);
    for (int64_t k0 = wave; k0 < items; k0 += 64 * waves_total) {
        const int64_t k = k0 + lane * waves_total;
        int r = 0, base = 0;
        bool live = false;
        if (k < items) {
            const int64_t blk = k / n;
            r = (int)(k - blk * n);
            base = (int)(blk * G);
            live = base < T.edges[r].npos;
        }
        uint64_t mask = wave_ballot(live);
        while (mask) {
            const int j = lowest_bit(mask);
            mask &= mask - 1;
            const int rj = __shfl(r, j), bj = __shfl(base, j);
            GrowEdge &E = T.edges[rj];
            if (__hip_atomic_load(&E.first_bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != INT_MAX) continue;
            const GrowCmd &C = T.cmds[rj];
            const int p = bj + sh.group;
            const int motion_pos = E.npos - E.extra;
            const bool plive = p < E.npos;
            float q[kDof];
            if (plive && p >= motion_pos) {
#pragma unroll
                for (int d = 0; d < kDof; ++d) q[d] = (float)E.dst[d];
            } else {
                double a[kDof], b[kDof];
#pragma unroll
                for (int d = 0; d < kDof; ++d) {
                    a[d] = E.a[d];
                    b[d] = E.b[d];
                }
                walk_state(a, b, E.nd, plive ? p : 0, q);
            }
            LaneSpheres<LPS> S;
            arm_fk<LPS, Shape::GS>(q, sh.slot, S);
            const float4 *obs = LPS == 7 ? C.obs_links : C.obs;
            const bool hit = collide_env<LPS, OPS, Exit::kAnyGroup, false>(S, obs, C.ns, C.nb, plive, sh.part);
            const uint64_t H = fold_groups<Shape::GS>(wave_ballot(hit && plive));
            if (H && lane == 0) atomicMin(&E.first_bad, bj + lowest_bit(H) / Shape::GS);
        }
    }
}

// Tick kernel 3 (one thread per request): append the new node of a valid
// motion, write the outcome, clear the tick counter.
__global__ void grow_commit_kernel(TickArgs T)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0) *T.max_pos = 0;
    if (r >= T.ncmd) return;
    const GrowCmd &C = T.cmds[r];
    const GrowEdge &E = T.edges[r];
    GrowRes &R = T.res[r];
    if (E.npos == 0 || E.first_bad != INT_MAX) {
        R.status = kTrapped;
        return;
    }
    const int idx = C.tree_n;
    for (int k = 0; k < kDof; ++k) {
        C.tree_s[(int64_t)idx * kDof + k] = E.dst[k];
        R.xstate[k] = E.dst[k];
    }
    C.tree_parent[idx] = E.nm;
    R.status = E.reach ? kReached : kAdvanced;
    R.xmotion = idx;
}

// ---------------------------------------------------------------------------
// RRT-Connect request state machine (host)
// ---------------------------------------------------------------------------
using Clock = std::chrono::steady_clock;

static double ms_since(Clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

struct Planner {
    // configuration
    ompl_amd_plan_request req;
    const Env *env = nullptr;
    double maxd = 0.0;
    // RNG of the state sampler (RNG::uniformReal over std::mt19937)
    std::mt19937 gen;
    std::uniform_real_distribution<> uni{0, 1};
    // trees: host mirrors of the node counts; parents/states live on device
    int n[2] = {0, 0};
    int cap = 0;
    double *tree_s[2] = {nullptr, nullptr};
    int32_t *tree_parent[2] = {nullptr, nullptr};
    // solve-loop state
    int iteration = 0;
    bool start_tree = true; // RRTConnect::startTree_
    int tgi_start = 1;
    int grow_tree = 0;      // tree index of the pending growTree (0 = start, 1 = goal)
    int phase = 0;          // 0 sample+grow, 1 first connect, 2 connect loop
    int added = -1, xmotion = -1;
    double rstate[kDof];
    // result
    int status = OMPL_AMD_PLAN_QUEUED;
    std::vector<double> path;
    int64_t nodes = 0;
    Clock::time_point t_submit, t_start, t_finish;
    bool cancel = false;

    double uniformReal(double lo, double hi) { return (hi - lo) * uni(gen) + lo; }
};

constexpr double kPi = 3.141592653589793238462643383279502884;

// Prepares the next growTree command of a planner (phase 0 samples first).
static void next_command(Planner &p, GrowCmd &c)
{
    if (p.phase == 0) {
        const int tree = p.start_tree ? 0 : 1;
        p.tgi_start = p.start_tree ? 1 : 0;
        p.start_tree = !p.start_tree;
        for (int k = 0; k < kDof; ++k) p.rstate[k] = p.uniformReal(-kPi, kPi);
        p.grow_tree = tree;
    }
    // phases 1/2 grow the other tree toward rstate with the current tgi_start
    c.obs = p.env->obs;
    c.obs_links = p.env->obs_links;
    c.ns = p.env->ns;
    c.nb = p.env->nb;
    c.tree_s = p.tree_s[p.grow_tree];
    c.tree_parent = p.tree_parent[p.grow_tree];
    c.tree_n = p.n[p.grow_tree];
    c.start_tree = p.tgi_start;
    c.lvs = p.req.longest_valid_segment;
    c.maxd = p.maxd;
    c.factor = p.req.factor;
    c.pad = 0;
    for (int k = 0; k < kDof; ++k) c.target[k] = p.rstate[k];
}

// Reconstructs the solution path (RRTConnect.cpp:303-340) from device trees.
static void build_path(Planner &p, int sm, int gm, hipStream_t s)
{
    std::vector<double> ss((size_t)p.n[0] * kDof), gs((size_t)p.n[1] * kDof);
    std::vector<int32_t> sp((size_t)p.n[0]), gp((size_t)p.n[1]);
    OA_HIP(hipMemcpyAsync(ss.data(), p.tree_s[0], ss.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    OA_HIP(hipMemcpyAsync(gs.data(), p.tree_s[1], gs.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    OA_HIP(hipMemcpyAsync(sp.data(), p.tree_parent[0], sp.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    OA_HIP(hipMemcpyAsync(gp.data(), p.tree_parent[1], gp.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    OA_HIP(hipStreamSynchronize(s));
    if (sp[sm] >= 0) sm = sp[sm];
    else gm = gp[gm];
    std::vector<int> chain1, chain2;
    for (int m = sm; m >= 0; m = sp[m]) chain1.push_back(m);
    for (int m = gm; m >= 0; m = gp[m]) chain2.push_back(m);
    p.path.clear();
    for (int i = (int)chain1.size() - 1; i >= 0; --i)
        p.path.insert(p.path.end(), ss.begin() + (size_t)chain1[i] * kDof, ss.begin() + (size_t)(chain1[i] + 1) * kDof);
    for (int m : chain2) p.path.insert(p.path.end(), gs.begin() + (size_t)m * kDof, gs.begin() + (size_t)(m + 1) * kDof);
}

// Applies a growTree outcome; returns true when the request finished.
static bool apply_result(Planner &p, const GrowRes &r, hipStream_t s)
{
    if (r.status != kTrapped) {
        p.n[p.grow_tree] = r.xmotion + 1;
        p.xmotion = r.xmotion;
    }
    auto end_iteration = [&]() {
        p.phase = 0;
        ++p.iteration;
        if (p.iteration >= p.req.max_iters) {
            p.status = OMPL_AMD_PLAN_FAILED;
            return true;
        }
        return false;
    };
    if (p.phase == 0) {
        if (r.status == kTrapped) return end_iteration();
        p.added = r.xmotion;
        if (r.status != kReached)
            for (int k = 0; k < kDof; ++k) p.rstate[k] = r.xstate[k];
        p.tgi_start = p.start_tree ? 1 : 0;
        p.grow_tree = p.start_tree ? 0 : 1; // otherTree
        p.phase = 1;
        return false;
    }
    // connect steps
    const int gsc = r.status;
    if (p.phase == 1 && gsc == kTrapped) p.tgi_start = !p.tgi_start;
    if (gsc == kAdvanced) {
        p.phase = 2;
        return false;
    }
    if (gsc == kReached) {
        const int sm = p.tgi_start ? p.xmotion : p.added;
        const int gm = p.tgi_start ? p.added : p.xmotion;
        build_path(p, sm, gm, s);
        p.status = OMPL_AMD_PLAN_SOLVED;
        ++p.iteration;
        return true;
    }
    return end_iteration();
}

static void init_planner(Planner &p, const ompl_amd_plan_request &req, double *s0, int32_t *p0, double *s1,
                         int32_t *p1, int cap, hipStream_t s)
{
    p.req = req;
    p.env = &req.env->env;
    p.maxd = req.range > 0.0 ? req.range : 0.2 * 2.0 * kPi * std::sqrt(7.0);
    p.gen.seed(req.seed);
    p.uni.reset();
    p.cap = cap;
    p.tree_s[0] = s0;
    p.tree_parent[0] = p0;
    p.tree_s[1] = s1;
    p.tree_parent[1] = p1;
    p.n[0] = p.n[1] = 1;
    const int32_t root = -1;
    OA_HIP(hipMemcpyAsync(s0, req.start, kDof * sizeof(double), hipMemcpyHostToDevice, s));
    OA_HIP(hipMemcpyAsync(s1, req.goal, kDof * sizeof(double), hipMemcpyHostToDevice, s));
    OA_HIP(hipMemcpyAsync(p0, &root, sizeof(int32_t), hipMemcpyHostToDevice, s));
    OA_HIP(hipMemcpyAsync(p1, &root, sizeof(int32_t), hipMemcpyHostToDevice, s));
    OA_HIP(hipStreamSynchronize(s));
    p.iteration = 0;
    p.start_tree = true;
    p.phase = 0;
    p.status = OMPL_AMD_PLAN_RUNNING;
}

static size_t footprint(const ompl_amd_plan_request &r)
{
    return (size_t)2 * (size_t)r.max_nodes * (kDof * sizeof(double) + sizeof(int32_t));
}

// ---------------------------------------------------------------------------
// server
// ---------------------------------------------------------------------------
struct Slot {
    double *s[2];
    int32_t *p[2];
};

struct Server {
    ompl_amd_server_config cfg;
    hipStream_t stream = nullptr;
    std::vector<Slot> slots;
    std::vector<int> free_slots;
    char *pool = nullptr;
    GrowCmd *cmd_host = nullptr, *cmd_dev = nullptr;
    GrowRes *res_host = nullptr, *res_dev = nullptr;
    GrowEdge *edges_dev = nullptr;
    int32_t *max_pos_dev = nullptr;
    unsigned check_blocks = 0;

    std::mutex mu;
    std::condition_variable cv, done_cv;
    std::deque<int64_t> queue;
    std::unordered_map<int64_t, Planner *> reqs;
    std::vector<std::pair<int64_t, int>> active; // (id, slot)
    int64_t next_id = 1;
    size_t used_bytes = 0;
    bool stop = false;
    std::thread th;
    std::atomic<int64_t> ticks{0}, steps{0};
    std::string fault;

    void run();
    void admit_locked();
};

void Server::admit_locked()
{
    while (!queue.empty() && !free_slots.empty()) {
        const int64_t id = queue.front();
        Planner *p = reqs[id];
        if (p->cancel) {
            queue.pop_front();
            p->status = OMPL_AMD_PLAN_CANCELLED;
            p->t_finish = Clock::now();
            continue;
        }
        const size_t fp = footprint(p->req);
        if (used_bytes + fp > cfg.hbm_budget_bytes) break; // FIFO: wait for memory
        queue.pop_front();
        const int slot = free_slots.back();
        free_slots.pop_back();
        used_bytes += fp;
        const Slot &S = slots[slot];
        init_planner(*p, p->req, S.s[0], S.p[0], S.s[1], S.p[1], p->req.max_nodes, stream);
        p->t_start = Clock::now();
        active.emplace_back(id, slot);
    }
}

void Server::run()
{
    try {
        OA_HIP(hipSetDevice(device()));
        for (;;) {
            std::vector<std::pair<int64_t, int>> batch;
            std::vector<Planner *> planners;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !queue.empty() || !active.empty(); });
                if (stop) return;
                // retire cancelled / expired / full requests, then admit
                for (auto q = queue.begin(); q != queue.end();) {
                    Planner *p = reqs[*q];
                    if (p->req.deadline_ms > 0.0 && ms_since(p->t_submit) > p->req.deadline_ms) {
                        p->status = OMPL_AMD_PLAN_DEADLINE;
                        p->t_finish = Clock::now();
                        q = queue.erase(q);
                        done_cv.notify_all();
                    } else {
                        ++q;
                    }
                }
                for (size_t i = 0; i < active.size();) {
                    Planner *p = reqs[active[i].first];
                    const bool expired = p->req.deadline_ms > 0.0 && ms_since(p->t_submit) > p->req.deadline_ms;
                    // a step appends at most one node to one tree
                    const bool full = p->n[0] >= p->cap || p->n[1] >= p->cap;
                    if (p->cancel || expired || full) {
                        p->status = p->cancel ? OMPL_AMD_PLAN_CANCELLED
                                              : (expired ? OMPL_AMD_PLAN_DEADLINE : OMPL_AMD_PLAN_TREE_FULL);
                        p->t_finish = Clock::now();
                        p->nodes = p->n[0] + p->n[1];
                        used_bytes -= footprint(p->req);
                        free_slots.push_back(active[i].second);
                        active[i] = active.back();
                        active.pop_back();
                        done_cv.notify_all();
                    } else {
                        ++i;
                    }
                }
                admit_locked();
                batch = active;
                for (auto &a : batch) planners.push_back(reqs[a.first]);
            }
            if (batch.empty()) continue;
            const int n = (int)batch.size();
            for (int i = 0; i < n; ++i) next_command(*planners[i], cmd_host[i]);
            OA_HIP(hipMemcpyAsync(cmd_dev, cmd_host, n * sizeof(GrowCmd), hipMemcpyHostToDevice, stream));
            {
                ProfScope ps("serve_tick", stream);
                TickArgs T{cmd_dev, edges_dev, res_dev, max_pos_dev, n};
                {
                    ProfScope p1("serve_prepare", stream);
                    hipLaunchKernelGGL(grow_prepare_kernel, dim3(n), dim3(kPrepThreads), 0, stream, T);
                }
                {
                    ProfScope p2("serve_check", stream);
                    const int L = cfg.check_lanes_per_state, O = cfg.check_obstacle_parts;
#define OA_CHECK(LL, OO)                                                                                            \
    else if (L == LL && O == OO) hipLaunchKernelGGL((grow_check_kernel<LL, OO>), dim3(check_blocks), dim3(256), 0, \
                                                    stream, T);
                    if (false) {
                    }
                    OA_CHECK(8, 2) OA_CHECK(8, 4) OA_CHECK(8, 8) OA_CHECK(7, 1) OA_CHECK(7, 2) OA_CHECK(7, 4)
                    OA_CHECK(7, 9)
#undef OA_CHECK
                }
                {
                    ProfScope p3("serve_commit", stream);
                    hipLaunchKernelGGL(grow_commit_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, T);
                }
                OA_HIP(hipGetLastError());
            }
            OA_HIP(hipMemcpyAsync(res_host, res_dev, n * sizeof(GrowRes), hipMemcpyDeviceToHost, stream));
            OA_HIP(hipStreamSynchronize(stream));
            ++ticks;
            steps += n;
            std::vector<int> finished;
            for (int i = 0; i < n; ++i)
                if (apply_result(*planners[i], res_host[i], stream)) finished.push_back(i);
            if (!finished.empty()) {
                std::lock_guard<std::mutex> lk(mu);
                for (int i : finished) {
                    Planner *p = planners[i];
                    p->t_finish = Clock::now();
                    p->nodes = p->n[0] + p->n[1];
                    used_bytes -= footprint(p->req);
                    for (size_t j = 0; j < active.size(); ++j)
                        if (active[j].first == batch[i].first) {
                            free_slots.push_back(active[j].second);
                            active[j] = active.back();
                            active.pop_back();
                            break;
                        }
                }
                done_cv.notify_all();
            }
        }
    } catch (const std::exception &e) {
        std::lock_guard<std::mutex> lk(mu);
        fault = e.what();
        for (auto &a : active) {
            reqs[a.first]->status = OMPL_AMD_PLAN_FAILED;
            reqs[a.first]->t_finish = Clock::now();
        }
        active.clear();
        done_cv.notify_all();
    }
}

static bool finished_status(int s)
{
    return s >= OMPL_AMD_PLAN_SOLVED;
}

static void fill_result(const Planner &p, ompl_amd_plan_result *r)
{
    r->status = p.status;
    r->iterations = p.iteration;
    r->path_len = (int)(p.path.size() / kDof);
    r->tree_nodes = p.nodes ? p.nodes : p.n[0] + p.n[1];
    auto ms = [](Clock::time_point a, Clock::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    const bool started = p.t_start.time_since_epoch().count() != 0;
    const bool done = finished_status(p.status);
    r->queued_ms = started ? ms(p.t_submit, p.t_start) : ms(p.t_submit, Clock::now());
    r->run_ms = started ? ms(p.t_start, done ? p.t_finish : Clock::now()) : 0.0;
    r->latency_ms = ms(p.t_submit, done ? p.t_finish : Clock::now());
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
                      unsigned factor, int &xmotion, double *xstate, ThreadResources &tr)
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
    const double *ns = &t.s[(size_t)nm * kDof];
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
    t.s.insert(t.s.end(), dst, dst + kDof);
    t.parent.push_back(nm);
    xmotion = t.size() - 1;
    for (int k = 0; k < kDof; ++k) xstate[k] = dst[k];
    return reach ? kReached : kAdvanced;
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
        const int cl = cfg->check_lanes_per_state, co = cfg->check_obstacle_parts;
        OA_REQUIRE(cl == 0 || cl == 7 || cl == 8, OMPL_AMD_ERR_ARG, "check_lanes_per_state must be 0, 7 or 8");
        OA_REQUIRE(cl == 7 ? (co == 0 || co == 1 || co == 2 || co == 4 || co == 9)
                           : (co == 0 || co == 2 || co == 4 || co == 8),
                   OMPL_AMD_ERR_ARG,
                   "check_obstacle_parts must be 0, 2, 4 or 8 (sphere slots) or 0, 1, 2, 4 or 9 (link layout)");
        ensure_init();
        auto *h = new ompl_amd_server_s();
        Server &S = h->srv;
        S.cfg = *cfg;
        // defaults: the link layout with two states per wave, the fastest tick check measured
        if (S.cfg.check_lanes_per_state == 0) S.cfg.check_lanes_per_state = 7;
        if (S.cfg.check_obstacle_parts == 0) S.cfg.check_obstacle_parts = S.cfg.check_lanes_per_state == 7 ? 4 : 8;
        OA_HIP(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
        const size_t per_tree = (size_t)cfg->max_nodes * (kDof * sizeof(double) + sizeof(int32_t));
        const size_t total = per_tree * 2 * (size_t)cfg->max_active;
        hipError_t e = hipMalloc(&S.pool, total);
        if (e != hipSuccess) {
            delete h;
            fail(OMPL_AMD_ERR_NOMEM, "server tree pool of %zu bytes: %s", total, hipGetErrorString(e));
        }
        for (int i = 0; i < cfg->max_active; ++i) {
            Slot sl;
            for (int t = 0; t < 2; ++t) {
                char *base = S.pool + per_tree * (2 * (size_t)i + t);
                sl.s[t] = (double *)base;
                sl.p[t] = (int32_t *)(base + (size_t)cfg->max_nodes * kDof * sizeof(double));
            }
            S.slots.push_back(sl);
            S.free_slots.push_back(cfg->max_active - 1 - i);
        }
        OA_HIP(hipHostMalloc((void **)&S.cmd_host, sizeof(GrowCmd) * cfg->max_active, hipHostMallocDefault));
        OA_HIP(hipHostMalloc((void **)&S.res_host, sizeof(GrowRes) * cfg->max_active, hipHostMallocDefault));
        OA_HIP(hipMalloc(&S.cmd_dev, sizeof(GrowCmd) * cfg->max_active));
        OA_HIP(hipMalloc(&S.res_dev, sizeof(GrowRes) * cfg->max_active));
        OA_HIP(hipMalloc(&S.edges_dev, sizeof(GrowEdge) * cfg->max_active));
        OA_HIP(hipMalloc(&S.max_pos_dev, sizeof(int32_t)));
        OA_HIP(hipMemset(S.max_pos_dev, 0, sizeof(int32_t)));
        S.check_blocks = (unsigned)num_cus() * 2; // 8 waves per CU
        S.th = std::thread([&S] { S.run(); });
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
        if (S.th.joinable()) S.th.join();
        for (auto &kv : S.reqs) delete kv.second;
        // teardown: release everything even if one release reports an error
        (void)hipFree(S.pool);
        (void)hipFree(S.cmd_dev);
        (void)hipFree(S.res_dev);
        (void)hipFree(S.edges_dev);
        (void)hipFree(S.max_pos_dev);
        (void)hipHostFree(S.cmd_host);
        (void)hipHostFree(S.res_host);
        (void)hipStreamDestroy(S.stream);
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
        Server &S = h->srv;
        auto *p = new Planner();
        p->req = *req;
        if (p->req.max_nodes <= 0 || p->req.max_nodes > S.cfg.max_nodes) p->req.max_nodes = S.cfg.max_nodes;
        p->t_submit = Clock::now();
        std::lock_guard<std::mutex> lk(S.mu);
        OA_REQUIRE(S.fault.empty(), OMPL_AMD_ERR_HIP, "server stopped after a fault: %s", S.fault.c_str());
        const int64_t rid = S.next_id++;
        S.reqs[rid] = p;
        if (footprint(p->req) > S.cfg.hbm_budget_bytes) {
            p->status = OMPL_AMD_PLAN_REJECTED;
            p->t_finish = p->t_submit;
        } else {
            S.queue.push_back(rid);
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
        if (out) std::copy(path.begin(), path.begin() + std::min<size_t>(path.size(), (size_t)cap * kDof), out);
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
            p->status = OMPL_AMD_PLAN_CANCELLED;
            p->t_finish = Clock::now();
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
        out[0] = S.ticks.load();
        out[1] = S.steps.load();
        out[2] = (int64_t)S.active.size();
        out[3] = (int64_t)S.queue.size();
        out[4] = (int64_t)S.used_bytes;
        out[5] = (int64_t)S.cfg.hbm_budget_bytes;
    });
}

int ompl_amd_plan_plain(const ompl_amd_plan_request *req, ompl_amd_plan_result *res, double *path, int cap,
                        int *path_len)
{
    return guarded([&] {
        OA_REQUIRE(req && res && req->env, OMPL_AMD_ERR_ARG, "null argument");
        const Clock::time_point t0 = Clock::now();
        const Env &env = req->env->env;
        ThreadResources &tr = thread_resources();
        std::mt19937 gen(req->seed);
        std::uniform_real_distribution<> uni{0, 1};
        const double maxd = req->range > 0.0 ? req->range : 0.2 * 2.0 * kPi * std::sqrt(7.0);
        HostTree ts, tg;
        ts.s.assign(req->start, req->start + kDof);
        ts.parent.push_back(-1);
        tg.s.assign(req->goal, req->goal + kDof);
        tg.parent.push_back(-1);
        bool start_tree = true, solved = false;
        int it = 0;
        double rstate[kDof], xstate[kDof];
        std::vector<double> out;
        for (it = 0; it < req->max_iters && !solved; ++it) {
            if (req->deadline_ms > 0.0 && ms_since(t0) > req->deadline_ms) break;
            HostTree &tree = start_tree ? ts : tg;
            bool tgi_start = start_tree;
            start_tree = !start_tree;
            HostTree &other = start_tree ? ts : tg;
            for (int k = 0; k < kDof; ++k) rstate[k] = (kPi - -kPi) * uni(gen) + -kPi;
            int xm = -1;
            int gs = plain_grow(env, tree, tgi_start, rstate, maxd, req->longest_valid_segment, req->factor, xm,
                                xstate, tr);
            if (gs == kTrapped) continue;
            const int added = xm;
            if (gs != kReached)
                for (int k = 0; k < kDof; ++k) rstate[k] = xstate[k];
            tgi_start = start_tree;
            int gsc = plain_grow(env, other, tgi_start, rstate, maxd, req->longest_valid_segment, req->factor, xm,
                                 xstate, tr);
            if (gsc == kTrapped) tgi_start = !tgi_start;
            while (gsc == kAdvanced)
                gsc = plain_grow(env, other, tgi_start, rstate, maxd, req->longest_valid_segment, req->factor, xm,
                                 xstate, tr);
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
                             : ((req->deadline_ms > 0.0 && ms > req->deadline_ms) ? OMPL_AMD_PLAN_DEADLINE
                                                                                  : OMPL_AMD_PLAN_FAILED);
        res->iterations = it;
        res->path_len = (int)(out.size() / kDof);
        res->tree_nodes = ts.size() + tg.size();
        res->queued_ms = 0.0;
        res->run_ms = ms;
        res->latency_ms = ms;
        if (path_len) *path_len = res->path_len;
        if (path) std::copy(out.begin(), out.begin() + std::min<size_t>(out.size(), (size_t)cap * kDof), path);
    });
}

} // extern "C"
