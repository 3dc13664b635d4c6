// C-ABI of environments, state validity and motion validation.
#include <algorithm>
#include <cmath>
#include <functional>
#include <memory>
#include <vector>

#include "runtime.hpp"

using namespace ompl_amd;

namespace {

Env &env_of(ompl_amd_env e)
{
    OA_REQUIRE(e, OMPL_AMD_ERR_ARG, "null environment");
    return e->env;
}

// Staging order of an obstacle list: increasing distance from the arm's base
// (the world origin) to the obstacle's nearest point, ties in the caller's
// order.  Every check result is "any obstacle hit", so the order leaves results
// unchanged; links near the base sweep the least volume, so the obstacles
// nearest to it are hit by the most configurations, and the checkers, which
// leave a state's obstacle loop at the first pass with a hit, resolve most
// invalid states in their first passes.
std::vector<int> nearest_first(int n, const std::function<double(int)> &gap)
{
    std::vector<double> key(n);
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) {
        key[i] = gap(i);
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key[x] < key[y]; });
    return order;
}

} // namespace

// Batched DiscreteMotionValidator::checkMotion(s1, s2) over device arrays.
void ompl_amd::check_motions_tuned(const Tuning &t, const Env &env, const double *a, const double *b, int64_t n,
                                   double lvs, unsigned factor, uint8_t *out, int32_t *checked, hipStream_t s,
                                   Workspace *own)
{
    OA_REQUIRE(n >= 0, OMPL_AMD_ERR_ARG, "negative edge count");
    if (n == 0) return;
    OA_REQUIRE(a && b && out, OMPL_AMD_ERR_ARG, "null edge or output pointer");
    OA_REQUIRE(lvs > 0.0 && factor >= 1, OMPL_AMD_ERR_ARG, "longest valid segment must be > 0 and factor >= 1");
    OA_REQUIRE(n < (int64_t)1 << 31, OMPL_AMD_ERR_ARG, "at most 2^31-1 edges per call");
    MotionArgs A{};
    A.env = env.view();
    A.a = a;
    A.b = b;
    A.n = n;
    A.lvs = lvs;
    A.factor = factor;
    A.out = out;
    A.checked = checked;
    const bool exact = checked != nullptr;
    const int LE = t.validity_lanes_per_state, OE = t.validity_obstacle_parts;
    const int LW = t.motion_lanes_per_state, OW = t.motion_obstacle_parts;
    const int64_t persistent_waves = (int64_t)num_cus() * t.motion_waves_per_cu;
    if (t.motion_strategy == 3 || t.motion_strategy == 4) {
        // Strategy 4 walks chunks with each position's FK computed once per
        // wave and shared across the obstacle partitions; the s2 pass shares
        // its FK when validity_shared_fk is set.  The survivors' booleans are
        // written by the s2 pass (1) and the walk (0 at an invalid position);
        // a results kernel runs after the walk when checked counts are asked
        // for, and clears the batch's counter set; otherwise phase 1 of the
        // walk clears the other set, which the next batch on this workspace
        // uses (a captured batch is recorded once per set, and its replays
        // alternate between the two).
        const bool shared = t.motion_strategy == 4;
        const bool eshared = t.validity_shared_fk != 0;
        std::unique_lock<std::mutex> lk; // held until this batch's launches are enqueued
        Workspace &w = own ? *own : workspace(s, (size_t)n, lk);
        const bool results = exact;
        int32_t *cur = w.counters + 8 + 2 * w.parity, *other = w.counters + 8 + 2 * (1 - w.parity);
        A.list = w.list;
        A.list_count = cur;
        A.max_m = cur + 1;
        A.reset_next = results ? nullptr : other;
        A.first_bad = w.first_bad;
        A.surv_nd = w.surv_nd;
        A.surv_ab = w.surv_ab;
        A.start = 1;
        launch_motion_endpoint(A, LE, OE, t.validity_block, s, eshared);
        launch_motion_chunks(A, exact, LW, OW, t.motion_block, persistent_waves, s, shared, results);
        if (!results) w.parity ^= 1;
    } else if (t.motion_strategy == 2) {
        A.start = 1;
        launch_motion_fused(A, exact, LE, OE, LW, OW, t.motion_block, s);
    } else if (t.motion_strategy == 1) {
        std::unique_lock<std::mutex> lk; // held until this batch's launches are enqueued
        Workspace &w = own ? *own : workspace(s, (size_t)n, lk);
        A.list = w.list;
        A.list_count = w.counters;
        A.queue = w.counters + 1;
        A.start = 1;
        launch_zero(w.counters, 4, s);
        launch_motion_endpoint(A, LE, OE, t.validity_block, s, t.validity_shared_fk != 0);
        launch_motion_walk(A, exact, LW, OW, t.motion_block, persistent_waves, s);
    } else {
        A.start = 0;
        launch_motion_walk(A, exact, LW, OW, t.motion_block, n, s);
    }
}

namespace {

void check_motions(const ompl_amd_env_s &e, const double *a, const double *b, int64_t n, double lvs, unsigned factor,
                   uint8_t *out, int32_t *checked, hipStream_t s)
{
    check_motions_tuned(tuning_for(e), e.env, a, b, n, lvs, factor, out, checked, s);
}

// Candidate launch shapes of the motion autotuner.
std::vector<Tuning> motion_candidates()
{
    std::vector<Tuning> c;
    const int shapes[][2] = {{8, 1}, {8, 2}, {8, 4}, {8, 8}, {7, 1}, {7, 2}, {7, 4}, {7, 9}, {4, 1}, {1, 1}};
    for (const auto &e : shapes)
        for (const auto &w : shapes) {
            const int LE = e[0], OE = e[1], LW = w[0], OW = w[1];
            // chunk-parallel walk of the survivors; shared-FK forms of the s2
            // pass (validity_shared_fk) and of the walk (strategy 4)
            // (shared FK exists for the sphere-slot shapes (8, parts > 1) only)
            const int es_max = (LE == 8 && OE > 1) ? 1 : 0;
            for (int wpc : {8, 16, 32})
                for (int es = 0; es <= es_max; ++es) {
                    c.push_back(Tuning{LE, OE, 256, LW, OW, 256, wpc, 3, es});
                    if (LW == 8 && OW > 1) c.push_back(Tuning{LE, OE, 256, LW, OW, 256, wpc, 4, es});
                }
            // fused one-launch form
            if (fused_supported(LE, OE, LW, OW))
                for (int block : {64, 256}) c.push_back(Tuning{LE, OE, block, LW, OW, block, 8, 2});
            // persistent queue, one wave per survivor
            for (int es = 0; es <= es_max; ++es) c.push_back(Tuning{LE, OE, 256, LW, OW, 256, 16, 1, es});
        }
    for (const auto &w : shapes) c.push_back(Tuning{8, 1, 256, w[0], w[1], 256, 8, 0});
    return c;
}

} // namespace

// A captured motion batch: the launches of one check_motions call recorded in
// HIP graphs (pointers, sizes and launch shape fixed at capture) with their own
// survivor workspace, so replays on any stream do not share state with eager
// calls.  No memset nodes: the strategies clear their counters with kernels.
// The batch is recorded twice, once per survivor counter set (each recording
// clears the other set for the next replay), and replays alternate between
// the two, as eager batches on one workspace do.
// A group of independent batches is recorded in one graph, batch 0 on the
// capturing stream and batch i > 0 on a side stream forked from it and joined
// back, so a replay runs the batches concurrently.
struct ompl_amd_motion_graph_s {
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int next = 0; // the recording the next replay launches
    std::vector<std::unique_ptr<Workspace>> ws; // one per batch
    std::vector<hipStream_t> side;     // batches 1.. are recorded on these
    std::vector<hipEvent_t> fork_join; // fork, then one join per side stream
    ~ompl_amd_motion_graph_s()
    {
        for (int i = 0; i < 2; ++i) {
            if (exec[i]) hipGraphExecDestroy(exec[i]);
            if (graph[i]) hipGraphDestroy(graph[i]);
        }
        for (auto &w : ws) {
            if (w->list) hipFree(w->list);
            if (w->first_bad) hipFree(w->first_bad);
            if (w->surv_nd) hipFree(w->surv_nd);
            if (w->surv_ab) hipFree(w->surv_ab);
            if (w->counters) hipFree(w->counters);
        }
        for (hipStream_t st : side) hipStreamDestroy(st);
        for (hipEvent_t ev : fork_join) hipEventDestroy(ev);
    }
};

namespace {

void graph_workspace(Workspace &w, int64_t n)
{
    const size_t cap = (size_t)n < 1024 ? 1024 : (size_t)n;
    OA_HIP(hipMalloc(&w.counters, 64 * sizeof(int32_t)));
    OA_HIP(hipMemset(w.counters, 0, 64 * sizeof(int32_t)));
    OA_HIP(hipMalloc(&w.list, cap * sizeof(int32_t)));
    OA_HIP(hipMalloc(&w.first_bad, cap * sizeof(int32_t)));
    OA_HIP(hipMalloc(&w.surv_nd, cap * sizeof(int32_t)));
    OA_HIP(hipMalloc(&w.surv_ab, cap * 14 * sizeof(double)));
    w.list_cap = cap;
    w.parity = 0;
}

// Records the group twice (once per survivor counter set of every batch's
// workspace) and instantiates both recordings.
void record_group(ompl_amd_motion_graph_s *g, const Env &env, const Tuning &t, int nb, const double *const *a,
                  const double *const *b, const int64_t *n, double lvs, unsigned factor, uint8_t *const *out,
                  int32_t *const *checked)
{
    for (int i = 0; i < nb; ++i) graph_workspace(*g->ws[(size_t)i], n[i]);
    for (int i = 1; i < nb; ++i) {
        hipStream_t st = nullptr;
        OA_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)); // default priority
        g->side.push_back(st);
    }
    for (int i = 0; i < nb; ++i) {
        hipEvent_t ev = nullptr;
        OA_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        g->fork_join.push_back(ev);
    }
    OA_HIP(hipDeviceSynchronize());
    // capture on the calling thread's own stream (the null stream cannot capture)
    hipStream_t cs = thread_resources().stream;
    for (int r = 0; r < 2; ++r) {
        // recording r uses counter set r of each workspace when its batch
        // alternates sets (check_motions_tuned flips ws.parity after recording)
        OA_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        try {
            if (nb > 1) {
                OA_HIP(hipEventRecord(g->fork_join[0], cs));
                for (int i = 1; i < nb; ++i) OA_HIP(hipStreamWaitEvent(g->side[(size_t)i - 1], g->fork_join[0], 0));
            }
            for (int i = 0; i < nb; ++i) {
                hipStream_t st = i == 0 ? cs : g->side[(size_t)i - 1];
                check_motions_tuned(t, env, a[i], b[i], n[i], lvs, factor, out[i], checked ? checked[i] : nullptr, st,
                                    g->ws[(size_t)i].get());
            }
            for (int i = 1; i < nb; ++i) {
                OA_HIP(hipEventRecord(g->fork_join[(size_t)i], g->side[(size_t)i - 1]));
                OA_HIP(hipStreamWaitEvent(cs, g->fork_join[(size_t)i], 0));
            }
        } catch (...) {
            hipGraph_t dropped = nullptr;
            hipStreamEndCapture(cs, &dropped);
            if (dropped) hipGraphDestroy(dropped);
            throw;
        }
        OA_HIP(hipStreamEndCapture(cs, &g->graph[r]));
        OA_HIP(hipGraphInstantiate(&g->exec[r], g->graph[r], nullptr, nullptr, 0));
    }
}

} // namespace

extern "C" {

int ompl_amd_env_create(const float *spheres, int n_spheres, const float *boxes, int n_boxes, ompl_amd_env *out)
{
    return guarded([&] {
        OA_REQUIRE(out, OMPL_AMD_ERR_ARG, "null output handle");
        OA_REQUIRE(n_spheres >= 0 && n_boxes >= 0, OMPL_AMD_ERR_ARG, "negative obstacle count");
        OA_REQUIRE(n_spheres == 0 || spheres, OMPL_AMD_ERR_ARG, "null sphere array");
        OA_REQUIRE(n_boxes == 0 || boxes, OMPL_AMD_ERR_ARG, "null box array");
        OA_REQUIRE((int64_t)n_spheres + 2 * (int64_t)n_boxes <= kMaxEnvWords, OMPL_AMD_ERR_ARG,
                   "environment too large: spheres + 2*boxes must be <= %d", kMaxEnvWords);
        ensure_init();
        std::vector<float4> h((size_t)n_spheres + 2 * (size_t)n_boxes + 1);
        for (int i = 0; i < n_spheres; ++i) {
            const float *o = spheres + 4 * i;
            OA_REQUIRE(std::isfinite(o[0]) && std::isfinite(o[1]) && std::isfinite(o[2]) && std::isfinite(o[3]),
                       OMPL_AMD_ERR_ARG, "sphere obstacle %d: non-finite value", i);
            OA_REQUIRE(o[3] >= 0.0f, OMPL_AMD_ERR_ARG, "negative sphere radius");
            const float rs = o[3] + kRobotR;
            h[i] = make_float4(o[0], o[1], o[2], rs * rs);
        }
        for (int i = 0; i < n_boxes; ++i) {
            const float *o = boxes + 6 * i;
            OA_REQUIRE(std::isfinite(o[0]) && std::isfinite(o[1]) && std::isfinite(o[2]) && std::isfinite(o[3]) &&
                           std::isfinite(o[4]) && std::isfinite(o[5]),
                       OMPL_AMD_ERR_ARG, "box obstacle %d: non-finite value", i);
            OA_REQUIRE(o[3] >= 0.0f && o[4] >= 0.0f && o[5] >= 0.0f, OMPL_AMD_ERR_ARG, "negative box extent");
            h[n_spheres + 2 * i] = make_float4(o[0] - o[3], o[1] - o[4], o[2] - o[5], 0.0f);
            h[n_spheres + 2 * i + 1] = make_float4(o[0] + o[3], o[1] + o[4], o[2] + o[5], 0.0f);
        }
        // h (caller's order) goes after the staging-order copy in one allocation
        const std::vector<int> so = nearest_first(n_spheres, [&](int i) {
            const float *o = spheres + 4 * i;
            return std::sqrt((double)o[0] * o[0] + (double)o[1] * o[1] + (double)o[2] * o[2]) - (double)o[3];
        });
        const std::vector<int> bo = nearest_first(n_boxes, [&](int i) {
            const float4 lo = h[n_spheres + 2 * i], hi = h[n_spheres + 2 * i + 1];
            const double dx = std::max({(double)lo.x, -(double)hi.x, 0.0});
            const double dy = std::max({(double)lo.y, -(double)hi.y, 0.0});
            const double dz = std::max({(double)lo.z, -(double)hi.z, 0.0});
            return std::sqrt(dx * dx + dy * dy + dz * dz);
        });
        std::vector<float4> both(2 * h.size());
        for (int i = 0; i < n_spheres; ++i) both[i] = h[so[i]];
        for (int i = 0; i < n_boxes; ++i) {
            both[n_spheres + 2 * i] = h[n_spheres + 2 * bo[i]];
            both[n_spheres + 2 * i + 1] = h[n_spheres + 2 * bo[i] + 1];
        }
        both[h.size() - 1] = h[h.size() - 1];
        std::copy(h.begin(), h.end(), both.begin() + h.size());
        auto *e = new ompl_amd_env_s();
        e->env.ns = n_spheres;
        e->env.nb = n_boxes;
        e->env.bytes = both.size() * sizeof(float4);
        hipError_t err = hipMalloc(&e->env.obs, e->env.bytes);
        if (err != hipSuccess) {
            delete e;
            fail(OMPL_AMD_ERR_NOMEM, "environment allocation failed: %s", hipGetErrorString(err));
        }
        e->env.obs_input = e->env.obs + h.size();
        OA_HIP(hipMemcpy(e->env.obs, both.data(), e->env.bytes, hipMemcpyHostToDevice));
        const size_t lb = links_bytes(e->env.view()) + sizeof(float4);
        err = hipMalloc(&e->env.obs_links, lb);
        if (err != hipSuccess) {
            (void)hipFree(e->env.obs);
            delete e;
            fail(OMPL_AMD_ERR_NOMEM, "environment allocation failed: %s", hipGetErrorString(err));
        }
        e->env.bytes += lb;
        ThreadResources &r = thread_resources();
        launch_stage_links(e->env.view(), e->env.obs_links, r.stream);
        OA_HIP(hipStreamSynchronize(r.stream));
        *out = e;
    });
}

int ompl_amd_env_destroy(ompl_amd_env e)
{
    return guarded([&] {
        if (!e) return;
        const int users = e->server_users.load();
        OA_REQUIRE(users == 0, OMPL_AMD_ERR_ARG,
                   "environment still in use by %d queued or running server request(s); destroy it after they finish",
                   users);
        if (e->env.obs) OA_HIP(hipFree(e->env.obs));
        if (e->env.obs_links) OA_HIP(hipFree(e->env.obs_links));
        delete e;
    });
}

int ompl_amd_env_bytes(ompl_amd_env e, size_t *bytes)
{
    return guarded([&] {
        OA_REQUIRE(bytes, OMPL_AMD_ERR_ARG, "null output");
        *bytes = env_of(e).bytes;
    });
}

int ompl_amd_validate_states_dev(ompl_amd_env e, const float *q, int64_t n, uint8_t *out, void *stream)
{
    return guarded([&] {
        const Env &env = env_of(e);
        OA_REQUIRE(n >= 0, OMPL_AMD_ERR_ARG, "negative state count");
        if (n == 0) return;
        OA_REQUIRE(q && out, OMPL_AMD_ERR_ARG, "null state or output pointer");
        const Tuning t = tuning_for(*e);
        launch_validate<float>(env.view(), q, n, out, t.validity_lanes_per_state, t.validity_obstacle_parts,
                               t.validity_block, resolve_stream(stream), t.validity_shared_fk != 0);
    });
}

int ompl_amd_validate_states_f64_dev(ompl_amd_env e, const double *q, int64_t n, uint8_t *out, void *stream)
{
    return guarded([&] {
        const Env &env = env_of(e);
        OA_REQUIRE(n >= 0, OMPL_AMD_ERR_ARG, "negative state count");
        if (n == 0) return;
        OA_REQUIRE(q && out, OMPL_AMD_ERR_ARG, "null state or output pointer");
        const Tuning t = tuning_for(*e);
        launch_validate<double>(env.view(), q, n, out, t.validity_lanes_per_state, t.validity_obstacle_parts,
                                t.validity_block, resolve_stream(stream), t.validity_shared_fk != 0);
    });
}

int ompl_amd_is_valid(ompl_amd_env e, const double *q, int *valid)
{
    return guarded([&] {
        const Env &env = env_of(e);
        OA_REQUIRE(q && valid, OMPL_AMD_ERR_ARG, "null state or output pointer");
        ThreadResources &r = thread_resources();
        State1 st;
        for (int j = 0; j < kDof; ++j) st.q[j] = q[j];
        launch_validate_single(env.view(), st, r.pinned_dev, r.stream);
        OA_HIP(hipStreamSynchronize(r.stream));
        *valid = r.pinned[0];
    });
}

int ompl_amd_arm_spheres_dev(const float *q, int64_t n, float *out, void *stream)
{
    return guarded([&] {
        OA_REQUIRE(n >= 0, OMPL_AMD_ERR_ARG, "negative state count");
        if (n == 0) return;
        OA_REQUIRE(q && out, OMPL_AMD_ERR_ARG, "null pointer");
        launch_arm_spheres(q, n, out, resolve_stream(stream));
    });
}

int ompl_amd_check_motions_dev(ompl_amd_env e, const double *a, const double *b, int64_t n, double lvs,
                               unsigned factor, uint8_t *out, int32_t *checked, void *stream)
{
    return guarded([&] {
        env_of(e);
        check_motions(*e, a, b, n, lvs, factor, out, checked, resolve_stream(stream));
    });
}

int ompl_amd_check_motions_last_valid_dev(ompl_amd_env e, const double *a, const double *b, int64_t n, double lvs,
                                          unsigned factor, uint8_t *out, double *last_t, void *stream)
{
    return guarded([&] {
        const Env &env = env_of(e);
        OA_REQUIRE(n >= 0, OMPL_AMD_ERR_ARG, "negative edge count");
        if (n == 0) return;
        OA_REQUIRE(a && b && out && last_t, OMPL_AMD_ERR_ARG, "null pointer");
        OA_REQUIRE(lvs > 0.0 && factor >= 1, OMPL_AMD_ERR_ARG, "longest valid segment must be > 0 and factor >= 1");
        OA_REQUIRE(n < (int64_t)1 << 31, OMPL_AMD_ERR_ARG, "at most 2^31-1 edges per call");
        const Tuning t = tuning_for(*e);
        MotionArgs A{};
        A.env = env.view();
        A.a = a;
        A.b = b;
        A.n = n;
        A.lvs = lvs;
        A.factor = factor;
        A.out = out;
        const hipStream_t st = resolve_stream(stream);
        std::unique_lock<std::mutex> lk; // held until the launches are enqueued
        Workspace &w = workspace(st, (size_t)n, lk);
        A.list = w.list;
        A.first_bad = w.first_bad;
        A.surv_nd = w.surv_nd;
        A.surv_ab = w.surv_ab;
        A.list_count = w.counters + 2;
        A.max_m = w.counters + 3;
        launch_motion_last_valid(A, last_t, t.motion_lanes_per_state, t.motion_obstacle_parts,
                                 (int64_t)num_cus() * t.motion_waves_per_cu, st);
    });
}

int ompl_amd_check_motion(ompl_amd_env e, const double *a, const double *b, double lvs, unsigned factor, int *valid)
{
    return guarded([&] {
        const Env &env = env_of(e);
        OA_REQUIRE(a && b && valid, OMPL_AMD_ERR_ARG, "null pointer");
        OA_REQUIRE(lvs > 0.0 && factor >= 1, OMPL_AMD_ERR_ARG, "longest valid segment must be > 0 and factor >= 1");
        ThreadResources &r = thread_resources();
        Edge1 E;
        for (int j = 0; j < kDof; ++j) {
            E.a[j] = a[j];
            E.b[j] = b[j];
        }
        launch_motion_single(env.view(), E, lvs, factor, r.pinned_dev, r.stream);
        OA_HIP(hipStreamSynchronize(r.stream));
        *valid = r.pinned[0];
    });
}

int ompl_amd_check_state_lists_dev(ompl_amd_env e, const double *q_soa, int64_t n_states, const int64_t *offsets,
                                   int64_t n_lists, uint8_t *out, int32_t *first_invalid, void *stream)
{
    return guarded([&] {
        const Env &env = env_of(e);
        OA_REQUIRE(n_states >= 0 && n_lists >= 0, OMPL_AMD_ERR_ARG, "negative sizes");
        if (n_lists == 0) return;
        OA_REQUIRE(offsets && out && (n_states == 0 || q_soa), OMPL_AMD_ERR_ARG, "null pointer");
        OA_REQUIRE(n_lists < (int64_t)1 << 31 && n_states < (int64_t)1 << 31, OMPL_AMD_ERR_ARG,
                   "at most 2^31-1 lists and states per call");
        const Tuning t = tuning_for(*e);
        const hipStream_t st = resolve_stream(stream);
        std::unique_lock<std::mutex> lk; // held until the launches are enqueued
        Workspace &w = workspace(st, (size_t)n_lists, lk);
        launch_state_lists(env.view(), q_soa, n_states, offsets, n_lists, out, first_invalid,
                           t.motion_lanes_per_state, t.motion_obstacle_parts,
                           (int64_t)num_cus() * t.motion_waves_per_cu, w.list, w.first_bad, w.counters + 2, st);
    });
}

int ompl_amd_check_states(ompl_amd_env e, const double *states, int count, int *valid, int *first_invalid)
{
    return guarded([&] {
        const Env &env = env_of(e);
        OA_REQUIRE(count >= 0 && valid && (count == 0 || states), OMPL_AMD_ERR_ARG, "bad arguments");
        ThreadResources &r = thread_resources();
        // staging: offsets [2] int64, result u8 + i32, then the states dof-major
        const size_t head = 64;
        stage_reserve(r, head + (size_t)count * kDof * sizeof(double));
        int64_t *off = reinterpret_cast<int64_t *>(r.stage);
        off[0] = 0;
        off[1] = count;
        double *q = reinterpret_cast<double *>(r.stage + head);
        for (int i = 0; i < count; ++i)
            for (int j = 0; j < kDof; ++j) q[(size_t)j * count + i] = states[(size_t)i * kDof + j];
        const Tuning t = tuning_for(*e);
        {
            std::unique_lock<std::mutex> lk; // held until the launches are enqueued
            Workspace &w = workspace(r.stream, 1, lk);
            launch_state_lists(env.view(), reinterpret_cast<const double *>(r.stage_dev + head), count,
                               reinterpret_cast<const int64_t *>(r.stage_dev), 1,
                               reinterpret_cast<uint8_t *>(r.stage_dev + 16),
                               reinterpret_cast<int32_t *>(r.stage_dev + 32), t.motion_lanes_per_state,
                               t.motion_obstacle_parts, (int64_t)num_cus() * t.motion_waves_per_cu, w.list,
                               w.first_bad, w.counters + 2, r.stream);
        }
        OA_HIP(hipStreamSynchronize(r.stream));
        *valid = reinterpret_cast<uint8_t *>(r.stage)[16];
        if (first_invalid && !*valid) *first_invalid = *reinterpret_cast<int32_t *>(r.stage + 32);
    });
}

int ompl_amd_check_motion_last_valid(ompl_amd_env e, const double *a, const double *b, double lvs, unsigned factor,
                                     int *valid, double *last_t)
{
    return guarded([&] {
        const Env &env = env_of(e);
        OA_REQUIRE(a && b && valid && last_t, OMPL_AMD_ERR_ARG, "null pointer");
        OA_REQUIRE(lvs > 0.0 && factor >= 1, OMPL_AMD_ERR_ARG, "longest valid segment must be > 0 and factor >= 1");
        ThreadResources &r = thread_resources();
        // endpoints staged through the pinned slot: [a(7), b(7), last_t] doubles
        double *pd = reinterpret_cast<double *>(r.pinned + 16);
        for (int j = 0; j < kDof; ++j) {
            pd[j] = a[j];
            pd[kDof + j] = b[j];
        }
        pd[2 * kDof] = *last_t;
        // the kernel reads dof-major [7][1] arrays: identical to a flat vector
        double *dev_pd = reinterpret_cast<double *>(r.pinned_dev + 16);
        uint8_t *dev_out = reinterpret_cast<uint8_t *>(r.pinned_dev);
        MotionArgs A{};
        A.env = env.view();
        A.a = dev_pd;
        A.b = dev_pd + kDof;
        A.n = 1;
        A.lvs = lvs;
        A.factor = factor;
        A.out = dev_out;
        {
            std::unique_lock<std::mutex> lk; // held until the launches are enqueued
            Workspace &w = workspace(r.stream, 1, lk);
            A.list = w.list;
            A.first_bad = w.first_bad;
            A.surv_nd = w.surv_nd;
            A.surv_ab = w.surv_ab;
            A.list_count = w.counters + 2;
            A.max_m = w.counters + 3;
            const Tuning t = tuning_for(*e);
            launch_motion_last_valid(A, dev_pd + 2 * kDof, t.motion_lanes_per_state, t.motion_obstacle_parts,
                                     (int64_t)num_cus() * t.motion_waves_per_cu, r.stream);
        }
        OA_HIP(hipStreamSynchronize(r.stream));
        *valid = reinterpret_cast<uint8_t *>(r.pinned)[0];
        *last_t = pd[2 * kDof];
    });
}

} // extern "C"

extern "C" int ompl_amd_env_set_tuning(ompl_amd_env e, const ompl_amd_tuning *t)
{
    return guarded([&] {
        env_of(e);
        if (t) check_tuning(tuning_from_c(*t));
        std::lock_guard<std::mutex> lk(e->tuning_mutex);
        e->has_tuning = t != nullptr;
        if (t) e->tuning = tuning_from_c(*t);
    });
}

extern "C" int ompl_amd_env_get_tuning(ompl_amd_env e, ompl_amd_tuning *t, int *own)
{
    return guarded([&] {
        env_of(e);
        OA_REQUIRE(t, OMPL_AMD_ERR_ARG, "null tuning");
        bool has = false;
        {
            std::lock_guard<std::mutex> lk(e->tuning_mutex);
            has = e->has_tuning;
        }
        *t = tuning_to_c(tuning_for(*e));
        if (own) *own = has ? 1 : 0;
    });
}

extern "C" int ompl_amd_autotune_motions(ompl_amd_env e, const double *a, const double *b, int64_t n, double lvs,
                                         unsigned factor, int reps, ompl_amd_tuning *best, double *best_us)
{
    return guarded([&] {
        const Env &env = env_of(e);
        OA_REQUIRE(n > 0 && a && b && reps >= 1, OMPL_AMD_ERR_ARG, "bad autotune arguments");
        ThreadResources &r = thread_resources();
        hipStream_t s = r.stream;
        uint8_t *out = nullptr;
        OA_HIP(hipMalloc(&out, (size_t)n));
        hipEvent_t e0, e1;
        OA_HIP(hipEventCreate(&e0));
        OA_HIP(hipEventCreate(&e1));
        const Tuning base = get_tuning();
        Tuning win = base;
        float win_ms = INFINITY;
        for (const Tuning &t : motion_candidates()) {
            for (int w = 0; w < 2; ++w) check_motions_tuned(t, env, a, b, n, lvs, factor, out, nullptr, s);
            OA_HIP(hipEventRecord(e0, s));
            for (int i = 0; i < reps; ++i) check_motions_tuned(t, env, a, b, n, lvs, factor, out, nullptr, s);
            OA_HIP(hipEventRecord(e1, s));
            OA_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            OA_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < win_ms) {
                win_ms = ms;
                win = t;
            }
        }
        OA_HIP(hipEventDestroy(e0));
        OA_HIP(hipEventDestroy(e1));
        OA_HIP(hipFree(out));
        set_tuning(win);
        if (best) {
            *best = tuning_to_c(win);
        }
        if (best_us) *best_us = 1e3 * (double)win_ms / reps;
    });
}

namespace {

// Device buffers and events of one stream-pipeline autotune, released on
// every exit.
struct PipelineScratch {
    uint8_t *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<hipEvent_t> join;
    ~PipelineScratch()
    {
        if (out) (void)hipFree(out);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        for (hipEvent_t e : join) (void)hipEventDestroy(e);
    }
};

// Launch shapes tried for batches on several streams: the s2 pass's and the
// walk's (lanes per state, obstacle parts) and the walk's waves per CU, all
// with the chunk-parallel walk (strategy 3).
std::vector<Tuning> stream_candidates(const Tuning &base)
{
    const int s2[][2] = {{7, 1}, {7, 2}, {7, 4}, {7, 9}, {8, 4}, {8, 8}};
    const int walk[][2] = {{7, 2}, {7, 4}, {7, 9}, {8, 8}};
    std::vector<Tuning> c;
    for (const auto &e : s2)
        for (const auto &w : walk)
            for (int wpc : {2, 4, 8, 16}) {
                Tuning t = base;
                t.validity_lanes_per_state = e[0];
                t.validity_obstacle_parts = e[1];
                t.motion_lanes_per_state = w[0];
                t.motion_obstacle_parts = w[1];
                t.motion_waves_per_cu = wpc;
                t.motion_strategy = 3;
                t.validity_shared_fk = 0;
                c.push_back(t);
            }
    return c;
}

} // namespace

extern "C" int ompl_amd_autotune_motions_streams(ompl_amd_env e, const double *a, const double *b, int64_t n,
                                                 double lvs, unsigned factor, void *const *streams, int n_streams,
                                                 int steps, ompl_amd_tuning *best, double *best_us)
{
    return guarded([&] {
        const Env &env = env_of(e);
        OA_REQUIRE(n > 0 && a && b && steps >= 1, OMPL_AMD_ERR_ARG, "bad autotune arguments");
        OA_REQUIRE(streams && n_streams >= 1 && n_streams <= 16, OMPL_AMD_ERR_ARG, "1 to 16 streams");
        {
            std::lock_guard<std::mutex> lk(e->tuning_mutex);
            OA_REQUIRE(!e->has_tuning, OMPL_AMD_ERR_ARG,
                       "the environment has its own launch shapes, which take the place of the library's");
        }
        std::vector<hipStream_t> st((size_t)n_streams);
        for (int j = 0; j < n_streams; ++j) st[(size_t)j] = resolve_stream(streams[j]);
        PipelineScratch S;
        OA_HIP(hipMalloc(&S.out, (size_t)n * (size_t)n_streams));
        OA_HIP(hipEventCreate(&S.e0));
        OA_HIP(hipEventCreate(&S.e1));
        for (int j = 0; j < n_streams; ++j) {
            hipEvent_t ev = nullptr;
            OA_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            S.join.push_back(ev);
        }
        // k batches round-robin on the streams, between events on the first,
        // which waits for the others at the end; microseconds per batch
        auto per_batch = [&](const Tuning &t, int k) {
            for (int i = 0; i < 2 * n_streams; ++i) { // warm: workspaces sized, kernels loaded
                const int j = i % n_streams;
                check_motions_tuned(t, env, a, b, n, lvs, factor, S.out + (size_t)j * (size_t)n, nullptr,
                                    st[(size_t)j]);
            }
            for (hipStream_t s : st) OA_HIP(hipStreamSynchronize(s));
            OA_HIP(hipEventRecord(S.e0, st[0]));
            for (int j = 1; j < n_streams; ++j) OA_HIP(hipStreamWaitEvent(st[(size_t)j], S.e0, 0));
            for (int i = 0; i < k; ++i) {
                const int j = i % n_streams;
                check_motions_tuned(t, env, a, b, n, lvs, factor, S.out + (size_t)j * (size_t)n, nullptr,
                                    st[(size_t)j]);
            }
            for (int j = 1; j < n_streams; ++j) {
                OA_HIP(hipEventRecord(S.join[(size_t)j], st[(size_t)j]));
                OA_HIP(hipStreamWaitEvent(st[0], S.join[(size_t)j], 0));
            }
            OA_HIP(hipEventRecord(S.e1, st[0]));
            OA_HIP(hipEventSynchronize(S.e1));
            float ms = 0.0f;
            OA_HIP(hipEventElapsedTime(&ms, S.e0, S.e1));
            return 1e3 * (double)ms / k;
        };
        std::vector<std::pair<double, Tuning>> timed;
        for (const Tuning &t : stream_candidates(get_tuning())) timed.emplace_back(per_batch(t, steps), t);
        std::stable_sort(timed.begin(), timed.end(),
                         [](const std::pair<double, Tuning> &x, const std::pair<double, Tuning> &y) {
                             return x.first < y.first;
                         });
        // the three fastest again over twice the batches
        double win_us = INFINITY;
        Tuning win = timed.front().second;
        for (size_t i = 0; i < timed.size() && i < 3; ++i) {
            const double us = per_batch(timed[i].second, 2 * steps);
            if (us < win_us) {
                win_us = us;
                win = timed[i].second;
            }
        }
        set_tuning(win);
        if (best) *best = tuning_to_c(win);
        if (best_us) *best_us = win_us;
    });
}

extern "C" {

int ompl_amd_motion_graph_create(ompl_amd_env e, const double *a, const double *b, int64_t n, double lvs,
                                 unsigned factor, uint8_t *out, int32_t *checked, ompl_amd_motion_graph *graph)
{
    return ompl_amd_motion_graph_create_group(e, 1, &a, &b, &n, lvs, factor, &out, checked ? &checked : nullptr,
                                              graph);
}

int ompl_amd_motion_graph_create_group(ompl_amd_env e, int n_batches, const double *const *a, const double *const *b,
                                       const int64_t *n, double lvs, unsigned factor, uint8_t *const *out,
                                       int32_t *const *checked, ompl_amd_motion_graph *graph)
{
    return guarded([&] {
        const Env &env = env_of(e);
        OA_REQUIRE(graph, OMPL_AMD_ERR_ARG, "null graph handle");
        OA_REQUIRE(n_batches >= 1 && n_batches <= 16, OMPL_AMD_ERR_ARG, "a captured group holds 1 to 16 batches");
        OA_REQUIRE(a && b && n && out, OMPL_AMD_ERR_ARG, "null batch arrays");
        for (int i = 0; i < n_batches; ++i) {
            OA_REQUIRE(n[i] >= 1, OMPL_AMD_ERR_ARG, "a captured batch needs at least one edge");
            OA_REQUIRE(a[i] && b[i] && out[i], OMPL_AMD_ERR_ARG, "null pointer in batch %d", i);
        }
        auto *g = new ompl_amd_motion_graph_s();
        for (int i = 0; i < n_batches; ++i) g->ws.push_back(std::make_unique<Workspace>());
        try {
            record_group(g, env, tuning_for(*e), n_batches, a, b, n, lvs, factor, out, checked);
        } catch (...) {
            delete g;
            throw;
        }
        *graph = g;
    });
}

int ompl_amd_motion_graph_launch(ompl_amd_motion_graph graph, void *stream)
{
    return guarded([&] {
        OA_REQUIRE(graph && graph->exec[0] && graph->exec[1], OMPL_AMD_ERR_ARG, "null graph");
        OA_HIP(hipGraphLaunch(graph->exec[graph->next], resolve_stream(stream)));
        graph->next ^= 1;
    });
}

int ompl_amd_motion_graph_destroy(ompl_amd_motion_graph graph)
{
    return guarded([&] {
        if (graph) OA_HIP(hipDeviceSynchronize());
        delete graph;
    });
}

} // extern "C"
