// Runtime state shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <atomic>
#include <mutex>
#include <vector>

#include "checker.hpp"
#include "common.hpp"

namespace ompl_amd {

void init_device(int dev);
// Takes the device address of the no-allocation entry point's scratch array
// (once per device, with that device current).
void init_noalloc_scratch(int dev);

// Per-stream scratch of the batched motion path: survivor list + counters.
struct Workspace {
    int32_t *list = nullptr;
    int32_t *first_bad = nullptr; // per survivor (chunk-parallel walk)
    int32_t *surv_nd = nullptr;   // per survivor segment count (chunk-parallel walk)
    double *surv_ab = nullptr;    // per survivor endpoints a, b (14 doubles)
    size_t list_cap = 0;
    // [0] survivor count, [1] walk queue (strategy 1, cleared per call);
    // strategies 3 and 4: two sets of (survivor count, max positions) at
    // [8..9] and [10..11], zero at creation.  A batch uses set `parity`; either
    // its results kernel clears that set, or (no results kernel) its phase-1
    // walk clears the other set and the next batch uses that one.
    int32_t *counters = nullptr;
    int parity = 0;
    // Held by a batch from choosing its counter set to enqueuing its launches
    // (and around resizing), so host threads that share a stream cannot pick
    // the same set or see arrays being replaced.
    std::mutex mu;
};
// The stream's workspace with room for list_entries survivors.  The caller
// must hold the returned workspace's mu while it uses the workspace (lock
// passes that lock back, already taken, before any resize).
Workspace &workspace(hipStream_t s, size_t list_entries, std::unique_lock<std::mutex> &lock);

// Per-thread stream and pinned result slot of the blocking single-call entry
// points (OMPL checkers must be thread-safe: base/StateValidityChecker.h:87-89).
struct ThreadResources {
    hipStream_t stream = nullptr;
    int32_t *pinned = nullptr;     // 4 KiB, host-coherent, device-mapped
    int32_t *pinned_dev = nullptr; // device address of the same memory
    char *stage = nullptr;         // growable host-coherent, device-mapped staging
    char *stage_dev = nullptr;
    size_t stage_cap = 0;
};
ThreadResources &thread_resources();
// Ensures r.stage holds `bytes` (callers synchronize r.stream before reuse).
void stage_reserve(ThreadResources &r, size_t bytes);

// Launch shapes (mirrors ompl_amd_tuning).
struct Tuning {
    int validity_lanes_per_state;
    int validity_obstacle_parts;
    int validity_block;
    int motion_lanes_per_state;
    int motion_obstacle_parts;
    int motion_block;
    int motion_waves_per_cu;
    int motion_strategy;
    int validity_shared_fk = 0;
};

Tuning get_tuning();
// Rejects (OA_REQUIRE) a tuning whose shapes or limits are not built.
void check_tuning(const Tuning &t);

struct Env;
// Batched DiscreteMotionValidator::checkMotion(s1, s2) over device arrays with
// launch shapes t; `own` supplies the survivor workspace (null: the stream's).
void check_motions_tuned(const Tuning &t, const Env &env, const double *a, const double *b, int64_t n, double lvs,
                         unsigned factor, uint8_t *out, int32_t *checked, hipStream_t s, Workspace *own = nullptr);
void set_tuning(const Tuning &t);
ompl_amd_tuning tuning_to_c(const Tuning &x);
Tuning tuning_from_c(const ompl_amd_tuning &t);

// Device-resident obstacle set.
struct Env {
    // ns spheres (x,y,z,(R+r)^2), then nb boxes as (lo),(hi); each list in
    // staging order (nearest to the arm's base first), for the checkers
    float4 *obs = nullptr;
    // the same words in the caller's order (same allocation), for the
    // reference-order work accounting only
    float4 *obs_input = nullptr;
    // the same obstacles in the link layout's staging (spheres (x,y,z,bound^2),
    // boxes, then the spheres' rr as floats), for kernels that read them from HBM
    float4 *obs_links = nullptr;
    int ns = 0;
    int nb = 0;
    size_t bytes = 0;
    EnvView view() const { return EnvView{obs, ns, nb}; }
    EnvView input_view() const { return EnvView{obs_input, ns, nb}; }
};

// A planning query on a batch PRM roadmap (prm.hip): start and goal become
// milestones 0 and 1 of a roadmap over n_samples uniform samples (n_streams
// generators seeded seed, seed + 1, ...) with k neighbours per milestone; the
// path is the shortest one between them over the valid edges.
struct PrmQuery {
    bool solved = false;
    int64_t milestones = 0, candidate_edges = 0, valid_edges = 0;
    std::vector<double> path; // row-major states
};
PrmQuery prm_query(ompl_amd_env env, const double start[7], const double goal[7], uint32_t seed, int n_streams,
                   int64_t n_samples, int k, double lvs, unsigned factor, hipStream_t s);
// Device bytes a prm_query of n_samples with k neighbours may hold at once (an upper estimate).
size_t prm_footprint(int64_t n_samples, int k);
// Shortest path between milestones 0 and 1 of a symmetric CSR roadmap.
std::vector<int32_t> roadmap_shortest_path(const std::vector<int64_t> &off, const std::vector<int32_t> &adj,
                                           const std::vector<double> &nodes);

// Largest obstacle set staged in LDS by one block (float4 words).
constexpr int kMaxEnvWords = 4096; // 64 KiB of dynamic LDS

} // namespace ompl_amd

struct ompl_amd_env_s {
    ompl_amd::Env env;
    // server requests (queued or running) that read this environment;
    // ompl_amd_env_destroy refuses while any remain
    std::atomic<int> server_users{0};
    // launch shapes of this environment's batched checks, when set; else the
    // library's current tuning (the fastest shape depends on the obstacle
    // count: many obstacle partitions suit a dense scene, one a sparse one)
    mutable std::mutex tuning_mutex;
    bool has_tuning = false;
    ompl_amd::Tuning tuning{};
};

namespace ompl_amd {
// The tuning the batched checks on environment e use.
Tuning tuning_for(const ompl_amd_env_s &e);
} // namespace ompl_amd
