// Host-visible argument blocks and launchers of the checker kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "arm_device.hpp"
#include "common.hpp"

namespace ompl_amd {

struct MotionArgs {
    EnvView env;
    const double *a; // FP64 [7][n] dof-major
    const double *b;
    int64_t n;
    double lvs;
    unsigned factor;
    uint8_t *out;
    int32_t *checked;    // optional
    int32_t *list;       // survivor list (endpoint output / walk input), optional
    int32_t *list_count; // number of entries in list
    int32_t *queue;      // dynamic work counter for the walk, optional
    int32_t *first_bad;  // per survivor: first invalid position (chunk-parallel walk)
    int32_t *max_m;      // largest bisection position count among survivors
    int32_t *surv_nd;    // per survivor: segment count nd
    double *surv_ab;     // per survivor: endpoints a then b (14 doubles, contiguous), optional
    int flat;             // chunk walk: the launch reserved LDS for the survivors' chunk offsets
    int32_t *reset_next;  // chunk walk without a results kernel: the counter set phase 1 clears
    int start;           // first walk position (0 = s2, 1 = skip s2)
};

struct Edge1 {
    double a[kDof];
    double b[kDof];
};

struct State1 {
    double q[kDof];
};

bool shape_supported(int lps, int ops);
// Bytes of an environment's link-layout copy, and the kernel that writes it.
size_t links_bytes(const EnvView &env);
void launch_stage_links(const EnvView &env, float4 *out, hipStream_t s);
bool fused_supported(int le, int oe, int lw, int ow);

template <typename T>
void launch_validate(const EnvView &env, const T *q, int64_t n, uint8_t *out, int lps, int ops, int block,
                     hipStream_t s, bool shared = false);
void launch_zero(int *p, int n, hipStream_t s);
void launch_motion_endpoint(const MotionArgs &A, int lps, int ops, int block, hipStream_t s, bool shared = false);
void launch_motion_walk(const MotionArgs &A, bool exact, int lps, int ops, int block, int64_t waves, hipStream_t s);
void launch_motion_chunks(const MotionArgs &A, bool exact, int lps, int ops, int block, int64_t waves,
                          hipStream_t s, bool shared, bool results);
void launch_motion_fused(const MotionArgs &A, bool exact, int le, int oe, int lw, int ow, int block, hipStream_t s);
void launch_motion_single(const EnvView &env, const Edge1 &E, double lvs, unsigned factor, int *result,
                          hipStream_t s);
void launch_validate_single(const EnvView &env, const State1 &st, int *result, hipStream_t s);
// Uses A.list, first_bad, surv_nd, surv_ab (room for A.n entries) and the
// adjacent counters list_count, max_m (cleared by the launch); waves: the
// persistent grid of the chunk phase.
void launch_motion_last_valid(const MotionArgs &A, double *last_t, int lps, int ops, int64_t waves, hipStream_t s);
// lps, ops: the launch shape (the walk's: lowest-invalid exit over a list's
// states); waves: the persistent grid of the chunk phase; unres / unres_fb:
// room for n_lists entries each; counters: two ints (cleared by the launch).
void launch_state_lists(const EnvView &env, const double *q_soa, int64_t n_states, const int64_t *offsets,
                        int64_t n_lists, uint8_t *out, int32_t *first_invalid, int lps, int ops, int64_t waves,
                        int32_t *unres, int32_t *unres_fb, int32_t *counters, hipStream_t s);
void launch_arm_spheres(const float *q, int64_t n, float *out, hipStream_t s);

} // namespace ompl_amd
