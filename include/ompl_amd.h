/*
 * ompl_amd.h -- C-ABI of libompl_amd, the MI355X (gfx950) implementation of
 * OMPL's data-parallel planning inner loop.  Plain pointers and sizes only.
 *
 * Every function returns OMPL_AMD_OK (0) or a negative status; the message of
 * the last failure on the calling thread is ompl_amd_last_error().  The C++
 * adapter ompl_amd/include/ompl_amd/nearest_neighbors_gpu.hpp (and checker /
 * validator bindings written the same way) rethrows failures as
 * ompl::Exception, the reference's error type (util/Exception.h:46).
 *
 * Pointer conventions: parameters named *_dev are device pointers (HBM);
 * all others are host pointers.  `stream` is a hipStream_t (NULL = HIP's
 * null stream).  *_dev entry points are asynchronous on `stream`;
 * host-pointer entry points return after the result is in host memory.
 *
 * State layouts: joint configurations are structure-of-arrays, dof-major:
 * q[j * n + i] is joint j of state i (coalesced loads: lane i reads q[j*n+i]).
 * Nearest-neighbour points are row-major [n][dim] FP32.
 */
#ifndef OMPL_AMD_H
#define OMPL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OMPL_AMD_OK 0
#define OMPL_AMD_ERR_ARG -1       /* invalid argument (null pointer, bad size) */
#define OMPL_AMD_ERR_HIP -2       /* HIP runtime error */
#define OMPL_AMD_ERR_NOMEM -3     /* device or host allocation failed */
#define OMPL_AMD_ERR_EMPTY -4     /* nearest() on an empty structure */
#define OMPL_AMD_ERR_ADMISSION -5 /* serving: request exceeds the HBM budget */
#define OMPL_AMD_ERR_DEADLINE -6  /* serving: request deadline expired */
#define OMPL_AMD_ERR_CANCELLED -7 /* serving: request cancelled */
#define OMPL_AMD_ERR_INTERNAL -8

/* ------------------------------------------------------------------ */
/* Library                                                             */
/* ------------------------------------------------------------------ */
int ompl_amd_init(int device);
const char *ompl_amd_last_error(void);
/* major*10000 + minor*100 + patch; bumped with every change of a signature or
 * struct layout, so a binding can refuse a library built from other sources. */
#define OMPL_AMD_VERSION 313
int ompl_amd_version(void);

/* Limits of this build and of the device it runs on. */
typedef struct {
    int arm_dof;              /* joints of the arm checker (7) */
    int arm_spheres;          /* its collision spheres (28) */
    int max_env_words;        /* spheres + 2 * boxes per environment */
    int max_segments;         /* segments per motion (longer motions are invalid) */
    int max_k;                /* k of nearest-neighbour queries */
    int max_nn_dim;           /* L2 dimension of nearest-neighbour points */
    int max_sample_dim;       /* dimension of uniform sampling */
    int device;               /* HIP device in use */
    int compute_units;
    char device_name[64];
} ompl_amd_caps;
/* Initialises the device like ompl_amd_init(0) when needed. */
int ompl_amd_capabilities(ompl_amd_caps *caps);
int ompl_amd_synchronize(void *stream);

/* Device memory helpers (thin wrappers over hipMalloc/hipMemcpy) for callers
 * that do not own a HIP runtime binding. */
int ompl_amd_malloc(void **ptr_dev, size_t bytes);
int ompl_amd_free(void *ptr_dev);
int ompl_amd_memcpy_h2d(void *dst_dev, const void *src, size_t bytes);
int ompl_amd_memcpy_d2h(void *dst, const void *src_dev, size_t bytes);

/* ------------------------------------------------------------------ */
/* Environments (obstacle sets, device resident)                       */
/* ------------------------------------------------------------------ */
typedef struct ompl_amd_env_s *ompl_amd_env;

/* spheres: n_spheres x (cx, cy, cz, radius); boxes: n_boxes x (cx, cy, cz,
 * hx, hy, hz) axis-aligned, half-extents.  Host pointers, copied to HBM.
 * Values must be finite, radii and half-extents >= 0.  The checkers visit
 * each list nearest to the arm's base first; results do not depend on the
 * order given here. */
int ompl_amd_env_create(const float *spheres, int n_spheres, const float *boxes, int n_boxes,
                        ompl_amd_env *out);
/* Fails with OMPL_AMD_ERR_ARG (and frees nothing) while queued or running
 * server requests still use the environment. */
int ompl_amd_env_destroy(ompl_amd_env env);
/* device bytes held by the environment */
int ompl_amd_env_bytes(ompl_amd_env env, size_t *bytes);

/* ------------------------------------------------------------------ */
/* State validity -- replaces ompl::base::StateValidityChecker::isValid */
/* (base/StateValidityChecker.h:111) for the 7-DOF sphere arm; the      */
/* FP64->FP32 cast follows vamp/Utils.h:26.                             */
/* ------------------------------------------------------------------ */
#define OMPL_AMD_ARM_DOF 7
#define OMPL_AMD_ARM_SPHERES 28

/* Batched: q_soa_dev FP32 [7][n]; out_dev[i] = 1 if state i is collision free. */
int ompl_amd_validate_states_dev(ompl_amd_env env, const float *q_soa_dev, int64_t n, uint8_t *out_dev,
                                 void *stream);
/* FP64 states (OMPL's StateType::values), cast per joint to FP32 on device. */
int ompl_amd_validate_states_f64_dev(ompl_amd_env env, const double *q_soa_dev, int64_t n, uint8_t *out_dev,
                                     void *stream);
/* Plain single call: one state (7 doubles), one launch, blocking. */
int ompl_amd_is_valid(ompl_amd_env env, const double *q, int *valid);
/* Robot sphere centres for FP32 states (debug/inspection): out_dev [n][28][3]. */
int ompl_amd_arm_spheres_dev(const float *q_soa_dev, int64_t n, float *out_dev, void *stream);

/* ------------------------------------------------------------------ */
/* Motion validation -- replaces DiscreteMotionValidator::checkMotion   */
/* (base/src/DiscreteMotionValidator.cpp:93-145, and :48-91 for the     */
/* lastValid overload).  nd = factor * ceil(|b - a| / lvs) in FP64      */
/* (base/src/StateSpace.cpp:851-854); interpolation in FP64             */
/* (base/spaces/src/RealVectorStateSpace.cpp:257-265), then FP32.       */
/* ------------------------------------------------------------------ */
/* a_soa_dev, b_soa_dev: FP64 [7][n] edge endpoints (a assumed valid).
 * out_dev[i] = 1 if the motion is valid.  checked_dev (optional) receives
 * the number of states the reference checks for that edge: s2, then the
 * bisection positions up to and including the first invalid one. */
int ompl_amd_check_motions_dev(ompl_amd_env env, const double *a_soa_dev, const double *b_soa_dev, int64_t n,
                               double longest_valid_segment, unsigned factor, uint8_t *out_dev,
                               int32_t *checked_dev, void *stream);
/* ompl_amd_check_motions_dev without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * survivor workspace is a fixed device array of the library (8.1 MiB per
 * device, held until the process ends), so n is at most 65,536 edges
 * (OMPL_AMD_ERR_ARG above that).  Every call on a device shares that array:
 * calls that overlap on one device race; calls on one stream are ordered.
 * Same results as ompl_amd_check_motions_dev. */
int ompl_amd_check_motions_noalloc_dev(ompl_amd_env env, const double *a_soa_dev, const double *b_soa_dev,
                                       int64_t n, double longest_valid_segment, unsigned factor, uint8_t *out_dev,
                                       int32_t *checked_dev, void *stream);
/* lastValid variant: last_t_dev[i] = (j-1)/nd for the first invalid j of the
 * sequential scan, (nd-1)/nd when only s2 is invalid; untouched when valid. */
int ompl_amd_check_motions_last_valid_dev(ompl_amd_env env, const double *a_soa_dev, const double *b_soa_dev,
                                          int64_t n, double longest_valid_segment, unsigned factor,
                                          uint8_t *out_dev, double *last_t_dev, void *stream);
/* Plain single call (host pointers, one launch, blocking): the GPU path a
 * planner gets when every si->checkMotion(s1, s2) becomes its own launch. */
int ompl_amd_check_motion(ompl_amd_env env, const double *a, const double *b, double longest_valid_segment,
                          unsigned factor, int *valid);
int ompl_amd_check_motion_last_valid(ompl_amd_env env, const double *a, const double *b,
                                     double longest_valid_segment, unsigned factor, int *valid, double *last_t);

/* Precomputed motion states -- replaces SpaceInformation::checkMotion(
 * states, count) and checkMotion(states, count, firstInvalidStateIndex)
 * (base/src/SpaceInformation.cpp:288-341).  List i is the states
 * [offsets_dev[i], offsets_dev[i+1]) of q_soa_dev (FP64 [7][n_states]);
 * out_dev[i] = 1 when every state of the list is valid (the reference's
 * bisection visit order does not change this result);
 * first_invalid_dev[i] (optional) = index within the list of the first
 * invalid state in sequential order, -1 when the list is valid.  Offsets are
 * clamped to [0, n_states]. */
int ompl_amd_check_state_lists_dev(ompl_amd_env env, const double *q_soa_dev, int64_t n_states,
                                   const int64_t *offsets_dev, int64_t n_lists, uint8_t *out_dev,
                                   int32_t *first_invalid_dev, void *stream);
/* Plain single call: one list of `count` states (host row-major [count][7]);
 * *first_invalid is written only when the list is invalid. */
int ompl_amd_check_states(ompl_amd_env env, const double *states, int count, int *valid, int *first_invalid);

/* Captured batch (HIP graph): records the launches of one
 * ompl_amd_check_motions_dev call with these exact pointers, sizes and the
 * current launch shape, with its own survivor workspace; each launch replays
 * the whole batch on `stream` (results as the eager call).  Replays of one
 * graph must not overlap each other.  No counterpart in the reference: it
 * removes the per-call launch cost for planners that re-check a fixed batch
 * slot (e.g. a PRM expansion buffer). */
typedef struct ompl_amd_motion_graph_s *ompl_amd_motion_graph;
int ompl_amd_motion_graph_create(ompl_amd_env env, const double *a_soa_dev, const double *b_soa_dev, int64_t n,
                                 double longest_valid_segment, unsigned factor, uint8_t *out_dev,
                                 int32_t *checked_dev, ompl_amd_motion_graph *graph);
/* Captured group of independent batches: batch i (a_soa_dev[i],
 * b_soa_dev[i], n[i] edges, out_dev[i], and checked_dev[i] unless checked_dev
 * is null) is recorded on its own stream, forked from the capturing stream and
 * joined back inside the same graph, with its own survivor workspace; a
 * launch runs the n_batches (1..16) batches concurrently and ends when all
 * have ended.  Same replay rules as a single captured batch. */
int ompl_amd_motion_graph_create_group(ompl_amd_env env, int n_batches, const double *const *a_soa_dev,
                                       const double *const *b_soa_dev, const int64_t *n,
                                       double longest_valid_segment, unsigned factor, uint8_t *const *out_dev,
                                       int32_t *const *checked_dev, ompl_amd_motion_graph *graph);
int ompl_amd_motion_graph_launch(ompl_amd_motion_graph graph, void *stream);
int ompl_amd_motion_graph_destroy(ompl_amd_motion_graph graph);

/* Launch-shape knobs used by the batched kernels (set by the autotuner).
 * A state is checked by lanes_per_state x obstacle_parts lanes: the first
 * splits the arm's 28 spheres, the second the obstacle list.  Built shapes
 * (lanes, parts): (1,1) (2,1) (4,1) (8,1) (8,2) (8,4) (8,8), and (7,9), the
 * link layout: one state per wave, each of 7 lanes holds one link's four
 * spheres and tests the link's bounding sphere against an obstacle before
 * the exact sphere tests (same results; environments whose staging would
 * exceed 64 KiB of LDS with the extra sphere radii run as (8,8)); block in
 * {64,128,256,512,1024}. */
typedef struct {
    int validity_lanes_per_state; /* validity and s2 (endpoint) checks */
    int validity_obstacle_parts;
    int validity_block;
    int motion_lanes_per_state;   /* interior (bisection) checks */
    int motion_obstacle_parts;
    int motion_block;
    int motion_waves_per_cu;      /* persistent kernels' occupancy */
    int motion_strategy;          /* 0: one wave walks each edge from s2;
                                     1: s2 pass, then a persistent queue walks
                                        each survivor with one wave;
                                     2: fused: each wave checks s2 of its edges
                                        and walks its own survivors (1 launch);
                                     3: s2 pass, then the survivors' bisection
                                        chunks run in parallel across the GPU,
                                        chunk-major, with early exit;
                                     4: as 3, the walk with each position's
                                        FK computed once per wave and shared
                                        across the obstacle partitions (shapes
                                        with parts > 1; chunks of 64/lanes
                                        positions) */
    int validity_shared_fk;       /* 1: batched validity and the s2 pass of
                                     the motion strategies compute each
                                     state's FK once per wave and share it
                                     across the obstacle partitions */
} ompl_amd_tuning;
int ompl_amd_get_tuning(ompl_amd_tuning *t);
int ompl_amd_set_tuning(const ompl_amd_tuning *t);
/* Autotuner: times every candidate launch shape (strategy x lanes x block x
 * occupancy) on the given device-resident batch with HIP events, installs the
 * fastest as the current tuning and returns it with its time per batch.  No
 * counterpart in the reference. */
int ompl_amd_autotune_motions(ompl_amd_env env, const double *a_soa_dev, const double *b_soa_dev, int64_t n,
                              double longest_valid_segment, unsigned factor, int reps, ompl_amd_tuning *best,
                              double *best_us);
/* Autotuner for batches issued round-robin on several streams (1 to 16 HIP
 * streams whose batches overlap, e.g. one per planner thread): the shape that
 * is fastest for one batch at a time keeps more persistent walk waves than is
 * best when the streams' batches share the device.  Each candidate (s2 pass
 * shapes (7,1) (7,2) (7,4) (7,9) (8,4) (8,8) x walk shapes (7,2) (7,4) (7,9)
 * (8,8) x 2/4/8/16 walk waves per CU, the chunk-parallel walk) runs `steps`
 * batches round-robin on the streams, timed by HIP events on streams[0]; the
 * three fastest run again over 2 x steps and the fastest is installed as the
 * current tuning and returned with its time per batch.  Fields outside the
 * grid keep the current tuning's values.  Refused for an environment with its
 * own shapes (ompl_amd_env_set_tuning), which would ignore the candidates.
 * No counterpart in the reference. */
int ompl_amd_autotune_motions_streams(ompl_amd_env env, const double *a_soa_dev, const double *b_soa_dev, int64_t n,
                                      double longest_valid_segment, unsigned factor, void *const *streams,
                                      int n_streams, int steps, ompl_amd_tuning *best, double *best_us);
/* Launch shapes of one environment: the batched checks on `env` (validity,
 * motions, last-valid motions, captured motion graphs at capture, the
 * no-allocation motion check) use `t` instead of the library's current tuning;
 * t == NULL makes the environment follow the library's tuning again.  The
 * fastest shape depends on the obstacle count (a dense scene wants many
 * obstacle partitions per state, a sparse one a single partition and more
 * waves per CU).  ompl_amd_env_get_tuning returns the shapes in force for
 * `env` and sets *own to 1 when they are the environment's own.  No
 * counterpart in the reference. */
int ompl_amd_env_set_tuning(ompl_amd_env env, const ompl_amd_tuning *t);
int ompl_amd_env_get_tuning(ompl_amd_env env, ompl_amd_tuning *t, int *own);

/* Work a scalar checker performs on the same inputs in the reference order
 * (s2, then bisection positions up to the first invalid state; per state the
 * robot spheres in order against sphere then box obstacles, stopping at the
 * first hit): out[0] states, out[1] sphere-pair tests, out[2] box-pair tests.
 * Used for roofline accounting; blocking; no counterpart in the reference. */
int ompl_amd_motion_work(ompl_amd_env env, const double *a_soa_dev, const double *b_soa_dev, int64_t n,
                         double longest_valid_segment, unsigned factor, int64_t out[3]);
int ompl_amd_validity_work(ompl_amd_env env, const float *q_soa_dev, int64_t n, int64_t out[3]);

/* Per-kernel timing with HIP events recorded around each launch on its own
 * stream (off by default).  Names: "validate_states", "motion_endpoint",
 * "motion_walk", "nn_knn", "nn_radius", "sample_uniform", ... */
int ompl_amd_profile_enable(int on);
int ompl_amd_profile_reset(void);
int ompl_amd_profile_query(const char *kernel, int64_t *launches, double *total_ms);

/* ------------------------------------------------------------------ */
/* State sampling -- RealVectorStateSampler::sampleUniform              */
/* (base/spaces/src/RealVectorStateSpace.cpp:45-52) over RNG::uniformReal */
/* (util/RandomNumbers.h:72-76): std::mt19937 + generate_canonical.     */
/* ------------------------------------------------------------------ */
/* n_streams independent mt19937 generators, stream s seeded with seeds[s];
 * stream s writes n_per_stream states of `dim` doubles, row-major, to
 * out_dev + s*n_per_stream*dim, exactly as n_per_stream successive
 * sampleUniform calls of an RNG with that local seed.  Also writes the FP32
 * dof-major copy to out_f32_soa_dev ([dim][n_streams*n_per_stream]) when
 * non-null. */
int ompl_amd_sample_uniform_dev(const uint32_t *seeds, int n_streams, int64_t n_per_stream, int dim,
                                const double *lo, const double *hi, double *out_dev, float *out_f32_soa_dev,
                                void *stream);

/* Valid-state sampling -- UniformValidStateSampler::sample
 * (base/samplers/src/UniformValidStateSampler.cpp:46-57) batched over
 * samplers.  Sampler s (generator seeded seeds[s], as above, bounds lo/hi)
 * makes n_calls successive sample() calls; a call draws uniform states until
 * one is valid for `env` or `attempts` draws were made (the reference
 * default is magic::MAX_VALID_SAMPLE_ATTEMPTS = 100,
 * tools/config/MagicConstants.h:90).  out_dev [n_streams][n_calls][7] FP64
 * receives each call's final state (the last draw when none was valid),
 * found_dev [n_streams][n_calls] = 1 when it is valid, draws_dev (optional)
 * the number of draws the call made.  Blocking. */
int ompl_amd_sample_valid_dev(ompl_amd_env env, const uint32_t *seeds, int n_streams, int64_t n_calls, int attempts,
                              const double *lo, const double *hi, double *out_dev, uint8_t *found_dev,
                              int32_t *draws_dev, void *stream);

/* Generators that persist between calls (the RNG of a long-lived sampler):
 * stream s starts as mt19937 seeded seeds[s]; every call continues where the
 * previous one stopped, exactly as successive calls on one reference
 * sampler.  rng_sample_valid leaves each generator just after the last draw
 * its calls consumed. */
typedef struct ompl_amd_rng_s *ompl_amd_rng;
int ompl_amd_rng_create(const uint32_t *seeds, int n_streams, ompl_amd_rng *out);
int ompl_amd_rng_destroy(ompl_amd_rng rng);
/* out_dev [n_streams][n_per_stream][dim]; out_f32_soa_dev (optional) [dim][n_streams*n_per_stream]. */
int ompl_amd_rng_sample_uniform_dev(ompl_amd_rng rng, int64_t n_per_stream, int dim, const double *lo,
                                    const double *hi, double *out_dev, float *out_f32_soa_dev, void *stream);
/* SE3StateSampler::sampleUniform (R^3 in [lo, hi] then RNG::quaternion,
 * util/src/RandomNumbers.cpp:266-279): rows (x, y, z, qx, qy, qz, qw),
 * out_dev FP64 [n_streams * n_per_stream][7], out_f32_dev (optional) FP32,
 * the nearest-neighbour point layout.  sin/cos are the device's FP64 ones. */
int ompl_amd_rng_sample_se3_dev(ompl_amd_rng rng, int64_t n_per_stream, const double *lo, const double *hi,
                                double *out_dev, float *out_f32_dev, void *stream);
int ompl_amd_rng_sample_valid_dev(ompl_amd_rng rng, ompl_amd_env env, int64_t n_calls, int attempts,
                                  const double *lo, const double *hi, double *out_dev, uint8_t *found_dev,
                                  int32_t *draws_dev, void *stream);

/* ------------------------------------------------------------------ */
/* Nearest neighbours -- ompl::NearestNeighbors<_T>                      */
/* (datastructures/NearestNeighbors.h:46-115), exact results in FP32,   */
/* ordered by (distance, index), radius inclusive (<=) as               */
/* NearestNeighborsLinear.h:135-142.                                    */
/* ------------------------------------------------------------------ */
#define OMPL_AMD_METRIC_L2 0  /* RealVectorStateSpace::distance */
#define OMPL_AMD_METRIC_SE3 1 /* (x,y,z,qx,qy,qz,qw): R3 + SO3 arc, weights 1,1 */

typedef struct ompl_amd_nn_s *ompl_amd_nn;

int ompl_amd_nn_create(int metric, int dim, ompl_amd_nn *out);
int ompl_amd_nn_destroy(ompl_amd_nn nn);
int ompl_amd_nn_clear(ompl_amd_nn nn);
/* Append n points (host row-major [n][dim]); ids are assigned 0,1,2,... in
 * insertion order (first id returned in *first_id). */
int ompl_amd_nn_add(ompl_amd_nn nn, const float *pts, int64_t n, int64_t *first_id);
int ompl_amd_nn_add_dev(ompl_amd_nn nn, const float *pts_dev, int64_t n, int64_t *first_id, void *stream);
/* Remove by id (lazy: the id is excluded from every later query). */
int ompl_amd_nn_remove(ompl_amd_nn nn, int64_t id, int *removed);
int ompl_amd_nn_size(ompl_amd_nn nn, int64_t *size);
/* Rebuild the acceleration structure after adds (called implicitly by queries). */
int ompl_amd_nn_build(ompl_amd_nn nn, void *stream);
/* k nearest of nq queries (device row-major [nq][dim]); out_idx_dev [nq][k]
 * (-1 past the end), out_dist_dev [nq][k] (+inf past the end). */
int ompl_amd_nn_knn_dev(ompl_amd_nn nn, const float *q_dev, int64_t nq, int k, int32_t *out_idx_dev,
                        float *out_dist_dev, void *stream);
/* Radius search, CSR results sorted by (distance, id) per query:
 * out_offsets_dev [nq+1] and *total are always written.  With null outputs the
 * call only counts.  With outputs of `capacity` entries the call counts and
 * writes in one pass over the points (a second pass when some query has more
 * than 64 results); if capacity < total it fails with OMPL_AMD_ERR_ARG after
 * writing the offsets and *total, and the outputs are untouched. */
int ompl_amd_nn_radius_dev(ompl_amd_nn nn, const float *q_dev, int64_t nq, float radius,
                           int64_t *out_offsets_dev, int32_t *out_idx_dev, float *out_dist_dev, int64_t capacity,
                           int64_t *total, void *stream);
/* Packed batches: the queries of several requests in one launch, each query
 * with its own k or radius -- NearestNeighbors::nearestK / nearestR
 * (datastructures/NearestNeighbors.h:97, 104) for many callers at once.
 * k-NN: k_dev [nq] (clamped to [0, kmax], kmax in [1, 64]); results are CSR:
 * out_offsets_dev [nq+1] is written (exclusive scan of the clamped k) and
 * out_idx_dev / out_dist_dev must hold nq * kmax entries (or the sum of k).
 * Radius: r_dev [nq] (a negative or NaN radius gives no results); otherwise
 * as ompl_amd_nn_radius_dev. */
int ompl_amd_nn_knn_packed_dev(ompl_amd_nn nn, const float *q_dev, int64_t nq, const int32_t *k_dev, int kmax,
                               int64_t *out_offsets_dev, int32_t *out_idx_dev, float *out_dist_dev, void *stream);
int ompl_amd_nn_radius_packed_dev(ompl_amd_nn nn, const float *q_dev, int64_t nq, const float *r_dev,
                                  int64_t *out_offsets_dev, int32_t *out_idx_dev, float *out_dist_dev,
                                  int64_t capacity, int64_t *total, void *stream);
/* Host convenience (blocking). */
int ompl_amd_nn_knn(ompl_amd_nn nn, const float *q, int64_t nq, int k, int32_t *out_idx, float *out_dist);
int ompl_amd_nn_knn_packed(ompl_amd_nn nn, const float *q, int64_t nq, const int32_t *k, int kmax,
                           int64_t *out_offsets, int32_t *out_idx, float *out_dist);
int ompl_amd_nn_radius_packed(ompl_amd_nn nn, const float *q, int64_t nq, const float *r, int64_t *out_offsets,
                              int32_t *out_idx, float *out_dist, int64_t capacity, int64_t *total);
int ompl_amd_nn_radius(ompl_amd_nn nn, const float *q, int64_t nq, float radius, int64_t *out_offsets,
                       int32_t *out_idx, float *out_dist, int64_t capacity, int64_t *total);
int ompl_amd_nn_nearest(ompl_amd_nn nn, const float *q, int64_t *out_id);

/* RRT* neighbourhood radius (geometric/planners/rrt/src/RRTstar.cpp:602-617,
 * 1147-1160): min(max_distance, r_rrt * (log(n+1)/(n+1))^(1/d)),
 * r_rrt = rewire * (2 (1 + 1/d) measure / unitNBallMeasure(d))^(1/d). */
double ompl_amd_rrtstar_radius(int dim, double measure, double rewire_factor, double max_distance, int64_t n);

/* ------------------------------------------------------------------ */
/* Planning server (continuous batching) -- RRT-Connect requests        */
/* (geometric/planners/rrt/src/RRTConnect.cpp:116-196, 245-330) on the  */
/* 7-DOF arm.  Each request's whole solve loop (sampler, trees, connect */
/* loop, path) is resident in a region of a device pool of              */
/* hbm_budget_bytes; the server runs every active request for a time    */
/* slice per launch (one workgroup per request) and admits queued       */
/* requests between slices while their region fits the pool.  A        */
/* request's result does not depend on the slices it ran in or the      */
/* requests it shared them with.  No counterpart in the reference (it   */
/* plans one query per thread).                                         */
/* ------------------------------------------------------------------ */
#define OMPL_AMD_PLAN_QUEUED 0
#define OMPL_AMD_PLAN_RUNNING 1
#define OMPL_AMD_PLAN_SOLVED 2
#define OMPL_AMD_PLAN_FAILED 3    /* max_iters reached without a solution */
#define OMPL_AMD_PLAN_DEADLINE 4  /* deadline_ms expired (queued or running) */
#define OMPL_AMD_PLAN_CANCELLED 5
#define OMPL_AMD_PLAN_REJECTED 6  /* footprint larger than the whole budget */
#define OMPL_AMD_PLAN_TREE_FULL 7 /* a tree reached max_nodes */
#define OMPL_AMD_PLAN_INVALID_START 8 /* no start state in bounds and valid */
#define OMPL_AMD_PLAN_INVALID_GOAL 9  /* no goal state in bounds and valid */

typedef struct ompl_amd_server_s *ompl_amd_server;

typedef struct {
    size_t hbm_budget_bytes; /* device pool holding the active requests' regions (allocated at
                                creation); a request is admitted when its region fits */
    int max_active;          /* concurrent requests */
    int max_nodes;           /* largest nodes per tree of a request */
    int check_obstacle_parts; /* lanes splitting the environment's obstacle list in a motion
                                 check (link layout): 1, 2, 4 or 9 (8, 4, 2 or 1 positions
                                 per wave); 0 = chosen per motion from its length */
    int check_lanes_per_state; /* 0 or 7 (the link layout with bounding-sphere culling) */
    int slice_us;            /* device time per slice before the host admits and retires
                                (at least one growTree step per request); 0 = 200 */
    int block_threads;       /* threads of each request's workgroup: 512 or 1024 (0 = 512) */
} ompl_amd_server_config;

typedef struct {
    ompl_amd_env env;
    double start[OMPL_AMD_ARM_DOF];
    double goal[OMPL_AMD_ARM_DOF];
    uint32_t seed;                /* local seed of the request's state sampler RNG */
    int max_iters;                /* RRT-Connect iterations (the termination condition) */
    double range;                 /* maxDistance_; <= 0: 0.2 * maximum extent (SelfConfig) */
    double longest_valid_segment; /* DiscreteMotionValidator resolution */
    unsigned factor;
    double deadline_ms;           /* <= 0: none */
    int max_nodes;                /* <= 0: the server's max_nodes; region size from
                                     ompl_amd_server_footprint */
    int planner;                  /* OMPL_AMD_PLANNER_RRT_CONNECT or OMPL_AMD_PLANNER_PRM */
    int prm_samples;              /* PRM: uniform samples of the roadmap (<= 0: 4,096) */
    int prm_k;                    /* PRM: neighbours per milestone (<= 0: 10, PRM's default) */
    /* RRT-Connect: several start / goal states (host rows of 7, copied at
     * submit, at most OMPL_AMD_MAX_ENDPOINTS each); n_starts > 0 replaces
     * `start`, n_goals > 0 replaces `goal`.  Start states out of bounds
     * ([-pi, pi] within machine epsilon) or invalid are skipped
     * (PlannerInputStates::nextStart, base/src/Planner.cpp:228-256); goal
     * states enter the goal tree in order as GoalStates hands them out
     * (RRTConnect.cpp:251-269, Planner.cpp:265-336), checked the same way. */
    const double *starts;
    int n_starts;
    const double *goals;
    int n_goals;
    int intermediate_states;      /* addIntermediateStates_ (RRTConnect.cpp:148-183) */
} ompl_amd_plan_request;

#define OMPL_AMD_MAX_ENDPOINTS 8

/* Planner of a request.  PRM (geometric/planners/prm/src/PRM.cpp:454-536,
 * 570-604, 696-727) is served as one batch roadmap per request: start and
 * goal become milestones 0 and 1 of a roadmap over prm_samples uniform
 * samples (64 generators seeded seed, seed + 1, ...; KStrategy with prm_k
 * neighbours, every milestone connected to its k nearest among all), and the
 * path is the shortest one between them over the valid edges (FP64 state
 * distances).  FAILED when start or goal is invalid or they are in different
 * components.  iterations = 1; tree_nodes = milestones; motion_checks =
 * candidate edges.  PRM requests run on the server's second stream beside
 * the RRT-Connect slices and reserve their roadmap's footprint in the pool. */
#define OMPL_AMD_PLANNER_RRT_CONNECT 0
#define OMPL_AMD_PLANNER_PRM 1

typedef struct {
    int status;
    int iterations;
    int path_len;       /* states in the solution path */
    int64_t tree_nodes; /* nodes in both trees */
    double queued_ms;   /* submit -> admission */
    double run_ms;      /* admission -> finish */
    double latency_ms;  /* submit -> finish */
    int64_t motion_checks; /* growTree steps that checked a motion (isValid + checkMotion
                              of a goal-tree step count once) */
} ompl_amd_plan_result;

int ompl_amd_server_create(const ompl_amd_server_config *cfg, ompl_amd_server *out);
int ompl_amd_server_destroy(ompl_amd_server server);
int ompl_amd_server_submit(ompl_amd_server server, const ompl_amd_plan_request *req, int64_t *request_id);
int ompl_amd_server_poll(ompl_amd_server server, int64_t request_id, ompl_amd_plan_result *res);
/* Blocks until the request finishes or timeout_ms passes (< 0: no timeout). */
int ompl_amd_server_wait(ompl_amd_server server, int64_t request_id, double timeout_ms, ompl_amd_plan_result *res);
/* Solution path, row-major states (up to cap states); *n = full length. */
int ompl_amd_server_path(ompl_amd_server server, int64_t request_id, double *out, int cap, int *n);
int ompl_amd_server_cancel(ompl_amd_server server, int64_t request_id);
/* Frees a finished request's record. */
int ompl_amd_server_release(ompl_amd_server server, int64_t request_id);
/* out: slices, growTree steps, active, queued, used bytes, budget bytes */
int ompl_amd_server_stats(ompl_amd_server server, int64_t out[6]);
/* Device time of all growTree steps run so far, summed over requests, in
 * microseconds: out[0] sampling, [1] nearest-neighbour scans, [2] motion
 * checks (steering included), [3] whole steps. */
int ompl_amd_server_phase_us(ompl_amd_server server, double out[4]);
/* Bytes of one request's region (trees, path, solve-loop state) for
 * max_nodes nodes per tree (0 if max_nodes < 2). */
size_t ompl_amd_server_footprint(int max_nodes);
/* The same RRT-Connect on the host: host nearest neighbours, one GPU launch
 * and synchronisation per isValid / checkMotion call (the plain path). */
int ompl_amd_plan_plain(const ompl_amd_plan_request *req, ompl_amd_plan_result *res, double *path, int cap,
                        int *path_len);

/* ------------------------------------------------------------------ */
/* Batched PRM roadmap -- the inner loop of PRM::growRoadmap /          */
/* addMilestone (geometric/planners/prm/src/PRM.cpp:364-385, 570-604)   */
/* with KStrategy (geometric/planners/prm/ConnectionStrategy.h:56-81)   */
/* for a whole batch: samples -> validity -> k nearest milestones ->    */
/* undirected candidate edges -> motion checks (lower index -> higher)  */
/* -> connected components.  Batch semantics: each milestone connects   */
/* to its k nearest among all milestones (the reference: among those    */
/* added before it).                                                    */
/* ------------------------------------------------------------------ */
typedef struct ompl_amd_prm_s *ompl_amd_prm;

typedef struct {
    int64_t samples;
    int64_t milestones;
    int64_t candidate_edges;
    int64_t valid_edges;
    int64_t components;
    int64_t k; /* neighbours per milestone used */
    double ms_sample, ms_validity, ms_knn, ms_candidates, ms_motions, ms_graph, ms_total;
} ompl_amd_prm_stats;

/* flags of ompl_amd_prm_build */
#define OMPL_AMD_PRM_AUTOTUNE 1u /* run the motion autotuner on (up to 32,768 of) the candidate edges first and
                                    keep the winner as the environment's own launch shapes
                                    (ompl_amd_env_set_tuning); the library's tuning is left as it was */
#define OMPL_AMD_PRM_KSTAR 2u    /* KStarStrategy (geometric/planners/prm/ConnectionStrategy.h:125-157, PRM*):
                                    k = ceil((e + e/7) * ln(milestones)), capped at 63; the k argument is
                                    ignored */

/* Samples U[-pi,pi]^7 from n_streams mt19937 generators seeded seed,
 * seed+1, ... (stream-major sample order; n_samples is rounded up to a
 * multiple of n_streams), checks them with the arm checker of `env`, and
 * builds the roadmap with k neighbours per milestone.  Milestone i is the
 * i-th valid sample.  Requires 2 * (n_samples + n_streams) * k < 2^31. */
int ompl_amd_prm_build(ompl_amd_env env, uint32_t seed, int n_streams, int64_t n_samples, int k,
                       double longest_valid_segment, unsigned factor, unsigned flags, ompl_amd_prm *out,
                       ompl_amd_prm_stats *stats);
int ompl_amd_prm_destroy(ompl_amd_prm prm);
/* Milestones (row-major FP64) and their sample indices. */
int ompl_amd_prm_milestones(ompl_amd_prm prm, double *nodes, int32_t *sample_index, int64_t cap);
/* Candidate edges (milestone indices u < v) and their motion validity. */
int ompl_amd_prm_edges(ompl_amd_prm prm, int32_t *u, int32_t *v, uint8_t *valid, int64_t cap);
/* Component label (smallest milestone index of the component) per milestone. */
int ompl_amd_prm_components(ompl_amd_prm prm, int32_t *label, int64_t cap);
/* Roadmap graph (replaces the Boost.Graph adjacency list of PRM.h:113-121, 382):
 * symmetric CSR of the valid edges, offsets [milestones + 1], neighbours
 * ascending per row; adjacency needs 2 * valid_edges entries (may be NULL
 * to read the offsets only). */
int ompl_amd_prm_graph(ompl_amd_prm prm, int64_t *offsets, int32_t *adjacency, int64_t adjacency_cap);

/* ------------------------------------------------------------------ */
/* Synthetic workloads (std::mt19937 + OMPL RNG semantics, host side)   */
/* ------------------------------------------------------------------ */
/* Obstacles: centre U[-1,1]^3 then sphere radius U[r_lo, r_hi] (or box
 * half-extents U[h_lo, h_hi] per axis); a draw is rejected while its centre is
 * within min_base_dist of the origin or it overlaps the arm's fixed base link
 * (link 0's spheres), so the base region stays free. */
int ompl_amd_workload_obstacles(uint32_t seed, int n_spheres, int n_boxes, double r_lo, double r_hi, double h_lo,
                                double h_hi, double min_base_dist, float *spheres_out, float *boxes_out);
/* n joint configs U[-pi,pi]^7, FP64 row-major [n][7] (sampleUniform order). */
int ompl_amd_workload_configs(uint32_t seed, int64_t n, double *out);
/* n edges: a U[-pi,pi]^7, direction uniform on S^6, length exp(N(mu, sigma))
 * clipped to [len_lo, len_hi], b = clamp(a + len*dir, -pi, pi); row-major. */
int ompl_amd_workload_edges(uint32_t seed, int64_t n, double mu, double sigma, double len_lo, double len_hi,
                            double *a_out, double *b_out);
/* n SE3 states: translation U[lo,hi]^3 then RNG::quaternion (x,y,z,w). */
int ompl_amd_workload_se3(uint32_t seed, int64_t n, double lo, double hi, float *out);

#ifdef __cplusplus
}
#endif
#endif
