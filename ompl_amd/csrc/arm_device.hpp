// Device-side arm model: FP32 forward kinematics of the 7-DOF sphere arm and
// the sphere/sphere, sphere/box collision predicates, with the obstacle set
// staged in LDS.  Operation order (fixed; -ffp-contract=off):
//
//   joint i (axis z for even i, y for odd i), (s, c) = sincos_pl(q_i):
//     z joint: X' = X*c + Y*s,  Y' = Y*c - X*s   (per component)
//     y joint: X' = X*c - Z*s,  Z' = X*s + Z*c
//   sphere j of link i: centre = p_i + Z_i * off_j, off = {0.0375, 0.1125,
//     0.1875, 0.2625} (link 0.3 m, fractions 1/8, 3/8, 5/8, 7/8), r = 0.05
//   p_{i+1} = p_i + Z_i * 0.3
//   sphere obstacle (o, R): d = o - centre, d2 = (dx*dx + dy*dy) + dz*dz,
//     collision iff d2 < rr, rr = (R + 0.05f) * (R + 0.05f)
//   box obstacle (lo = b - h, hi = b + h): p = min(max(centre, lo), hi),
//     d = centre - p, d2 = (dx*dx + dy*dy) + dz*dz, collision iff d2 < 0.05f*0.05f
//
// The reference's own FP32 checker (VAMP, vamp/VampStateValidityChecker.h:37-41)
// is an absent submodule; this arm and these predicates are the project's
// specification of the checker behind StateValidityChecker::isValid.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fp32_math.hpp"

namespace ompl_amd {

constexpr int kDof = 7;
constexpr int kSpheres = 28;
constexpr float kLink = 0.3f;
constexpr float kRobotR = 0.05f;
constexpr float kFar = 1e30f; // centre of an unused sphere slot: never collides

__device__ __forceinline__ float sphere_offset(int j)
{
    return j == 0 ? 0.0375f : (j == 1 ? 0.1125f : (j == 2 ? 0.1875f : 0.2625f));
}

// Obstacles in LDS: [0, ns) spheres (x, y, z, rr); then nb boxes as pairs
// (lo.xyz, 0), (hi.xyz, 0).
struct EnvView {
    const float4 *obs; // device (HBM) copy in the same layout
    int ns;
    int nb;
};

__device__ __forceinline__ int env_words(const EnvView &e)
{
    return e.ns + 2 * e.nb;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  This permutation gives XCD x a contiguous run of logical blocks
// (hardware block b = 8 i + x -> the i-th block of XCD x's run; the first
// nb % 8 XCDs get one block more), so that neighbouring states, which share
// cache lines of the dof-major arrays, are fetched into one L2 only.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb)
{
    const int64_t q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// Link layout (lanes_per_state 7): bounding sphere of one link's four robot
// spheres, centred at p + Z*0.15 (the middle of the sphere offsets), radius
// kLinkReach + r with kLinkReach = 0.2625 - 0.15 the largest offset from that
// centre.  A sphere obstacle (o, rr = (R + r)^2) can touch a sphere of the
// link only if |o - centre| < sqrt(rr) + kLinkReach; the staged bound widens
// that by kBoundAbs + kBoundRel * sqrt(rr) (and a relative 1e-6 on the
// square), far above the FP32 rounding of the centres and distances (a few
// 1e-7 m for a 2.1 m arm), so a pair the bound rejects is never one the exact
// test accepts: culling changes no result.
constexpr float kLinkCentre = 0.15f;
constexpr float kLinkReach = 0.1125f;
constexpr float kBoundAbs = 1e-4f;
constexpr float kBoundRel = 1e-5f;
// box: distance from the centre to the box below kLinkReach + r + kBoundAbs
constexpr float kBoxBound = kLinkReach + kRobotR + kBoundAbs;

__device__ __forceinline__ float sphere_bound2(float rr)
{
    const float b = sqrtf(rr) * (1.0f + kBoundRel) + (kLinkReach + kBoundAbs);
    return b * b * 1.000001f;
}

// Obstacles per lane per pass of the link layout's bound tests: 16 with 9
// obstacle partitions (one pass covers 144 obstacles of a type: a state of
// the 128 + 128 benchmark environment takes one sphere and one box pass),
// else 8 (shorter passes let colliding states leave sooner; measured: the
// (7, 4) s2 pass 18.4 us with 8, 18.9 with 16; the (7, 9) walk 19.1 us with
// 8, 18.0 with 16; with obstacles nearest the base first, 4 per pass gave the
// (7, 4) s2 pass 16.5 us against 16.3 with 8).
__host__ __device__ constexpr int link_pass(int ops)
{
    return ops >= 9 ? 16 : 8;
}

__host__ __device__ constexpr int pad_to(int n, int m)
{
    return (n + m - 1) / m * m;
}

// Cooperative copy of the environment into LDS by the whole block.  With
// LINKS (link layout) the spheres are staged as (x, y, z, bound^2), the sphere
// and box lists are each padded to a multiple of PAD entries (spheres (0, 0,
// 0, -1), boxes at 1e30: their bound tests always fail), and the spheres'
// exact rr follow the boxes as a float array; layout [pad(ns) spheres]
// [2 pad(nb) box words][ns rr].
template <bool LINKS = false, int PAD = 1>
__device__ __forceinline__ void stage_env(const EnvView &e, float4 *lds)
{
    if constexpr (!LINKS) {
        const int nw = env_words(e);
        for (int i = threadIdx.x; i < nw; i += blockDim.x) lds[i] = e.obs[i];
    } else {
        const int nsp = pad_to(e.ns, PAD), nbp = pad_to(e.nb, PAD), total = nsp + 2 * nbp;
        float *rr = reinterpret_cast<float *>(lds + nsp + 2 * nbp);
        // source word of staging word i in e.obs, or -1 for padding
        auto src = [&](int i) -> int {
            if (i < nsp) return i < e.ns ? i : -1;
            return i - nsp < 2 * e.nb ? e.ns + (i - nsp) : -1;
        };
        auto put = [&](int i, int w, float4 o) {
            if (i < nsp) {
                if (w >= 0) {
                    rr[i] = o.w;
                    o.w = sphere_bound2(o.w);
                } else {
                    o = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
                }
            } else if (w < 0) {
                o = make_float4(1e30f, 1e30f, 1e30f, 0.0f);
            }
            lds[i] = o;
        };
        // two words per thread and step, both read before either is written
        // (the reads are unconditional -- word 0 stands in for padding; the
        // array always holds at least one word -- so they share one round trip)
        for (int i0 = threadIdx.x; i0 < total; i0 += 2 * blockDim.x) {
            const int i1 = i0 + blockDim.x;
            const int w0 = src(i0), w1 = i1 < total ? src(i1) : -1;
            const float4 o0 = e.obs[w0 < 0 ? 0 : w0], o1 = e.obs[w1 < 0 ? 0 : w1];
            put(i0, w0, o0);
            if (i1 < total) put(i1, w1, o1);
        }
    }
    __syncthreads();
}

// Sphere centres held by one lane: the lane is member g of a group of LPS
// lanes that together hold the 28 spheres of one state; slot m holds sphere
// m*LPS + g (link (m*LPS+g)/4, sphere (m*LPS+g)%4).
template <int LPS>
struct LaneSpheres {
    static constexpr int N = (kSpheres + LPS - 1) / LPS;
    float x[N], y[N], z[N];
};

// Link layout: the lane holds the four spheres of one link (slot m = sphere
// 4g + m of link g) and the centre of their bounding sphere.
template <>
struct LaneSpheres<7> {
    static constexpr int N = 4;
    float x[N], y[N], z[N];
    float cx, cy, cz;
};

// FK of the link layout (GS lanes per state, GS >= 8 a power of two): lane
// i < 7 of the state's group evaluates sincos of joint i once and the group
// reads the seven results (readlane when the group is the whole wave); the
// rotation chain is the one arm_fk_slots evaluates, and the lane keeps the
// frame (p, Z) of its link g, so its sphere centres are the same bits as
// arm_fk_slots's.
template <int GS>
__device__ __forceinline__ void arm_fk_links(const float q[kDof], int g, LaneSpheres<7> &S)
{
    const int j = threadIdx.x & 7;
    float qj = q[0];
#pragma unroll
    for (int i = 1; i < kDof; ++i) {
        qj = (j == i) ? q[i] : qj;
        asm volatile("" : "+v"(qj)); // keep the selects (not an indexed private-memory load)
    }
    float sm, cm;
    sincos_pl(qj, sm, cm);
    float sj[kDof], cj[kDof];
#pragma unroll
    for (int i = 0; i < kDof; ++i) {
        if constexpr (GS == 64) {
            sj[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sm), i));
            cj[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cm), i));
        } else {
            const int src = (threadIdx.x & 63 & ~(GS - 1)) + i;
            sj[i] = __shfl(sm, src);
            cj[i] = __shfl(cm, src);
        }
    }
    float X0 = 1.0f, X1 = 0.0f, X2 = 0.0f;
    float Y0 = 0.0f, Y1 = 1.0f, Y2 = 0.0f;
    float Z0 = 0.0f, Z1 = 0.0f, Z2 = 1.0f;
    float p0 = 0.0f, p1 = 0.0f, p2 = 0.0f;
    float P0 = 0.0f, P1 = 0.0f, P2 = 0.0f, W0 = 0.0f, W1 = 0.0f, W2 = 1.0f;
#pragma unroll
    for (int i = 0; i < kDof; ++i) {
        const float s = sj[i], c = cj[i];
        if ((i & 1) == 0) {
            float nx0 = X0 * c + Y0 * s, ny0 = Y0 * c - X0 * s;
            float nx1 = X1 * c + Y1 * s, ny1 = Y1 * c - X1 * s;
            float nx2 = X2 * c + Y2 * s, ny2 = Y2 * c - X2 * s;
            X0 = nx0; X1 = nx1; X2 = nx2;
            Y0 = ny0; Y1 = ny1; Y2 = ny2;
        } else {
            float nx0 = X0 * c - Z0 * s, nz0 = X0 * s + Z0 * c;
            float nx1 = X1 * c - Z1 * s, nz1 = X1 * s + Z1 * c;
            float nx2 = X2 * c - Z2 * s, nz2 = X2 * s + Z2 * c;
            X0 = nx0; X1 = nx1; X2 = nx2;
            Z0 = nz0; Z1 = nz1; Z2 = nz2;
        }
        const bool mine = g == i;
        P0 = mine ? p0 : P0;
        P1 = mine ? p1 : P1;
        P2 = mine ? p2 : P2;
        W0 = mine ? Z0 : W0;
        W1 = mine ? Z1 : W1;
        W2 = mine ? Z2 : W2;
        p0 = p0 + Z0 * kLink;
        p1 = p1 + Z1 * kLink;
        p2 = p2 + Z2 * kLink;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const float off = sphere_offset(m);
        S.x[m] = P0 + W0 * off;
        S.y[m] = P1 + W1 * off;
        S.z[m] = P2 + W2 * off;
    }
    S.cx = P0 + W0 * kLinkCentre;
    S.cy = P1 + W1 * kLinkCentre;
    S.cz = P2 + W2 * kLinkCentre;
}

// With LPS = 8 the eight lanes of an aligned lane octet always hold the same
// state (every layout puts a state's sphere slots on lanes 8k .. 8k+7), so
// lane 8k + j evaluates sincos of joint j once and the octet shares the seven
// results by shuffles instead of each lane evaluating all seven; the values
// are the same bits either way.
template <int LPS>
__device__ __forceinline__ void arm_fk_slots(const float q[kDof], int g, LaneSpheres<LPS> &S)
{
    float sj[kDof], cj[kDof];
    if constexpr (LPS == 8) {
        const int lane = threadIdx.x & 63;
        const int j = lane & 7;
        float qj = q[0];
#pragma unroll
        for (int i = 1; i < kDof; ++i) qj = (j == i) ? q[i] : qj;
        float sm, cm;
        sincos_pl(qj, sm, cm);
        const int base = lane & ~7;
#pragma unroll
        for (int i = 0; i < kDof; ++i) {
            sj[i] = __shfl(sm, base + i);
            cj[i] = __shfl(cm, base + i);
        }
    } else {
#pragma unroll
        for (int i = 0; i < kDof; ++i) sincos_pl(q[i], sj[i], cj[i]);
    }
    float X0 = 1.0f, X1 = 0.0f, X2 = 0.0f;
    float Y0 = 0.0f, Y1 = 1.0f, Y2 = 0.0f;
    float Z0 = 0.0f, Z1 = 0.0f, Z2 = 1.0f;
    float p0 = 0.0f, p1 = 0.0f, p2 = 0.0f;
#pragma unroll
    for (int m = 0; m < LaneSpheres<LPS>::N; ++m) {
        S.x[m] = kFar;
        S.y[m] = kFar;
        S.z[m] = kFar;
    }
#pragma unroll
    for (int i = 0; i < kDof; ++i) {
        const float s = sj[i], c = cj[i];
        if ((i & 1) == 0) {
            float nx0 = X0 * c + Y0 * s, ny0 = Y0 * c - X0 * s;
            float nx1 = X1 * c + Y1 * s, ny1 = Y1 * c - X1 * s;
            float nx2 = X2 * c + Y2 * s, ny2 = Y2 * c - X2 * s;
            X0 = nx0; X1 = nx1; X2 = nx2;
            Y0 = ny0; Y1 = ny1; Y2 = ny2;
        } else {
            float nx0 = X0 * c - Z0 * s, nz0 = X0 * s + Z0 * c;
            float nx1 = X1 * c - Z1 * s, nz1 = X1 * s + Z1 * c;
            float nx2 = X2 * c - Z2 * s, nz2 = X2 * s + Z2 * c;
            X0 = nx0; X1 = nx1; X2 = nx2;
            Z0 = nz0; Z1 = nz1; Z2 = nz2;
        }
#pragma unroll
        for (int m = 0; m < LaneSpheres<LPS>::N; ++m) {
            const int s_lo = m * LPS, s_hi = m * LPS + LPS - 1;
            if (s_hi < 4 * i || s_lo > 4 * i + 3) continue; // resolved at compile time
            const int sph = m * LPS + g;
            if (sph < kSpheres && (sph >> 2) == i) {
                const float off = sphere_offset(sph & 3);
                S.x[m] = p0 + Z0 * off;
                S.y[m] = p1 + Z1 * off;
                S.z[m] = p2 + Z2 * off;
            }
        }
        p0 = p0 + Z0 * kLink;
        p1 = p1 + Z1 * kLink;
        p2 = p2 + Z2 * kLink;
    }
}

// FK of the lane's spheres (slot or link g); GS: lanes per state of the link
// layout.
template <int LPS, int GS = 64>
__device__ __forceinline__ void arm_fk(const float q[kDof], int g, LaneSpheres<LPS> &S)
{
    if constexpr (LPS == 7)
        arm_fk_links<GS>(q, g, S);
    else
        arm_fk_slots<LPS>(q, g, S);
}

template <int LPS>
__device__ __forceinline__ bool hits_sphere(const LaneSpheres<LPS> &S, float4 o)
{
    bool hit = false;
#pragma unroll
    for (int m = 0; m < LaneSpheres<LPS>::N; ++m) {
        float dx = o.x - S.x[m], dy = o.y - S.y[m], dz = o.z - S.z[m];
        float d2 = (dx * dx + dy * dy) + dz * dz;
        hit |= d2 < o.w;
    }
    return hit;
}

// min(max(x, lo), hi) for finite lo <= hi (every stored box: coordinates are
// checked finite and extents non-negative when an environment is created) in
// one instruction: v_med3_f32 returns the same bits as the min/max pair for
// every x (a NaN x gives a NaN distance either way), without the
// canonicalising moves IEEE-mode min/max need.
__device__ __forceinline__ float clamp_box(float x, float lo, float hi)
{
    return __builtin_amdgcn_fmed3f(x, lo, hi);
}

template <int LPS>
__device__ __forceinline__ bool hits_box(const LaneSpheres<LPS> &S, float4 lo, float4 hi)
{
    bool hit = false;
#pragma unroll
    for (int m = 0; m < LaneSpheres<LPS>::N; ++m) {
        float px = clamp_box(S.x[m], lo.x, hi.x);
        float py = clamp_box(S.y[m], lo.y, hi.y);
        float pz = clamp_box(S.z[m], lo.z, hi.z);
        float dx = S.x[m] - px, dy = S.y[m] - py, dz = S.z[m] - pz;
        float d2 = (dx * dx + dy * dy) + dz * dz;
        hit |= d2 < kRobotR * kRobotR;
    }
    return hit;
}

// Wave-level helpers (wave64).
__device__ __forceinline__ uint64_t wave_ballot(bool p)
{
    return __ballot(p);
}

// Bits set at the first lane of every LPS-lane group.
template <int LPS>
__host__ __device__ constexpr uint64_t leader_mask()
{
    uint64_t m = 0;
    for (int i = 0; i < 64; i += LPS) m |= (uint64_t)1 << i;
    return m;
}

// Folds a per-lane mask onto group leaders: bit (g*LPS) set iff any lane of
// group g is set.
template <int LPS>
__device__ __forceinline__ uint64_t fold_groups(uint64_t m)
{
#pragma unroll
    for (int sh = 1; sh < LPS; sh <<= 1) m |= m >> sh;
    return m & leader_mask<LPS>();
}

// Exit policies of the obstacle loop.
//   kAllGroups: independent states per group; stop when every group has hit
//               (its state is invalid) -- batched isValid.
//   kAnyGroup:  all groups belong to one edge; stop at the first hit.
//   kLowestGroup: all groups belong to one edge; stop once the lowest group
//               with a hit has no unresolved group below it (exact position
//               of the first invalid state in check order).
enum class Exit { kAllGroups, kAnyGroup, kLowestGroup };

// Lane roles inside a state group of LPS*OPS lanes: the lane holds sphere slot
// `slot` (LaneSpheres<LPS>) and visits obstacle partition `part` (obstacles
// part, part + OPS, part + 2*OPS, ...).
template <int LPS, int OPS>
struct GroupShape {
    // lanes per state; the link layout (LPS 7) rounds 7*OPS up to a power of
    // two, lanes with part >= OPS idle
    static constexpr int GS = LPS != 7 ? LPS * OPS : (OPS == 1 ? 8 : OPS == 2 ? 16 : OPS <= 4 ? 32 : 64);
    static constexpr int STATES = 64 / GS;  // states per wave
    int slot, part, group;
    __device__ __forceinline__ explicit GroupShape(int lane)
    {
        const int sub = lane % GS;
        slot = sub % LPS;
        part = sub / LPS;
        group = lane / GS;
    }
};

// Shared forward kinematics.  The FK layout computes 64/LPS states per wave
// (state f = lane / LPS, sphere slot lane % LPS); the collision layout of
// shape (LPS, OPS) checks 64/(LPS*OPS) of them at a time, so pass t of OPS
// passes checks FK states t*STATES + group.  A collision lane takes the
// centres of its (state, slot) from the FK lane that computed them, so each
// state's FK is computed once instead of once per obstacle partition.
template <int LPS, int OPS>
__device__ __forceinline__ void spheres_from_fk(const LaneSpheres<LPS> &F, int t, LaneSpheres<LPS> &S)
{
    const GroupShape<LPS, OPS> sh(threadIdx.x & 63);
    const int src = (t * GroupShape<LPS, OPS>::STATES + sh.group) * LPS + sh.slot;
#pragma unroll
    for (int m = 0; m < LaneSpheres<LPS>::N; ++m) {
        S.x[m] = __shfl(F.x[m], src);
        S.y[m] = __shfl(F.y[m], src);
        S.z[m] = __shfl(F.z[m], src);
    }
}

// Checks the lane's spheres against its obstacle partition of the LDS
// environment.  `live` lanes take part; dead lanes count as resolved.  The
// obstacle loop runs a wave-uniform trip count (two obstacles of the lane's
// partition per iteration, indices clamped to the last obstacle, whose
// repeated test is harmless) and ends each iteration with a ballot.  Returns
// this lane's hit flag (the state's result is the OR over its group).
// Link layout collision: the lane tests its link's bounding sphere against K
// obstacles of its partition (obstacles part, part + OPS, ...) per pass,
// then runs the exact four-sphere test on the obstacles whose bound it
// touches, one per lane per round until no lane has one left; after each
// pass the wave leaves once its exit policy is met.  LDS holds the
// link-layout staging of stage_env<true>.  PADTO (PADDED only): the staging's
// list padding when it is a multiple of a whole pass rather than one pass
// (several obstacle partitionings reading one staging); 0 = one pass.
// stop (kAnyGroup only, may be null): an LDS flag shared by the waves of a
// block checking one motion; a wave that finds a hit sets it, and a wave that
// sees it set leaves after its current pass (the motion is invalid anyway).
template <int OPS, Exit E, bool PADDED, int PADTO = 0>
__device__ __forceinline__ bool collide_links(const LaneSpheres<7> &S, const float4 *lds, int ns, int nb, bool live,
                                              int part, int *stop = nullptr)
{
    constexpr int GS = GroupShape<7, OPS>::GS;
    constexpr int K = link_pass(OPS);
    constexpr int PADM = PADTO ? PADTO : K * OPS;
    static_assert(PADM % (K * OPS) == 0, "staging padding must hold whole passes");
    // PADDED: lists padded to whole passes (stage_env<true, PADM>), no index
    // clamps; else the unpadded layout, indices clamped to the last entry
    const int nsp = PADDED ? pad_to(ns, PADM) : ns, nbp = PADDED ? pad_to(nb, PADM) : nb;
    auto sidx = [&](int o) { return PADDED ? o : min(o, ns - 1); };
    auto bidx = [&](int o) { return PADDED ? o : min(o, nb - 1); };
    const bool act = live && part < OPS;
    part = part < OPS ? part : 0; // idle lanes read partition 0 (their candidates are dropped)
    bool hit = false;
    const uint64_t live_leaders = fold_groups<GS>(wave_ballot(live));
    if (live_leaders == 0) return false;
    int stopped = 0; // the shared flag as read at the start of the current pass
    auto done = [&]() -> bool {
        const uint64_t H = fold_groups<GS>(wave_ballot(hit && live));
        if (E == Exit::kAllGroups) return (H & live_leaders) == live_leaders;
        if (E == Exit::kAnyGroup) {
            if (stop && H != 0 && (threadIdx.x & 63) == 0)
                __hip_atomic_store(stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            return H != 0 || stopped != 0;
        }
        if (H == 0) return false;
        const uint64_t low = H & (~H + 1);
        return ((low - 1) & live_leaders) == 0;
    };
    // read before a pass's bound tests, tested after them
    auto peek = [&]() {
        if (E == Exit::kAnyGroup && stop) stopped = __hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    const float *rr = reinterpret_cast<const float *>(lds + nsp + 2 * nbp);
    const float4 *boxes = lds + nsp;
    // One pass: K obstacles of the lane's partition, bound tests first, then
    // exact tests of the candidates, lowest first, until no lane has any.
    // The bound tests read their obstacles KL at a time, all reads issued
    // before the first test (the scheduler barrier keeps them together): a
    // wave left alone on its SIMD, which the last states of a launch are,
    // then waits for one LDS round trip per KL obstacles instead of one each.
    constexpr int KL = 4, KB = 2;
    auto sphere_pass = [&](int ob) {
        unsigned cand = 0;
#pragma unroll
        for (int j0 = 0; j0 < K; j0 += KL) {
            float4 o[KL];
#pragma unroll
            for (int j = 0; j < KL; ++j) o[j] = lds[sidx(ob + part + (j0 + j) * OPS)];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < KL; ++j) {
                // bound tests need only be conservative: fused multiply-adds
                const float dx = o[j].x - S.cx, dy = o[j].y - S.cy, dz = o[j].z - S.cz;
                const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                cand |= (unsigned)(d2 < o[j].w) << (j0 + j);
            }
        }
        if (!act) cand = 0;
        while (wave_ballot(cand != 0)) {
            if (cand) {
                const int o = sidx(ob + part + __builtin_ctz(cand) * OPS);
                cand &= cand - 1;
                float4 e = lds[o];
                e.w = rr[o];
                hit |= hits_sphere<7>(S, e);
            }
        }
    };
    auto box_pass = [&](int ob) {
        unsigned cand = 0;
#pragma unroll
        for (int j0 = 0; j0 < K; j0 += KB) {
            float3 lo[KB], hi[KB];
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                const int o = bidx(ob + part + (j0 + j) * OPS);
                const float4 l = boxes[2 * o], h = boxes[2 * o + 1];
                lo[j] = make_float3(l.x, l.y, l.z);
                hi[j] = make_float3(h.x, h.y, h.z);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                const float dx = S.cx - clamp_box(S.cx, lo[j].x, hi[j].x);
                const float dy = S.cy - clamp_box(S.cy, lo[j].y, hi[j].y);
                const float dz = S.cz - clamp_box(S.cz, lo[j].z, hi[j].z);
                const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                cand |= (unsigned)(d2 < kBoxBound * kBoxBound) << (j0 + j);
            }
        }
        if (!act) cand = 0;
        while (wave_ballot(cand != 0)) {
            if (cand) {
                const int o = bidx(ob + part + __builtin_ctz(cand) * OPS);
                cand &= cand - 1;
                hit |= hits_box<7>(S, boxes[2 * o], boxes[2 * o + 1]);
            }
        }
    };
    // Box and sphere passes alternate, boxes first (a box of half-extent h
    // holds the sphere of radius h), each list nearest to the base first (the
    // environment's staging order), so that a colliding configuration usually
    // meets its first hit in an early pass.
    for (int ob = 0; ob < nb || ob < ns; ob += K * OPS) {
        if (ob < nb) {
            peek();
            box_pass(ob);
            if (done()) return hit;
        }
        if (ob < ns) {
            peek();
            sphere_pass(ob);
            if (done()) return hit;
        }
    }
    return hit;
}

// Link layout (7,4) under the all-groups exit: two states per wave, lanes
// 0-31 and 32-63.  Once one state is resolved (a hit found, or not live) and
// the other is not, the resolved state's lanes take the other state's link
// spheres (from the lane 32 apart, which holds the same link and partition)
// and check every other pass of it from then on (its own lanes at passes
// ob, ob + 2P, ..., the helpers at ob + P, ob + 3P, ...), so a state that
// scans the whole list -- a valid one does -- is not left with half of its
// wave idle.  Every obstacle is still tested for each state until that state
// is resolved; returns the lane's state's result.
__device__ __forceinline__ bool collide_links_pair(const LaneSpheres<7> &S0, const float4 *lds, int ns, int nb,
                                                   bool live, int part)
{
    constexpr int OPS = 4, K = link_pass(OPS), P = K * OPS;
    constexpr int KL = 4, KB = 2; // LDS reads in flight per bound-test group (spheres, boxes)
    const int nsp = pad_to(ns, P), nbp = pad_to(nb, P);
    const int g = (threadIdx.x & 63) >> 5;
    const bool has_link = part < OPS;
    part = has_link ? part : 0;
    const uint64_t lm = wave_ballot(live);
    bool res0 = (lm & 0xffffffffull) == 0, res1 = (lm >> 32) == 0; // resolved: hit, or not live
    bool hit0 = false, hit1 = false;
    if (res0 && res1) return false;
    LaneSpheres<7> S = S0;
    bool helper = false, coop = false, act = live && has_link, hit = false;
    uint64_t w1 = 0xffffffff00000000ull; // lanes whose working state is state 1
    const float *rr = reinterpret_cast<const float *>(lds + nsp + 2 * nbp);
    const float4 *boxes = lds + nsp;
    auto join = [&]() {
        const int u = res0 ? 1 : 0; // the unresolved state
        helper = g != u;
#pragma unroll
        for (int m = 0; m < LaneSpheres<7>::N; ++m) {
            const float x = __shfl_xor(S.x[m], 32), y = __shfl_xor(S.y[m], 32), z = __shfl_xor(S.z[m], 32);
            if (helper) {
                S.x[m] = x;
                S.y[m] = y;
                S.z[m] = z;
            }
        }
        const float cx = __shfl_xor(S.cx, 32), cy = __shfl_xor(S.cy, 32), cz = __shfl_xor(S.cz, 32);
        if (helper) {
            S.cx = cx;
            S.cy = cy;
            S.cz = cz;
        }
        act = has_link;
        hit = false;
        w1 = u == 1 ? ~0ull : 0ull;
        coop = true;
    };
    auto update = [&]() -> bool {
        const uint64_t m = wave_ballot(hit);
        if (m & ~w1) hit0 = res0 = true;
        if (m & w1) hit1 = res1 = true;
        return res0 && res1;
    };
    // one pass of the lane's partition starting at list index `base`
    auto sphere_pass = [&](int base, bool valid) {
        unsigned cand = 0;
#pragma unroll
        for (int j0 = 0; j0 < K; j0 += KL) {
            float4 o[KL];
#pragma unroll
            for (int j = 0; j < KL; ++j) o[j] = lds[base + part + (j0 + j) * OPS];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < KL; ++j) {
                const float dx = o[j].x - S.cx, dy = o[j].y - S.cy, dz = o[j].z - S.cz;
                const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                cand |= (unsigned)(d2 < o[j].w) << (j0 + j);
            }
        }
        if (!(act && valid)) cand = 0;
        while (wave_ballot(cand != 0)) {
            if (cand) {
                const int o = base + part + __builtin_ctz(cand) * OPS;
                cand &= cand - 1;
                float4 e = lds[o];
                e.w = rr[o];
                hit |= hits_sphere<7>(S, e);
            }
        }
    };
    auto box_pass = [&](int base, bool valid) {
        unsigned cand = 0;
#pragma unroll
        for (int j0 = 0; j0 < K; j0 += KB) {
            float3 lo[KB], hi[KB];
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                const int o = base + part + (j0 + j) * OPS;
                const float4 l = boxes[2 * o], h = boxes[2 * o + 1];
                lo[j] = make_float3(l.x, l.y, l.z);
                hi[j] = make_float3(h.x, h.y, h.z);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                const float dx = S.cx - clamp_box(S.cx, lo[j].x, hi[j].x);
                const float dy = S.cy - clamp_box(S.cy, lo[j].y, hi[j].y);
                const float dz = S.cz - clamp_box(S.cz, lo[j].z, hi[j].z);
                const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                cand |= (unsigned)(d2 < kBoxBound * kBoxBound) << (j0 + j);
            }
        }
        if (!(act && valid)) cand = 0;
        while (wave_ballot(cand != 0)) {
            if (cand) {
                const int o = base + part + __builtin_ctz(cand) * OPS;
                cand &= cand - 1;
                hit |= hits_box<7>(S, boxes[2 * o], boxes[2 * o + 1]);
            }
        }
    };
    if (res0 != res1) join();
    // box and sphere passes alternate, boxes first, as in collide_links
    for (int ob = 0; ob < nb || ob < ns;) {
        const int step = coop ? 2 * P : P;
        const int mine = ob + (helper ? P : 0); // helpers take the next pass
        if (ob < nb) {
            const bool v = mine < nbp; // else no pass left for this lane: read in range, test nothing
            box_pass(v ? mine : ob, v);
            if (update()) break;
        }
        if (ob < ns) {
            const bool v = mine < nsp;
            sphere_pass(v ? mine : ob, v);
            if (update()) break;
        }
        ob += step;
        if (!coop && res0 != res1) join();
    }
    return g == 0 ? hit0 : hit1;
}

// PADDED (link layout only): the obstacles are in the padded LDS staging of
// stage_env<true, link_pass(OPS) * OPS>; false for the unpadded copy in HBM.
template <int LPS, int OPS, Exit E, bool PADDED = true>
__device__ __forceinline__ bool collide_env(const LaneSpheres<LPS> &S, const float4 *lds, int ns, int nb, bool live,
                                            int part)
{
    if constexpr (LPS == 7 && OPS == 4 && E == Exit::kAllGroups && PADDED)
        return collide_links_pair(S, lds, ns, nb, live, part);
    else if constexpr (LPS == 7)
        return collide_links<OPS, E, PADDED>(S, lds, ns, nb, live, part);
    constexpr int GS = LPS * OPS;
    bool hit = false;
    const uint64_t live_leaders = fold_groups<GS>(wave_ballot(live));
    auto done = [&](bool h) -> bool {
        const uint64_t H = fold_groups<GS>(wave_ballot(h && live));
        if (E == Exit::kAllGroups) return (H & live_leaders) == live_leaders;
        if (E == Exit::kAnyGroup) return H != 0;
        if (H == 0) return false;
        const uint64_t low = H & (~H + 1);
        return ((low - 1) & live_leaders) == 0;
    };
    if (live_leaders == 0) return hit;
    for (int ob = 0; ob < ns; ob += 2 * OPS) {
        const int o0 = min(ob + part, ns - 1), o1 = min(ob + OPS + part, ns - 1);
        const float4 s0 = lds[o0], s1 = lds[o1];
        hit |= hits_sphere<LPS>(S, s0) | hits_sphere<LPS>(S, s1);
        if (done(hit)) return hit;
    }
    const float4 *boxes = lds + ns;
    for (int ob = 0; ob < nb; ob += 2 * OPS) {
        const int o0 = min(ob + part, nb - 1), o1 = min(ob + OPS + part, nb - 1);
        const float4 l0 = boxes[2 * o0], h0 = boxes[2 * o0 + 1];
        const float4 l1 = boxes[2 * o1], h1 = boxes[2 * o1 + 1];
        hit |= hits_box<LPS>(S, l0, h0) | hits_box<LPS>(S, l1, h1);
        if (done(hit)) return hit;
    }
    return hit;
}

} // namespace ompl_amd
