// Batched motion validation that allocates nothing: the survivor workspace of
// ompl_amd_check_motions_dev (survivor list, per-survivor first invalid
// position, segment count and FP64 endpoints, and the counters) lives in a
// file-scope device array instead of memory the library allocates, for callers
// that cannot allocate (code under HIP graph capture, tight serving loops).
//
// The array's device address is taken once per device at library set-up
// (hipGetSymbolAddress with that device current) and only that address is
// used from host code.  Every call on a device shares that device's array:
// calls that overlap on one device (on different streams, or from different
// threads) race; calls on one stream are ordered.  The counters are cleared by
// a fill kernel at the start of each call.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "runtime.hpp"

namespace ompl_amd {

constexpr int64_t kNoAllocMaxEdges = 65536;
constexpr size_t kNoAllocCounterBytes = 256; // 64 int32 counters
constexpr size_t kNoAllocListBytes = (size_t)kNoAllocMaxEdges * sizeof(int32_t);
constexpr size_t kNoAllocEndpointBytes = (size_t)kNoAllocMaxEdges * 14 * sizeof(double);
// counters | list | first_bad | surv_nd | surv_ab  (every part 256-byte aligned)
constexpr size_t kNoAllocScratchBytes = kNoAllocCounterBytes + 3 * kNoAllocListBytes + kNoAllocEndpointBytes;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::vector<unsigned char *> g_scratch_ptr; // device address of `scratch`, per device
} // namespace

// Called once per device by the library's initialisation, with that device
// current.
void init_noalloc_scratch(int dev)
{
    void *p = nullptr;
    OA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
    g_scratch_ptr[(size_t)dev] = static_cast<unsigned char *>(p);
}

static unsigned char *scratch_ptr(int dev)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    OA_REQUIRE(dev >= 0 && dev < (int)g_scratch_ptr.size() && g_scratch_ptr[(size_t)dev], OMPL_AMD_ERR_INTERNAL,
               "no-allocation scratch of device %d not set up", dev);
    return g_scratch_ptr[(size_t)dev];
}

} // namespace ompl_amd

using namespace ompl_amd;

extern "C" int ompl_amd_check_motions_noalloc_dev(ompl_amd_env e, const double *a, const double *b, int64_t n,
                                                  double lvs, unsigned factor, uint8_t *out, int32_t *checked,
                                                  void *stream)
{
    return guarded([&] {
        OA_REQUIRE(e, OMPL_AMD_ERR_ARG, "null environment");
        OA_REQUIRE(n >= 0, OMPL_AMD_ERR_ARG, "negative edge count");
        OA_REQUIRE(n <= kNoAllocMaxEdges, OMPL_AMD_ERR_ARG,
                   "%lld edges: the no-allocation motion check takes at most %lld edges per call",
                   (long long)n, (long long)kNoAllocMaxEdges);
        if (n == 0) return;
        const hipStream_t s = resolve_stream(stream);
        unsigned char *base = scratch_ptr(device());
        Workspace w;
        w.counters = reinterpret_cast<int32_t *>(base);
        w.list = reinterpret_cast<int32_t *>(base + kNoAllocCounterBytes);
        w.first_bad = reinterpret_cast<int32_t *>(base + kNoAllocCounterBytes + kNoAllocListBytes);
        w.surv_nd = reinterpret_cast<int32_t *>(base + kNoAllocCounterBytes + 2 * kNoAllocListBytes);
        w.surv_ab = reinterpret_cast<double *>(base + kNoAllocCounterBytes + 3 * kNoAllocListBytes);
        w.list_cap = (size_t)kNoAllocMaxEdges;
        w.parity = 0;
        launch_zero(w.counters, (int)(kNoAllocCounterBytes / sizeof(int32_t)), s);
        check_motions_tuned(tuning_for(*e), e->env, a, b, n, lvs, factor, out, checked, s, &w);
    });
}
