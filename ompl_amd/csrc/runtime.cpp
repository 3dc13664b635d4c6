// Library runtime: initialisation, error reporting, device-memory helpers,
// per-stream workspaces and per-thread resources of the blocking entry points.
#include "runtime.hpp"

#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace ompl_amd {

namespace {
thread_local std::string g_last_error;
std::once_flag g_init_once;
int g_device = 0;
int g_cus = 256;
hipStream_t g_stream = nullptr;
std::mutex g_ws_mutex;
std::map<hipStream_t, Workspace *> g_ws;
Tuning g_tuning{8, 2, 256, 8, 1, 256, 8, 3};
std::mutex g_tuning_mutex;
} // namespace

void set_last_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

const char *last_error()
{
    return g_last_error.c_str();
}

void fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Error(code, buf);
}

static void do_init(int dev)
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        fail(OMPL_AMD_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    OA_REQUIRE(dev >= 0 && dev < count, OMPL_AMD_ERR_ARG, "device %d out of range (have %d)", dev, count);
    OA_HIP(hipSetDevice(dev));
    hipDeviceProp_t prop;
    OA_HIP(hipGetDeviceProperties(&prop, dev));
    g_device = dev;
    g_cus = prop.multiProcessorCount;
    OA_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    init_noalloc_scratch(dev);
}

void init_device(int dev)
{
    static int init_err = 0;
    static std::string init_msg;
    std::call_once(g_init_once, [&] {
        try {
            do_init(dev);
        } catch (const Error &e) {
            init_err = e.code;
            init_msg = e.what();
        }
    });
    if (init_err) fail(init_err, "%s", init_msg.c_str());
    OA_HIP(hipSetDevice(g_device));
}

void ensure_init()
{
    init_device(0);
}

int device()
{
    return g_device;
}

int num_cus()
{
    return g_cus;
}

hipStream_t default_stream()
{
    ensure_init();
    return g_stream;
}

Workspace &workspace(hipStream_t s, size_t list_entries, std::unique_lock<std::mutex> &lock)
{
    Workspace *w = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_ws_mutex);
        Workspace *&slot = g_ws[s];
        if (!slot) {
            slot = new Workspace();
            OA_HIP(hipMalloc(&slot->counters, 64 * sizeof(int32_t)));
            OA_HIP(hipMemset(slot->counters, 0, 64 * sizeof(int32_t)));
            OA_HIP(hipDeviceSynchronize());
        }
        w = slot;
    }
    // workspaces are never freed: the pointer stays valid without g_ws_mutex
    lock = std::unique_lock<std::mutex>(w->mu);
    if (w->list_cap < list_entries) {
        // the stream may still use the old arrays
        OA_HIP(hipStreamSynchronize(s));
        if (w->list) OA_HIP(hipFree(w->list));
        if (w->first_bad) OA_HIP(hipFree(w->first_bad));
        if (w->surv_nd) OA_HIP(hipFree(w->surv_nd));
        if (w->surv_ab) OA_HIP(hipFree(w->surv_ab));
        size_t cap = list_entries < 1024 ? 1024 : list_entries;
        OA_HIP(hipMalloc(&w->list, cap * sizeof(int32_t)));
        OA_HIP(hipMalloc(&w->first_bad, cap * sizeof(int32_t)));
        OA_HIP(hipMalloc(&w->surv_nd, cap * sizeof(int32_t)));
        OA_HIP(hipMalloc(&w->surv_ab, cap * 14 * sizeof(double)));
        w->list_cap = cap;
    }
    return *w;
}

ThreadResources &thread_resources()
{
    thread_local ThreadResources r;
    if (!r.stream) {
        ensure_init();
        OA_HIP(hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking));
        OA_HIP(hipHostMalloc((void **)&r.pinned, 4096, hipHostMallocCoherent | hipHostMallocMapped));
        OA_HIP(hipHostGetDevicePointer((void **)&r.pinned_dev, r.pinned, 0));
    }
    return r;
}

void stage_reserve(ThreadResources &r, size_t bytes)
{
    if (bytes <= r.stage_cap) return;
    size_t cap = r.stage_cap ? r.stage_cap : 4096;
    while (cap < bytes) cap *= 2;
    if (r.stage) OA_HIP(hipHostFree(r.stage));
    r.stage = nullptr;
    r.stage_cap = 0;
    OA_HIP(hipHostMalloc((void **)&r.stage, cap, hipHostMallocCoherent | hipHostMallocMapped));
    OA_HIP(hipHostGetDevicePointer((void **)&r.stage_dev, r.stage, 0));
    r.stage_cap = cap;
}

namespace {
struct ProfRec {
    std::string name;
    hipEvent_t start, stop;
};
std::mutex g_prof_mutex;
bool g_prof_on = false;
std::vector<ProfRec> g_prof;
std::vector<hipEvent_t> g_event_pool;

hipEvent_t take_event()
{
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    OA_HIP(hipEventCreate(&e));
    return e;
}
} // namespace

ProfScope::ProfScope(const char *name, hipStream_t s) : stream(s)
{
    std::lock_guard<std::mutex> lk(g_prof_mutex);
    if (!g_prof_on) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) return; // not in graphs
    ProfRec r{name, take_event(), take_event()};
    OA_HIP(hipEventRecord(r.start, s));
    idx = (int)g_prof.size();
    g_prof.push_back(r);
}

ProfScope::~ProfScope()
{
    if (idx < 0) return;
    std::lock_guard<std::mutex> lk(g_prof_mutex);
    if (idx < (int)g_prof.size()) hipEventRecord(g_prof[idx].stop, stream);
}

Tuning get_tuning()
{
    std::lock_guard<std::mutex> lk(g_tuning_mutex);
    return g_tuning;
}

void check_tuning(const Tuning &t)
{
    auto pow2 = [](int v, int lo, int hi) { return v >= lo && v <= hi && (v & (v - 1)) == 0; };
    OA_REQUIRE(shape_supported(t.validity_lanes_per_state, t.validity_obstacle_parts), OMPL_AMD_ERR_ARG,
               "validity shape (lanes %d, obstacle parts %d) not built", t.validity_lanes_per_state,
               t.validity_obstacle_parts);
    OA_REQUIRE(shape_supported(t.motion_lanes_per_state, t.motion_obstacle_parts), OMPL_AMD_ERR_ARG,
               "motion shape (lanes %d, obstacle parts %d) not built", t.motion_lanes_per_state,
               t.motion_obstacle_parts);
    OA_REQUIRE(pow2(t.validity_block, 64, 1024) && pow2(t.motion_block, 64, 1024), OMPL_AMD_ERR_ARG,
               "block must be a power of two in [64, 1024]");
    OA_REQUIRE(t.motion_waves_per_cu >= 1 && t.motion_waves_per_cu <= 32, OMPL_AMD_ERR_ARG,
               "motion_waves_per_cu must be in [1, 32]");
    OA_REQUIRE(t.motion_strategy >= 0 && t.motion_strategy <= 4, OMPL_AMD_ERR_ARG,
               "motion_strategy must be in [0, 4]");
    OA_REQUIRE(t.validity_shared_fk == 0 || t.validity_shared_fk == 1, OMPL_AMD_ERR_ARG,
               "validity_shared_fk must be 0 or 1");
    if (t.motion_strategy == 2)
        OA_REQUIRE(fused_supported(t.validity_lanes_per_state, t.validity_obstacle_parts, t.motion_lanes_per_state,
                                   t.motion_obstacle_parts),
                   OMPL_AMD_ERR_ARG, "fused strategy: shape pair not built");
}

void set_tuning(const Tuning &t)
{
    check_tuning(t);
    std::lock_guard<std::mutex> lk(g_tuning_mutex);
    g_tuning = t;
}

Tuning tuning_for(const ompl_amd_env_s &e)
{
    {
        std::lock_guard<std::mutex> lk(e.tuning_mutex);
        if (e.has_tuning) return e.tuning;
    }
    return get_tuning();
}

ompl_amd_tuning tuning_to_c(const Tuning &x)
{
    ompl_amd_tuning t;
    t.validity_lanes_per_state = x.validity_lanes_per_state;
    t.validity_obstacle_parts = x.validity_obstacle_parts;
    t.validity_block = x.validity_block;
    t.motion_lanes_per_state = x.motion_lanes_per_state;
    t.motion_obstacle_parts = x.motion_obstacle_parts;
    t.motion_block = x.motion_block;
    t.motion_waves_per_cu = x.motion_waves_per_cu;
    t.motion_strategy = x.motion_strategy;
    t.validity_shared_fk = x.validity_shared_fk;
    return t;
}

Tuning tuning_from_c(const ompl_amd_tuning &t)
{
    return Tuning{t.validity_lanes_per_state, t.validity_obstacle_parts, t.validity_block,
                  t.motion_lanes_per_state,   t.motion_obstacle_parts,   t.motion_block,
                  t.motion_waves_per_cu,      t.motion_strategy,         t.validity_shared_fk};
}

} // namespace ompl_amd

using namespace ompl_amd;

extern "C" {

int ompl_amd_init(int dev)
{
    return guarded([&] { init_device(dev); });
}

const char *ompl_amd_last_error(void)
{
    return last_error();
}

int ompl_amd_version(void)
{
    return OMPL_AMD_VERSION;
}

int ompl_amd_capabilities(ompl_amd_caps *caps)
{
    return guarded([&] {
        OA_REQUIRE(caps, OMPL_AMD_ERR_ARG, "null argument");
        ensure_init();
        ompl_amd_caps c{};
        c.arm_dof = OMPL_AMD_ARM_DOF;
        c.arm_spheres = OMPL_AMD_ARM_SPHERES;
        c.max_env_words = kMaxEnvWords;
        c.max_segments = 1 << 24;
        c.max_k = 64;
        c.max_nn_dim = 14;
        c.max_sample_dim = 16;
        c.device = device();
        c.compute_units = num_cus();
        hipDeviceProp_t prop;
        OA_HIP(hipGetDeviceProperties(&prop, device()));
        snprintf(c.device_name, sizeof(c.device_name), "%s", prop.gcnArchName);
        *caps = c;
    });
}

int ompl_amd_synchronize(void *stream)
{
    return guarded([&] { OA_HIP(hipStreamSynchronize(resolve_stream(stream))); });
}

int ompl_amd_malloc(void **ptr, size_t bytes)
{
    return guarded([&] {
        OA_REQUIRE(ptr, OMPL_AMD_ERR_ARG, "null output pointer");
        ensure_init();
        hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
        if (e == hipErrorOutOfMemory) fail(OMPL_AMD_ERR_NOMEM, "hipMalloc(%zu) out of memory", bytes);
        OA_HIP(e);
    });
}

int ompl_amd_free(void *ptr)
{
    return guarded([&] {
        ensure_init();
        if (ptr) OA_HIP(hipFree(ptr));
    });
}

int ompl_amd_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    return guarded([&] {
        ensure_init();
        OA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    });
}

int ompl_amd_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    return guarded([&] {
        ensure_init();
        OA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    });
}

int ompl_amd_profile_enable(int on)
{
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_prof_mutex);
        g_prof_on = on != 0;
    });
}

int ompl_amd_profile_reset(void)
{
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_prof_mutex);
        for (auto &r : g_prof) {
            OA_HIP(hipEventSynchronize(r.stop));
            g_event_pool.push_back(r.start);
            g_event_pool.push_back(r.stop);
        }
        g_prof.clear();
    });
}

int ompl_amd_profile_query(const char *name, int64_t *launches, double *total_ms)
{
    return guarded([&] {
        OA_REQUIRE(name && launches && total_ms, OMPL_AMD_ERR_ARG, "null argument");
        std::lock_guard<std::mutex> lk(g_prof_mutex);
        int64_t n = 0;
        double ms = 0.0;
        for (auto &r : g_prof) {
            if (r.name != name) continue;
            OA_HIP(hipEventSynchronize(r.stop));
            float t = 0.0f;
            OA_HIP(hipEventElapsedTime(&t, r.start, r.stop));
            ms += t;
            ++n;
        }
        *launches = n;
        *total_ms = ms;
    });
}

int ompl_amd_get_tuning(ompl_amd_tuning *t)
{
    return guarded([&] {
        OA_REQUIRE(t, OMPL_AMD_ERR_ARG, "null tuning");
        *t = tuning_to_c(get_tuning());
    });
}

int ompl_amd_set_tuning(const ompl_amd_tuning *t)
{
    return guarded([&] {
        OA_REQUIRE(t, OMPL_AMD_ERR_ARG, "null tuning");
        set_tuning(tuning_from_c(*t));
    });
}

} // extern "C"
