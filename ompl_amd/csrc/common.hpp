// Shared host-side plumbing of libompl_amd: status codes, the thread-local
// last-error string and HIP error propagation.  The C-ABI returns an int
// status; the C++ adapters turn a non-zero status into ompl::Exception
// (util/Exception.h:46 of the reference).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/ompl_amd.h"

namespace ompl_amd {

void set_last_error(const char *fmt, ...);
const char *last_error();

// Error raised inside the library; converted to a status code at the C-ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const char *fmt, ...);

#define OA_HIP(expr)                                                                                 \
    do {                                                                                             \
        hipError_t oa_err_ = (expr);                                                                 \
        if (oa_err_ != hipSuccess)                                                                   \
            ::ompl_amd::fail(OMPL_AMD_ERR_HIP, "%s failed: %s (%s:%d)", #expr,                       \
                             hipGetErrorString(oa_err_), __FILE__, __LINE__);                        \
    } while (0)

#define OA_REQUIRE(cond, code, ...)                                                                  \
    do {                                                                                             \
        if (!(cond)) ::ompl_amd::fail((code), __VA_ARGS__);                                          \
    } while (0)

// Runs f() and maps exceptions to a status code + last-error string.
template <class F>
int guarded(F &&f)
{
    try {
        f();
        return OMPL_AMD_OK;
    } catch (const Error &e) {
        set_last_error("%s", e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return OMPL_AMD_ERR_NOMEM;
    } catch (const std::exception &e) {
        set_last_error("%s", e.what());
        return OMPL_AMD_ERR_INTERNAL;
    }
}

// The library's device and its default stream (created by ompl_amd_init).
int device();
hipStream_t default_stream();
void ensure_init();
// A NULL stream argument is HIP's null stream (as for HIP runtime calls).
inline hipStream_t resolve_stream(void *s)
{
    ensure_init();
    return reinterpret_cast<hipStream_t>(s);
}
void ensure_init();

// Number of compute units of the library's device (for persistent grids).
int num_cus();

// Optional per-kernel timing with HIP events recorded on the launch stream
// around each launch (enabled by ompl_amd_profile_enable; off by default).
struct ProfScope {
    ProfScope(const char *name, hipStream_t s);
    ~ProfScope();
    int idx = -1;
    hipStream_t stream;
};

} // namespace ompl_amd
