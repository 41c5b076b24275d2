// Shared device-side helpers of the RRTMGP HIP kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <limits>
#include <string>
#include <type_traits>

#include "rrtmgp_hip.h"

namespace rrtmgp_hip
{
using Bool = signed char;   // RTE_USE_CBOOL semantics of the reference (include/types.h:13-17)

// Error state shared by all entry points (thread-local message, status codes of rrtmgp_hip.h).
void set_error(const std::string& msg);
int check_hip(hipError_t e, const char* what);

#define RRTMGP_TRY(expr)                                                        \
    do { const int rc_ = (expr); if (rc_ != RRTMGP_OK) return rc_; } while (0)

#define RRTMGP_HIP(call) RRTMGP_TRY(::rrtmgp_hip::check_hip((call), #call))

// Launch-status check: launch errors surface as hipGetLastError.
#define RRTMGP_LAUNCHED(name) RRTMGP_TRY(::rrtmgp_hip::check_hip(hipGetLastError(), name))

// Optional per-launch timing (rrtmgp_hip_timing_* of rrtmgp_hip.h).  When timing is on, a
// LaunchTimer records a HIP event on the launch stream at construction and another at
// destruction, so one timer must wrap exactly one launch in its own scope.  Off: no HIP calls.
class LaunchTimer
{
    public:
        LaunchTimer(const char* name, hipStream_t stream);
        ~LaunchTimer();
        LaunchTimer(const LaunchTimer&) = delete;
        LaunchTimer& operator=(const LaunchTimer&) = delete;

    private:
        const char* name_;
        hipStream_t stream_;
        hipEvent_t start_ = nullptr;
};

// While a TimingSuspend lives, launches on this thread are not timed (no events are created),
// whatever rrtmgp_hip_timing_enable says.
class TimingSuspend
{
    public:
        TimingSuspend();
        ~TimingSuspend();
        TimingSuspend(const TimingSuspend&) = delete;
        TimingSuspend& operator=(const TimingSuspend&) = delete;

    private:
        bool prev_;
};

#define RRTMGP_CAT2(a, b) a##b
#define RRTMGP_CAT(a, b) RRTMGP_CAT2(a, b)
#define RRTMGP_TIMED(name, stream) \
    ::rrtmgp_hip::LaunchTimer RRTMGP_CAT(launch_timer_, __LINE__)((name), (stream))

inline int arg_error(const std::string& msg)
{
    set_error(msg);
    return RRTMGP_ERR_ARG;
}

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

inline unsigned grid1(int64_t n, int block)
{
    return static_cast<unsigned>((n + block - 1) / block);
}

// Caching device allocator of the library (rrtmgp_hip_cache_* of rrtmgp_hip.h): blocks are
// reused by later requests on the same device and stream; see runtime.hip.
int cache_alloc(void** p, size_t bytes, hipStream_t stream);
int cache_free(void* p, hipStream_t stream);
int retain_pool_memory();

// Scratch of a launcher: taken from the cache, returned right after the launches that use it
// are queued (stream order keeps a later reuse on the same stream behind them).
template<typename T>
int scratch_alloc(T** p, size_t n, hipStream_t s)
{
    void* v = nullptr;
    RRTMGP_TRY(cache_alloc(&v, n * sizeof(T), s));
    *p = static_cast<T*>(v);
    return RRTMGP_OK;
}

template<typename T>
void scratch_free(T* p, hipStream_t s)
{
    if (p) (void)cache_free(p, s);
}

// Precision-dependent constants.
template<typename Real> struct Limits;
template<> struct Limits<float>
{
    static constexpr float eps = FLT_EPSILON;
    static constexpr float tiny = FLT_MIN;
    static constexpr float k_min = 1.e-4f;     // two-stream k floor in single precision
};
template<> struct Limits<double>
{
    static constexpr double eps = DBL_EPSILON;
    static constexpr double tiny = DBL_MIN;
    static constexpr double k_min = 1.e-12;
};

// Cut-off of the optical-property rules: combine_abs_and_rayleigh's ssa = tau_ray / tau for
// tau > 2 eps (src/Gas_optics_rrtmgp.cpp:378), and the max(tau, eps) guards of cloud and aerosol
// optics (src/Cloud_optics.cpp:159-171).  It is the FP64 epsilon in both precisions: with FP32's
// (1.2e-7) every layer with tau < 2.4e-7, which near a 1 Pa model top is most g-points of the top
// ~10 layers, would count as purely absorbing (ssa = 0), moving FP32 heating rates there by
// ~0.5 K/day and the pressure gradients of the adjoint by tens of percent.
constexpr double prop_eps = DBL_EPSILON;

// 8 XCDs with private L2 each: work-group i is dispatched to XCD i % 8.  Remap the flat
// work-group id so that consecutive work units (which share k-table slices) land on one XCD.
constexpr int n_xcd = 8;

__device__ __forceinline__ int64_t xcd_remap(int64_t flat_id, int64_t n_units)
{
    const int64_t per_xcd = (n_units + n_xcd - 1) / n_xcd;
    return (flat_id % n_xcd) * per_xcd + flat_id / n_xcd;
}

inline unsigned xcd_grid(int64_t n_units)
{
    return static_cast<unsigned>(((n_units + n_xcd - 1) / n_xcd) * n_xcd);
}


// Compute units of the current device (cached after the first call; 256 if the query fails).
int device_cu_count();
}  // namespace rrtmgp_hip
