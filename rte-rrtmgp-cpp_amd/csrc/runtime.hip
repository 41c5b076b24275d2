// Runtime helpers and error state of the C ABI (include/rrtmgp_hip.h).
#include <algorithm>
#include <cstring>
#include <string>

#include "common.h"

namespace rrtmgp_hip
{
namespace
{
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

int check_hip(hipError_t e, const char* what)
{
    if (e == hipSuccess) return RRTMGP_OK;
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return RRTMGP_ERR_HIP;
}

int retain_pool_memory()
{
    static thread_local int done_device = -1;
    int dev = 0;
    RRTMGP_HIP(hipGetDevice(&dev));
    if (done_device == dev) return RRTMGP_OK;
    hipMemPool_t pool;
    RRTMGP_HIP(hipDeviceGetDefaultMemPool(&pool, dev));
    uint64_t threshold = UINT64_MAX;
    RRTMGP_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold));
    done_device = dev;
    return RRTMGP_OK;
}
}  // namespace rrtmgp_hip

using namespace rrtmgp_hip;

extern "C" {

const char* rrtmgp_hip_last_error(void) { return g_last_error.c_str(); }

int rrtmgp_hip_device_count(int* n)
{
    RRTMGP_HIP(hipGetDeviceCount(n));
    return RRTMGP_OK;
}

int rrtmgp_hip_set_device(int device)
{
    RRTMGP_HIP(hipSetDevice(device));
    return RRTMGP_OK;
}

int rrtmgp_hip_device_name(char* buf, int len)
{
    int dev = 0;
    RRTMGP_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    RRTMGP_HIP(hipGetDeviceProperties(&p, dev));
    std::string s = std::string(p.name) + " " + p.gcnArchName;
    std::strncpy(buf, s.c_str(), len - 1);
    buf[len - 1] = '\0';
    return RRTMGP_OK;
}

int rrtmgp_hip_malloc(void** ptr, size_t bytes)
{
    RRTMGP_HIP(hipMalloc(ptr, bytes));
    return RRTMGP_OK;
}

int rrtmgp_hip_free(void* ptr)
{
    RRTMGP_HIP(hipFree(ptr));
    return RRTMGP_OK;
}

int rrtmgp_hip_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream)
{
    RRTMGP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    RRTMGP_HIP(hipStreamSynchronize(as_stream(stream)));
    return RRTMGP_OK;
}

int rrtmgp_hip_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream)
{
    RRTMGP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    RRTMGP_HIP(hipStreamSynchronize(as_stream(stream)));
    return RRTMGP_OK;
}

int rrtmgp_hip_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream)
{
    RRTMGP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    return RRTMGP_OK;
}

int rrtmgp_hip_memset(void* ptr, int value, size_t bytes, void* stream)
{
    RRTMGP_HIP(hipMemsetAsync(ptr, value, bytes, as_stream(stream)));
    return RRTMGP_OK;
}

int rrtmgp_hip_stream_create(void** stream)
{
    hipStream_t s;
    RRTMGP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return RRTMGP_OK;
}

int rrtmgp_hip_stream_destroy(void* stream)
{
    RRTMGP_HIP(hipStreamDestroy(as_stream(stream)));
    return RRTMGP_OK;
}

int rrtmgp_hip_stream_synchronize(void* stream)
{
    RRTMGP_HIP(hipStreamSynchronize(as_stream(stream)));
    return RRTMGP_OK;
}

int rrtmgp_hip_device_synchronize(void)
{
    RRTMGP_HIP(hipDeviceSynchronize());
    return RRTMGP_OK;
}

int rrtmgp_hip_event_create(void** event)
{
    hipEvent_t e;
    RRTMGP_HIP(hipEventCreate(&e));
    *event = e;
    return RRTMGP_OK;
}

int rrtmgp_hip_event_destroy(void* event)
{
    RRTMGP_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return RRTMGP_OK;
}

int rrtmgp_hip_event_record(void* event, void* stream)
{
    RRTMGP_HIP(hipEventRecord(static_cast<hipEvent_t>(event), as_stream(stream)));
    return RRTMGP_OK;
}

int rrtmgp_hip_event_elapsed_ms(void* start, void* stop, float* ms)
{
    RRTMGP_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
    RRTMGP_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    return RRTMGP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// Range validation (count of out-of-range values; NaN counts as out of range).
namespace rrtmgp_hip
{
template<typename Real>
__global__ void count_outside_kernel(int64_t n, const Real* __restrict__ a, Real lo, Real hi, int check_hi,
                                     unsigned long long* __restrict__ count)
{
    unsigned long long local = 0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    {
        const Real v = a[i];
        const bool ok = (v >= lo) && (!check_hi || v <= hi);
        local += ok ? 0ull : 1ull;
    }
    if (local) atomicAdd(count, local);
}

template<typename Real>
int count_outside(int64_t n, const Real* arr, Real lo, Real hi, int64_t* count, void* stream)
{
    hipStream_t s = as_stream(stream);
    unsigned long long* d = nullptr;
    RRTMGP_HIP(hipMallocAsync(reinterpret_cast<void**>(&d), sizeof(unsigned long long), s));
    RRTMGP_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), s));
    const int check_hi = (hi == hi) ? 1 : 0;
    const unsigned blocks = unsigned(std::min<int64_t>(4096, (n + 255) / 256 + 1));
    hipLaunchKernelGGL(count_outside_kernel<Real>, dim3(blocks), dim3(256), 0, s, n, arr, lo, hi, check_hi, d);
    RRTMGP_LAUNCHED("count_outside_kernel");
    unsigned long long h = 0;
    RRTMGP_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s));
    RRTMGP_HIP(hipFreeAsync(d, s));
    RRTMGP_HIP(hipStreamSynchronize(s));
    *count = int64_t(h);
    return RRTMGP_OK;
}
}  // namespace rrtmgp_hip

extern "C" int rrtmgp_hip_count_outside_f32(int64_t n, const float* arr, float lo, float hi, int64_t* count, void* stream)
{
    return rrtmgp_hip::count_outside<float>(n, arr, lo, hi, count, stream);
}

extern "C" int rrtmgp_hip_count_outside_f64(int64_t n, const double* arr, double lo, double hi, int64_t* count, void* stream)
{
    return rrtmgp_hip::count_outside<double>(n, arr, lo, hi, count, stream);
}
