// Runtime helpers and error state of the C ABI (include/rrtmgp_hip.h).
#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "cache_policy.h"
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

int device_cu_count()
{
    constexpr int max_dev = 64;
    static std::atomic<int> cus[max_dev] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= max_dev) return 256;
    int n = cus[dev].load();
    if (n == 0)
    {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cus[dev].store(n);
    }
    return n;
}

namespace
{
struct TimedLaunch { const char* name; hipEvent_t start, stop; };
std::mutex g_timing_mutex;
std::atomic<bool> g_timing_on{false};
std::vector<TimedLaunch> g_timed;       // launches recorded since timing was switched on
std::vector<hipEvent_t> g_event_pool;   // events of cleared records, reused
struct KernelTotal { std::string name; int64_t launches; double total_ms; };
std::vector<KernelTotal> g_totals;      // built by rrtmgp_hip_timing_collect

hipEvent_t pooled_event()
{
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    if (!g_event_pool.empty())
    {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void clear_records()
{
    for (auto& t : g_timed) { g_event_pool.push_back(t.start); g_event_pool.push_back(t.stop); }
    g_timed.clear();
    g_totals.clear();
}
}  // namespace

namespace
{
thread_local bool g_timing_suspended = false;
}

TimingSuspend::TimingSuspend() : prev_(g_timing_suspended) { g_timing_suspended = true; }
TimingSuspend::~TimingSuspend() { g_timing_suspended = prev_; }

LaunchTimer::LaunchTimer(const char* name, hipStream_t stream) : name_(name), stream_(stream)
{
    if (g_timing_suspended || !g_timing_on.load(std::memory_order_relaxed)) return;
    start_ = pooled_event();
    if (start_ && hipEventRecord(start_, stream_) != hipSuccess) start_ = nullptr;
}

LaunchTimer::~LaunchTimer()
{
    if (!start_) return;
    hipEvent_t stop = pooled_event();
    if (!stop) return;
    (void)hipEventRecord(stop, stream_);
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    g_timed.push_back({name_, start_, stop});
}

namespace
{
// Caching allocator: the reuse rules are those of CachePolicy (cache_policy.h).  Nothing is
// returned to the driver until rrtmgp_hip_cache_release or an out-of-memory retry.
std::mutex g_cache_mutex;
CachePolicy g_cache;

void destroy_fences(const CacheBlock& b)
{
    for (void* e : b.fence)
        if (e) (void)hipEventDestroy(static_cast<hipEvent_t>(e));
}

int release_free_blocks()
{
    std::vector<CacheBlock> blocks;
    {
        std::lock_guard<std::mutex> lock(g_cache_mutex);
        blocks = g_cache.drain_free();
    }
    if (blocks.empty()) return RRTMGP_OK;
    int cur = 0;
    RRTMGP_HIP(hipGetDevice(&cur));
    for (const auto& b : blocks)
    {
        RRTMGP_HIP(hipSetDevice(b.device));
        RRTMGP_HIP(hipDeviceSynchronize());
        destroy_fences(b);
        RRTMGP_HIP(hipFree(b.ptr));
    }
    RRTMGP_HIP(hipSetDevice(cur));
    return RRTMGP_OK;
}

// Event recorded on a stream, for the fence of a block freed on another stream than its owner.
int fence_event(hipStream_t s, void** ev)
{
    hipEvent_t e = nullptr;
    RRTMGP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const hipError_t r = hipEventRecord(e, s);
    if (r != hipSuccess)
    {
        (void)hipEventDestroy(e);
        return check_hip(r, "hipEventRecord (cache fence)");
    }
    *ev = e;
    return RRTMGP_OK;
}
}  // namespace

int cache_alloc(void** p, size_t bytes, hipStream_t stream)
{
    *p = nullptr;
    if (bytes == 0) return RRTMGP_OK;
    bytes = CachePolicy::round_block(bytes);
    int dev = 0;
    RRTMGP_HIP(hipGetDevice(&dev));
    CacheBlock b;
    bool reused = false;
    {
        std::lock_guard<std::mutex> lock(g_cache_mutex);
        reused = g_cache.take(dev, stream, bytes, b);
    }
    if (reused)
    {
        // A fenced block: this stream waits until the uses queued before its free have ended.
        int rc = RRTMGP_OK;
        for (void* e : b.fence)
            if (e && rc == RRTMGP_OK) rc = check_hip(hipStreamWaitEvent(stream, static_cast<hipEvent_t>(e), 0),
                                                     "hipStreamWaitEvent (cache fence)");
        destroy_fences(b);
        if (rc != RRTMGP_OK) return rc;
        *p = b.ptr;
        return RRTMGP_OK;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory)
    {
        (void)hipGetLastError();
        RRTMGP_TRY(release_free_blocks());
        e = hipMalloc(p, bytes);
    }
    RRTMGP_TRY(check_hip(e, "hipMalloc (cached scratch)"));
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    g_cache.add_live(*p, bytes, dev, stream);
    return RRTMGP_OK;
}

int cache_free(void* p, hipStream_t stream)
{
    if (!p) return RRTMGP_OK;
    CacheBlock b;
    bool cross = false;
    {
        std::lock_guard<std::mutex> lock(g_cache_mutex);
        if (!g_cache.find_live(p, stream, b, cross))
            return arg_error("cache_free: pointer was not allocated by the cache");
    }
    void* f0 = nullptr;
    void* f1 = nullptr;
    if (cross)
    {
        // Freed on another stream than the one it was allocated on: fence both.
        int cur = 0;
        RRTMGP_HIP(hipGetDevice(&cur));
        if (cur != b.device) RRTMGP_HIP(hipSetDevice(b.device));
        int rc = fence_event(static_cast<hipStream_t>(b.owner), &f0);
        if (rc == RRTMGP_OK) rc = fence_event(stream, &f1);
        if (cur != b.device) (void)hipSetDevice(cur);
        if (rc != RRTMGP_OK)
        {
            if (f0) (void)hipEventDestroy(static_cast<hipEvent_t>(f0));
            return rc;
        }
    }
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    g_cache.finish_release(p, f0, f1);
    return RRTMGP_OK;
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

int rrtmgp_hip_cache_malloc(void** ptr, size_t bytes, void* stream)
{
    if (!ptr) return arg_error("cache_malloc: NULL output");
    return cache_alloc(ptr, bytes, as_stream(stream));
}

int rrtmgp_hip_cache_free(void* ptr, void* stream) { return cache_free(ptr, as_stream(stream)); }

int rrtmgp_hip_cache_release(void) { return release_free_blocks(); }

int rrtmgp_hip_cache_stats(size_t* live_bytes, size_t* cached_bytes)
{
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    if (live_bytes) *live_bytes = g_cache.live_bytes();
    if (cached_bytes) *cached_bytes = g_cache.cached_bytes();
    return RRTMGP_OK;
}

int rrtmgp_hip_timing_enable(int on)
{
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    if (on) clear_records();
    g_timing_on.store(on != 0);
    return RRTMGP_OK;
}

int rrtmgp_hip_timing_collect(int* n_kernels)
{
    if (!n_kernels) return arg_error("timing_collect: NULL output");
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    g_totals.clear();
    for (const auto& t : g_timed)
    {
        RRTMGP_HIP(hipEventSynchronize(t.stop));
        float ms = 0.f;
        RRTMGP_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
        auto it = std::find_if(g_totals.begin(), g_totals.end(),
                               [&](const KernelTotal& k) { return k.name == t.name; });
        if (it == g_totals.end()) g_totals.push_back({t.name, 1, double(ms)});
        else { it->launches += 1; it->total_ms += ms; }
    }
    *n_kernels = int(g_totals.size());
    return RRTMGP_OK;
}

int rrtmgp_hip_timing_get(int i, char* name, int len, int64_t* launches, double* total_ms)
{
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    if (i < 0 || i >= int(g_totals.size())) return arg_error("timing_get: index out of range");
    if (name && len > 0)
    {
        std::strncpy(name, g_totals[size_t(i)].name.c_str(), size_t(len - 1));
        name[len - 1] = '\0';
    }
    if (launches) *launches = g_totals[size_t(i)].launches;
    if (total_ms) *total_ms = g_totals[size_t(i)].total_ms;
    return RRTMGP_OK;
}

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

int rrtmgp_hip_get_device(int* device)
{
    if (!device) return arg_error("get_device: NULL output");
    RRTMGP_HIP(hipGetDevice(device));
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
    // Blocks owned by the stream are handed to the default stream first (after the stream's
    // work has ended), so that no later fence is recorded on a destroyed stream.
    RRTMGP_HIP(hipStreamSynchronize(as_stream(stream)));
    {
        std::lock_guard<std::mutex> lock(g_cache_mutex);
        g_cache.forget_stream(stream);
    }
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
    RRTMGP_TRY(scratch_alloc(&d, 1, s));
    RRTMGP_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), s));
    const int check_hi = (hi == hi) ? 1 : 0;
    const unsigned blocks = unsigned(std::min<int64_t>(4096, (n + 255) / 256 + 1));
    {
        RRTMGP_TIMED("count_outside_kernel", s);
        hipLaunchKernelGGL(count_outside_kernel<Real>, dim3(blocks), dim3(256), 0, s, n, arr, lo, hi, check_hi, d);
        RRTMGP_LAUNCHED("count_outside_kernel");
    }
    unsigned long long h = 0;
    RRTMGP_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s));
    RRTMGP_HIP(hipStreamSynchronize(s));
    scratch_free(d, s);
    *count = int64_t(h);
    return RRTMGP_OK;
}
}  // namespace rrtmgp_hip

namespace rrtmgp_hip
{
// All range checks of one gas_optics call in one launch: blockIdx.y selects the array.  Block
// (0, 0) also copies play(1,1) and play(1,nlay), which decide the surface layer.
constexpr int MAX_CHECKED = 8;
template<typename Real>
struct CheckSet
{
    int n;
    const Real* a[MAX_CHECKED];
    int64_t size[MAX_CHECKED];
    Real lo[MAX_CHECKED], hi[MAX_CHECKED];
    int check_hi[MAX_CHECKED];
};

template<typename Real>
__global__ void check_ranges_kernel(CheckSet<Real> c, const Real* __restrict__ play, int64_t ncol, int64_t nlay,
                                    unsigned long long* __restrict__ counts, Real* __restrict__ ends)
{
    const int k = blockIdx.y;
    const Real* a = c.a[k];
    const Real lo = c.lo[k], hi = c.hi[k];
    const int chk = c.check_hi[k];
    unsigned long long local = 0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < c.size[k]; i += int64_t(gridDim.x) * blockDim.x)
    {
        const Real v = a[i];
        local += ((v >= lo) && (!chk || v <= hi)) ? 0ull : 1ull;
    }
    if (local) atomicAdd(counts + k, local);
    if (blockIdx.x == 0 && k == 0 && threadIdx.x == 0 && play)
    {
        ends[0] = play[0];
        ends[1] = play[ncol * (nlay - 1)];
    }
}

template<typename Real>
int check_ranges(int n, const Real* const* arrays, const int64_t* sizes, const Real* lo, const Real* hi,
                 int64_t* counts, const Real* play, int64_t ncol, int64_t nlay, Real* play_ends, void* stream)
{
    if (n < 1 || n > MAX_CHECKED) return arg_error("check_ranges: 1 to 8 arrays");
    if (!arrays || !sizes || !lo || !hi || !counts) return arg_error("check_ranges: NULL argument");
    if (play && (!play_ends || ncol < 1 || nlay < 1)) return arg_error("check_ranges: play needs extents and output");
    hipStream_t s = as_stream(stream);
    CheckSet<Real> c{};
    c.n = n;
    int64_t most = 1;
    for (int k = 0; k < n; ++k)
    {
        if (sizes[k] > 0 && !arrays[k]) return arg_error("check_ranges: NULL array");
        c.a[k] = arrays[k]; c.size[k] = sizes[k]; c.lo[k] = lo[k]; c.hi[k] = hi[k];
        c.check_hi[k] = (hi[k] == hi[k]) ? 1 : 0;
        most = std::max(most, sizes[k]);
    }
    // One scratch block: n counters, then the two pressures.
    unsigned long long* d = nullptr;
    RRTMGP_TRY(scratch_alloc(&d, MAX_CHECKED + 2, s));
    RRTMGP_HIP(hipMemsetAsync(d, 0, MAX_CHECKED * sizeof(unsigned long long), s));
    Real* ends = reinterpret_cast<Real*>(d + MAX_CHECKED);
    const unsigned bx = unsigned(std::min<int64_t>(1024, (most + 255) / 256));
    {
        RRTMGP_TIMED("check_ranges_kernel", s);
        hipLaunchKernelGGL(check_ranges_kernel<Real>, dim3(bx, unsigned(n)), dim3(256), 0, s, c, play, ncol, nlay,
                           d, ends);
        RRTMGP_LAUNCHED("check_ranges_kernel");
    }
    unsigned long long h[MAX_CHECKED + 2];
    RRTMGP_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s));
    RRTMGP_HIP(hipStreamSynchronize(s));
    scratch_free(d, s);
    for (int k = 0; k < n; ++k) counts[k] = int64_t(h[k]);
    if (play) std::memcpy(play_ends, h + MAX_CHECKED, 2 * sizeof(Real));
    return RRTMGP_OK;
}
}  // namespace rrtmgp_hip

extern "C" int rrtmgp_hip_check_ranges_f32(int n, const float* const* arrays, const int64_t* sizes, const float* lo,
                                           const float* hi, int64_t* counts, const float* play, int64_t ncol,
                                           int64_t nlay, float* play_ends, void* stream)
{
    return rrtmgp_hip::check_ranges<float>(n, arrays, sizes, lo, hi, counts, play, ncol, nlay, play_ends, stream);
}

extern "C" int rrtmgp_hip_check_ranges_f64(int n, const double* const* arrays, const int64_t* sizes,
                                           const double* lo, const double* hi, int64_t* counts, const double* play,
                                           int64_t ncol, int64_t nlay, double* play_ends, void* stream)
{
    return rrtmgp_hip::check_ranges<double>(n, arrays, sizes, lo, hi, counts, play, ncol, nlay, play_ends, stream);
}

extern "C" int rrtmgp_hip_count_outside_f32(int64_t n, const float* arr, float lo, float hi, int64_t* count, void* stream)
{
    return rrtmgp_hip::count_outside<float>(n, arr, lo, hi, count, stream);
}

extern "C" int rrtmgp_hip_count_outside_f64(int64_t n, const double* arr, double lo, double hi, int64_t* count, void* stream)
{
    return rrtmgp_hip::count_outside<double>(n, arr, lo, hi, count, stream);
}
