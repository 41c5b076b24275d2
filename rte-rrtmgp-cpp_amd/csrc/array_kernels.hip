// Device kernels behind Array_gpu::fill and Array_gpu::subset (rte-rrtmgp-cpp_amd/include/Array.h),
// the counterparts of the reference's fill_kernel and subset_kernel (include/Array.h:296-335,
// launched at :501-511 and :583-606).  Extents and offsets are 64-bit; the element type only
// matters through its size, so one kernel per size serves every Array_gpu<T,N>.
#include <algorithm>
#include <cstring>

#include "common.h"

namespace rrtmgp_hip
{
constexpr int SUBSET_MAX_DIMS = 8;

template<int B> struct Word;
template<> struct Word<1> { using T = uint8_t; };
template<> struct Word<2> { using T = uint16_t; };
template<> struct Word<4> { using T = uint32_t; };
template<> struct Word<8> { using T = uint64_t; };

template<typename W>
__global__ void fill_kernel(int64_t n, W v, W* __restrict__ dst)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        dst[i] = v;
}

// Subset geometry, all 0-based: source element of destination multi-index d is
//     sum_k (spread_k ? 0 : d_k + first_k) * src_stride_k,
// with first_k = start_k - offset_k - 1 (the reference's 1-based start and Array offset).
struct SubsetGeom
{
    int ndim;
    int64_t dst_dims[SUBSET_MAX_DIMS];
    int64_t src_stride[SUBSET_MAX_DIMS];
    int64_t first[SUBSET_MAX_DIMS];
    int spread[SUBSET_MAX_DIMS];
};

template<typename W>
__global__ void subset_kernel(int64_t n, SubsetGeom g, const W* __restrict__ src, W* __restrict__ dst)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    {
        int64_t rem = i, s = 0;
        for (int k = 0; k < g.ndim; ++k)
        {
            const int64_t d = rem % g.dst_dims[k];
            rem /= g.dst_dims[k];
            if (!g.spread[k]) s += (d + g.first[k]) * g.src_stride[k];
        }
        dst[i] = src[s];
    }
}

inline unsigned grid_stride_blocks(int64_t n)
{
    const int64_t b = (n + 255) / 256;
    return unsigned(std::min<int64_t>(b, int64_t(device_cu_count()) * 32));
}

template<int B>
int fill_t(void* dst, int64_t n, const void* value, hipStream_t s)
{
    using W = typename Word<B>::T;
    W v;
    std::memcpy(&v, value, B);
    RRTMGP_TIMED("fill_kernel", s);
    hipLaunchKernelGGL(fill_kernel<W>, dim3(grid_stride_blocks(n)), dim3(256), 0, s, n, v, static_cast<W*>(dst));
    RRTMGP_LAUNCHED("fill_kernel");
    return RRTMGP_OK;
}

template<int B>
int subset_t(int64_t n, const SubsetGeom& g, const void* src, void* dst, hipStream_t s)
{
    using W = typename Word<B>::T;
    RRTMGP_TIMED("subset_kernel", s);
    hipLaunchKernelGGL(subset_kernel<W>, dim3(grid_stride_blocks(n)), dim3(256), 0, s, n, g,
                       static_cast<const W*>(src), static_cast<W*>(dst));
    RRTMGP_LAUNCHED("subset_kernel");
    return RRTMGP_OK;
}
}  // namespace rrtmgp_hip

using namespace rrtmgp_hip;

extern "C" int rrtmgp_hip_fill(void* dst, int64_t n, int elem_size, const void* value, void* stream)
{
    if (n < 0) return arg_error("fill: negative element count");
    if (n == 0) return RRTMGP_OK;
    if (!dst || !value) return arg_error("fill: NULL pointer");
    hipStream_t s = as_stream(stream);
    switch (elem_size)
    {
        case 1: return fill_t<1>(dst, n, value, s);
        case 2: return fill_t<2>(dst, n, value, s);
        case 4: return fill_t<4>(dst, n, value, s);
        case 8: return fill_t<8>(dst, n, value, s);
        default: return arg_error("fill: element size must be 1, 2, 4 or 8 bytes");
    }
}

extern "C" int rrtmgp_hip_subset(int elem_size, int ndim, const void* src, const int64_t* src_dims,
                                 const int64_t* src_offsets, const int64_t* starts, void* dst,
                                 const int64_t* dst_dims, void* stream)
{
    if (ndim < 1 || ndim > SUBSET_MAX_DIMS) return arg_error("subset: 1 to 8 dimensions");
    if (!src_dims || !starts || !dst_dims) return arg_error("subset: NULL geometry");
    SubsetGeom g{};
    g.ndim = ndim;
    int64_t n = 1, stride = 1;
    for (int k = 0; k < ndim; ++k)
    {
        const int64_t sd = src_dims[k], dd = dst_dims[k];
        const int64_t off = src_offsets ? src_offsets[k] : 0;
        if (sd < 1 || dd < 0) return arg_error("subset: bad extents");
        g.dst_dims[k] = dd;
        g.src_stride[k] = stride;
        g.spread[k] = (sd == 1) ? 1 : 0;   // a dimension of length 1 is spread (Array.h:230-233)
        g.first[k] = starts[k] - off - 1;
        // Every read stays inside the source (the reference does not check; a fault here would be).
        if (!g.spread[k] && dd > 0 && (g.first[k] < 0 || g.first[k] + dd > sd))
            return arg_error("subset: range " + std::to_string(k + 1) + " lies outside the source array");
        n *= dd;
        stride *= sd;
    }
    if (n == 0) return RRTMGP_OK;
    if (!src || !dst) return arg_error("subset: NULL array");
    hipStream_t s = as_stream(stream);
    switch (elem_size)
    {
        case 1: return subset_t<1>(n, g, src, dst, s);
        case 2: return subset_t<2>(n, g, src, dst, s);
        case 4: return subset_t<4>(n, g, src, dst, s);
        case 8: return subset_t<8>(n, g, src, dst, s);
        default: return arg_error("subset: element size must be 1, 2, 4 or 8 bytes");
    }
}
