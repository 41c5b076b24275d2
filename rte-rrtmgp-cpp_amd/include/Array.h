// Host Array<T,N> and device Array_gpu<T,N>: Fortran-order (first index fastest), 1-based
// operator()({i,j,...}) with optional offsets, as the reference's include/Array.h:42-281,340-634.
// Differences by design: the cell count and every offset are 64-bit (the reference's int count
// overflows at 65536 x 128 x 256 cells), and device memory goes through the C ABI of
// rrtmgp_hip.h instead of cudaMalloc.
#ifndef RRTMGP_AMD_ARRAY_H
#define RRTMGP_AMD_ARRAY_H

#include <algorithm>
#include <array>
#include <cstdint>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "rrtmgp_hip.h"

namespace rrtmgp_cpp
{
// Stream used by all class-API launches of this thread (NULL = default stream).
void* stream();
void set_stream(void* s);

inline void hip_check(int rc)
{
    if (rc == RRTMGP_OK) return;
    const std::string msg = rrtmgp_hip_last_error();
    if (rc == RRTMGP_ERR_RANGE) throw std::range_error(msg);
    if (rc == RRTMGP_ERR_ARG) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}
}  // namespace rrtmgp_cpp

template<int N>
inline int64_t product(const std::array<int, N>& dims)
{
    int64_t n = 1;
    for (int d : dims) n *= d;
    return n;
}

template<typename T, int N>
class Array
{
    public:
        Array() : dims{}, offsets{}, ncells(0) {}

        explicit Array(const std::array<int, N>& dims_in) :
            dims(dims_in), offsets{}, ncells(product<N>(dims_in)), data(size_t(ncells))
        {}

        Array(std::vector<T>&& data_in, const std::array<int, N>& dims_in) :
            dims(dims_in), offsets{}, ncells(product<N>(dims_in)), data(std::move(data_in))
        {
            if (int64_t(data.size()) != ncells) throw std::runtime_error("Array: size mismatch");
        }

        Array(const std::vector<T>& data_in, const std::array<int, N>& dims_in) :
            dims(dims_in), offsets{}, ncells(product<N>(dims_in)), data(data_in)
        {
            if (int64_t(data.size()) != ncells) throw std::runtime_error("Array: size mismatch");
        }

        void set_dims(const std::array<int, N>& dims_in)
        {
            if (ncells != 0) throw std::runtime_error("Only arrays of size 0 can be resized");
            dims = dims_in;
            ncells = product<N>(dims);
            data.resize(size_t(ncells));
        }

        void set_offsets(const std::array<int, N>& o) { offsets = o; }

        int64_t calc_index(const std::array<int, N>& idx) const
        {
            int64_t i = 0, stride = 1;
            for (int k = 0; k < N; ++k)
            {
                i += stride * (idx[k] - offsets[k] - 1);
                stride *= dims[k];
            }
            return i;
        }

        T& operator()(const std::array<int, N>& idx) { return data[size_t(calc_index(idx))]; }
        const T& operator()(const std::array<int, N>& idx) const { return data[size_t(calc_index(idx))]; }

        T* ptr() { return data.data(); }
        const T* ptr() const { return data.data(); }
        std::vector<T>& v() { return data; }
        const std::vector<T>& v() const { return data; }
        int dim(int i) const { return dims[i - 1]; }
        int64_t size() const { return ncells; }
        const std::array<int, N>& get_dims() const { return dims; }
        T max() const { return *std::max_element(data.begin(), data.end()); }
        T min() const { return *std::min_element(data.begin(), data.end()); }

        // subset({{ {lo1, hi1}, {lo2, hi2}, ... }}), 1-based inclusive (Array.h:220-250).
        Array<T, N> subset(const std::array<std::pair<int, int>, N>& r) const
        {
            std::array<int, N> nd;
            for (int k = 0; k < N; ++k) nd[k] = r[k].second - r[k].first + 1;
            Array<T, N> out(nd);
            std::array<int, N> idx;
            for (int64_t c = 0; c < out.size(); ++c)
            {
                int64_t rem = c;
                for (int k = 0; k < N; ++k)
                {
                    idx[k] = int(rem % nd[k]) + r[k].first;
                    rem /= nd[k];
                }
                out.data[size_t(c)] = (*this)(idx);
            }
            return out;
        }

    private:
        std::array<int, N> dims;
        std::array<int, N> offsets;
        int64_t ncells;
        std::vector<T> data;

        template<typename, int> friend class Array;
};

template<typename T, int N>
class Array_gpu
{
    public:
        Array_gpu() : dims{}, ncells(0), data_ptr(nullptr), owning(true) {}

        explicit Array_gpu(const std::array<int, N>& dims_in) :
            dims(dims_in), ncells(product<N>(dims_in)), data_ptr(nullptr), owning(true)
        {
            allocate(true);
        }

        explicit Array_gpu(const Array<T, N>& a) :
            dims(a.get_dims()), ncells(a.size()), data_ptr(nullptr), owning(true)
        {
            allocate(false);
            if (ncells)
                rrtmgp_cpp::hip_check(rrtmgp_hip_memcpy_h2d(data_ptr, a.ptr(), bytes(), rrtmgp_cpp::stream()));
        }

        // Non-owning view of device memory (Array.h:460-470).
        Array_gpu(T* ptr, const std::array<int, N>& dims_in) :
            dims(dims_in), ncells(product<N>(dims_in)), data_ptr(ptr), owning(false)
        {}

        Array_gpu(const Array_gpu& o) : dims(o.dims), ncells(o.ncells), data_ptr(nullptr), owning(true)
        {
            allocate(false);
            if (ncells)
                rrtmgp_cpp::hip_check(rrtmgp_hip_memcpy_d2d(data_ptr, o.data_ptr, bytes(), rrtmgp_cpp::stream()));
        }

        Array_gpu(Array_gpu&& o) noexcept :
            dims(o.dims), ncells(o.ncells), data_ptr(o.data_ptr), owning(o.owning)
        {
            o.data_ptr = nullptr; o.ncells = 0; o.dims = {};
        }

        Array_gpu& operator=(Array_gpu&& o) noexcept
        {
            if (this != &o)
            {
                release();
                dims = o.dims; ncells = o.ncells; data_ptr = o.data_ptr; owning = o.owning;
                o.data_ptr = nullptr; o.ncells = 0; o.dims = {};
            }
            return *this;
        }

        Array_gpu& operator=(const Array_gpu& o)
        {
            if (this == &o) return *this;
            if (ncells == 0 && owning)
            {
                dims = o.dims; ncells = o.ncells;
                allocate(false);
            }
            else if (ncells != o.ncells)
                throw std::runtime_error("Array_gpu: copy-assign requires equal size");
            if (ncells)
                rrtmgp_cpp::hip_check(rrtmgp_hip_memcpy_d2d(data_ptr, o.data_ptr, bytes(), rrtmgp_cpp::stream()));
            return *this;
        }

        Array_gpu& operator=(const Array<T, N>& a)
        {
            if (ncells == 0 && owning)
            {
                dims = a.get_dims(); ncells = a.size();
                allocate(false);
            }
            else if (ncells != a.size())
                throw std::runtime_error("Array_gpu: assign from host requires equal size");
            if (ncells)
                rrtmgp_cpp::hip_check(rrtmgp_hip_memcpy_h2d(data_ptr, a.ptr(), bytes(), rrtmgp_cpp::stream()));
            return *this;
        }

        ~Array_gpu() { release(); }

        void set_dims(const std::array<int, N>& dims_in)
        {
            if (ncells != 0) throw std::runtime_error("Only arrays of size 0 can be resized");
            dims = dims_in;
            ncells = product<N>(dims);
            allocate(true);
        }

        Array<T, N> to_host() const
        {
            Array<T, N> a(dims);
            if (ncells)
                rrtmgp_cpp::hip_check(rrtmgp_hip_memcpy_d2h(a.ptr(), data_ptr, bytes(), rrtmgp_cpp::stream()));
            return a;
        }

        // Scalar read of one element (a device-to-host copy, as Array.h:551-558).
        T operator()(const std::array<int, N>& idx) const
        {
            int64_t i = 0, stride = 1;
            for (int k = 0; k < N; ++k) { i += stride * (idx[k] - 1); stride *= dims[k]; }
            T v;
            rrtmgp_cpp::hip_check(rrtmgp_hip_memcpy_d2h(&v, data_ptr + i, sizeof(T), rrtmgp_cpp::stream()));
            return v;
        }

        T* ptr() { return data_ptr; }
        const T* ptr() const { return data_ptr; }
        int dim(int i) const { return dims[i - 1]; }
        int64_t size() const { return ncells; }
        const std::array<int, N>& get_dims() const { return dims; }
        size_t bytes() const { return size_t(ncells) * sizeof(T); }

    private:
        void allocate(bool zero)
        {
            if (ncells == 0) { data_ptr = nullptr; return; }
            void* p = nullptr;
            rrtmgp_cpp::hip_check(rrtmgp_hip_cache_malloc(&p, bytes(), rrtmgp_cpp::stream()));
            data_ptr = static_cast<T*>(p);
            if (zero) rrtmgp_cpp::hip_check(rrtmgp_hip_memset(data_ptr, 0, bytes(), rrtmgp_cpp::stream()));
        }

        void release()
        {
            if (owning && data_ptr) (void)rrtmgp_hip_cache_free(data_ptr, rrtmgp_cpp::stream());
            data_ptr = nullptr;
        }

        std::array<int, N> dims;
        int64_t ncells;
        T* data_ptr;
        bool owning;
};

#endif
