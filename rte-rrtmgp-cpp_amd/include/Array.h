// Host Array<T,N> and device Array_gpu<T,N>: Fortran-order (first index fastest), 1-based
// operator()({i,j,...}) with optional offsets, as the reference's include/Array.h:42-281,340-634.
// Differences by design: the cell count and every offset are 64-bit (the reference's int count
// overflows at 65536 x 128 x 256 cells), and device memory goes through the C ABI of
// rrtmgp_hip.h instead of cudaMalloc.  As in the reference, Array_gpu(dims) leaves its memory
// uninitialised (Tools_gpu::allocate_gpu, include/tools_gpu.h:72-99): callers fill() what
// they accumulate into.
#ifndef RRTMGP_AMD_ARRAY_H
#define RRTMGP_AMD_ARRAY_H

#include <algorithm>
#include <array>
#include <cstdint>
#include <fstream>
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

template<typename T, int N> class Array_gpu;

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

        // Host copy of a device array (Array.h:138-147).
        explicit Array(const Array_gpu<T, N>& a);

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
        bool is_empty() const { return ncells == 0; }
        const std::array<int, N>& get_offsets() const { return offsets; }

        // subset({{ {lo1, hi1}, {lo2, hi2}, ... }}), 1-based inclusive (Array.h:220-250); a
        // dimension of length 1 is spread over its whole range.
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
                    idx[k] = (dims[k] == 1) ? 1 + offsets[k] : int(rem % nd[k]) + r[k].first;
                    rem /= nd[k];
                }
                out.data[size_t(c)] = (*this)(idx);
            }
            return out;
        }

        void fill(const T value) { std::fill(data.begin(), data.end(), value); }

        // Raw values to <name>.bin (Array.h:256-268).
        void dump(const std::string& name) const
        {
            const std::string file_name = name + ".bin";
            std::ofstream f(file_name, std::ios::out | std::ios::trunc | std::ios::binary);
            if (!f) throw std::runtime_error("Cannot write file \"" + file_name + "\"");
            f.write(reinterpret_cast<const char*>(data.data()), std::streamsize(ncells * int64_t(sizeof(T))));
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
        Array_gpu() : dims{}, offsets{}, ncells(0), data_ptr(nullptr), owning(true) {}

        // Uninitialised device array (the reference's Array_gpu(dims), Array.h:433-443).
        explicit Array_gpu(const std::array<int, N>& dims_in) :
            dims(dims_in), offsets{}, ncells(product<N>(dims_in)), data_ptr(nullptr), owning(true)
        {
            allocate();
        }

        explicit Array_gpu(const Array<T, N>& a) :
            dims(a.get_dims()), offsets(a.get_offsets()), ncells(a.size()), data_ptr(nullptr), owning(true)
        {
            allocate();
            if (ncells)
                rrtmgp_cpp::hip_check(rrtmgp_hip_memcpy_h2d(data_ptr, a.ptr(), bytes(), rrtmgp_cpp::stream()));
        }

        // Non-owning view of device memory (Array.h:460-470).
        Array_gpu(T* ptr, const std::array<int, N>& dims_in) :
            dims(dims_in), offsets{}, ncells(product<N>(dims_in)), data_ptr(ptr), owning(false)
        {}

        // A copy of a view is a view of the same memory; a copy of an owning array is a deep copy
        // (Array.h:407-424).
        Array_gpu(const Array_gpu& o) :
            dims(o.dims), offsets(o.offsets), ncells(o.ncells), data_ptr(o.owning ? nullptr : o.data_ptr),
            owning(o.owning)
        {
            if (!owning) return;
            allocate();
            if (ncells)
                rrtmgp_cpp::hip_check(rrtmgp_hip_memcpy_d2d(data_ptr, o.data_ptr, bytes(), rrtmgp_cpp::stream()));
        }

        Array_gpu(Array_gpu&& o) noexcept :
            dims(o.dims), offsets(o.offsets), ncells(o.ncells), data_ptr(o.data_ptr), owning(o.owning)
        {
            o.data_ptr = nullptr; o.ncells = 0; o.dims = {}; o.offsets = {}; o.owning = true;
        }

        Array_gpu& operator=(Array_gpu&& o) noexcept
        {
            if (this != &o)
            {
                release();
                dims = o.dims; offsets = o.offsets; ncells = o.ncells; data_ptr = o.data_ptr; owning = o.owning;
                o.data_ptr = nullptr; o.ncells = 0; o.dims = {}; o.offsets = {}; o.owning = true;
            }
            return *this;
        }

        Array_gpu& operator=(const Array_gpu& o)
        {
            if (this == &o) return *this;
            if (ncells == 0 && data_ptr == nullptr)
            {
                dims = o.dims; offsets = o.offsets; ncells = o.ncells;
                if (!o.owning) { owning = false; data_ptr = o.data_ptr; return *this; }
                owning = true;
                allocate();
            }
            else if (ncells != o.ncells)
                throw std::runtime_error("Array_gpu: copy-assign requires equal size");
            if (ncells)
                rrtmgp_cpp::hip_check(rrtmgp_hip_memcpy_d2d(data_ptr, o.data_ptr, bytes(), rrtmgp_cpp::stream()));
            return *this;
        }

        Array_gpu& operator=(const Array<T, N>& a)
        {
            if (ncells == 0 && data_ptr == nullptr)
            {
                dims = a.get_dims(); offsets = a.get_offsets(); ncells = a.size(); owning = true;
                allocate();
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
            offsets = {};
            ncells = product<N>(dims);
            owning = true;
            allocate();
        }

        void set_offsets(const std::array<int, N>& o) { offsets = o; }
        const std::array<int, N>& get_offsets() const { return offsets; }

        Array<T, N> to_host() const
        {
            Array<T, N> a(dims);
            a.set_offsets(offsets);
            if (ncells)
                rrtmgp_cpp::hip_check(rrtmgp_hip_memcpy_d2h(a.ptr(), data_ptr, bytes(), rrtmgp_cpp::stream()));
            return a;
        }

        // Scalar read of one element (a device-to-host copy, as Array.h:551-558).
        T operator()(const std::array<int, N>& idx) const
        {
            T v;
            rrtmgp_cpp::hip_check(rrtmgp_hip_memcpy_d2h(&v, data_ptr + calc_index(idx), sizeof(T), rrtmgp_cpp::stream()));
            return v;
        }

        // One element from another array (Array.h:521-528) / from the host (:530-535).
        void copy(const std::array<int, N>& idx, const Array_gpu<T, N>& in, const std::array<int, N>& idx_in) const
        {
            rrtmgp_cpp::hip_check(rrtmgp_hip_memcpy_d2d(data_ptr + calc_index(idx), in.data_ptr + in.calc_index(idx_in),
                                                        sizeof(T), rrtmgp_cpp::stream()));
        }

        void insert(const std::array<int, N>& idx, const T value) const
        {
            rrtmgp_cpp::hip_check(rrtmgp_hip_memcpy_h2d(data_ptr + calc_index(idx), &value, sizeof(T), rrtmgp_cpp::stream()));
        }

        // Every element set to value on the device (Array.h:501-511).
        void fill(const T value)
        {
            static_assert(sizeof(T) == 1 || sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8,
                          "Array_gpu::fill: element size");
            rrtmgp_cpp::hip_check(rrtmgp_hip_fill(data_ptr, ncells, int(sizeof(T)), &value, rrtmgp_cpp::stream()));
        }

        // subset({{ {lo1, hi1}, ... }}): a new array of the 1-based inclusive ranges, dimensions of
        // length 1 spread over the range (Array.h:563-607).
        Array_gpu<T, N> subset(const std::array<std::pair<int, int>, N>& ranges) const
        {
            std::array<int, N> sub;
            std::array<int, N> corners;
            for (int k = 0; k < N; ++k)
            {
                sub[k] = ranges[k].second - ranges[k].first + 1;
                corners[k] = ranges[k].first;
            }
            Array_gpu<T, N> a_sub(sub);
            subset_into(a_sub, corners);
            return a_sub;
        }

        // The subset of this array at 1-based corner block_corners written into a_sub, which is
        // also returned by value (a deep copy of an owning a_sub, a view of a view), as
        // Array.h:585-607 does.
        Array_gpu<T, N> subset_copy(Array_gpu<T, N>& a_sub, const std::array<int, N>& block_corners) const
        {
            subset_into(a_sub, block_corners);
            return a_sub;
        }

        // Raw values to <name>.bin (Array.h:609-623).
        void dump(const std::string& name) const { to_host().dump(name); }

        T* ptr() { return data_ptr; }
        const T* ptr() const { return data_ptr; }
        int dim(int i) const { return dims[i - 1]; }
        int64_t size() const { return ncells; }
        bool is_empty() const { return ncells == 0; }
        bool is_view() const { return !owning; }
        const std::array<int, N>& get_dims() const { return dims; }
        size_t bytes() const { return size_t(ncells) * sizeof(T); }

    private:
        int64_t calc_index(const std::array<int, N>& idx) const
        {
            int64_t i = 0, stride = 1;
            for (int k = 0; k < N; ++k) { i += stride * (idx[k] - offsets[k] - 1); stride *= dims[k]; }
            return i;
        }

        void subset_into(Array_gpu<T, N>& a_sub, const std::array<int, N>& corners) const
        {
            std::array<int64_t, N> sd, so, st, dd;
            for (int k = 0; k < N; ++k)
            {
                sd[k] = dims[k]; so[k] = offsets[k]; st[k] = corners[k]; dd[k] = a_sub.dims[k];
            }
            rrtmgp_cpp::hip_check(rrtmgp_hip_subset(int(sizeof(T)), N, data_ptr, sd.data(), so.data(), st.data(),
                                                    a_sub.data_ptr, dd.data(), rrtmgp_cpp::stream()));
        }

        void allocate()
        {
            if (ncells == 0) { data_ptr = nullptr; return; }
            void* p = nullptr;
            rrtmgp_cpp::hip_check(rrtmgp_hip_cache_malloc(&p, bytes(), rrtmgp_cpp::stream()));
            data_ptr = static_cast<T*>(p);
        }

        void release()
        {
            if (owning && data_ptr) (void)rrtmgp_hip_cache_free(data_ptr, rrtmgp_cpp::stream());
            data_ptr = nullptr;
        }

        std::array<int, N> dims;
        std::array<int, N> offsets;
        int64_t ncells;
        T* data_ptr;
        bool owning;

        template<typename, int> friend class Array_gpu;
};

template<typename T, int N>
Array<T, N>::Array(const Array_gpu<T, N>& a) : Array(a.to_host()) {}

#endif
