// Device gas concentrations (reference src_cuda/Gas_concs.cu:29-80, checks of
// src/Gas_concs.cpp:54-97).
#include "Gas_concs.h"

#include <stdexcept>

#include "hip_dispatch.h"

namespace
{
template<typename A>
void check_range(const std::string& name, const A& a)
{
    for (auto v : a.v())
        if (!(v >= Float(0.) && v <= Float(1.)))
            throw std::range_error("Gas concentration " + name + " is out of range");
}
}

void Gas_concs_gpu::set_vmr(const std::string& name, const Float data)
{
    if (!(data >= Float(0.) && data <= Float(1.)))
        throw std::range_error("Gas concentration " + name + " is out of range");
    Array<Float,2> d({1, 1});
    d({1, 1}) = data;
    gas_concs_map.erase(name);
    gas_concs_map.emplace(name, Array_gpu<Float,2>(d));
}

void Gas_concs_gpu::set_vmr(const std::string& name, const Array<Float,1>& data)
{
    check_range(name, data);
    Array<Float,2> d(data.v(), {1, data.dim(1)});
    gas_concs_map.erase(name);
    gas_concs_map.emplace(name, Array_gpu<Float,2>(d));
}

void Gas_concs_gpu::set_vmr(const std::string& name, const Array<Float,2>& data)
{
    check_range(name, data);
    gas_concs_map.erase(name);
    gas_concs_map.emplace(name, Array_gpu<Float,2>(data));
}

void Gas_concs_gpu::set_vmr(const std::string& name, const Array_gpu<Float,2>& data)
{
    int64_t bad = 0;
    rrtmgp_cpp::hip_check(K<Float>::count_outside(data.size(), data.ptr(), Float(0.), Float(1.), &bad,
                                                  rrtmgp_cpp::stream()));
    if (bad) throw std::range_error("Gas concentration " + name + " is out of range");
    gas_concs_map.erase(name);
    gas_concs_map.emplace(name, Array_gpu<Float,2>(data));
}

// Column subset (src/Gas_concs.cpp:29-42): profiles and scalars are shared.
Gas_concs_gpu::Gas_concs_gpu(const Gas_concs_gpu& ref, const int start, const int size)
{
    for (const auto& g : ref.gas_concs_map)
    {
        if (g.second.dim(1) == 1)
            gas_concs_map.emplace(g.first, Array_gpu<Float,2>(g.second));
        else
        {
            Array_gpu<Float,2> sub({size, g.second.dim(2)});
            rrtmgp_cpp::hip_check(K<Float>::copy_cols(g.second.ptr(), g.second.dim(1), start - 1, sub.ptr(),
                                                      size, 0, size, g.second.dim(2), rrtmgp_cpp::stream()));
            gas_concs_map.emplace(g.first, std::move(sub));
        }
    }
}

const Array_gpu<Float,2>& Gas_concs_gpu::get_vmr(const std::string& name) const
{
    return gas_concs_map.at(name);
}

Bool Gas_concs_gpu::exists(const std::string& name) const
{
    return gas_concs_map.count(name) != 0;
}
