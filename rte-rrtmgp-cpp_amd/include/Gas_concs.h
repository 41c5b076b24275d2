// Gas concentrations on the device (reference include/Gas_concs.h:74-91): name ->
// (ncol|1, nlay|1) volume mixing ratios, checked to lie in [0, 1] (src/Gas_concs.cpp:54-97).
#ifndef RRTMGP_AMD_GAS_CONCS_H
#define RRTMGP_AMD_GAS_CONCS_H

#include <map>
#include <string>

#include "Array.h"
#include "types.h"

class Gas_concs_gpu
{
    public:
        Gas_concs_gpu() = default;
        Gas_concs_gpu(const Gas_concs_gpu& gas_concs_ref, const int start, const int size);

        void set_vmr(const std::string& name, const Float data);
        void set_vmr(const std::string& name, const Array<Float,1>& data);
        void set_vmr(const std::string& name, const Array<Float,2>& data);
        void set_vmr(const std::string& name, const Array_gpu<Float,2>& data);

        const Array_gpu<Float,2>& get_vmr(const std::string& name) const;
        Bool exists(const std::string& name) const;

    private:
        std::map<std::string, Array_gpu<Float,2>> gas_concs_map;
};

#endif
