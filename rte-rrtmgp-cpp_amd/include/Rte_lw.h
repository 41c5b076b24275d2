// Longwave radiative transfer driver (reference include/Rte_lw.h:62-80).  A g-point flux array
// with dims (ncol, nlev, 1) selects broadband output (the CPU semantics of src/Rte_lw.cpp:137;
// the reference GPU class throws instead).
#ifndef RRTMGP_AMD_RTE_LW_H
#define RRTMGP_AMD_RTE_LW_H

#include <memory>

#include "Array.h"
#include "types.h"

class Optical_props_arry_gpu;
class Source_func_lw_gpu;

class Rte_lw_gpu
{
    public:
        void rte_lw(
                const std::unique_ptr<Optical_props_arry_gpu>& optical_props,
                const Bool top_at_1,
                const Source_func_lw_gpu& sources,
                const Array_gpu<Float,2>& sfc_emis,
                const Array_gpu<Float,2>& inc_flux,
                Array_gpu<Float,3>& gpt_flux_up,
                Array_gpu<Float,3>& gpt_flux_dn,
                const int n_gauss_angles);

        // Band and broadband fluxes summed inside the solver (no g-point flux arrays); either
        // output pair may be empty.  Not in the reference, whose drivers reduce g-point fluxes
        // with Fluxes_byband (src/Fluxes.cpp:142-197).  With p_lev (ncol, nlev) and a non-empty
        // heating_rate (ncol, nlay), heating rates [K/day] from each layer's net-flux convergence
        // formed per g-point and summed in FP64 (rrtmgp_hip_lw_solver_noscat_bnd_*).
        void rte_lw_fluxes(
                const std::unique_ptr<Optical_props_arry_gpu>& optical_props,
                const Bool top_at_1,
                const Source_func_lw_gpu& sources,
                const Array_gpu<Float,2>& sfc_emis,
                const Array_gpu<Float,2>& inc_flux,
                Array_gpu<Float,3>& bnd_flux_up, Array_gpu<Float,3>& bnd_flux_dn,
                Array_gpu<Float,2>& flux_up, Array_gpu<Float,2>& flux_dn,
                const int n_gauss_angles,
                const Array_gpu<Float,2>* p_lev = nullptr, Array_gpu<Float,2>* heating_rate = nullptr);

        void expand_and_transpose(
                const std::unique_ptr<Optical_props_arry_gpu>& ops,
                const Array_gpu<Float,2> arr_in,
                Array_gpu<Float,2>& arr_out);
};

#endif
