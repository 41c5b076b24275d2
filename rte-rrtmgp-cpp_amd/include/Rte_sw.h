// Shortwave two-stream radiative transfer driver (reference include/Rte_sw.h:62-81).
#ifndef RRTMGP_AMD_RTE_SW_H
#define RRTMGP_AMD_RTE_SW_H

#include <memory>

#include "Array.h"
#include "types.h"

class Optical_props_arry_gpu;

class Rte_sw_gpu
{
    public:
        void rte_sw(
                const std::unique_ptr<Optical_props_arry_gpu>& optical_props,
                const Bool top_at_1,
                const Array_gpu<Float,1>& mu0,
                const Array_gpu<Float,2>& inc_flux_dir,
                const Array_gpu<Float,2>& sfc_alb_dir,
                const Array_gpu<Float,2>& sfc_alb_dif,
                const Array_gpu<Float,2>& inc_flux_dif,
                Array_gpu<Float,3>& gpt_flux_up,
                Array_gpu<Float,3>& gpt_flux_dn,
                Array_gpu<Float,3>& gpt_flux_dir);

        // Band and broadband fluxes summed inside the solver (no g-point flux arrays); either
        // output set may be empty.  flux_dn is the total (diffuse + direct) downward flux.  With
        // p_lev and a non-empty heating_rate (ncol, nlay), heating rates [K/day] from each layer's
        // absorbed flux formed per g-point and summed in FP64 (rrtmgp_hip_sw_solver_2stream_bnd_*).
        // With coalbedo the ssa array holds the co-albedo 1 - ssa
        // (rrtmgp_hip_sw_solver_2stream_bnd_coalbedo_*).
        void rte_sw_fluxes(
                const std::unique_ptr<Optical_props_arry_gpu>& optical_props,
                const Bool top_at_1,
                const Array_gpu<Float,1>& mu0,
                const Array_gpu<Float,2>& inc_flux_dir,
                const Array_gpu<Float,2>& sfc_alb_dir,
                const Array_gpu<Float,2>& sfc_alb_dif,
                const Array_gpu<Float,2>& inc_flux_dif,
                Array_gpu<Float,3>& bnd_flux_up, Array_gpu<Float,3>& bnd_flux_dn,
                Array_gpu<Float,3>& bnd_flux_dir,
                Array_gpu<Float,2>& flux_up, Array_gpu<Float,2>& flux_dn, Array_gpu<Float,2>& flux_dir,
                const Array_gpu<Float,2>* p_lev = nullptr, Array_gpu<Float,2>* heating_rate = nullptr,
                bool coalbedo = false);

        void expand_and_transpose(
                const std::unique_ptr<Optical_props_arry_gpu>& ops,
                const Array_gpu<Float,2> arr_in,
                Array_gpu<Float,2>& arr_out);
};

#endif
