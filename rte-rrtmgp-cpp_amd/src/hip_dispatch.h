// Maps the class layer's Float onto the _f32 / _f64 entry points of rrtmgp_hip.h.
#ifndef RRTMGP_AMD_HIP_DISPATCH_H
#define RRTMGP_AMD_HIP_DISPATCH_H

#include "rrtmgp_hip.h"

template<typename T> struct K;

#define RRTMGP_K_ENTRIES(S)                                              \
    static constexpr auto gas_optics_lw = rrtmgp_hip_gas_optics_lw_##S;  \
    static constexpr auto kdist_pack = rrtmgp_hip_kdist_pack_##S;  \
    static constexpr auto gas_optics_lw_bands = rrtmgp_hip_gas_optics_lw_bands_##S; \
    static constexpr auto gas_optics_lw_tiles = rrtmgp_hip_gas_optics_lw_tiles_##S; \
    static constexpr auto gas_optics_lw_factored = rrtmgp_hip_gas_optics_lw_factored_##S; \
    static constexpr auto lw_solver_noscat_bnd_factored = rrtmgp_hip_lw_solver_noscat_bnd_factored_##S; \
    static constexpr auto gas_optics_sw_tiles = rrtmgp_hip_gas_optics_sw_tiles_##S; \
    static constexpr auto gas_optics_sw_bands = rrtmgp_hip_gas_optics_sw_bands_##S; \
    static constexpr auto gas_optics_lw_adjoint = rrtmgp_hip_gas_optics_lw_adjoint_##S; \
    static constexpr auto gas_optics_sw_adjoint = rrtmgp_hip_gas_optics_sw_adjoint_##S; \
    static constexpr auto get_col_dry_adjoint = rrtmgp_hip_get_col_dry_adjoint_##S; \
    static constexpr auto lw_solver_noscat_adjoint = rrtmgp_hip_lw_solver_noscat_adjoint_##S; \
    static constexpr auto sw_solver_2stream_adjoint = rrtmgp_hip_sw_solver_2stream_adjoint_##S; \
    static constexpr auto adjoint_col_chain = rrtmgp_hip_adjoint_col_chain_##S; \
    static constexpr auto adjoint_levels = rrtmgp_hip_adjoint_levels_##S; \
    static constexpr auto gpt_to_band = rrtmgp_hip_gpt_to_band_##S; \
    static constexpr auto gas_optics_sw = rrtmgp_hip_gas_optics_sw_##S;  \
    static constexpr auto gas_optics_sw_coalbedo = rrtmgp_hip_gas_optics_sw_coalbedo_##S; \
    static constexpr auto interpolation = rrtmgp_hip_interpolation_##S;  \
    static constexpr auto compute_tau_absorption = rrtmgp_hip_compute_tau_absorption_##S; \
    static constexpr auto compute_tau_rayleigh = rrtmgp_hip_compute_tau_rayleigh_##S;     \
    static constexpr auto combine_abs_and_rayleigh = rrtmgp_hip_combine_abs_and_rayleigh_##S; \
    static constexpr auto compute_planck_source = rrtmgp_hip_compute_planck_source_##S;   \
    static constexpr auto get_col_dry = rrtmgp_hip_get_col_dry_##S;      \
    static constexpr auto fill_col_gas = rrtmgp_hip_fill_col_gas_##S;    \
    static constexpr auto cloud_optics = rrtmgp_hip_cloud_optics_##S;    \
    static constexpr auto aerosol_optics = rrtmgp_hip_aerosol_optics_##S; \
    static constexpr auto increment = rrtmgp_hip_increment_##S;          \
    static constexpr auto inc_bybnd = rrtmgp_hip_inc_bybnd_##S;          \
    static constexpr auto inc_bybnd_coalbedo = rrtmgp_hip_inc_bybnd_coalbedo_##S; \
    static constexpr auto delta_scale_2str = rrtmgp_hip_delta_scale_2str_##S; \
    static constexpr auto apply_bc = rrtmgp_hip_apply_bc_##S;            \
    static constexpr auto expand_and_transpose = rrtmgp_hip_expand_and_transpose_##S; \
    static constexpr auto lw_secants_array = rrtmgp_hip_lw_secants_array_##S; \
    static constexpr auto lw_solver_noscat = rrtmgp_hip_lw_solver_noscat_##S; \
    static constexpr auto sw_solver_2stream = rrtmgp_hip_sw_solver_2stream_##S; \
    static constexpr auto lw_solver_noscat_bnd = rrtmgp_hip_lw_solver_noscat_bnd_##S; \
    static constexpr auto sw_solver_2stream_bnd = rrtmgp_hip_sw_solver_2stream_bnd_##S; \
    static constexpr auto sw_solver_2stream_bnd_coalbedo = rrtmgp_hip_sw_solver_2stream_bnd_coalbedo_##S; \
    static constexpr auto sum_broadband = rrtmgp_hip_sum_broadband_##S;  \
    static constexpr auto net_broadband = rrtmgp_hip_net_broadband_##S;  \
    static constexpr auto sum_byband = rrtmgp_hip_sum_byband_##S;        \
    static constexpr auto net_byband_full = rrtmgp_hip_net_byband_full_##S; \
    static constexpr auto heating_rate = rrtmgp_hip_heating_rate_##S;    \
    static constexpr auto count_outside = rrtmgp_hip_count_outside_##S;  \
    static constexpr auto check_ranges = rrtmgp_hip_check_ranges_##S;    \
    static constexpr auto spread_col = rrtmgp_hip_spread_col_##S;        \
    static constexpr auto scale_by_col = rrtmgp_hip_scale_by_col_##S;    \
    static constexpr auto copy_cols = rrtmgp_hip_copy_cols_##S;

template<> struct K<float> { RRTMGP_K_ENTRIES(f32) };
template<> struct K<double> { RRTMGP_K_ENTRIES(f64) };

#endif
