// Column-radiation drivers on the GPU (reference include_test/Radiation_solver.h:33-233,
// src_test/Radiation_solver.cu:419-950).  The call sequence is the reference's
// (gas optics -> cloud optics -> add_to -> rte -> flux reduction); the cloud add_to is folded
// into the gas-optics pass, the SW cloud delta-scaling into the cloud-optics pass, and band and
// broadband fluxes are summed inside the solvers (the CPU semantics of broadband output), so no
// g-point flux arrays exist.  Workspaces persist across calls; columns are processed in blocks
// of n_col_block.
#ifndef RRTMGP_AMD_RADIATION_SOLVER_H
#define RRTMGP_AMD_RADIATION_SOLVER_H

#include <memory>

#include "Aerosol_optics.h"
#include "Array.h"
#include "Cloud_optics.h"
#include "Fluxes.h"
#include "Gas_concs.h"
#include "Gas_optics_rrtmgp.h"
#include "Optical_props.h"
#include "Source_functions.h"
#include "types.h"

class Radiation_solver_longwave_gpu
{
    public:
        Radiation_solver_longwave_gpu(std::shared_ptr<Gas_optics_rrtmgp_gpu> kdist,
                                      std::shared_ptr<Cloud_optics_gpu> cloud_optics);

        // Arrays are device arrays of the full domain: (ncol, nlay|nlev) profiles, t_sfc (ncol),
        // emis_sfc (nbnd, ncol); col_dry may be empty (computed from h2o); cloud arrays are used
        // when switch_cloud_optics.  Outputs (ncol, nlev[, nbnd]).  lw_heating_rate (ncol, nlay),
        // when given, receives heating rates [K/day] from the layers' net-flux convergence formed
        // inside the solver and summed in FP64 (an addition: the reference has no heating rates).
        void solve_gpu(
                const bool switch_cloud_optics, const bool switch_output_bnd_fluxes,
                const Gas_concs_gpu& gas_concs,
                const Array_gpu<Float,2>& p_lay, const Array_gpu<Float,2>& p_lev,
                const Array_gpu<Float,2>& t_lay, const Array_gpu<Float,2>& t_lev,
                const Array_gpu<Float,2>& col_dry, const Array_gpu<Float,1>& t_sfc,
                const Array_gpu<Float,2>& emis_sfc,
                const Array_gpu<Float,2>& lwp, const Array_gpu<Float,2>& iwp,
                const Array_gpu<Float,2>& rel, const Array_gpu<Float,2>& rei,
                Array_gpu<Float,2>& lw_flux_up, Array_gpu<Float,2>& lw_flux_dn,
                Array_gpu<Float,2>& lw_flux_net,
                Array_gpu<Float,3>& lw_bnd_flux_up, Array_gpu<Float,3>& lw_bnd_flux_dn,
                Array_gpu<Float,3>& lw_bnd_flux_net,
                Array_gpu<Float,2>* lw_heating_rate = nullptr);

        void set_n_col_block(int n) { n_col_block = n; }
        int get_n_gpt() const { return kdist->get_ngpt(); }
        int get_n_bnd() const { return kdist->get_nband(); }
        Gas_optics_rrtmgp_gpu& get_kdist() { return *kdist; }
        Cloud_optics_gpu* get_cloud_optics() { return cloud_optics.get(); }

    private:
        std::shared_ptr<Gas_optics_rrtmgp_gpu> kdist;
        std::shared_ptr<Cloud_optics_gpu> cloud_optics;
        int n_col_block = 1 << 30;
        int ws_ncol = -1, ws_nlay = -1, ws_bnd = -1;
        std::unique_ptr<Optical_props_arry_gpu> optical_props;
        std::unique_ptr<Optical_props_1scl_gpu> cloud_props;
        std::unique_ptr<Source_func_lw_gpu> sources;
        Array_gpu<Float,2> blk_flux_up, blk_flux_dn;
        Array_gpu<Float,3> blk_bnd_up, blk_bnd_dn;
        Array_gpu<Float,2> blk_hr;
        Array_gpu<Float,2> col_dry_ws;
};

class Radiation_solver_shortwave_gpu
{
    public:
        Radiation_solver_shortwave_gpu(std::shared_ptr<Gas_optics_rrtmgp_gpu> kdist,
                                       std::shared_ptr<Cloud_optics_gpu> cloud_optics,
                                       std::shared_ptr<Aerosol_optics_gpu> aerosol_optics = nullptr);

        // sfc_alb_dir/dif (nbnd, ncol), tsi_scaling and mu0 (ncol); sw_heating_rate as for the
        // longwave solver.
        void solve_gpu(
                const bool switch_cloud_optics, const bool switch_output_bnd_fluxes,
                const bool switch_delta_cloud,
                const Gas_concs_gpu& gas_concs,
                const Array_gpu<Float,2>& p_lay, const Array_gpu<Float,2>& p_lev,
                const Array_gpu<Float,2>& t_lay, const Array_gpu<Float,2>& t_lev,
                const Array_gpu<Float,2>& col_dry,
                const Array_gpu<Float,2>& sfc_alb_dir, const Array_gpu<Float,2>& sfc_alb_dif,
                const Array_gpu<Float,1>& tsi_scaling, const Array_gpu<Float,1>& mu0,
                const Array_gpu<Float,2>& lwp, const Array_gpu<Float,2>& iwp,
                const Array_gpu<Float,2>& rel, const Array_gpu<Float,2>& rei,
                Array_gpu<Float,2>& sw_flux_up, Array_gpu<Float,2>& sw_flux_dn,
                Array_gpu<Float,2>& sw_flux_dn_dir, Array_gpu<Float,2>& sw_flux_net,
                Array_gpu<Float,3>& sw_bnd_flux_up, Array_gpu<Float,3>& sw_bnd_flux_dn,
                Array_gpu<Float,3>& sw_bnd_flux_dn_dir, Array_gpu<Float,3>& sw_bnd_flux_net,
                Array_gpu<Float,2>* sw_heating_rate = nullptr);

        // With aerosols (src_test/Radiation_solver.cu:794-810): after the clouds, the aerosol
        // optical properties from rh (ncol, nlay) and aerosol_concs (aermr01..aermr11), delta-scaled
        // when switch_delta_aerosol, are added to the gas + cloud properties.  The aerosol
        // concentrations of a column block are those of its own columns (the reference's driver
        // takes columns 1..n of every block, src_test/Radiation_solver.cpp:769, which only
        // matters for column-varying aerosols and blocked runs).
        void solve_gpu(
                const bool switch_cloud_optics, const bool switch_aerosol_optics,
                const bool switch_output_bnd_fluxes, const bool switch_delta_cloud,
                const bool switch_delta_aerosol,
                const Gas_concs_gpu& gas_concs,
                const Array_gpu<Float,2>& p_lay, const Array_gpu<Float,2>& p_lev,
                const Array_gpu<Float,2>& t_lay, const Array_gpu<Float,2>& t_lev,
                const Array_gpu<Float,2>& col_dry,
                const Array_gpu<Float,2>& sfc_alb_dir, const Array_gpu<Float,2>& sfc_alb_dif,
                const Array_gpu<Float,1>& tsi_scaling, const Array_gpu<Float,1>& mu0,
                const Array_gpu<Float,2>& lwp, const Array_gpu<Float,2>& iwp,
                const Array_gpu<Float,2>& rel, const Array_gpu<Float,2>& rei,
                const Array_gpu<Float,2>& rh, const Aerosol_concs_gpu& aerosol_concs,
                Array_gpu<Float,2>& sw_flux_up, Array_gpu<Float,2>& sw_flux_dn,
                Array_gpu<Float,2>& sw_flux_dn_dir, Array_gpu<Float,2>& sw_flux_net,
                Array_gpu<Float,3>& sw_bnd_flux_up, Array_gpu<Float,3>& sw_bnd_flux_dn,
                Array_gpu<Float,3>& sw_bnd_flux_dn_dir, Array_gpu<Float,3>& sw_bnd_flux_net,
                Array_gpu<Float,2>* sw_heating_rate = nullptr);

        void set_n_col_block(int n) { n_col_block = n; }
        int get_n_gpt() const { return kdist->get_ngpt(); }
        int get_n_bnd() const { return kdist->get_nband(); }
        Float get_tsi() const { return kdist->get_tsi(); }
        Gas_optics_rrtmgp_gpu& get_kdist() { return *kdist; }
        Cloud_optics_gpu* get_cloud_optics() { return cloud_optics.get(); }
        Aerosol_optics_gpu* get_aerosol_optics() { return aerosol_optics.get(); }

    private:
        std::shared_ptr<Gas_optics_rrtmgp_gpu> kdist;
        std::shared_ptr<Cloud_optics_gpu> cloud_optics;
        std::shared_ptr<Aerosol_optics_gpu> aerosol_optics;
        int n_col_block = 1 << 30;
        int ws_ncol = -1, ws_nlay = -1, ws_bnd = -1;
        std::unique_ptr<Optical_props_arry_gpu> optical_props;
        std::unique_ptr<Optical_props_2str_gpu> cloud_props;
        std::unique_ptr<Optical_props_2str_gpu> aerosol_props;
        Array_gpu<Float,2> toa_src;
        Array_gpu<Float,2> blk_flux_up, blk_flux_dn, blk_flux_dir;
        Array_gpu<Float,3> blk_bnd_up, blk_bnd_dn, blk_bnd_dir;
        Array_gpu<Float,2> blk_hr;
        Array_gpu<Float,2> col_dry_ws;
};

#endif
