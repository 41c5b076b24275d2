// Reverse-mode (adjoint) drivers of the column-radiation path.  No counterpart in the reference,
// whose drivers (src_test/Radiation_solver.cu:419-950) are forward only.
//
// For cotangents of the broadband fluxes (ncol, nlev), these compute the gradients of
//     L = <g_flux_up, flux_up> + <g_flux_dn, flux_dn> [+ <g_flux_dir, flux_dn_dir>]
// of the same all-sky (or clear-sky) problem that Radiation_solver_{longwave,shortwave}_gpu
// solves, with respect to T_lay, T_lev, T_sfc (LW), ln p_lay, p_lev, the volume mixing ratios of
// every gas of the k-distribution and the surface emissivity / albedos.  Gas optics is recomputed
// band by band (rrtmgp_hip_gas_optics_{lw,sw}_bands_*), so no (ncol, nlay, ngpt) array is kept.
// Clouds are held fixed.  col_dry is computed from h2o and p_lev as in the forward driver.
#pragma once

#include "Array.h"
#include "Cloud_optics.h"
#include "Gas_concs.h"
#include "Gas_optics_rrtmgp.h"

struct Lw_adjoint_out
{
    Array_gpu<Float,2> d_t_lay;      // (ncol, nlay)
    Array_gpu<Float,2> d_t_lev;      // (ncol, nlev)
    Array_gpu<Float,1> d_t_sfc;      // (ncol)
    Array_gpu<Float,2> d_lnp_lay;    // (ncol, nlay): with respect to ln p_lay
    Array_gpu<Float,2> d_p_lev;      // (ncol, nlev): through col_dry
    Array_gpu<Float,3> d_vmr;        // (ncol, nlay, ngas): the k-distribution's gas order
    Array_gpu<Float,2> d_emis_sfc;   // (nbnd, ncol)
};

struct Sw_adjoint_out
{
    Array_gpu<Float,2> d_t_lay;
    Array_gpu<Float,2> d_lnp_lay;
    Array_gpu<Float,2> d_p_lev;
    Array_gpu<Float,3> d_vmr;
    Array_gpu<Float,2> d_sfc_alb_dir;   // (nbnd, ncol)
    Array_gpu<Float,2> d_sfc_alb_dif;   // (nbnd, ncol)
};

// Output arrays are allocated when empty and must otherwise have the shapes above; they are
// overwritten.  Cloud arrays (lwp, iwp, rel, rei) empty or cloud_optics NULL: clear sky.
void lw_adjoint_gpu(Gas_optics_rrtmgp_gpu& kdist, Cloud_optics_gpu* cloud_optics, const Gas_concs_gpu& gas_concs,
                    const Array_gpu<Float,2>& p_lay, const Array_gpu<Float,2>& p_lev,
                    const Array_gpu<Float,2>& t_lay, const Array_gpu<Float,2>& t_lev,
                    const Array_gpu<Float,1>& t_sfc, const Array_gpu<Float,2>& emis_sfc,
                    const Array_gpu<Float,2>& lwp, const Array_gpu<Float,2>& iwp,
                    const Array_gpu<Float,2>& rel, const Array_gpu<Float,2>& rei,
                    const Array_gpu<Float,2>& g_flux_up, const Array_gpu<Float,2>& g_flux_dn,
                    Lw_adjoint_out& out);

void sw_adjoint_gpu(Gas_optics_rrtmgp_gpu& kdist, Cloud_optics_gpu* cloud_optics, bool delta_cloud,
                    const Gas_concs_gpu& gas_concs,
                    const Array_gpu<Float,2>& p_lay, const Array_gpu<Float,2>& p_lev,
                    const Array_gpu<Float,2>& t_lay,
                    const Array_gpu<Float,2>& sfc_alb_dir, const Array_gpu<Float,2>& sfc_alb_dif,
                    const Array_gpu<Float,1>& tsi_scaling, const Array_gpu<Float,1>& mu0,
                    const Array_gpu<Float,2>& lwp, const Array_gpu<Float,2>& iwp,
                    const Array_gpu<Float,2>& rel, const Array_gpu<Float,2>& rei,
                    const Array_gpu<Float,2>& g_flux_up, const Array_gpu<Float,2>& g_flux_dn,
                    const Array_gpu<Float,2>& g_flux_dir, Sw_adjoint_out& out);
