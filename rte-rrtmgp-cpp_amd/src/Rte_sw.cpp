// Shortwave driver (reference src_cuda/Rte_sw.cu:116-161, CPU src/Rte_sw.cpp:97-145).
#include "Rte_sw.h"

#include "Optical_props.h"
#include "hip_dispatch.h"

using rrtmgp_cpp::hip_check;
using rrtmgp_cpp::stream;

void Rte_sw_gpu::rte_sw(
        const std::unique_ptr<Optical_props_arry_gpu>& optical_props,
        const Bool top_at_1,
        const Array_gpu<Float,1>& mu0,
        const Array_gpu<Float,2>& inc_flux_dir,
        const Array_gpu<Float,2>& sfc_alb_dir,
        const Array_gpu<Float,2>& sfc_alb_dif,
        const Array_gpu<Float,2>& inc_flux_dif,
        Array_gpu<Float,3>& gpt_flux_up,
        Array_gpu<Float,3>& gpt_flux_dn,
        Array_gpu<Float,3>& gpt_flux_dir)
{
    const int ncol = optical_props->get_ncol();
    const int nlay = optical_props->get_nlay();
    const int ngpt = optical_props->get_ngpt();

    Array_gpu<Float,2> sfc_alb_dir_gpt({ncol, ngpt});
    Array_gpu<Float,2> sfc_alb_dif_gpt({ncol, ngpt});
    expand_and_transpose(optical_props, sfc_alb_dir, sfc_alb_dir_gpt);
    expand_and_transpose(optical_props, sfc_alb_dif, sfc_alb_dif_gpt);

    const bool has_dif_bc = inc_flux_dif.size() > 0;
    const bool do_broadband = (gpt_flux_up.dim(3) == 1) && (ngpt > 1);
    hip_check(K<Float>::sw_solver_2stream(
            ncol, nlay, ngpt, top_at_1, optical_props->get_tau().ptr(), optical_props->get_ssa().ptr(),
            optical_props->get_g().ptr(), mu0.ptr(), sfc_alb_dir_gpt.ptr(), sfc_alb_dif_gpt.ptr(),
            inc_flux_dir.ptr(),
            do_broadband ? nullptr : gpt_flux_up.ptr(), do_broadband ? nullptr : gpt_flux_dn.ptr(),
            do_broadband ? nullptr : gpt_flux_dir.ptr(), has_dif_bc,
            has_dif_bc ? inc_flux_dif.ptr() : nullptr, do_broadband,
            do_broadband ? gpt_flux_up.ptr() : nullptr, do_broadband ? gpt_flux_dn.ptr() : nullptr,
            do_broadband ? gpt_flux_dir.ptr() : nullptr, stream()));
}

void Rte_sw_gpu::rte_sw_fluxes(
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
        const Array_gpu<Float,2>* p_lev, Array_gpu<Float,2>* heating_rate, bool coalbedo)
{
    const int ncol = optical_props->get_ncol();
    const int nlay = optical_props->get_nlay();
    const int ngpt = optical_props->get_ngpt();
    const int nbnd = optical_props->get_nband();
    const bool hr = heating_rate != nullptr && heating_rate->size() > 0;
    if (hr && (!p_lev || p_lev->dim(1) != ncol || p_lev->dim(2) != nlay + 1 || heating_rate->dim(1) != ncol ||
               heating_rate->dim(2) != nlay))
        throw std::runtime_error("rte_sw: heating rates need p_lev (ncol, nlev) and heating_rate (ncol, nlay)");
    if (sfc_alb_dir.dim(1) != nbnd || sfc_alb_dir.dim(2) != ncol || sfc_alb_dif.dim(1) != nbnd ||
        sfc_alb_dif.dim(2) != ncol)
        throw std::runtime_error("rte_sw: surface albedos must be (nband, ncol)");

    // The band-summed solver reads the albedos per band: nothing is expanded per call.
    const bool has_dif_bc = inc_flux_dif.size() > 0;
    const bool bnd = bnd_flux_up.size() > 0;
    const bool bb = flux_up.size() > 0;
    auto solver = coalbedo ? K<Float>::sw_solver_2stream_bnd_coalbedo : K<Float>::sw_solver_2stream_bnd;
    hip_check(solver(
            ncol, nlay, ngpt, top_at_1, optical_props->get_tau().ptr(), optical_props->get_ssa().ptr(),
            optical_props->get_g().ptr(), mu0.ptr(), sfc_alb_dir.ptr(), sfc_alb_dif.ptr(),
            inc_flux_dir.ptr(), has_dif_bc, has_dif_bc ? inc_flux_dif.ptr() : nullptr, nbnd,
            optical_props->get_band_lims_gpoint().ptr(), bnd ? bnd_flux_up.ptr() : nullptr,
            bnd ? bnd_flux_dn.ptr() : nullptr, bnd ? bnd_flux_dir.ptr() : nullptr,
            bb ? flux_up.ptr() : nullptr, bb ? flux_dn.ptr() : nullptr, bb ? flux_dir.ptr() : nullptr,
            hr ? p_lev->ptr() : nullptr, hr ? heating_rate->ptr() : nullptr, stream()));
}

void Rte_sw_gpu::expand_and_transpose(
        const std::unique_ptr<Optical_props_arry_gpu>& ops,
        const Array_gpu<Float,2> arr_in,
        Array_gpu<Float,2>& arr_out)
{
    hip_check(K<Float>::expand_and_transpose(arr_in.dim(2), ops->get_nband(), ops->get_ngpt(),
                                             ops->get_band_lims_gpoint_gpu().ptr(), arr_in.ptr(),
                                             arr_out.ptr(), stream()));
}
