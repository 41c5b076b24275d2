// Flux reductions (reference src_cuda/Fluxes.cu:31-136 / src/Fluxes.cpp:87-197).
#include "Fluxes.h"

#include "Optical_props.h"
#include "hip_dispatch.h"

using rrtmgp_cpp::hip_check;
using rrtmgp_cpp::stream;

Fluxes_broadband_gpu::Fluxes_broadband_gpu(const int ncol, const int nlev) :
    flux_up({ncol, nlev}), flux_dn({ncol, nlev}), flux_dn_dir({ncol, nlev}), flux_net({ncol, nlev})
{}

void Fluxes_broadband_gpu::reduce(const Array_gpu<Float,3>& up, const Array_gpu<Float,3>& dn,
                                  const std::unique_ptr<Optical_props_arry_gpu>&, const Bool)
{
    const int ncol = up.dim(1), nlev = up.dim(2), ngpt = up.dim(3);
    hip_check(K<Float>::sum_broadband(ncol, nlev, ngpt, up.ptr(), flux_up.ptr(), stream()));
    hip_check(K<Float>::sum_broadband(ncol, nlev, ngpt, dn.ptr(), flux_dn.ptr(), stream()));
    hip_check(K<Float>::net_broadband(ncol, nlev, flux_dn.ptr(), flux_up.ptr(), flux_net.ptr(), stream()));
}

void Fluxes_broadband_gpu::reduce(const Array_gpu<Float,3>& up, const Array_gpu<Float,3>& dn,
                                  const Array_gpu<Float,3>& dir,
                                  const std::unique_ptr<Optical_props_arry_gpu>& p, const Bool top_at_1)
{
    reduce(up, dn, p, top_at_1);
    hip_check(K<Float>::sum_broadband(up.dim(1), up.dim(2), up.dim(3), dir.ptr(), flux_dn_dir.ptr(), stream()));
}

Fluxes_byband_gpu::Fluxes_byband_gpu(const int ncol, const int nlev, const int nbnd) :
    Fluxes_broadband_gpu(ncol, nlev), bnd_flux_up({ncol, nlev, nbnd}), bnd_flux_dn({ncol, nlev, nbnd}),
    bnd_flux_dn_dir({ncol, nlev, nbnd}), bnd_flux_net({ncol, nlev, nbnd})
{}

void Fluxes_byband_gpu::reduce(const Array_gpu<Float,3>& up, const Array_gpu<Float,3>& dn,
                               const std::unique_ptr<Optical_props_arry_gpu>& p, const Bool top_at_1)
{
    const int ncol = up.dim(1), nlev = up.dim(2);
    const int ngpt = p->get_ngpt(), nbnd = p->get_nband();
    const int* bl = p->get_band_lims_gpoint_gpu().ptr();
    Fluxes_broadband_gpu::reduce(up, dn, p, top_at_1);
    hip_check(K<Float>::sum_byband(ncol, nlev, ngpt, nbnd, bl, up.ptr(), bnd_flux_up.ptr(), stream()));
    hip_check(K<Float>::sum_byband(ncol, nlev, ngpt, nbnd, bl, dn.ptr(), bnd_flux_dn.ptr(), stream()));
    hip_check(K<Float>::net_broadband(ncol, nlev * nbnd, bnd_flux_dn.ptr(), bnd_flux_up.ptr(),
                                      bnd_flux_net.ptr(), stream()));
}

void Fluxes_byband_gpu::reduce(const Array_gpu<Float,3>& up, const Array_gpu<Float,3>& dn,
                               const Array_gpu<Float,3>& dir,
                               const std::unique_ptr<Optical_props_arry_gpu>& p, const Bool top_at_1)
{
    Fluxes_broadband_gpu::reduce(up, dn, dir, p, top_at_1);
    reduce(up, dn, p, top_at_1);
    hip_check(K<Float>::sum_byband(up.dim(1), up.dim(2), p->get_ngpt(), p->get_nband(),
                                   p->get_band_lims_gpoint_gpu().ptr(), dir.ptr(), bnd_flux_dn_dir.ptr(), stream()));
}

void compute_heating_rate(const Array_gpu<Float,2>& up, const Array_gpu<Float,2>& dn,
                          const Array_gpu<Float,2>& p_lev, Array_gpu<Float,2>& hr)
{
    hip_check(K<Float>::heating_rate(up.dim(1), up.dim(2) - 1, up.ptr(), dn.ptr(), p_lev.ptr(), hr.ptr(), stream()));
}
