// Cloud optics (reference src_cuda/Cloud_optics.cu:130-329, CPU src/Cloud_optics.cpp:29-232).
#include "Cloud_optics.h"

#include "hip_dispatch.h"

using rrtmgp_cpp::hip_check;
using rrtmgp_cpp::stream;

Cloud_optics_gpu::Cloud_optics_gpu(
        const Array<Float,2>& band_lims_wvn,
        const Float radliq_lwr, const Float radliq_upr, const Float radliq_fac,
        const Float radice_lwr, const Float radice_upr, const Float radice_fac,
        const Array<Float,2>& lut_extliq, const Array<Float,2>& lut_ssaliq, const Array<Float,2>& lut_asyliq,
        const Array<Float,3>& lut_extice, const Array<Float,3>& lut_ssaice, const Array<Float,3>& lut_asyice) :
    Optical_props_gpu(band_lims_wvn)
{
    const int nsize_liq = lut_extliq.dim(1);
    const int nsize_ice = lut_extice.dim(1);
    liq_nsteps = nsize_liq;
    ice_nsteps = nsize_ice;
    liq_step_size = (radliq_upr - radliq_lwr) / (nsize_liq - Float(1.));
    ice_step_size = (radice_upr - radice_lwr) / (nsize_ice - Float(1.));
    this->radliq_lwr = radliq_lwr; this->radliq_upr = radliq_upr;
    this->radice_lwr = radice_lwr; this->radice_upr = radice_upr;
    lut_extliq_gpu = lut_extliq; lut_ssaliq_gpu = lut_ssaliq; lut_asyliq_gpu = lut_asyliq;

    // Intermediately rough ice particles, icergh = 2 (src/Cloud_optics.cpp:61-68).
    constexpr int icergh = 2;
    Array<Float,2> ext({lut_extice.dim(1), lut_extice.dim(2)});
    Array<Float,2> ssa({lut_ssaice.dim(1), lut_ssaice.dim(2)});
    Array<Float,2> asy({lut_asyice.dim(1), lut_asyice.dim(2)});
    for (int ibnd = 1; ibnd <= lut_extice.dim(2); ++ibnd)
        for (int isize = 1; isize <= lut_extice.dim(1); ++isize)
        {
            ext({isize, ibnd}) = lut_extice({isize, ibnd, icergh});
            ssa({isize, ibnd}) = lut_ssaice({isize, ibnd, icergh});
            asy({isize, ibnd}) = lut_asyice({isize, ibnd, icergh});
        }
    lut_extice_gpu = ext; lut_ssaice_gpu = ssa; lut_asyice_gpu = asy;
    (void)radliq_fac; (void)radice_fac;
}

void Cloud_optics_gpu::run(const Array_gpu<Float,2>& clwp, const Array_gpu<Float,2>& ciwp,
                           const Array_gpu<Float,2>& reliq, const Array_gpu<Float,2>& reice,
                           int two_stream, int delta, Array_gpu<Float,3>& tau, Float* ssa, Float* g)
{
    hip_check(K<Float>::cloud_optics(
            clwp.dim(1), clwp.dim(2), get_nband(), liq_nsteps, liq_step_size, radliq_lwr,
            lut_extliq_gpu.ptr(), lut_ssaliq_gpu.ptr(), lut_asyliq_gpu.ptr(), ice_nsteps, ice_step_size,
            radice_lwr, lut_extice_gpu.ptr(), lut_ssaice_gpu.ptr(), lut_asyice_gpu.ptr(), clwp.ptr(),
            ciwp.ptr(), reliq.ptr(), reice.ptr(), two_stream, delta, tau.ptr(), ssa, g, stream()));
}

void Cloud_optics_gpu::cloud_optics(const Array_gpu<Float,2>& clwp, const Array_gpu<Float,2>& ciwp,
                                    const Array_gpu<Float,2>& reliq, const Array_gpu<Float,2>& reice,
                                    Optical_props_1scl_gpu& p)
{
    run(clwp, ciwp, reliq, reice, 0, 0, p.get_tau(), nullptr, nullptr);
}

void Cloud_optics_gpu::cloud_optics(const Array_gpu<Float,2>& clwp, const Array_gpu<Float,2>& ciwp,
                                    const Array_gpu<Float,2>& reliq, const Array_gpu<Float,2>& reice,
                                    Optical_props_2str_gpu& p)
{
    run(clwp, ciwp, reliq, reice, 1, 0, p.get_tau(), p.get_ssa().ptr(), p.get_g().ptr());
}

void Cloud_optics_gpu::cloud_optics_delta_scaled(const Array_gpu<Float,2>& clwp, const Array_gpu<Float,2>& ciwp,
                                                 const Array_gpu<Float,2>& reliq, const Array_gpu<Float,2>& reice,
                                                 Optical_props_2str_gpu& p)
{
    run(clwp, ciwp, reliq, reice, 1, 1, p.get_tau(), p.get_ssa().ptr(), p.get_g().ptr());
}

void Cloud_optics_gpu::cloud_optics_coalbedo(const Array_gpu<Float,2>& clwp, const Array_gpu<Float,2>& ciwp,
                                             const Array_gpu<Float,2>& reliq, const Array_gpu<Float,2>& reice,
                                             bool delta_scale, Optical_props_2str_gpu& p)
{
    run(clwp, ciwp, reliq, reice, 2, delta_scale ? 1 : 0, p.get_tau(), p.get_ssa().ptr(), p.get_g().ptr());
}
