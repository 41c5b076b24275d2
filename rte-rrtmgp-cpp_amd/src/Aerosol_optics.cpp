// Aerosol optics (reference src_cuda/Aerosol_optics.cu, CPU src/Aerosol_optics.cpp:7-224).
// fill_aerosols_3d (:153-165) is not needed: (1, nlay) profiles are read with a zero column
// stride by the kernel instead of being expanded to (ncol, nlay).
#include "Aerosol_optics.h"

#include <stdexcept>
#include <string>

#include "hip_dispatch.h"

using rrtmgp_cpp::hip_check;
using rrtmgp_cpp::stream;

Aerosol_optics_gpu::Aerosol_optics_gpu(
        const Array<Float,2>& band_lims_wvn, const Array<Float,1>& rh_upper,
        const Array<Float,2>& mext_phobic, const Array<Float,2>& ssa_phobic, const Array<Float,2>& g_phobic,
        const Array<Float,3>& mext_philic, const Array<Float,3>& ssa_philic, const Array<Float,3>& g_philic) :
    Optical_props_gpu(band_lims_wvn)
{
    const int nbnd = band_lims_wvn.dim(2);
    nhum = rh_upper.dim(1);
    nphobic = mext_phobic.dim(2);
    nphilic = mext_philic.dim(3);
    if (mext_phobic.dim(1) != nbnd || mext_philic.dim(1) != nbnd || mext_philic.dim(2) != nhum)
        throw std::runtime_error("Aerosol_optics: table dimensions do not match the bands and humidity classes");
    rh_upper_gpu = rh_upper;
    mext_phobic_gpu = mext_phobic; ssa_phobic_gpu = ssa_phobic; g_phobic_gpu = g_phobic;
    mext_philic_gpu = mext_philic; ssa_philic_gpu = ssa_philic; g_philic_gpu = g_philic;
}

void Aerosol_optics_gpu::compute(const Aerosol_concs_gpu& concs, const Array_gpu<Float,2>& rh,
                                 const Array_gpu<Float,2>& plev, bool delta, Optical_props_2str_gpu& p,
                                 bool coalbedo)
{
    const int ncol = rh.dim(1), nlay = rh.dim(2);
    if (plev.dim(1) != ncol || plev.dim(2) != nlay + 1)
        throw std::runtime_error("Aerosol_optics: plev must be (ncol, nlay+1)");
    if (p.get_ncol() != ncol || p.get_nlay() != nlay || p.get_tau().dim(3) != get_nband())
        throw std::runtime_error("Aerosol_optics: optical properties are not (ncol, nlay, nbnd)");
    rrtmgp_gas_vmr mix{};
    mix.ngas = 11;
    for (int i = 1; i <= 11; ++i)
    {
        const std::string name = (i < 10 ? "aermr0" : "aermr") + std::to_string(i);
        if (!concs.exists(name)) throw std::runtime_error("Aerosol_optics: " + name + " is not set");
        const Array_gpu<Float,2>& v = concs.get_vmr(name);
        mix.vmr[i - 1] = v.ptr();
        mix.dim1[i - 1] = v.dim(1);
        mix.dim2[i - 1] = v.dim(2);
    }
    hip_check(K<Float>::aerosol_optics(
            ncol, nlay, get_nband(), nhum, nphobic, nphilic, &mix, rh.ptr(), plev.ptr(), rh_upper_gpu.ptr(),
            mext_phobic_gpu.ptr(), ssa_phobic_gpu.ptr(), g_phobic_gpu.ptr(), mext_philic_gpu.ptr(),
            ssa_philic_gpu.ptr(), g_philic_gpu.ptr(), delta ? 1 : 0, coalbedo ? 1 : 0, p.get_tau().ptr(),
            p.get_ssa().ptr(), p.get_g().ptr(), stream()));
}

void Aerosol_optics_gpu::aerosol_optics(Aerosol_concs_gpu& concs, const Array_gpu<Float,2>& rh,
                                        const Array_gpu<Float,2>& plev, Optical_props_2str_gpu& p)
{
    compute(concs, rh, plev, false, p);
}
