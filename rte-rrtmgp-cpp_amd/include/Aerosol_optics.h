// Aerosol optics lookup tables on the GPU (reference include/Aerosol_optics.h:17-101).
// Aerosol concentrations are held like gas concentrations (Aerosol_concs_gpu = Gas_concs_gpu, as
// in the reference): mass mixing ratios aermr01..aermr11, each (ncol, nlay) or a (1, nlay) profile.
#ifndef RRTMGP_AMD_AEROSOL_OPTICS_H
#define RRTMGP_AMD_AEROSOL_OPTICS_H

#include "Array.h"
#include "Gas_concs.h"
#include "Optical_props.h"
#include "types.h"

using Aerosol_concs_gpu = Gas_concs_gpu;

class Aerosol_optics_gpu : public Optical_props_gpu
{
    public:
        // Tables as load_and_init_aerosol_optics reads them (src_test/Radiation_solver.cpp:329-366):
        // rh_upper (nhum), mext/ssa/g_phobic (nbnd, nphobic), mext/ssa/g_philic (nbnd, nhum, nphilic).
        Aerosol_optics_gpu(
                const Array<Float,2>& band_lims_wvn, const Array<Float,1>& rh_upper,
                const Array<Float,2>& mext_phobic, const Array<Float,2>& ssa_phobic, const Array<Float,2>& g_phobic,
                const Array<Float,3>& mext_philic, const Array<Float,3>& ssa_philic, const Array<Float,3>& g_philic);

        // rh (ncol, nlay), plev (ncol, nlay+1); optical_props (ncol, nlay, nbnd).
        void aerosol_optics(
                Aerosol_concs_gpu& aerosol_concs,
                const Array_gpu<Float,2>& rh, const Array_gpu<Float,2>& plev,
                Optical_props_2str_gpu& optical_props);

        // The same, with delta_scale() applied in the same pass when delta_scale is set, and the
        // co-albedo 1 - ssa written in place of ssa when coalbedo is set; the concentrations are
        // only read.
        void compute(
                const Aerosol_concs_gpu& aerosol_concs,
                const Array_gpu<Float,2>& rh, const Array_gpu<Float,2>& plev,
                bool delta_scale, Optical_props_2str_gpu& optical_props, bool coalbedo = false);

    private:

        int nhum, nphobic, nphilic;
        Array_gpu<Float,1> rh_upper_gpu;
        Array_gpu<Float,2> mext_phobic_gpu, ssa_phobic_gpu, g_phobic_gpu;
        Array_gpu<Float,3> mext_philic_gpu, ssa_philic_gpu, g_philic_gpu;
};

#endif
