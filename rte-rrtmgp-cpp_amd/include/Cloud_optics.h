// Cloud optics lookup tables on the GPU (reference include/Cloud_optics.h:84-128).
#ifndef RRTMGP_AMD_CLOUD_OPTICS_H
#define RRTMGP_AMD_CLOUD_OPTICS_H

#include "Array.h"
#include "Optical_props.h"
#include "types.h"

class Cloud_optics_gpu : public Optical_props_gpu
{
    public:
        Cloud_optics_gpu(
                const Array<Float,2>& band_lims_wvn,
                const Float radliq_lwr, const Float radliq_upr, const Float radliq_fac,
                const Float radice_lwr, const Float radice_upr, const Float radice_fac,
                const Array<Float,2>& lut_extliq, const Array<Float,2>& lut_ssaliq, const Array<Float,2>& lut_asyliq,
                const Array<Float,3>& lut_extice, const Array<Float,3>& lut_ssaice, const Array<Float,3>& lut_asyice);

        void cloud_optics(
                const Array_gpu<Float,2>& clwp, const Array_gpu<Float,2>& ciwp,
                const Array_gpu<Float,2>& reliq, const Array_gpu<Float,2>& reice,
                Optical_props_1scl_gpu& optical_props);

        void cloud_optics(
                const Array_gpu<Float,2>& clwp, const Array_gpu<Float,2>& ciwp,
                const Array_gpu<Float,2>& reliq, const Array_gpu<Float,2>& reice,
                Optical_props_2str_gpu& optical_props);

        // Two-stream variant with delta_scale() applied in the same pass.
        void cloud_optics_delta_scaled(
                const Array_gpu<Float,2>& clwp, const Array_gpu<Float,2>& ciwp,
                const Array_gpu<Float,2>& reliq, const Array_gpu<Float,2>& reice,
                Optical_props_2str_gpu& optical_props);

        // Two-stream variant writing the co-albedo 1 - ssa in place of ssa (delta-scaled when
        // delta_scale), for the shortwave driver's co-albedo chain.
        void cloud_optics_coalbedo(
                const Array_gpu<Float,2>& clwp, const Array_gpu<Float,2>& ciwp,
                const Array_gpu<Float,2>& reliq, const Array_gpu<Float,2>& reice,
                bool delta_scale, Optical_props_2str_gpu& optical_props);

    private:
        void run(const Array_gpu<Float,2>& clwp, const Array_gpu<Float,2>& ciwp,
                 const Array_gpu<Float,2>& reliq, const Array_gpu<Float,2>& reice,
                 int two_stream, int delta, Array_gpu<Float,3>& tau, Float* ssa, Float* g);

        int liq_nsteps, ice_nsteps;
        Float liq_step_size, ice_step_size;
        Float radliq_lwr, radliq_upr, radice_lwr, radice_upr;
        Array_gpu<Float,2> lut_extliq_gpu, lut_ssaliq_gpu, lut_asyliq_gpu;
        Array_gpu<Float,2> lut_extice_gpu, lut_ssaice_gpu, lut_asyice_gpu;
};

#endif
