// Longwave source-function holder (reference include/Source_functions.h:73-99).
#ifndef RRTMGP_AMD_SOURCE_FUNCTIONS_H
#define RRTMGP_AMD_SOURCE_FUNCTIONS_H

#include "Array.h"
#include "Optical_props.h"
#include "types.h"

class Source_func_lw_gpu : public Optical_props_gpu
{
    public:
        Source_func_lw_gpu(const int n_col, const int n_lay, const Optical_props_gpu& optical_props);
        // Factored form (no reference counterpart): every source is a g-point's Planck fraction
        // times its band's Planck function, and the product is left unevaluated.  planck_frac
        // (n_col, n_lay, n_gpt) holds the fractions, band_lay_source (n_col, n_lay, n_band) and
        // band_lev_source (n_col, n_lay + 1, n_band) the band Planck functions; lay_source and
        // lev_source_inc/dec stay empty.  Gas_optics_rrtmgp_gpu::gas_optics_with_clouds and
        // Rte_lw_gpu::rte_lw_fluxes take either form; the longwave driver uses this one.
        Source_func_lw_gpu(const int n_col, const int n_lay, const Optical_props_gpu& optical_props,
                           bool factored);

        bool is_factored() const { return factored; }
        Array_gpu<Float,3>& get_planck_frac() { return planck_frac; }
        Array_gpu<Float,3>& get_band_lay_source() { return band_lay_source; }
        Array_gpu<Float,3>& get_band_lev_source() { return band_lev_source; }
        const Array_gpu<Float,3>& get_planck_frac() const { return planck_frac; }
        const Array_gpu<Float,3>& get_band_lay_source() const { return band_lay_source; }
        const Array_gpu<Float,3>& get_band_lev_source() const { return band_lev_source; }

        Array_gpu<Float,2>& get_sfc_source()     { return sfc_source;     }
        Array_gpu<Float,2>& get_sfc_source_jac() { return sfc_source_jac; }
        Array_gpu<Float,3>& get_lay_source()     { return lay_source;     }
        Array_gpu<Float,3>& get_lev_source_inc() { return lev_source_inc; }
        Array_gpu<Float,3>& get_lev_source_dec() { return lev_source_dec; }

        const Array_gpu<Float,2>& get_sfc_source()     const { return sfc_source;     }
        const Array_gpu<Float,2>& get_sfc_source_jac() const { return sfc_source_jac; }
        const Array_gpu<Float,3>& get_lay_source()     const { return lay_source;     }
        const Array_gpu<Float,3>& get_lev_source_inc() const { return lev_source_inc; }
        const Array_gpu<Float,3>& get_lev_source_dec() const { return lev_source_dec; }

    private:
        Array_gpu<Float,2> sfc_source;
        Array_gpu<Float,2> sfc_source_jac;
        Array_gpu<Float,3> lay_source;
        Array_gpu<Float,3> lev_source_inc;
        Array_gpu<Float,3> lev_source_dec;
        bool factored = false;
        Array_gpu<Float,3> planck_frac;
        Array_gpu<Float,3> band_lay_source;
        Array_gpu<Float,3> band_lev_source;
};

#endif
