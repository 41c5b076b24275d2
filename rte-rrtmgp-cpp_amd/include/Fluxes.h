// Flux reductions (reference include/Fluxes.h:100-211): broadband and by-band sums of g-point
// fluxes; by-band uses the CPU semantics of mo_fluxes_byband_kernels.F90 (inclusive 1-based
// band limits).
#ifndef RRTMGP_AMD_FLUXES_H
#define RRTMGP_AMD_FLUXES_H

#include <memory>
#include <stdexcept>

#include "Array.h"
#include "types.h"

class Optical_props_arry_gpu;

class Fluxes_gpu
{
    public:
        virtual ~Fluxes_gpu() {}
        virtual void reduce(const Array_gpu<Float,3>& gpt_flux_up, const Array_gpu<Float,3>& gpt_flux_dn,
                            const std::unique_ptr<Optical_props_arry_gpu>& optical_props, const Bool top_at_1) = 0;
        virtual void reduce(const Array_gpu<Float,3>& gpt_flux_up, const Array_gpu<Float,3>& gpt_flux_dn,
                            const Array_gpu<Float,3>& gpt_flux_dn_dir,
                            const std::unique_ptr<Optical_props_arry_gpu>& optical_props, const Bool top_at_1) = 0;
};

class Fluxes_broadband_gpu : public Fluxes_gpu
{
    public:
        Fluxes_broadband_gpu(const int ncol, const int nlev);
        virtual ~Fluxes_broadband_gpu() {}

        virtual void reduce(const Array_gpu<Float,3>& gpt_flux_up, const Array_gpu<Float,3>& gpt_flux_dn,
                            const std::unique_ptr<Optical_props_arry_gpu>& optical_props, const Bool top_at_1);
        virtual void reduce(const Array_gpu<Float,3>& gpt_flux_up, const Array_gpu<Float,3>& gpt_flux_dn,
                            const Array_gpu<Float,3>& gpt_flux_dn_dir,
                            const std::unique_ptr<Optical_props_arry_gpu>& optical_props, const Bool top_at_1);

        Array_gpu<Float,2>& get_flux_up    () { return flux_up;     }
        Array_gpu<Float,2>& get_flux_dn    () { return flux_dn;     }
        Array_gpu<Float,2>& get_flux_dn_dir() { return flux_dn_dir; }
        Array_gpu<Float,2>& get_flux_net   () { return flux_net;    }

        virtual Array_gpu<Float,3>& get_bnd_flux_up    () { throw std::runtime_error("Band fluxes are not available"); }
        virtual Array_gpu<Float,3>& get_bnd_flux_dn    () { throw std::runtime_error("Band fluxes are not available"); }
        virtual Array_gpu<Float,3>& get_bnd_flux_dn_dir() { throw std::runtime_error("Band fluxes are not available"); }
        virtual Array_gpu<Float,3>& get_bnd_flux_net   () { throw std::runtime_error("Band fluxes are not available"); }

    private:
        Array_gpu<Float,2> flux_up;
        Array_gpu<Float,2> flux_dn;
        Array_gpu<Float,2> flux_dn_dir;
        Array_gpu<Float,2> flux_net;
};

class Fluxes_byband_gpu : public Fluxes_broadband_gpu
{
    public:
        Fluxes_byband_gpu(const int ncol, const int nlev, const int nbnd);
        virtual ~Fluxes_byband_gpu() {}

        virtual void reduce(const Array_gpu<Float,3>& gpt_flux_up, const Array_gpu<Float,3>& gpt_flux_dn,
                            const std::unique_ptr<Optical_props_arry_gpu>& optical_props, const Bool top_at_1);
        virtual void reduce(const Array_gpu<Float,3>& gpt_flux_up, const Array_gpu<Float,3>& gpt_flux_dn,
                            const Array_gpu<Float,3>& gpt_flux_dn_dir,
                            const std::unique_ptr<Optical_props_arry_gpu>& optical_props, const Bool top_at_1);

        Array_gpu<Float,3>& get_bnd_flux_up    () { return bnd_flux_up;     }
        Array_gpu<Float,3>& get_bnd_flux_dn    () { return bnd_flux_dn;     }
        Array_gpu<Float,3>& get_bnd_flux_dn_dir() { return bnd_flux_dn_dir; }
        Array_gpu<Float,3>& get_bnd_flux_net   () { return bnd_flux_net;    }

    private:
        Array_gpu<Float,3> bnd_flux_up;
        Array_gpu<Float,3> bnd_flux_dn;
        Array_gpu<Float,3> bnd_flux_dn_dir;
        Array_gpu<Float,3> bnd_flux_net;
};

// Heating rate [K/day] per layer from level fluxes and pressures (no reference counterpart);
// the net-flux and pressure differences are formed in double precision.
void compute_heating_rate(const Array_gpu<Float,2>& flux_up, const Array_gpu<Float,2>& flux_dn,
                          const Array_gpu<Float,2>& p_lev, Array_gpu<Float,2>& heating_rate);

#endif
