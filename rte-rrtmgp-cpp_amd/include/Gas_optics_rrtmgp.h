// RRTMGP gas optics on the GPU, drop-in for Gas_optics_rrtmgp_gpu
// (reference include/Gas_optics_rrtmgp.h:291-551).  The constructors take the coefficient
// tables in the layout of the reference reader (src_test/Radiation_solver.cpp:70-284) and run
// the host-side initialisation of src/Gas_optics_rrtmgp.cpp:391-742 before uploading.
// gas_optics() runs one fused device pass (interpolation, absorption, Planck or Rayleigh)
// instead of the reference's chain of materializing kernels.
#ifndef RRTMGP_AMD_GAS_OPTICS_RRTMGP_H
#define RRTMGP_AMD_GAS_OPTICS_RRTMGP_H

#include <memory>
#include <string>

#include "Array.h"
#include "Gas_optics.h"
#include "types.h"

class Gas_concs_gpu;
class Source_func_lw_gpu;
class Optical_props_arry_gpu;

class Gas_optics_rrtmgp_gpu : public Gas_optics_gpu
{
    public:
        // Longwave variant (:293-325).
        Gas_optics_rrtmgp_gpu(
                const Gas_concs_gpu& available_gases,
                const Array<std::string,1>& gas_names,
                const Array<int,3>& key_species,
                const Array<int,2>& band2gpt,
                const Array<Float,2>& band_lims_wavenum,
                const Array<Float,1>& press_ref,
                const Float press_ref_trop,
                const Array<Float,1>& temp_ref,
                const Float temp_ref_p,
                const Float temp_ref_t,
                const Array<Float,3>& vmr_ref,
                const Array<Float,4>& kmajor,
                const Array<Float,3>& kminor_lower,
                const Array<Float,3>& kminor_upper,
                const Array<std::string,1>& gas_minor,
                const Array<std::string,1>& identifier_minor,
                const Array<std::string,1>& minor_gases_lower,
                const Array<std::string,1>& minor_gases_upper,
                const Array<int,2>& minor_limits_gpt_lower,
                const Array<int,2>& minor_limits_gpt_upper,
                const Array<Bool,1>& minor_scales_with_density_lower,
                const Array<Bool,1>& minor_scales_with_density_upper,
                const Array<std::string,1>& scaling_gas_lower,
                const Array<std::string,1>& scaling_gas_upper,
                const Array<Bool,1>& scale_by_complement_lower,
                const Array<Bool,1>& scale_by_complement_upper,
                const Array<int,1>& kminor_start_lower,
                const Array<int,1>& kminor_start_upper,
                const Array<Float,2>& totplnk,
                const Array<Float,4>& planck_frac,
                const Array<Float,3>& rayl_lower,
                const Array<Float,3>& rayl_upper);

        // Shortwave variant (:328-364).
        Gas_optics_rrtmgp_gpu(
                const Gas_concs_gpu& available_gases,
                const Array<std::string,1>& gas_names,
                const Array<int,3>& key_species,
                const Array<int,2>& band2gpt,
                const Array<Float,2>& band_lims_wavenum,
                const Array<Float,1>& press_ref,
                const Float press_ref_trop,
                const Array<Float,1>& temp_ref,
                const Float temp_ref_p,
                const Float temp_ref_t,
                const Array<Float,3>& vmr_ref,
                const Array<Float,4>& kmajor,
                const Array<Float,3>& kminor_lower,
                const Array<Float,3>& kminor_upper,
                const Array<std::string,1>& gas_minor,
                const Array<std::string,1>& identifier_minor,
                const Array<std::string,1>& minor_gases_lower,
                const Array<std::string,1>& minor_gases_upper,
                const Array<int,2>& minor_limits_gpt_lower,
                const Array<int,2>& minor_limits_gpt_upper,
                const Array<Bool,1>& minor_scales_with_density_lower,
                const Array<Bool,1>& minor_scales_with_density_upper,
                const Array<std::string,1>& scaling_gas_lower,
                const Array<std::string,1>& scaling_gas_upper,
                const Array<Bool,1>& scale_by_complement_lower,
                const Array<Bool,1>& scale_by_complement_upper,
                const Array<int,1>& kminor_start_lower,
                const Array<int,1>& kminor_start_upper,
                const Array<Float,1>& solar_src_quiet,
                const Array<Float,1>& solar_src_facular,
                const Array<Float,1>& solar_src_sunspot,
                const Float tsi_default,
                const Float mg_default,
                const Float sb_default,
                const Array<Float,3>& rayl_lower,
                const Array<Float,3>& rayl_upper);

        static void get_col_dry(
                Array_gpu<Float,2>& col_dry,
                const Array_gpu<Float,2>& vmr_h2o,
                const Array_gpu<Float,2>& plev);

        bool source_is_internal() const { return (totplnk.size() > 0) && (planck_frac.size() > 0); }
        bool source_is_external() const { return (solar_source.size() > 0); }

        Float get_press_ref_min() const { return press_ref_min; }
        Float get_press_ref_max() const { return press_ref_max; }
        Float get_temp_min() const { return temp_ref_min; }
        Float get_temp_max() const { return temp_ref_max; }

        int get_nflav() const { return flavor.dim(2); }
        int get_neta() const { return kmajor_neta; }
        int get_npres() const { return kmajor_npres; }
        int get_ntemp() const { return kmajor_ntemp; }
        int get_nPlanckTemp() const { return totplnk.dim(1); }
        int get_ngas() const { return this->gas_names.dim(1); }

        Float get_tsi() const;

        // Longwave variant (:389-398).
        void gas_optics(
                const Array_gpu<Float,2>& play, const Array_gpu<Float,2>& plev,
                const Array_gpu<Float,2>& tlay, const Array_gpu<Float,1>& tsfc,
                const Gas_concs_gpu& gas_desc, std::unique_ptr<Optical_props_arry_gpu>& optical_props,
                Source_func_lw_gpu& sources, const Array_gpu<Float,2>& col_dry,
                const Array_gpu<Float,2>& tlev);

        // Shortwave variant (:401-408).
        void gas_optics(
                const Array_gpu<Float,2>& play, const Array_gpu<Float,2>& plev,
                const Array_gpu<Float,2>& tlay, const Gas_concs_gpu& gas_desc,
                std::unique_ptr<Optical_props_arry_gpu>& optical_props,
                Array_gpu<Float,2>& toa_src, const Array_gpu<Float,2>& col_dry);

        // Fused variants that also add band-resolved cloud properties (an add_to folded into
        // the gas-optics write).  cloud may be nullptr.  The shortwave one with coalbedo writes
        // the co-albedo 1 - ssa in place of ssa, and reads the cloud's as co-albedo too
        // (rrtmgp_hip_gas_optics_sw_coalbedo_*; plain gas optics only, not with a surrogate).
        void gas_optics_with_clouds(
                const Array_gpu<Float,2>& play, const Array_gpu<Float,2>& plev,
                const Array_gpu<Float,2>& tlay, const Array_gpu<Float,1>& tsfc,
                const Gas_concs_gpu& gas_desc, Optical_props_arry_gpu& optical_props,
                Source_func_lw_gpu& sources, const Array_gpu<Float,2>& col_dry,
                const Array_gpu<Float,2>& tlev, const Optical_props_arry_gpu* cloud);
        void gas_optics_with_clouds(
                const Array_gpu<Float,2>& play, const Array_gpu<Float,2>& plev,
                const Array_gpu<Float,2>& tlay, const Gas_concs_gpu& gas_desc,
                Optical_props_arry_gpu& optical_props, Array_gpu<Float,2>& toa_src,
                const Array_gpu<Float,2>& col_dry, const Optical_props_arry_gpu* cloud,
                bool coalbedo = false);

        // gas concentrations as the rrtmgp_gas_vmr descriptor of the kernels, in this
        // k-distribution's gas order
        void vmr_descriptor(const Gas_concs_gpu& gas_desc, int ncol, int nlay, rrtmgp_gas_vmr& out) const
        { make_vmr_desc(gas_desc, ncol, nlay, out); }

        // Reference CPU semantics: inputs outside the table range throw std::range_error
        // (src/Gas_optics_rrtmgp.cpp:813-826).  Off skips the checks, as the reference GPU path.
        void set_check_ranges(bool on) { check_ranges = on; }

        // Host copies of derived tables, for inspection and tests.
        const Array<std::string,1>& get_gas_names() const { return gas_names; }
        const Array<int,2>& get_flavor() const { return flavor; }
        const Array<int,2>& get_gpoint_flavor() const { return gpoint_flavor; }
        const rrtmgp_kdist_dev& device_tables() const { return dev; }
        const Array_gpu<Float,1>& solar_source_device() const { return solar_source_gpu; }
        const Array<int,2>& get_minor_limits_gpt(bool lower) const { return lower ? minor_limits_gpt_lower : minor_limits_gpt_upper; }
        const Array<int,1>& get_kminor_start(bool lower) const { return lower ? kminor_start_lower : kminor_start_upper; }
        const Array<int,1>& get_idx_minor(bool lower) const { return lower ? idx_minor_lower : idx_minor_upper; }
        const Array<int,1>& get_idx_minor_scaling(bool lower) const { return lower ? idx_minor_scaling_lower : idx_minor_scaling_upper; }
        const Array<Float,3>& get_kminor(bool lower) const { return lower ? kminor_lower : kminor_upper; }
        const Array<Float,4>& get_kmajor() const { return kmajor; }

        // MLP gas-optics surrogate (FP32 builds; no reference counterpart).  With a network set,
        // gas_optics / gas_optics_with_clouds evaluate it (rrtmgp_hip_mlp_gas_optics_*) and
        // recompute the (64-column tile, layer) units outside its training envelope with the plain
        // kernels.  The network's device arrays belong to the caller and must outlive its use;
        // nullptr restores the plain gas optics.  Throws std::invalid_argument in FP64 builds or
        // when the output count is not 2 ngpt.
        void set_surrogate(const rrtmgp_mlp_net* net);
        bool has_surrogate() const { return use_mlp; }
        // (tile, layer) units the last surrogate call recomputed with the plain kernels.
        int64_t last_fallback_units() const { return n_fallback; }

    private:
        bool use_mlp = false;
        rrtmgp_mlp_net mlp{};
        Array_gpu<int,1> gpt2band_gpu;      // 0-based band of each g-point
        Array_gpu<int,1> fallback_gpu;      // count + list of (tile, layer) units
        int64_t n_fallback = 0;
        int64_t run_fallback_list(int ncol, int nlay);

        Array<Float,2> totplnk;
        Array<Float,4> planck_frac;
        Float totplnk_delta = 0;
        Float temp_ref_min, temp_ref_max;
        Float press_ref_min, press_ref_max;
        Float press_ref_trop_log;
        Float press_ref_log_delta;
        Float temp_ref_delta;
        int kmajor_ntemp = 0, kmajor_neta = 0, kmajor_npres = 0;
        bool check_ranges = true;

        Array<Float,1> press_ref, press_ref_log, temp_ref;
        Array<std::string,1> gas_names;
        Array<Float,3> vmr_ref;
        Array<int,2> flavor;
        Array<int,2> gpoint_flavor;
        Array<Float,4> kmajor;
        Array<Float,3> kminor_lower, kminor_upper;
        Array<int,2> minor_limits_gpt_lower, minor_limits_gpt_upper;
        Array<Bool,1> minor_scales_with_density_lower, minor_scales_with_density_upper;
        Array<Bool,1> scale_by_complement_lower, scale_by_complement_upper;
        Array<int,1> kminor_start_lower, kminor_start_upper;
        Array<int,1> idx_minor_lower, idx_minor_upper;
        Array<int,1> idx_minor_scaling_lower, idx_minor_scaling_upper;
        Array<int,1> is_key;
        Array<Float,1> solar_source_quiet, solar_source_facular, solar_source_sunspot, solar_source;
        Array<Float,4> krayl;
        int idx_h2o = -1;

        // Device copies.
        Array_gpu<Float,1> solar_source_gpu, press_ref_log_gpu, temp_ref_gpu;
        Array_gpu<Float,2> totplnk_gpu;
        Array_gpu<Float,4> planck_frac_gpu, kmajor_gpu, krayl_gpu;
        Array_gpu<Float,3> vmr_ref_gpu, kminor_lower_gpu, kminor_upper_gpu;
        Array_gpu<int,2> flavor_gpu, gpoint_flavor_gpu, minor_limits_gpt_lower_gpu, minor_limits_gpt_upper_gpu;
        Array_gpu<Bool,1> minor_scales_with_density_lower_gpu, minor_scales_with_density_upper_gpu;
        Array_gpu<Bool,1> scale_by_complement_lower_gpu, scale_by_complement_upper_gpu;
        Array_gpu<int,1> kminor_start_lower_gpu, kminor_start_upper_gpu;
        Array_gpu<int,1> idx_minor_lower_gpu, idx_minor_upper_gpu;
        Array_gpu<int,1> idx_minor_scaling_lower_gpu, idx_minor_scaling_upper_gpu;
        std::shared_ptr<void> go_pack;    // band-packed tables (rrtmgp_hip_kdist_pack)
        rrtmgp_kdist_dev dev;

        void init_abs_coeffs(
                const Gas_concs_gpu& available_gases,
                const Array<std::string,1>& gas_names,
                const Array<int,3>& key_species,
                const Array<int,2>& band2gpt,
                const Array<Float,2>& band_lims_wavenum,
                const Array<Float,1>& press_ref,
                const Array<Float,1>& temp_ref,
                const Float press_ref_trop,
                const Float temp_ref_p,
                const Float temp_ref_t,
                const Array<Float,3>& vmr_ref,
                const Array<Float,4>& kmajor,
                const Array<Float,3>& kminor_lower,
                const Array<Float,3>& kminor_upper,
                const Array<std::string,1>& gas_minor,
                const Array<std::string,1>& identifier_minor,
                const Array<std::string,1>& minor_gases_lower,
                const Array<std::string,1>& minor_gases_upper,
                const Array<int,2>& minor_limits_gpt_lower,
                const Array<int,2>& minor_limits_gpt_upper,
                const Array<Bool,1>& minor_scales_with_density_lower,
                const Array<Bool,1>& minor_scales_with_density_upper,
                const Array<std::string,1>& scaling_gas_lower,
                const Array<std::string,1>& scaling_gas_upper,
                const Array<Bool,1>& scale_by_complement_lower,
                const Array<Bool,1>& scale_by_complement_upper,
                const Array<int,1>& kminor_start_lower,
                const Array<int,1>& kminor_start_upper,
                const Array<Float,3>& rayl_lower,
                const Array<Float,3>& rayl_upper);

        void upload();
        void set_solar_variability(const Float mg_index, const Float sb_index);
        void make_vmr_desc(const Gas_concs_gpu& gas_desc, int ncol, int nlay, rrtmgp_gas_vmr& out) const;
        // The range checks of src/Gas_optics_rrtmgp.cpp:805-826 (when enabled) and the surface
        // layer of :1333, from one device pass and one read-back; returns sfc_lay.
        int check_inputs(const Array_gpu<Float,2>& play, const Array_gpu<Float,2>& plev,
                         const Array_gpu<Float,2>& tlay, const Array_gpu<Float,2>& col_dry,
                         const Array_gpu<Float,2>* tlev, const Array_gpu<Float,1>* tsfc) const;
};

#endif
