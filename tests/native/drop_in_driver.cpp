// A caller of the reference's GPU class API, written as the reference's own driver is
// (src_test/test_rte_rrtmgp.cpp:171-639 for reading the case and the switches,
// src_test/Radiation_solver.cu:419-580 and :683-893 for the per-column-block call sequence), and
// compiled unchanged against this build's headers and libraries.  It reads the case and the
// coefficient sets from bundles (include/Coefficient_bundle.h: the reference's NetCDF names and
// dimension orders), runs LW and SW in column blocks cut with Array_gpu::subset, and writes the
// fluxes with Array_gpu::dump into the output directory.
//
// usage: drop_in_driver <case dir> <lw coef> <sw coef> <lw cloud> <sw cloud> <out dir> <n_col_block>
#include <cstdio>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "Array.h"
#include "Cloud_optics.h"
#include "Coefficient_bundle.h"
#include "Fluxes.h"
#include "Gas_concs.h"
#include "Gas_optics_rrtmgp.h"
#include "Optical_props.h"
#include "Rte_lw.h"
#include "Rte_sw.h"
#include "Source_functions.h"
#include "rrtmgp_hip.h"

namespace
{
template<typename T>
void must(int rc) { rrtmgp_cpp::hip_check(rc); }

// src_test/test_rte_rrtmgp.cpp:30-67: a scalar, a (lay) profile or a (lay, y, x) field.
void read_and_set_vmr(const std::string& gas, int n_col, int n_lay, const Coefficient_bundle& in,
                      Gas_concs_gpu& gas_concs)
{
    const std::string name = "vmr_" + gas;
    if (!in.variable_exists(name)) return;
    const auto dims = in.get_variable_dimensions(name);
    if (dims.empty())
        gas_concs.set_vmr(gas, in.get_variable<Float>(name));
    else if (dims.size() == 1)
        gas_concs.set_vmr(gas, Array<Float,1>(in.get_variable<Float>(name, {n_lay}), {n_lay}));
    else
        gas_concs.set_vmr(gas, Array<Float,2>(in.get_variable<Float>(name, {n_lay, 1, n_col}), {n_col, n_lay}));
}

// Columns [col_s, col_s + n) of a block result into the full array (the reference's
// Subset_kernels_cuda::get_from_subset, here rrtmgp_hip_copy_cols).
void get_from_subset(int n_col, int64_t rows, int n_col_in, int col_s, Float* full, const Float* sub)
{
#ifdef RTE_USE_SP
    must<float>(rrtmgp_hip_copy_cols_f32(sub, n_col_in, 0, full, n_col, col_s - 1, n_col_in, rows, rrtmgp_cpp::stream()));
#else
    must<double>(rrtmgp_hip_copy_cols_f64(sub, n_col_in, 0, full, n_col, col_s - 1, n_col_in, rows, rrtmgp_cpp::stream()));
#endif
}

void scaling_to_subset(int n_col, int n_gpt, Array_gpu<Float,2>& toa_src, const Array_gpu<Float,1>& tsi_scaling)
{
#ifdef RTE_USE_SP
    must<float>(rrtmgp_hip_scale_by_col_f32(n_col, n_gpt, toa_src.ptr(), tsi_scaling.ptr(), rrtmgp_cpp::stream()));
#else
    must<double>(rrtmgp_hip_scale_by_col_f64(n_col, n_gpt, toa_src.ptr(), tsi_scaling.ptr(), rrtmgp_cpp::stream()));
#endif
}
}  // namespace

int main(int argc, char** argv)
{
    if (argc != 8)
    {
        std::fprintf(stderr, "usage: %s case lw_coef sw_coef lw_cloud sw_cloud out_dir n_col_block\n", argv[0]);
        return 2;
    }
    try
    {
        const std::string out = argv[6];
        const int n_col_block = std::stoi(argv[7]);
        const Coefficient_bundle input_nc(argv[1]);
        const int n_col = input_nc.get_dimension_size("x") * input_nc.get_dimension_size("y");
        const int n_lay = input_nc.get_dimension_size("lay");
        const int n_lev = input_nc.get_dimension_size("lev");

        Array<Float,2> p_lay_h(input_nc.get_variable<Float>("p_lay", {n_lay, 1, n_col}), {n_col, n_lay});
        Array<Float,2> t_lay_h(input_nc.get_variable<Float>("t_lay", {n_lay, 1, n_col}), {n_col, n_lay});
        Array<Float,2> p_lev_h(input_nc.get_variable<Float>("p_lev", {n_lev, 1, n_col}), {n_col, n_lev});
        Array<Float,2> t_lev_h(input_nc.get_variable<Float>("t_lev", {n_lev, 1, n_col}), {n_col, n_lev});
        Array<Float,2> lwp_h(input_nc.get_variable<Float>("lwp", {n_lay, 1, n_col}), {n_col, n_lay});
        Array<Float,2> iwp_h(input_nc.get_variable<Float>("iwp", {n_lay, 1, n_col}), {n_col, n_lay});
        Array<Float,2> rel_h(input_nc.get_variable<Float>("rel", {n_lay, 1, n_col}), {n_col, n_lay});
        Array<Float,2> rei_h(input_nc.get_variable<Float>("rei", {n_lay, 1, n_col}), {n_col, n_lay});

        Gas_concs_gpu gas_concs;
        for (const char* g : {"h2o", "co2", "o3", "n2o", "co", "ch4", "o2", "n2"})
            read_and_set_vmr(g, n_col, n_lay, input_nc, gas_concs);

        const Array_gpu<Float,2> p_lay(p_lay_h), t_lay(t_lay_h), p_lev(p_lev_h), t_lev(t_lev_h);
        const Array_gpu<Float,2> lwp(lwp_h), iwp(iwp_h), rel(rel_h), rei(rei_h);
        const Array_gpu<Float,2> col_dry;   // empty: computed per block from h2o
        const Bool top_at_1 = p_lay({1, 1}) < p_lay({1, n_lay});

        // ---------------- longwave (Radiation_solver.cu:419-580) ----------------
        {
            std::unique_ptr<Gas_optics_rrtmgp_gpu> kdist_gpu = load_and_init_gas_optics(gas_concs, argv[2]);
            std::unique_ptr<Cloud_optics_gpu> cloud_optics_gpu = load_and_init_cloud_optics(argv[4]);
            const int n_bnd = kdist_gpu->get_nband();
            Array<Float,2> emis_h(input_nc.get_variable<Float>("emis_sfc", {1, n_col, n_bnd}), {n_bnd, n_col});
            Array<Float,1> t_sfc_h(input_nc.get_variable<Float>("t_sfc", {1, n_col}), {n_col});
            const Array_gpu<Float,2> emis_sfc(emis_h);
            const Array_gpu<Float,1> t_sfc(t_sfc_h);

            Array_gpu<Float,2> lw_flux_up({n_col, n_lev}), lw_flux_dn({n_col, n_lev}), lw_flux_net({n_col, n_lev});
            Array_gpu<Float,3> lw_bnd_flux_up({n_col, n_lev, n_bnd}), lw_bnd_flux_dn({n_col, n_lev, n_bnd});
            Rte_lw_gpu rte_lw;

            for (int col_s = 1; col_s <= n_col; col_s += n_col_block)
            {
                const int col_e = std::min(n_col, col_s + n_col_block - 1);
                const int n_col_in = col_e - col_s + 1;
                std::unique_ptr<Optical_props_arry_gpu> optical_props =
                        std::make_unique<Optical_props_1scl_gpu>(n_col_in, n_lay, *kdist_gpu);
                Source_func_lw_gpu sources(n_col_in, n_lay, *kdist_gpu);
                Optical_props_1scl_gpu cloud_props(n_col_in, n_lay, *cloud_optics_gpu);

                Gas_concs_gpu gas_concs_subset(gas_concs, col_s, n_col_in);
                auto p_lev_subset = p_lev.subset({{ {col_s, col_e}, {1, n_lev} }});
                Array_gpu<Float,2> col_dry_subset({n_col_in, n_lay});
                if (col_dry.size() == 0)
                    Gas_optics_rrtmgp_gpu::get_col_dry(col_dry_subset, gas_concs_subset.get_vmr("h2o"), p_lev_subset);

                kdist_gpu->gas_optics(
                        p_lay.subset({{ {col_s, col_e}, {1, n_lay} }}), p_lev_subset,
                        t_lay.subset({{ {col_s, col_e}, {1, n_lay} }}), t_sfc.subset({{ {col_s, col_e} }}),
                        gas_concs_subset, optical_props, sources, col_dry_subset,
                        t_lev.subset({{ {col_s, col_e}, {1, n_lev} }}));

                cloud_optics_gpu->cloud_optics(
                        lwp.subset({{ {col_s, col_e}, {1, n_lay} }}), iwp.subset({{ {col_s, col_e}, {1, n_lay} }}),
                        rel.subset({{ {col_s, col_e}, {1, n_lay} }}), rei.subset({{ {col_s, col_e}, {1, n_lay} }}),
                        cloud_props);
                add_to(dynamic_cast<Optical_props_1scl_gpu&>(*optical_props), cloud_props);

                Array_gpu<Float,3> gpt_flux_up({n_col_in, n_lev, kdist_gpu->get_ngpt()});
                Array_gpu<Float,3> gpt_flux_dn({n_col_in, n_lev, kdist_gpu->get_ngpt()});
                rte_lw.rte_lw(optical_props, top_at_1, sources,
                              emis_sfc.subset({{ {1, n_bnd}, {col_s, col_e} }}), Array_gpu<Float,2>(),
                              gpt_flux_up, gpt_flux_dn, 1);

                Fluxes_broadband_gpu fluxes(n_col_in, n_lev);
                fluxes.reduce(gpt_flux_up, gpt_flux_dn, optical_props, top_at_1);
                get_from_subset(n_col, n_lev, n_col_in, col_s, lw_flux_up.ptr(), fluxes.get_flux_up().ptr());
                get_from_subset(n_col, n_lev, n_col_in, col_s, lw_flux_dn.ptr(), fluxes.get_flux_dn().ptr());
                get_from_subset(n_col, n_lev, n_col_in, col_s, lw_flux_net.ptr(), fluxes.get_flux_net().ptr());

                Fluxes_byband_gpu bnd_fluxes(n_col_in, n_lev, n_bnd);
                bnd_fluxes.reduce(gpt_flux_up, gpt_flux_dn, optical_props, top_at_1);
                get_from_subset(n_col, int64_t(n_lev) * n_bnd, n_col_in, col_s, lw_bnd_flux_up.ptr(),
                                bnd_fluxes.get_bnd_flux_up().ptr());
                get_from_subset(n_col, int64_t(n_lev) * n_bnd, n_col_in, col_s, lw_bnd_flux_dn.ptr(),
                                bnd_fluxes.get_bnd_flux_dn().ptr());
            }
            lw_flux_up.dump(out + "/lw_flux_up");
            lw_flux_dn.dump(out + "/lw_flux_dn");
            lw_flux_net.dump(out + "/lw_flux_net");
            lw_bnd_flux_up.dump(out + "/lw_bnd_flux_up");
            lw_bnd_flux_dn.dump(out + "/lw_bnd_flux_dn");
        }

        // ---------------- shortwave (Radiation_solver.cu:683-893) ----------------
        {
            std::unique_ptr<Gas_optics_rrtmgp_gpu> kdist_gpu = load_and_init_gas_optics(gas_concs, argv[3]);
            std::unique_ptr<Cloud_optics_gpu> cloud_optics_gpu = load_and_init_cloud_optics(argv[5]);
            const int n_bnd = kdist_gpu->get_nband(), n_gpt = kdist_gpu->get_ngpt();
            Array<Float,1> mu0_h(input_nc.get_variable<Float>("mu0", {1, n_col}), {n_col});
            Array<Float,2> alb_dir_h(input_nc.get_variable<Float>("sfc_alb_dir", {1, n_col, n_bnd}), {n_bnd, n_col});
            Array<Float,2> alb_dif_h(input_nc.get_variable<Float>("sfc_alb_dif", {1, n_col, n_bnd}), {n_bnd, n_col});
            Array<Float,1> tsi_scaling_h({n_col});
            tsi_scaling_h.fill(input_nc.variable_exists("tsi_scaling") ? input_nc.get_variable<Float>("tsi_scaling")
                                                                       : Float(1.));
            const Array_gpu<Float,1> mu0(mu0_h), tsi_scaling(tsi_scaling_h);
            const Array_gpu<Float,2> sfc_alb_dir(alb_dir_h), sfc_alb_dif(alb_dif_h);

            Array_gpu<Float,2> sw_flux_up({n_col, n_lev}), sw_flux_dn({n_col, n_lev});
            Array_gpu<Float,2> sw_flux_dn_dir({n_col, n_lev}), sw_flux_net({n_col, n_lev});
            Rte_sw_gpu rte_sw;

            for (int col_s = 1; col_s <= n_col; col_s += n_col_block)
            {
                const int col_e = std::min(n_col, col_s + n_col_block - 1);
                const int n_col_in = col_e - col_s + 1;
                std::unique_ptr<Optical_props_arry_gpu> optical_props =
                        std::make_unique<Optical_props_2str_gpu>(n_col_in, n_lay, *kdist_gpu);
                std::unique_ptr<Optical_props_2str_gpu> cloud_props =
                        std::make_unique<Optical_props_2str_gpu>(n_col_in, n_lay, *cloud_optics_gpu);

                Gas_concs_gpu gas_concs_subset(gas_concs, col_s, n_col_in);
                auto p_lev_subset = p_lev.subset({{ {col_s, col_e}, {1, n_lev} }});
                Array_gpu<Float,2> col_dry_subset({n_col_in, n_lay});
                Gas_optics_rrtmgp_gpu::get_col_dry(col_dry_subset, gas_concs_subset.get_vmr("h2o"), p_lev_subset);

                Array_gpu<Float,2> toa_src_subset({n_col_in, n_gpt});
                kdist_gpu->gas_optics(
                        p_lay.subset({{ {col_s, col_e}, {1, n_lay} }}), p_lev_subset,
                        t_lay.subset({{ {col_s, col_e}, {1, n_lay} }}), gas_concs_subset, optical_props,
                        toa_src_subset, col_dry_subset);
                auto tsi_scaling_subset = tsi_scaling.subset({{ {col_s, col_e} }});
                scaling_to_subset(n_col_in, n_gpt, toa_src_subset, tsi_scaling_subset);

                cloud_optics_gpu->cloud_optics(
                        lwp.subset({{ {col_s, col_e}, {1, n_lay} }}), iwp.subset({{ {col_s, col_e}, {1, n_lay} }}),
                        rel.subset({{ {col_s, col_e}, {1, n_lay} }}), rei.subset({{ {col_s, col_e}, {1, n_lay} }}),
                        *cloud_props);
                cloud_props->delta_scale();
                add_to(dynamic_cast<Optical_props_2str_gpu&>(*optical_props), *cloud_props);

                Array_gpu<Float,3> gpt_flux_up({n_col_in, n_lev, n_gpt});
                Array_gpu<Float,3> gpt_flux_dn({n_col_in, n_lev, n_gpt});
                Array_gpu<Float,3> gpt_flux_dn_dir({n_col_in, n_lev, n_gpt});
                rte_sw.rte_sw(optical_props, top_at_1, mu0.subset({{ {col_s, col_e} }}), toa_src_subset,
                              sfc_alb_dir.subset({{ {1, n_bnd}, {col_s, col_e} }}),
                              sfc_alb_dif.subset({{ {1, n_bnd}, {col_s, col_e} }}), Array_gpu<Float,2>(),
                              gpt_flux_up, gpt_flux_dn, gpt_flux_dn_dir);

                Fluxes_broadband_gpu fluxes(n_col_in, n_lev);
                fluxes.reduce(gpt_flux_up, gpt_flux_dn, gpt_flux_dn_dir, optical_props, top_at_1);
                get_from_subset(n_col, n_lev, n_col_in, col_s, sw_flux_up.ptr(), fluxes.get_flux_up().ptr());
                get_from_subset(n_col, n_lev, n_col_in, col_s, sw_flux_dn.ptr(), fluxes.get_flux_dn().ptr());
                get_from_subset(n_col, n_lev, n_col_in, col_s, sw_flux_dn_dir.ptr(), fluxes.get_flux_dn_dir().ptr());
                get_from_subset(n_col, n_lev, n_col_in, col_s, sw_flux_net.ptr(), fluxes.get_flux_net().ptr());
            }
            sw_flux_up.dump(out + "/sw_flux_up");
            sw_flux_dn.dump(out + "/sw_flux_dn");
            sw_flux_dn_dir.dump(out + "/sw_flux_dn_dir");
            sw_flux_net.dump(out + "/sw_flux_net");
        }
        std::cout << "drop_in_driver: ok" << std::endl;
    }
    catch (const std::exception& e)
    {
        std::cerr << "EXCEPTION: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
