// Column-radiation drivers (reference src_test/Radiation_solver.cu:419-950 / .cpp:384-913).
#include "Radiation_solver.h"

#include <algorithm>

#include "Rte_lw.h"
#include "Rte_sw.h"
#include "hip_dispatch.h"

using rrtmgp_cpp::hip_check;
using rrtmgp_cpp::stream;

namespace
{
template<int N>
Array_gpu<Float,N> view(const Float* p, const std::array<int,N>& dims)
{
    return Array_gpu<Float,N>(const_cast<Float*>(p), dims);
}

// Block [c0, c0+n) of a (ncol, n2) array: the array itself when the block spans all columns,
// otherwise a copy into ws.
Array_gpu<Float,2> block_of(const Array_gpu<Float,2>& full, int c0, int n, Array_gpu<Float,2>& ws)
{
    if (full.size() == 0) return Array_gpu<Float,2>();
    if (n == full.dim(1)) return view<2>(full.ptr(), full.get_dims());
    if (ws.dim(1) != n || ws.dim(2) != full.dim(2)) ws = Array_gpu<Float,2>({n, full.dim(2)});
    hip_check(K<Float>::copy_cols(full.ptr(), full.dim(1), c0, ws.ptr(), n, 0, n, full.dim(2), stream()));
    return view<2>(ws.ptr(), ws.get_dims());
}

// Write a (n, n2[, n3]) block result into columns [c0, c0+n) of the full output.
void scatter(const Float* blk, int n, int64_t rows, Float* full, int ncol, int c0)
{
    if (blk == full) return;
    hip_check(K<Float>::copy_cols(blk, n, 0, full, ncol, c0, n, rows, stream()));
}

struct BlockInputs
{
    Array_gpu<Float,2> p_lay, p_lev, t_lay, t_lev, col_dry, lwp, iwp, rel, rei, rh;
};
}  // namespace


Radiation_solver_longwave_gpu::Radiation_solver_longwave_gpu(
        std::shared_ptr<Gas_optics_rrtmgp_gpu> kdist, std::shared_ptr<Cloud_optics_gpu> cloud_optics) :
    kdist(std::move(kdist)), cloud_optics(std::move(cloud_optics))
{}

void Radiation_solver_longwave_gpu::solve_gpu(
        const bool switch_cloud_optics, const bool switch_output_bnd_fluxes,
        const Gas_concs_gpu& gas_concs,
        const Array_gpu<Float,2>& p_lay, const Array_gpu<Float,2>& p_lev,
        const Array_gpu<Float,2>& t_lay, const Array_gpu<Float,2>& t_lev,
        const Array_gpu<Float,2>& col_dry, const Array_gpu<Float,1>& t_sfc,
        const Array_gpu<Float,2>& emis_sfc,
        const Array_gpu<Float,2>& lwp, const Array_gpu<Float,2>& iwp,
        const Array_gpu<Float,2>& rel, const Array_gpu<Float,2>& rei,
        Array_gpu<Float,2>& lw_flux_up, Array_gpu<Float,2>& lw_flux_dn, Array_gpu<Float,2>& lw_flux_net,
        Array_gpu<Float,3>& lw_bnd_flux_up, Array_gpu<Float,3>& lw_bnd_flux_dn,
        Array_gpu<Float,3>& lw_bnd_flux_net, Array_gpu<Float,2>* lw_heating_rate)
{
    const bool hr = lw_heating_rate != nullptr && lw_heating_rate->size() > 0;
    const int ncol = p_lay.dim(1), nlay = p_lay.dim(2), nlev = nlay + 1;
    const int ngpt = kdist->get_ngpt(), nbnd = kdist->get_nband();
    if (switch_cloud_optics && !cloud_optics) throw std::runtime_error("cloud optics requested but not configured");
    // Orientation from the first column (src_test/Radiation_solver.cpp:407).
    const Bool top_at_1 = p_lay({1, 1}) < p_lay({1, nlay});
    const int nblk = std::max(1, std::min(n_col_block, ncol));
    if (ws_ncol != nblk || ws_nlay != nlay || ws_bnd != int(switch_output_bnd_fluxes))
    {
        optical_props = std::make_unique<Optical_props_1scl_gpu>(nblk, nlay, *kdist);
        // Factored Planck sources (fractions per g-point, band Planck functions per layer and
        // level): two (ncol, nlay, ngpt) arrays fewer through HBM than the three source arrays.
        sources = std::make_unique<Source_func_lw_gpu>(nblk, nlay, *kdist, true);
        cloud_props.reset();
        if (cloud_optics) cloud_props = std::make_unique<Optical_props_1scl_gpu>(nblk, nlay, *cloud_optics);
        // Block flux workspaces, used when a block does not span all columns.
        const bool partial = nblk < ncol;
        blk_flux_up = partial ? Array_gpu<Float,2>({nblk, nlev}) : Array_gpu<Float,2>();
        blk_flux_dn = partial ? Array_gpu<Float,2>({nblk, nlev}) : Array_gpu<Float,2>();
        const bool pb = partial && switch_output_bnd_fluxes;
        blk_bnd_up = pb ? Array_gpu<Float,3>({nblk, nlev, nbnd}) : Array_gpu<Float,3>();
        blk_bnd_dn = pb ? Array_gpu<Float,3>({nblk, nlev, nbnd}) : Array_gpu<Float,3>();
        ws_ncol = nblk; ws_nlay = nlay; ws_bnd = int(switch_output_bnd_fluxes);
    }

    BlockInputs ws;
    Rte_lw_gpu rte;
    for (int c0 = 0; c0 < ncol; c0 += nblk)
    {
        const int n = std::min(nblk, ncol - c0);
        if (n != nblk)
        {
            // Residual block: fresh containers of the residual size.
            Radiation_solver_longwave_gpu tail(kdist, cloud_optics);
            Gas_concs_gpu gsub(gas_concs, c0 + 1, n);
            Array_gpu<Float,2> pl = block_of(p_lay, c0, n, ws.p_lay), pv = block_of(p_lev, c0, n, ws.p_lev);
            Array_gpu<Float,2> tl = block_of(t_lay, c0, n, ws.t_lay), tv = block_of(t_lev, c0, n, ws.t_lev);
            Array_gpu<Float,2> cd = block_of(col_dry, c0, n, ws.col_dry);
            Array_gpu<Float,2> a = block_of(lwp, c0, n, ws.lwp), b = block_of(iwp, c0, n, ws.iwp);
            Array_gpu<Float,2> c = block_of(rel, c0, n, ws.rel), d = block_of(rei, c0, n, ws.rei);
            Array_gpu<Float,1> ts = Array_gpu<Float,1>(const_cast<Float*>(t_sfc.ptr()) + c0, {n});
            Array_gpu<Float,2> em = view<2>(emis_sfc.ptr() + int64_t(nbnd) * c0, {nbnd, n});
            Array_gpu<Float,2> fu({n, nlev}), fd({n, nlev}), fn({n, nlev});
            Array_gpu<Float,3> bu, bd, bn;
            if (switch_output_bnd_fluxes)
            {
                bu = Array_gpu<Float,3>({n, nlev, nbnd}); bd = Array_gpu<Float,3>({n, nlev, nbnd});
                bn = Array_gpu<Float,3>({n, nlev, nbnd});
            }
            Array_gpu<Float,2> th = hr ? Array_gpu<Float,2>({n, nlay}) : Array_gpu<Float,2>();
            tail.solve_gpu(switch_cloud_optics, switch_output_bnd_fluxes, gsub, pl, pv, tl, tv, cd, ts, em,
                           a, b, c, d, fu, fd, fn, bu, bd, bn, hr ? &th : nullptr);
            scatter(fu.ptr(), n, nlev, lw_flux_up.ptr(), ncol, c0);
            scatter(fd.ptr(), n, nlev, lw_flux_dn.ptr(), ncol, c0);
            if (hr) scatter(th.ptr(), n, nlay, lw_heating_rate->ptr(), ncol, c0);
            if (switch_output_bnd_fluxes)
            {
                scatter(bu.ptr(), n, int64_t(nlev) * nbnd, lw_bnd_flux_up.ptr(), ncol, c0);
                scatter(bd.ptr(), n, int64_t(nlev) * nbnd, lw_bnd_flux_dn.ptr(), ncol, c0);
            }
            continue;
        }
        const bool whole = (n == ncol);
        std::unique_ptr<Gas_concs_gpu> gsub;
        if (!whole) gsub = std::make_unique<Gas_concs_gpu>(gas_concs, c0 + 1, n);
        const Gas_concs_gpu& gc = whole ? gas_concs : *gsub;
        Array_gpu<Float,2> pl = block_of(p_lay, c0, n, ws.p_lay), pv = block_of(p_lev, c0, n, ws.p_lev);
        Array_gpu<Float,2> tl = block_of(t_lay, c0, n, ws.t_lay), tv = block_of(t_lev, c0, n, ws.t_lev);
        Array_gpu<Float,1> ts(const_cast<Float*>(t_sfc.ptr()) + c0, {n});
        Array_gpu<Float,2> em = view<2>(emis_sfc.ptr() + int64_t(nbnd) * c0, {nbnd, n});

        Array_gpu<Float,2> cd;
        if (col_dry.size() == 0)
        {
            if (col_dry_ws.dim(1) != n || col_dry_ws.dim(2) != nlay) col_dry_ws = Array_gpu<Float,2>({n, nlay});
            Gas_optics_rrtmgp_gpu::get_col_dry(col_dry_ws, gc.get_vmr("h2o"), pv);
            cd = view<2>(col_dry_ws.ptr(), col_dry_ws.get_dims());
        }
        else
            cd = block_of(col_dry, c0, n, ws.col_dry);

        const Optical_props_arry_gpu* cloud = nullptr;
        if (switch_cloud_optics)
        {
            cloud_optics->cloud_optics(block_of(lwp, c0, n, ws.lwp), block_of(iwp, c0, n, ws.iwp),
                                       block_of(rel, c0, n, ws.rel), block_of(rei, c0, n, ws.rei), *cloud_props);
            cloud = cloud_props.get();
        }
        kdist->gas_optics_with_clouds(pl, pv, tl, ts, gc, *optical_props, *sources, cd, tv, cloud);

        // Band and broadband fluxes summed inside the solver, straight into the outputs when the
        // block spans all columns.
        Array_gpu<Float,2>& fu = whole ? lw_flux_up : blk_flux_up;
        Array_gpu<Float,2>& fd = whole ? lw_flux_dn : blk_flux_dn;
        Array_gpu<Float,3> no_bnd;
        Array_gpu<Float,3>& bu = switch_output_bnd_fluxes ? (whole ? lw_bnd_flux_up : blk_bnd_up) : no_bnd;
        Array_gpu<Float,3>& bd = switch_output_bnd_fluxes ? (whole ? lw_bnd_flux_dn : blk_bnd_dn) : no_bnd;
        Array_gpu<Float,2>* hb = nullptr;
        if (hr)
        {
            if (!whole && (blk_hr.dim(1) != n || blk_hr.dim(2) != nlay)) blk_hr = Array_gpu<Float,2>({n, nlay});
            hb = whole ? lw_heating_rate : &blk_hr;
        }
        rte.rte_lw_fluxes(optical_props, top_at_1, *sources, em, Array_gpu<Float,2>(), bu, bd, fu, fd, 1, &pv, hb);
        if (!whole)
        {
            scatter(fu.ptr(), n, nlev, lw_flux_up.ptr(), ncol, c0);
            scatter(fd.ptr(), n, nlev, lw_flux_dn.ptr(), ncol, c0);
            if (hr) scatter(blk_hr.ptr(), n, nlay, lw_heating_rate->ptr(), ncol, c0);
            if (switch_output_bnd_fluxes)
            {
                const int64_t rows = int64_t(nlev) * nbnd;
                scatter(bu.ptr(), n, rows, lw_bnd_flux_up.ptr(), ncol, c0);
                scatter(bd.ptr(), n, rows, lw_bnd_flux_dn.ptr(), ncol, c0);
            }
        }
    }
    hip_check(K<Float>::net_broadband(ncol, nlev, lw_flux_dn.ptr(), lw_flux_up.ptr(), lw_flux_net.ptr(), stream()));
    if (switch_output_bnd_fluxes)
        hip_check(K<Float>::net_broadband(ncol, int64_t(nlev) * nbnd, lw_bnd_flux_dn.ptr(), lw_bnd_flux_up.ptr(),
                                          lw_bnd_flux_net.ptr(), stream()));
}


Radiation_solver_shortwave_gpu::Radiation_solver_shortwave_gpu(
        std::shared_ptr<Gas_optics_rrtmgp_gpu> kdist, std::shared_ptr<Cloud_optics_gpu> cloud_optics,
        std::shared_ptr<Aerosol_optics_gpu> aerosol_optics) :
    kdist(std::move(kdist)), cloud_optics(std::move(cloud_optics)), aerosol_optics(std::move(aerosol_optics))
{}

void Radiation_solver_shortwave_gpu::solve_gpu(
        const bool switch_cloud_optics, const bool switch_output_bnd_fluxes, const bool switch_delta_cloud,
        const Gas_concs_gpu& gas_concs,
        const Array_gpu<Float,2>& p_lay, const Array_gpu<Float,2>& p_lev,
        const Array_gpu<Float,2>& t_lay, const Array_gpu<Float,2>& t_lev,
        const Array_gpu<Float,2>& col_dry,
        const Array_gpu<Float,2>& sfc_alb_dir, const Array_gpu<Float,2>& sfc_alb_dif,
        const Array_gpu<Float,1>& tsi_scaling, const Array_gpu<Float,1>& mu0,
        const Array_gpu<Float,2>& lwp, const Array_gpu<Float,2>& iwp,
        const Array_gpu<Float,2>& rel, const Array_gpu<Float,2>& rei,
        Array_gpu<Float,2>& sw_flux_up, Array_gpu<Float,2>& sw_flux_dn,
        Array_gpu<Float,2>& sw_flux_dn_dir, Array_gpu<Float,2>& sw_flux_net,
        Array_gpu<Float,3>& sw_bnd_flux_up, Array_gpu<Float,3>& sw_bnd_flux_dn,
        Array_gpu<Float,3>& sw_bnd_flux_dn_dir, Array_gpu<Float,3>& sw_bnd_flux_net,
        Array_gpu<Float,2>* sw_heating_rate)
{
    solve_gpu(switch_cloud_optics, false, switch_output_bnd_fluxes, switch_delta_cloud, false, gas_concs, p_lay,
              p_lev, t_lay, t_lev, col_dry, sfc_alb_dir, sfc_alb_dif, tsi_scaling, mu0, lwp, iwp, rel, rei,
              Array_gpu<Float,2>(), Aerosol_concs_gpu(), sw_flux_up, sw_flux_dn, sw_flux_dn_dir, sw_flux_net,
              sw_bnd_flux_up, sw_bnd_flux_dn, sw_bnd_flux_dn_dir, sw_bnd_flux_net, sw_heating_rate);
}

void Radiation_solver_shortwave_gpu::solve_gpu(
        const bool switch_cloud_optics, const bool switch_aerosol_optics, const bool switch_output_bnd_fluxes,
        const bool switch_delta_cloud, const bool switch_delta_aerosol,
        const Gas_concs_gpu& gas_concs,
        const Array_gpu<Float,2>& p_lay, const Array_gpu<Float,2>& p_lev,
        const Array_gpu<Float,2>& t_lay, const Array_gpu<Float,2>& t_lev,
        const Array_gpu<Float,2>& col_dry,
        const Array_gpu<Float,2>& sfc_alb_dir, const Array_gpu<Float,2>& sfc_alb_dif,
        const Array_gpu<Float,1>& tsi_scaling, const Array_gpu<Float,1>& mu0,
        const Array_gpu<Float,2>& lwp, const Array_gpu<Float,2>& iwp,
        const Array_gpu<Float,2>& rel, const Array_gpu<Float,2>& rei,
        const Array_gpu<Float,2>& rh, const Aerosol_concs_gpu& aerosol_concs,
        Array_gpu<Float,2>& sw_flux_up, Array_gpu<Float,2>& sw_flux_dn,
        Array_gpu<Float,2>& sw_flux_dn_dir, Array_gpu<Float,2>& sw_flux_net,
        Array_gpu<Float,3>& sw_bnd_flux_up, Array_gpu<Float,3>& sw_bnd_flux_dn,
        Array_gpu<Float,3>& sw_bnd_flux_dn_dir, Array_gpu<Float,3>& sw_bnd_flux_net,
        Array_gpu<Float,2>* sw_heating_rate)
{
    const bool hr = sw_heating_rate != nullptr && sw_heating_rate->size() > 0;
    const int ncol = p_lay.dim(1), nlay = p_lay.dim(2), nlev = nlay + 1;
    const int ngpt = kdist->get_ngpt(), nbnd = kdist->get_nband();
    if (switch_cloud_optics && !cloud_optics) throw std::runtime_error("cloud optics requested but not configured");
    if (switch_aerosol_optics && !aerosol_optics)
        throw std::runtime_error("aerosol optics requested but not configured");
    if (switch_aerosol_optics && (rh.dim(1) != ncol || rh.dim(2) != nlay))
        throw std::runtime_error("aerosol optics: rh must be (ncol, nlay)");
    const Bool top_at_1 = p_lay({1, 1}) < p_lay({1, nlay});
    const int nblk = std::max(1, std::min(n_col_block, ncol));

    if (ws_ncol != nblk || ws_nlay != nlay || ws_bnd != int(switch_output_bnd_fluxes))
    {
        optical_props = std::make_unique<Optical_props_2str_gpu>(nblk, nlay, *kdist);
        cloud_props.reset();
        if (cloud_optics) cloud_props = std::make_unique<Optical_props_2str_gpu>(nblk, nlay, *cloud_optics);
        aerosol_props.reset();
        if (aerosol_optics) aerosol_props = std::make_unique<Optical_props_2str_gpu>(nblk, nlay, *aerosol_optics);
        toa_src = Array_gpu<Float,2>({nblk, ngpt});
        const bool partial = nblk < ncol;
        blk_flux_up = partial ? Array_gpu<Float,2>({nblk, nlev}) : Array_gpu<Float,2>();
        blk_flux_dn = partial ? Array_gpu<Float,2>({nblk, nlev}) : Array_gpu<Float,2>();
        blk_flux_dir = partial ? Array_gpu<Float,2>({nblk, nlev}) : Array_gpu<Float,2>();
        const bool pb = partial && switch_output_bnd_fluxes;
        blk_bnd_up = pb ? Array_gpu<Float,3>({nblk, nlev, nbnd}) : Array_gpu<Float,3>();
        blk_bnd_dn = pb ? Array_gpu<Float,3>({nblk, nlev, nbnd}) : Array_gpu<Float,3>();
        blk_bnd_dir = pb ? Array_gpu<Float,3>({nblk, nlev, nbnd}) : Array_gpu<Float,3>();
        ws_ncol = nblk; ws_nlay = nlay; ws_bnd = int(switch_output_bnd_fluxes);
    }

    BlockInputs ws;
    Rte_sw_gpu rte;
    for (int c0 = 0; c0 < ncol; c0 += nblk)
    {
        const int n = std::min(nblk, ncol - c0);
        if (n != nblk)
        {
            Radiation_solver_shortwave_gpu tail(kdist, cloud_optics, aerosol_optics);
            Gas_concs_gpu gsub(gas_concs, c0 + 1, n);
            Aerosol_concs_gpu asub = switch_aerosol_optics ? Aerosol_concs_gpu(aerosol_concs, c0 + 1, n)
                                                           : Aerosol_concs_gpu();
            Array_gpu<Float,2> rh_ws;
            Array_gpu<Float,2> rhb = switch_aerosol_optics ? block_of(rh, c0, n, rh_ws) : Array_gpu<Float,2>();
            Array_gpu<Float,2> pl = block_of(p_lay, c0, n, ws.p_lay), pv = block_of(p_lev, c0, n, ws.p_lev);
            Array_gpu<Float,2> tl = block_of(t_lay, c0, n, ws.t_lay), tv = block_of(t_lev, c0, n, ws.t_lev);
            Array_gpu<Float,2> cd = block_of(col_dry, c0, n, ws.col_dry);
            Array_gpu<Float,2> a = block_of(lwp, c0, n, ws.lwp), b = block_of(iwp, c0, n, ws.iwp);
            Array_gpu<Float,2> c = block_of(rel, c0, n, ws.rel), d = block_of(rei, c0, n, ws.rei);
            Array_gpu<Float,2> ad = view<2>(sfc_alb_dir.ptr() + int64_t(nbnd) * c0, {nbnd, n});
            Array_gpu<Float,2> af = view<2>(sfc_alb_dif.ptr() + int64_t(nbnd) * c0, {nbnd, n});
            Array_gpu<Float,1> tsi(const_cast<Float*>(tsi_scaling.ptr()) + c0, {n});
            Array_gpu<Float,1> m0(const_cast<Float*>(mu0.ptr()) + c0, {n});
            Array_gpu<Float,2> fu({n, nlev}), fd({n, nlev}), fdir({n, nlev}), fn({n, nlev});
            Array_gpu<Float,3> bu, bd, bdir, bn;
            if (switch_output_bnd_fluxes)
            {
                bu = Array_gpu<Float,3>({n, nlev, nbnd}); bd = Array_gpu<Float,3>({n, nlev, nbnd});
                bdir = Array_gpu<Float,3>({n, nlev, nbnd}); bn = Array_gpu<Float,3>({n, nlev, nbnd});
            }
            Array_gpu<Float,2> th = hr ? Array_gpu<Float,2>({n, nlay}) : Array_gpu<Float,2>();
            tail.solve_gpu(switch_cloud_optics, switch_aerosol_optics, switch_output_bnd_fluxes, switch_delta_cloud,
                           switch_delta_aerosol, gsub, pl, pv, tl, tv, cd, ad, af, tsi, m0, a, b, c, d, rhb, asub,
                           fu, fd, fdir, fn, bu, bd, bdir, bn, hr ? &th : nullptr);
            if (hr) scatter(th.ptr(), n, nlay, sw_heating_rate->ptr(), ncol, c0);
            scatter(fu.ptr(), n, nlev, sw_flux_up.ptr(), ncol, c0);
            scatter(fd.ptr(), n, nlev, sw_flux_dn.ptr(), ncol, c0);
            scatter(fdir.ptr(), n, nlev, sw_flux_dn_dir.ptr(), ncol, c0);
            if (switch_output_bnd_fluxes)
            {
                const int64_t rows = int64_t(nlev) * nbnd;
                scatter(bu.ptr(), n, rows, sw_bnd_flux_up.ptr(), ncol, c0);
                scatter(bd.ptr(), n, rows, sw_bnd_flux_dn.ptr(), ncol, c0);
                scatter(bdir.ptr(), n, rows, sw_bnd_flux_dn_dir.ptr(), ncol, c0);
            }
            continue;
        }
        const bool whole = (n == ncol);
        std::unique_ptr<Gas_concs_gpu> gsub;
        if (!whole) gsub = std::make_unique<Gas_concs_gpu>(gas_concs, c0 + 1, n);
        const Gas_concs_gpu& gc = whole ? gas_concs : *gsub;
        Array_gpu<Float,2> pl = block_of(p_lay, c0, n, ws.p_lay), pv = block_of(p_lev, c0, n, ws.p_lev);
        Array_gpu<Float,2> tl = block_of(t_lay, c0, n, ws.t_lay);
        Array_gpu<Float,2> ad = view<2>(sfc_alb_dir.ptr() + int64_t(nbnd) * c0, {nbnd, n});
        Array_gpu<Float,2> af = view<2>(sfc_alb_dif.ptr() + int64_t(nbnd) * c0, {nbnd, n});
        Array_gpu<Float,1> tsi(const_cast<Float*>(tsi_scaling.ptr()) + c0, {n});
        Array_gpu<Float,1> m0(const_cast<Float*>(mu0.ptr()) + c0, {n});

        Array_gpu<Float,2> cd;
        if (col_dry.size() == 0)
        {
            if (col_dry_ws.dim(1) != n || col_dry_ws.dim(2) != nlay) col_dry_ws = Array_gpu<Float,2>({n, nlay});
            Gas_optics_rrtmgp_gpu::get_col_dry(col_dry_ws, gc.get_vmr("h2o"), pv);
            cd = view<2>(col_dry_ws.ptr(), col_dry_ws.get_dims());
        }
        else
            cd = block_of(col_dry, c0, n, ws.col_dry);

        // With the plain gas optics the whole chain carries the co-albedo 1 - ssa in the ssa
        // arrays (FP32 keeps absorption far below 6e-8 of extinction only that way); a surrogate
        // writes ssa.
        const bool co = !kdist->has_surrogate();
        const Optical_props_arry_gpu* cloud = nullptr;
        if (switch_cloud_optics)
        {
            auto a = block_of(lwp, c0, n, ws.lwp), b = block_of(iwp, c0, n, ws.iwp);
            auto c = block_of(rel, c0, n, ws.rel), d = block_of(rei, c0, n, ws.rei);
            if (co) cloud_optics->cloud_optics_coalbedo(a, b, c, d, switch_delta_cloud, *cloud_props);
            else if (switch_delta_cloud) cloud_optics->cloud_optics_delta_scaled(a, b, c, d, *cloud_props);
            else cloud_optics->cloud_optics(a, b, c, d, *cloud_props);
            cloud = cloud_props.get();
        }
        kdist->gas_optics_with_clouds(pl, pv, tl, gc, *optical_props, toa_src, cd, cloud, co);
        if (switch_aerosol_optics)
        {
            std::unique_ptr<Aerosol_concs_gpu> asub;
            if (!whole) asub = std::make_unique<Aerosol_concs_gpu>(aerosol_concs, c0 + 1, n);
            Array_gpu<Float,2> rhb = block_of(rh, c0, n, ws.rh);
            const Aerosol_concs_gpu& ac = whole ? aerosol_concs : *asub;
            aerosol_optics->compute(ac, rhb, pv, switch_delta_aerosol, *aerosol_props, co);
            auto& ops = dynamic_cast<Optical_props_2str_gpu&>(*optical_props);
            if (co) add_to_coalbedo(ops, *aerosol_props);
            else add_to(ops, *aerosol_props);
        }
        // TSI scaling of the incoming flux (src_test/Radiation_solver.cpp:739-743).
        hip_check(K<Float>::scale_by_col(n, ngpt, toa_src.ptr(), tsi.ptr(), stream()));

        Array_gpu<Float,2>& fu = whole ? sw_flux_up : blk_flux_up;
        Array_gpu<Float,2>& fd = whole ? sw_flux_dn : blk_flux_dn;
        Array_gpu<Float,2>& fr = whole ? sw_flux_dn_dir : blk_flux_dir;
        Array_gpu<Float,3> no_bnd;
        Array_gpu<Float,3>& bu = switch_output_bnd_fluxes ? (whole ? sw_bnd_flux_up : blk_bnd_up) : no_bnd;
        Array_gpu<Float,3>& bd = switch_output_bnd_fluxes ? (whole ? sw_bnd_flux_dn : blk_bnd_dn) : no_bnd;
        Array_gpu<Float,3>& br = switch_output_bnd_fluxes ? (whole ? sw_bnd_flux_dn_dir : blk_bnd_dir) : no_bnd;
        Array_gpu<Float,2>* hb = nullptr;
        if (hr)
        {
            if (!whole && (blk_hr.dim(1) != n || blk_hr.dim(2) != nlay)) blk_hr = Array_gpu<Float,2>({n, nlay});
            hb = whole ? sw_heating_rate : &blk_hr;
        }
        rte.rte_sw_fluxes(optical_props, top_at_1, m0, toa_src, ad, af, Array_gpu<Float,2>(), bu, bd, br, fu, fd, fr,
                          &pv, hb, co);
        if (!whole)
        {
            if (hr) scatter(blk_hr.ptr(), n, nlay, sw_heating_rate->ptr(), ncol, c0);
            scatter(fu.ptr(), n, nlev, sw_flux_up.ptr(), ncol, c0);
            scatter(fd.ptr(), n, nlev, sw_flux_dn.ptr(), ncol, c0);
            scatter(fr.ptr(), n, nlev, sw_flux_dn_dir.ptr(), ncol, c0);
            if (switch_output_bnd_fluxes)
            {
                const int64_t rows = int64_t(nlev) * nbnd;
                scatter(bu.ptr(), n, rows, sw_bnd_flux_up.ptr(), ncol, c0);
                scatter(bd.ptr(), n, rows, sw_bnd_flux_dn.ptr(), ncol, c0);
                scatter(br.ptr(), n, rows, sw_bnd_flux_dn_dir.ptr(), ncol, c0);
            }
        }
    }
    hip_check(K<Float>::net_broadband(ncol, nlev, sw_flux_dn.ptr(), sw_flux_up.ptr(), sw_flux_net.ptr(), stream()));
    if (switch_output_bnd_fluxes)
        hip_check(K<Float>::net_broadband(ncol, int64_t(nlev) * nbnd, sw_bnd_flux_dn.ptr(), sw_bnd_flux_up.ptr(),
                                          sw_bnd_flux_net.ptr(), stream()));
}
