// Reverse-mode drivers of the column-radiation path (no reference counterpart; the forward
// sequence they differentiate is src_test/Radiation_solver.cu:489-580 LW and :763-893 SW).
#include "Radiation_adjoint.h"

#include <stdexcept>
#include <vector>

#include "Optical_props.h"
#include "hip_dispatch.h"

using rrtmgp_cpp::hip_check;
using rrtmgp_cpp::stream;

namespace
{
// An output array: allocated (and zeroed) when empty, else checked for shape.
template<int N>
void ensure(Array_gpu<Float,N>& a, const std::array<int,N>& dims)
{
    if (a.size() == 0) { a = Array_gpu<Float,N>(dims); a.fill(Float(0)); return; }
    for (int i = 0; i < N; ++i)
        if (a.get_dims()[i] != dims[i]) throw std::invalid_argument("adjoint output has the wrong shape");
}

template<int N>
void zero(Array_gpu<Float,N>& a)
{
    if (a.size()) hip_check(rrtmgp_hip_memset(a.ptr(), 0, size_t(a.size()) * sizeof(Float), stream()));
}

// (nbnd, ncol) band values -> (ncol, ngpt) g-point values.
Array_gpu<Float,2> expand(const Gas_optics_rrtmgp_gpu& kd, const Array_gpu<Float,2>& bnd, int ncol)
{
    Array_gpu<Float,2> out({ncol, kd.get_ngpt()});
    hip_check(K<Float>::expand_and_transpose(ncol, kd.get_nband(), kd.get_ngpt(),
                                             kd.get_band_lims_gpoint_gpu().ptr(), bnd.ptr(), out.ptr(), stream()));
    return out;
}

// Largest number of g-points in one band.
int max_band_gpts(const Gas_optics_rrtmgp_gpu& kd)
{
    const Array<int,2>& bl = kd.get_band_lims_gpoint();
    int m = 0;
    for (int b = 1; b <= kd.get_nband(); ++b) m = std::max(m, bl({2, b}) - bl({1, b}) + 1);
    return m;
}

// The adjoint differentiates the plain table gas optics.  With an MLP surrogate set, the forward
// solve_gpu runs the network instead, so a gradient from here would belong to another function.
void refuse_surrogate(const Gas_optics_rrtmgp_gpu& kdist)
{
    if (kdist.has_surrogate())
        throw std::invalid_argument("adjoint: the k-distribution has an MLP gas-optics surrogate set, whose "
                                    "forward the adjoint does not differentiate; clear it (set_surrogate(NULL)) "
                                    "before computing gradients");
}

bool cloudy(const Cloud_optics_gpu* c, const Array_gpu<Float,2>& lwp)
{
    return c != nullptr && lwp.size() > 0;
}
}  // namespace


void lw_adjoint_gpu(Gas_optics_rrtmgp_gpu& kdist, Cloud_optics_gpu* cloud_optics, const Gas_concs_gpu& gas_concs,
                    const Array_gpu<Float,2>& p_lay, const Array_gpu<Float,2>& p_lev,
                    const Array_gpu<Float,2>& t_lay, const Array_gpu<Float,2>& t_lev,
                    const Array_gpu<Float,1>& t_sfc, const Array_gpu<Float,2>& emis_sfc,
                    const Array_gpu<Float,2>& lwp, const Array_gpu<Float,2>& iwp,
                    const Array_gpu<Float,2>& rel, const Array_gpu<Float,2>& rei,
                    const Array_gpu<Float,2>& g_flux_up, const Array_gpu<Float,2>& g_flux_dn,
                    Lw_adjoint_out& out)
{
    if (!kdist.source_is_internal()) throw std::invalid_argument("lw_adjoint_gpu needs a longwave k-distribution");
    refuse_surrogate(kdist);
    const int ncol = p_lay.dim(1), nlay = p_lay.dim(2), nlev = nlay + 1;
    const int ngpt = kdist.get_ngpt(), nbnd = kdist.get_nband(), ngas = kdist.get_ngas();
    const int64_t ncl = int64_t(ncol) * nlay;
    const rrtmgp_kdist_dev& kd = kdist.device_tables();
    rrtmgp_gas_vmr vmr;
    kdist.vmr_descriptor(gas_concs, ncol, nlay, vmr);
    const Bool top_at_1 = p_lay({1, 1}) < p_lay({1, nlay});
    const int sfc_lay = p_lay({1, 1}) > p_lay({1, nlay}) ? 1 : nlay;

    ensure<2>(out.d_t_lay, {ncol, nlay}); ensure<2>(out.d_t_lev, {ncol, nlev}); ensure<1>(out.d_t_sfc, {ncol});
    ensure<2>(out.d_lnp_lay, {ncol, nlay}); ensure<2>(out.d_p_lev, {ncol, nlev});
    ensure<3>(out.d_vmr, {ncol, nlay, ngas}); ensure<2>(out.d_emis_sfc, {nbnd, ncol});

    Array_gpu<Float,2> col_dry({ncol, nlay});
    const Array_gpu<Float,2>& h2o = gas_concs.get_vmr("h2o");
    Gas_optics_rrtmgp_gpu::get_col_dry(col_dry, h2o, p_lev);
    std::unique_ptr<Optical_props_1scl_gpu> cl;
    if (cloudy(cloud_optics, lwp))
    {
        cl = std::make_unique<Optical_props_1scl_gpu>(ncol, nlay, *cloud_optics);
        cloud_optics->cloud_optics(lwp, iwp, rel, rei, *cl);
    }

    // One Gaussian angle, as the forward driver (src_test/Radiation_solver.cpp:533): D = 1.66, w = 0.5.
    Array_gpu<Float,2> emis_gpt = expand(kdist, emis_sfc, ncol);
    Array<Float,2> sec_h({ncol, ngpt});
    for (auto& x : sec_h.v()) x = Float(1.66);
    Array_gpu<Float,2> secants(sec_h);
    Array<Float,1> w_h({1});
    w_h({1}) = Float(0.5);
    Array_gpu<Float,1> weights(w_h);

    Array_gpu<Float,3> d_col({ncol, nlay, ngas + 2});
    Array_gpu<Float,2> d_lo({ncol, nlay}), d_hi({ncol, nlay}), d_col_dry({ncol, nlay}), d_emis_gpt({ncol, ngpt});
    zero(d_col); zero(d_lo); zero(d_hi); zero(out.d_t_lay); zero(out.d_t_sfc); zero(out.d_lnp_lay); zero(out.d_p_lev);

    const int mb = max_band_gpts(kdist);
    Array_gpu<Float,3> tau({ncol, nlay, mb}), lay({ncol, nlay, mb}), linc({ncol, nlay, mb}), ldec({ncol, nlay, mb});
    Array_gpu<Float,3> dtau({ncol, nlay, mb}), dlay({ncol, nlay, mb}), dlinc({ncol, nlay, mb}), dldec({ncol, nlay, mb});
    Array_gpu<Float,2> sfc({ncol, mb}), sfcj({ncol, mb}), dsfc({ncol, mb});
    const Array<int,2>& bl = kdist.get_band_lims_gpoint();
    for (int b = 0; b < nbnd; ++b)
    {
        const int g0 = bl({1, b + 1}) - 1, ngb = bl({2, b + 1}) - g0;
        hip_check(K<Float>::gas_optics_lw_bands(&kd, b, 1, ncol, nlay, p_lay.ptr(), t_lay.ptr(), t_lev.ptr(),
                                                t_sfc.ptr(), col_dry.ptr(), &vmr, sfc_lay,
                                                cl ? cl->get_tau().ptr() : nullptr, tau.ptr(), lay.ptr(),
                                                linc.ptr(), ldec.ptr(), sfc.ptr(), sfcj.ptr(), stream()));
        hip_check(K<Float>::lw_solver_noscat_adjoint(
                ncol, nlay, ngb, top_at_1, 1, secants.ptr() + int64_t(ncol) * g0, weights.ptr(), tau.ptr(),
                lay.ptr(), linc.ptr(), ldec.ptr(), emis_gpt.ptr() + int64_t(ncol) * g0, sfc.ptr(), nullptr,
                g_flux_up.ptr(), g_flux_dn.ptr(), dtau.ptr(), dlay.ptr(), dlinc.ptr(), dldec.ptr(), dsfc.ptr(),
                d_emis_gpt.ptr() + int64_t(ncol) * g0, stream()));
        hip_check(K<Float>::gas_optics_lw_adjoint(
                &kd, b, ncol, nlay, p_lay.ptr(), t_lay.ptr(), t_lev.ptr(), t_sfc.ptr(), col_dry.ptr(), &vmr,
                sfc_lay, dtau.ptr(), dlay.ptr(), dlinc.ptr(), dldec.ptr(), dsfc.ptr(), d_col.ptr(),
                out.d_t_lay.ptr(), out.d_lnp_lay.ptr(), d_lo.ptr(), d_hi.ptr(), out.d_t_sfc.ptr(), stream()));
    }
    hip_check(K<Float>::adjoint_levels(ncol, nlay, d_lo.ptr(), d_hi.ptr(), out.d_t_lev.ptr(), stream()));
    hip_check(K<Float>::adjoint_col_chain(ncol, nlay, &vmr, col_dry.ptr(), d_col.ptr(), out.d_vmr.ptr(),
                                          d_col_dry.ptr(), stream()));
    hip_check(K<Float>::get_col_dry_adjoint(ncol, nlay, h2o.ptr(), h2o.dim(1), h2o.dim(2), p_lev.ptr(),
                                            d_col_dry.ptr(), out.d_vmr.ptr() + ncl * (kd.idx_h2o - 1),
                                            out.d_p_lev.ptr(), stream()));
    hip_check(K<Float>::gpt_to_band(ncol, nbnd, kd.band_lims_gpt, d_emis_gpt.ptr(), out.d_emis_sfc.ptr(), stream()));
}


void sw_adjoint_gpu(Gas_optics_rrtmgp_gpu& kdist, Cloud_optics_gpu* cloud_optics, bool delta_cloud,
                    const Gas_concs_gpu& gas_concs,
                    const Array_gpu<Float,2>& p_lay, const Array_gpu<Float,2>& p_lev,
                    const Array_gpu<Float,2>& t_lay,
                    const Array_gpu<Float,2>& sfc_alb_dir, const Array_gpu<Float,2>& sfc_alb_dif,
                    const Array_gpu<Float,1>& tsi_scaling, const Array_gpu<Float,1>& mu0,
                    const Array_gpu<Float,2>& lwp, const Array_gpu<Float,2>& iwp,
                    const Array_gpu<Float,2>& rel, const Array_gpu<Float,2>& rei,
                    const Array_gpu<Float,2>& g_flux_up, const Array_gpu<Float,2>& g_flux_dn,
                    const Array_gpu<Float,2>& g_flux_dir, Sw_adjoint_out& out)
{
    if (!kdist.source_is_external()) throw std::invalid_argument("sw_adjoint_gpu needs a shortwave k-distribution");
    refuse_surrogate(kdist);
    const int ncol = p_lay.dim(1), nlay = p_lay.dim(2), nlev = nlay + 1;
    const int ngpt = kdist.get_ngpt(), nbnd = kdist.get_nband(), ngas = kdist.get_ngas();
    const int64_t ncl = int64_t(ncol) * nlay;
    const rrtmgp_kdist_dev& kd = kdist.device_tables();
    rrtmgp_gas_vmr vmr;
    kdist.vmr_descriptor(gas_concs, ncol, nlay, vmr);
    const Bool top_at_1 = p_lay({1, 1}) < p_lay({1, nlay});

    ensure<2>(out.d_t_lay, {ncol, nlay}); ensure<2>(out.d_lnp_lay, {ncol, nlay}); ensure<2>(out.d_p_lev, {ncol, nlev});
    ensure<3>(out.d_vmr, {ncol, nlay, ngas});
    ensure<2>(out.d_sfc_alb_dir, {nbnd, ncol}); ensure<2>(out.d_sfc_alb_dif, {nbnd, ncol});

    Array_gpu<Float,2> col_dry({ncol, nlay});
    const Array_gpu<Float,2>& h2o = gas_concs.get_vmr("h2o");
    Gas_optics_rrtmgp_gpu::get_col_dry(col_dry, h2o, p_lev);
    std::unique_ptr<Optical_props_2str_gpu> cl;
    if (cloudy(cloud_optics, lwp))
    {
        cl = std::make_unique<Optical_props_2str_gpu>(ncol, nlay, *cloud_optics);
        if (delta_cloud) cloud_optics->cloud_optics_delta_scaled(lwp, iwp, rel, rei, *cl);
        else cloud_optics->cloud_optics(lwp, iwp, rel, rei, *cl);
    }
    // Incoming direct flux per g-point: solar source times the TSI scaling
    // (src_test/Radiation_solver.cpp:739-743).
    Array_gpu<Float,2> toa({ncol, ngpt});
    hip_check(K<Float>::spread_col(ncol, ngpt, kdist.solar_source_device().ptr(), toa.ptr(), stream()));
    hip_check(K<Float>::scale_by_col(ncol, ngpt, toa.ptr(), tsi_scaling.ptr(), stream()));
    Array_gpu<Float,2> ad_gpt = expand(kdist, sfc_alb_dir, ncol), af_gpt = expand(kdist, sfc_alb_dif, ncol);

    Array_gpu<Float,3> d_col({ncol, nlay, ngas + 2});
    Array_gpu<Float,2> d_col_dry({ncol, nlay}), d_ad({ncol, ngpt}), d_af({ncol, ngpt});
    zero(d_col); zero(out.d_t_lay); zero(out.d_lnp_lay); zero(out.d_p_lev);

    const int mb = max_band_gpts(kdist);
    Array_gpu<Float,3> tau({ncol, nlay, mb}), ssa({ncol, nlay, mb}), g({ncol, nlay, mb});
    Array_gpu<Float,3> dtau({ncol, nlay, mb}), dssa({ncol, nlay, mb}), dg({ncol, nlay, mb});
    const Float* ct = cl ? cl->get_tau().ptr() : nullptr;
    const Float* cs = cl ? cl->get_ssa().ptr() : nullptr;
    const Float* cg = cl ? cl->get_g().ptr() : nullptr;
    const Array<int,2>& bl = kdist.get_band_lims_gpoint();
    for (int b = 0; b < nbnd; ++b)
    {
        const int g0 = bl({1, b + 1}) - 1, ngb = bl({2, b + 1}) - g0;
        const int64_t o = int64_t(ncol) * g0;
        hip_check(K<Float>::gas_optics_sw_bands(&kd, b, 1, ncol, nlay, p_lay.ptr(), t_lay.ptr(), col_dry.ptr(), &vmr,
                                                ct, cs, cg, tau.ptr(), ssa.ptr(), g.ptr(), stream()));
        hip_check(K<Float>::sw_solver_2stream_adjoint(
                ncol, nlay, ngb, top_at_1, tau.ptr(), ssa.ptr(), g.ptr(), mu0.ptr(), ad_gpt.ptr() + o,
                af_gpt.ptr() + o, toa.ptr() + o, nullptr, g_flux_up.ptr(), g_flux_dn.ptr(), g_flux_dir.ptr(),
                dtau.ptr(), dssa.ptr(), dg.ptr(), d_ad.ptr() + o, d_af.ptr() + o, stream()));
        hip_check(K<Float>::gas_optics_sw_adjoint(&kd, b, ncol, nlay, p_lay.ptr(), t_lay.ptr(), col_dry.ptr(), &vmr,
                                                  ct, cs, cg, dtau.ptr(), dssa.ptr(), dg.ptr(), d_col.ptr(),
                                                  out.d_t_lay.ptr(), out.d_lnp_lay.ptr(), stream()));
    }
    hip_check(K<Float>::adjoint_col_chain(ncol, nlay, &vmr, col_dry.ptr(), d_col.ptr(), out.d_vmr.ptr(),
                                          d_col_dry.ptr(), stream()));
    hip_check(K<Float>::get_col_dry_adjoint(ncol, nlay, h2o.ptr(), h2o.dim(1), h2o.dim(2), p_lev.ptr(),
                                            d_col_dry.ptr(), out.d_vmr.ptr() + ncl * (kd.idx_h2o - 1),
                                            out.d_p_lev.ptr(), stream()));
    hip_check(K<Float>::gpt_to_band(ncol, nbnd, kd.band_lims_gpt, d_ad.ptr(), out.d_sfc_alb_dir.ptr(), stream()));
    hip_check(K<Float>::gpt_to_band(ncol, nbnd, kd.band_lims_gpt, d_af.ptr(), out.d_sfc_alb_dif.ptr(), stream()));
}
