// Longwave driver (reference src_cuda/Rte_lw.cu:81-152, CPU src/Rte_lw.cpp:90-162).
#include "Rte_lw.h"

#include <map>
#include <mutex>

#include "Optical_props.h"
#include "Source_functions.h"
#include "hip_dispatch.h"

using rrtmgp_cpp::hip_check;
using rrtmgp_cpp::stream;

namespace
{
// Gaussian quadrature secants and weights for 1-4 angles (src/Rte_lw.cpp:100-113).
const Float gauss_Ds[4][4] = {{1.66, 0., 0., 0.},
                              {1.18350343, 2.81649655, 0., 0.},
                              {1.09719858, 1.69338507, 4.70941630, 0.},
                              {1.06056257, 1.38282560, 2.40148179, 7.15513024}};
const Float gauss_wts[4][4] = {{0.5, 0., 0., 0.},
                               {0.3180413817, 0.1819586183, 0., 0.},
                               {0.2009319137, 0.2292411064, 0.0698269799, 0.},
                               {0.1355069134, 0.2034645680, 0.1298475476, 0.0311809710}};
}

namespace
{
// The Gauss tables on the device, [secants (4 x 4) | weights (4 x 4)], row n-1 holding the
// n-point quadrature: uploaded once per device and kept for the life of the process (32 values),
// so that a solve neither allocates nor copies them.
const Float* gauss_tables_device()
{
    static std::mutex m;
    static std::map<int, Float*> tables;
    int dev = 0;
    hip_check(rrtmgp_hip_get_device(&dev));
    std::lock_guard<std::mutex> lock(m);
    auto it = tables.find(dev);
    if (it != tables.end()) return it->second;
    Float h[32];
    for (int n = 0; n < 4; ++n)
        for (int i = 0; i < 4; ++i)
        {
            h[4 * n + i] = gauss_Ds[n][i];
            h[16 + 4 * n + i] = gauss_wts[n][i];
        }
    void* p = nullptr;
    hip_check(rrtmgp_hip_malloc(&p, sizeof(h)));
    hip_check(rrtmgp_hip_memcpy_h2d(p, h, sizeof(h), nullptr));
    tables[dev] = static_cast<Float*>(p);
    return static_cast<Float*>(p);
}

// Per-call solver inputs: surface emissivity by g-point, secants (ncol, ngpt, nmus) and weights.
struct LwPrep
{
    Array_gpu<Float,2> sfc_emis_gpt;
    Array_gpu<Float,1> weights;
    Array_gpu<Float,3> secants;
};

LwPrep lw_prep(Rte_lw_gpu& rte, const std::unique_ptr<Optical_props_arry_gpu>& optical_props,
               const Array_gpu<Float,2>& sfc_emis, const int n_gauss_angles)
{
    const int ncol = optical_props->get_ncol();
    const int ngpt = optical_props->get_ngpt();
    if (n_gauss_angles < 1 || n_gauss_angles > 4)
        throw std::runtime_error("n_gauss_angles must be between 1 and 4");
    LwPrep p;
    p.sfc_emis_gpt = Array_gpu<Float,2>({ncol, ngpt});
    rte.expand_and_transpose(optical_props, sfc_emis, p.sfc_emis_gpt);

    // Gauss tables (max_gauss_pts = 4) as the reference keeps them (src_cuda/Rte_lw.cu:91-124):
    // column nmus holds the secants / weights of an nmus-point quadrature.
    const int nmus = n_gauss_angles;
    Array<Float,2> ds_h({4, 4});
    Array<Float,1> weights_h({nmus});
    for (int n = 0; n < 4; ++n)
        for (int imu = 0; imu < 4; ++imu)
            ds_h({imu + 1, n + 1}) = gauss_Ds[n][imu];
    for (int imu = 0; imu < nmus; ++imu)
        weights_h({imu + 1}) = gauss_wts[nmus - 1][imu];
    Array_gpu<Float,2> ds(ds_h);
    p.weights = Array_gpu<Float,1>(weights_h);
    p.secants = Array_gpu<Float,3>({ncol, ngpt, nmus});
    hip_check(K<Float>::lw_secants_array(ncol, ngpt, nmus, 4, ds.ptr(), p.secants.ptr(), stream()));
    return p;
}
}  // namespace

void Rte_lw_gpu::rte_lw(
        const std::unique_ptr<Optical_props_arry_gpu>& optical_props,
        const Bool top_at_1,
        const Source_func_lw_gpu& sources,
        const Array_gpu<Float,2>& sfc_emis,
        const Array_gpu<Float,2>& inc_flux,
        Array_gpu<Float,3>& gpt_flux_up,
        Array_gpu<Float,3>& gpt_flux_dn,
        const int n_gauss_angles)
{
    const int ncol = optical_props->get_ncol();
    const int nlay = optical_props->get_nlay();
    const int ngpt = optical_props->get_ngpt();
    if (sources.is_factored()) throw std::runtime_error("rte_lw: g-point fluxes need unfactored sources");
    const LwPrep p = lw_prep(*this, optical_props, sfc_emis, n_gauss_angles);

    // CPU semantics: a (ncol, nlev, 1) output array selects broadband fluxes (src/Rte_lw.cpp:137).
    const bool do_broadband = (gpt_flux_up.dim(3) == 1) && (ngpt > 1);
    const Float* inc = (inc_flux.size() == 0) ? nullptr : inc_flux.ptr();
    hip_check(K<Float>::lw_solver_noscat(
            ncol, nlay, ngpt, top_at_1, n_gauss_angles, p.secants.ptr(), p.weights.ptr(),
            optical_props->get_tau().ptr(), sources.get_lay_source().ptr(),
            sources.get_lev_source_inc().ptr(), sources.get_lev_source_dec().ptr(),
            p.sfc_emis_gpt.ptr(), sources.get_sfc_source().ptr(), inc,
            do_broadband ? nullptr : gpt_flux_up.ptr(), do_broadband ? nullptr : gpt_flux_dn.ptr(),
            do_broadband, do_broadband ? gpt_flux_up.ptr() : nullptr,
            do_broadband ? gpt_flux_dn.ptr() : nullptr, 0, nullptr, nullptr, stream()));
}

void Rte_lw_gpu::rte_lw_fluxes(
        const std::unique_ptr<Optical_props_arry_gpu>& optical_props,
        const Bool top_at_1,
        const Source_func_lw_gpu& sources,
        const Array_gpu<Float,2>& sfc_emis,
        const Array_gpu<Float,2>& inc_flux,
        Array_gpu<Float,3>& bnd_flux_up, Array_gpu<Float,3>& bnd_flux_dn,
        Array_gpu<Float,2>& flux_up, Array_gpu<Float,2>& flux_dn,
        const int n_gauss_angles,
        const Array_gpu<Float,2>* p_lev, Array_gpu<Float,2>* heating_rate)
{
    const int ncol = optical_props->get_ncol();
    const int nlay = optical_props->get_nlay();
    const int ngpt = optical_props->get_ngpt();
    const int nbnd = optical_props->get_nband();
    const bool hr = heating_rate != nullptr && heating_rate->size() > 0;
    if (hr && (!p_lev || p_lev->dim(1) != ncol || p_lev->dim(2) != nlay + 1 || heating_rate->dim(1) != ncol ||
               heating_rate->dim(2) != nlay))
        throw std::runtime_error("rte_lw: heating rates need p_lev (ncol, nlev) and heating_rate (ncol, nlay)");
    const Float* plev = hr ? p_lev->ptr() : nullptr;
    Float* hrp = hr ? heating_rate->ptr() : nullptr;
    if (n_gauss_angles < 1 || n_gauss_angles > 4)
        throw std::runtime_error("n_gauss_angles must be between 1 and 4");
    if (sfc_emis.dim(1) != nbnd || sfc_emis.dim(2) != ncol)
        throw std::runtime_error("rte_lw: sfc_emis must be (nband, ncol)");
    // The band-summed solvers read the emissivity per band and one secant per angle.
    const Float* gt = gauss_tables_device();
    const Float* secants = gt + 4 * (n_gauss_angles - 1);
    const Float* weights = gt + 16 + 4 * (n_gauss_angles - 1);
    const Float* inc = (inc_flux.size() == 0) ? nullptr : inc_flux.ptr();
    const bool bnd = bnd_flux_up.size() > 0;
    const bool bb = flux_up.size() > 0;
    if (sources.is_factored())
    {
        hip_check(K<Float>::lw_solver_noscat_bnd_factored(
                ncol, nlay, ngpt, top_at_1, n_gauss_angles, secants, weights,
                optical_props->get_tau().ptr(), sources.get_planck_frac().ptr(), sources.get_band_lay_source().ptr(),
                sources.get_band_lev_source().ptr(), sfc_emis.ptr(), sources.get_sfc_source().ptr(), inc, nbnd,
                optical_props->get_band_lims_gpoint().ptr(), bnd ? bnd_flux_up.ptr() : nullptr,
                bnd ? bnd_flux_dn.ptr() : nullptr, bb ? flux_up.ptr() : nullptr, bb ? flux_dn.ptr() : nullptr,
                plev, hrp, stream()));
        return;
    }
    hip_check(K<Float>::lw_solver_noscat_bnd(
            ncol, nlay, ngpt, top_at_1, n_gauss_angles, secants, weights,
            optical_props->get_tau().ptr(), sources.get_lay_source().ptr(),
            sources.get_lev_source_inc().ptr(), sources.get_lev_source_dec().ptr(),
            sfc_emis.ptr(), sources.get_sfc_source().ptr(), inc, nbnd,
            optical_props->get_band_lims_gpoint().ptr(), bnd ? bnd_flux_up.ptr() : nullptr,
            bnd ? bnd_flux_dn.ptr() : nullptr, bb ? flux_up.ptr() : nullptr,
            bb ? flux_dn.ptr() : nullptr, plev, hrp, stream()));
}

void Rte_lw_gpu::expand_and_transpose(
        const std::unique_ptr<Optical_props_arry_gpu>& ops,
        const Array_gpu<Float,2> arr_in,
        Array_gpu<Float,2>& arr_out)
{
    hip_check(K<Float>::expand_and_transpose(arr_in.dim(2), ops->get_nband(), ops->get_ngpt(),
                                             ops->get_band_lims_gpoint_gpu().ptr(), arr_in.ptr(),
                                             arr_out.ptr(), stream()));
}
