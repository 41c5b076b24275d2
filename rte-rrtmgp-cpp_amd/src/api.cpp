// C facade over the class API (include/rrtmgp_api.h).
#include "rrtmgp_api.h"

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "Aerosol_optics.h"
#include "Cloud_optics.h"
#include "Fluxes.h"
#include "Gas_concs.h"
#include "Gas_optics_rrtmgp.h"
#include "Optical_props.h"
#include "Radiation_adjoint.h"
#include "Radiation_solver.h"
#include "Rte_lw.h"
#include "Rte_sw.h"
#include "Source_functions.h"
#include "hip_dispatch.h"

namespace
{
thread_local std::string api_error;

template<typename F>
int guarded(F&& f)
{
    try { f(); return RRTMGP_OK; }
    catch (const std::range_error& e) { api_error = e.what(); return RRTMGP_ERR_RANGE; }
    catch (const std::invalid_argument& e) { api_error = e.what(); return RRTMGP_ERR_ARG; }
    catch (const std::exception& e) { api_error = e.what(); return RRTMGP_ERR_RUNTIME; }
}

Array<std::string,1> names(const char* s, int n)
{
    std::vector<std::string> out;
    if (n > 0)
    {
        std::string cur;
        for (const char* p = s; ; ++p)
        {
            if (*p == '\n' || *p == '\0')
            {
                out.push_back(cur);
                cur.clear();
                if (*p == '\0') break;
            }
            else cur.push_back(*p);
        }
    }
    if (int(out.size()) != n) throw std::invalid_argument("name list length mismatch");
    return Array<std::string,1>(out, {n});
}

template<int N>
Array<Float,N> real_arr(const double* p, const std::array<int,N>& dims)
{
    Array<Float,N> a(dims);
    if (a.size() && !p) throw std::invalid_argument("missing table");
    for (int64_t i = 0; i < a.size(); ++i) a.v()[size_t(i)] = Float(p[i]);
    return a;
}

template<int N>
Array<int,N> int_arr(const int* p, const std::array<int,N>& dims)
{
    Array<int,N> a(dims);
    for (int64_t i = 0; i < a.size(); ++i) a.v()[size_t(i)] = p[i];
    return a;
}

Array<Bool,1> bool_arr(const signed char* p, int n)
{
    Array<Bool,1> a({n});
    for (int i = 0; i < n; ++i) a.v()[size_t(i)] = p[i];
    return a;
}

struct Kdist { std::shared_ptr<Gas_optics_rrtmgp_gpu> k; bool is_lw; };
struct Cloud { std::shared_ptr<Cloud_optics_gpu> c; };
struct Aerosol { std::shared_ptr<Aerosol_optics_gpu> a; };
struct Solver { std::unique_ptr<Radiation_solver_longwave_gpu> lw; std::unique_ptr<Radiation_solver_shortwave_gpu> sw; };

template<int N>
Array_gpu<Float,N> view(const rrtmgp_real* p, const std::array<int,N>& dims)
{
    if (!p) return Array_gpu<Float,N>();
    return Array_gpu<Float,N>(const_cast<Float*>(p), dims);
}

Kdist& K_(void* h) { if (!h) throw std::invalid_argument("NULL k-distribution handle"); return *static_cast<Kdist*>(h); }
}  // namespace

extern "C" {

int rrtmgp_api_precision(void) { return int(sizeof(Float)); }
const char* rrtmgp_api_last_error(void) { return api_error.c_str(); }
int rrtmgp_api_set_stream(void* s) { rrtmgp_cpp::set_stream(s); return RRTMGP_OK; }

int rrtmgp_api_kdist_create(const rrtmgp_kdist_file* f, void** out)
{
    return guarded([&] {
        const int nb = f->nbnd, ng = f->ngpt, nt = f->ntemp, np = f->npres, ne = f->neta;
        Gas_concs_gpu avail;
        const auto avail_names = names(f->available_gases, f->n_available);
        for (const auto& g : avail_names.v()) avail.set_vmr(g, Float(0.));
        auto gas = names(f->gas_names, f->ngas);
        auto ks = int_arr<3>(f->key_species, {2, 2, nb});
        auto b2g = int_arr<2>(f->band2gpt, {2, nb});
        auto blw = real_arr<2>(f->band_lims_wavenum, {2, nb});
        auto pr = real_arr<1>(f->press_ref, {np});
        auto tr = real_arr<1>(f->temp_ref, {nt});
        auto vr = real_arr<3>(f->vmr_ref, {2, f->ngas + 1, nt});
        auto km = real_arr<4>(f->kmajor, {ng, ne, np + 1, nt});
        auto gm = names(f->gas_minor, f->nminor_ident);
        auto im = names(f->identifier_minor, f->nminor_ident);
        const int nl = f->nminor_lower, nu = f->nminor_upper;
        auto kl = real_arr<3>(f->kminor_lower, {f->ncontrib_lower, ne, nt});
        auto ku = real_arr<3>(f->kminor_upper, {f->ncontrib_upper, ne, nt});
        auto mgl = names(f->minor_gases_lower, nl), mgu = names(f->minor_gases_upper, nu);
        auto mll = int_arr<2>(f->minor_limits_gpt_lower, {2, nl}), mlu = int_arr<2>(f->minor_limits_gpt_upper, {2, nu});
        auto dl = bool_arr(f->minor_scales_with_density_lower, nl), du = bool_arr(f->minor_scales_with_density_upper, nu);
        auto sgl = names(f->scaling_gas_lower, nl), sgu = names(f->scaling_gas_upper, nu);
        auto cl = bool_arr(f->scale_by_complement_lower, nl), cu = bool_arr(f->scale_by_complement_upper, nu);
        auto ksl = int_arr<1>(f->kminor_start_lower, {nl}), ksu = int_arr<1>(f->kminor_start_upper, {nu});
        std::unique_ptr<Kdist> kd(new Kdist);
        kd->is_lw = f->nplanck > 0;
        if (kd->is_lw)
        {
            kd->k = std::make_shared<Gas_optics_rrtmgp_gpu>(
                avail, gas, ks, b2g, blw, pr, Float(f->press_ref_trop), tr, Float(f->temp_ref_p),
                Float(f->temp_ref_t), vr, km, kl, ku, gm, im, mgl, mgu, mll, mlu, dl, du, sgl, sgu, cl, cu,
                ksl, ksu, real_arr<2>(f->totplnk, {f->nplanck, nb}), real_arr<4>(f->planck_frac, {ng, ne, np + 1, nt}),
                Array<Float,3>(), Array<Float,3>());
        }
        else
        {
            kd->k = std::make_shared<Gas_optics_rrtmgp_gpu>(
                avail, gas, ks, b2g, blw, pr, Float(f->press_ref_trop), tr, Float(f->temp_ref_p),
                Float(f->temp_ref_t), vr, km, kl, ku, gm, im, mgl, mgu, mll, mlu, dl, du, sgl, sgu, cl, cu,
                ksl, ksu, real_arr<1>(f->solar_src_quiet, {ng}), real_arr<1>(f->solar_src_facular, {ng}),
                real_arr<1>(f->solar_src_sunspot, {ng}), Float(f->tsi_default), Float(f->mg_default),
                Float(f->sb_default), real_arr<3>(f->rayl_lower, {ng, ne, nt}), real_arr<3>(f->rayl_upper, {ng, ne, nt}));
        }
        *out = kd.release();
    });
}

int rrtmgp_api_kdist_destroy(void* h) { delete static_cast<Kdist*>(h); return RRTMGP_OK; }

int rrtmgp_api_kdist_info(void* h, int64_t* info)
{
    return guarded([&] {
        auto& k = *K_(h).k;
        const rrtmgp_kdist_dev& d = k.device_tables();
        const int64_t v[] = {d.ngas, d.nflav, d.neta, d.npres, d.ntemp, d.ngpt, d.nbnd, d.nplanck,
                             d.nminor_lower, d.nminor_upper, d.idx_h2o, K_(h).is_lw ? 1 : 0};
        std::memcpy(info, v, sizeof(v));
    });
}

int rrtmgp_api_kdist_get_int(void* h, const char* name, int* out, int64_t cap, int64_t* n)
{
    return guarded([&] {
        auto& k = *K_(h).k;
        const std::string s(name);
        std::vector<int> v;
        auto take = [&](const auto& a) { v.assign(a.v().begin(), a.v().end()); };
        if (s == "flavor") take(k.get_flavor());
        else if (s == "gpoint_flavor") take(k.get_gpoint_flavor());
        else if (s == "minor_limits_gpt_lower") take(k.get_minor_limits_gpt(true));
        else if (s == "minor_limits_gpt_upper") take(k.get_minor_limits_gpt(false));
        else if (s == "kminor_start_lower") take(k.get_kminor_start(true));
        else if (s == "kminor_start_upper") take(k.get_kminor_start(false));
        else if (s == "idx_minor_lower") take(k.get_idx_minor(true));
        else if (s == "idx_minor_upper") take(k.get_idx_minor(false));
        else if (s == "idx_minor_scaling_lower") take(k.get_idx_minor_scaling(true));
        else if (s == "idx_minor_scaling_upper") take(k.get_idx_minor_scaling(false));
        else if (s == "gpt2band") take(k.get_gpoint_bands());
        else throw std::invalid_argument("unknown table " + s);
        if (int64_t(v.size()) > cap) throw std::invalid_argument("output buffer too small");
        std::copy(v.begin(), v.end(), out);
        *n = int64_t(v.size());
    });
}

int rrtmgp_api_kdist_get_real(void* h, const char* name, double* out, int64_t cap, int64_t* n)
{
    return guarded([&] {
        auto& k = *K_(h).k;
        const std::string s(name);
        std::vector<double> v;
        auto take = [&](const auto& a) { v.assign(a.v().begin(), a.v().end()); };
        if (s == "kmajor") take(k.get_kmajor());
        else if (s == "kminor_lower") take(k.get_kminor(true));
        else if (s == "kminor_upper") take(k.get_kminor(false));
        else if (s == "tsi") v.push_back(k.get_tsi());
        else if (s == "press_ref_min") v.push_back(k.get_press_ref_min());
        else if (s == "press_ref_max") v.push_back(k.get_press_ref_max());
        else if (s == "temp_ref_min") v.push_back(k.get_temp_min());
        else if (s == "temp_ref_max") v.push_back(k.get_temp_max());
        else throw std::invalid_argument("unknown table " + s);
        if (int64_t(v.size()) > cap) throw std::invalid_argument("output buffer too small");
        std::copy(v.begin(), v.end(), out);
        *n = int64_t(v.size());
    });
}

int rrtmgp_api_kdist_gas_names(void* h, char* buf, int len)
{
    return guarded([&] {
        std::string s;
        const auto& g = K_(h).k->get_gas_names();
        for (int i = 1; i <= g.dim(1); ++i) { if (i > 1) s += "\n"; s += g({i}); }
        if (int(s.size()) + 1 > len) throw std::invalid_argument("buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    });
}

int rrtmgp_api_kdist_set_check_ranges(void* h, int on)
{
    return guarded([&] { K_(h).k->set_check_ranges(on != 0); });
}

int rrtmgp_api_kdist_set_surrogate(void* h, const rrtmgp_mlp_net* net)
{
    return guarded([&] { K_(h).k->set_surrogate(net); });
}

int rrtmgp_api_kdist_last_fallback(void* h, int64_t* n)
{
    return guarded([&] { *n = K_(h).k->last_fallback_units(); });
}

int rrtmgp_api_gas_concs_create(const rrtmgp_api_gases* g, void** out)
{
    return guarded([&] {
        std::unique_ptr<Gas_concs_gpu> gc(new Gas_concs_gpu);
        auto nm = names(g->names, g->n);
        for (int i = 0; i < g->n; ++i)
        {
            const int d1 = int(g->dims[2 * i]), d2 = int(g->dims[2 * i + 1]);
            gc->set_vmr(nm({i + 1}), view<2>(static_cast<const rrtmgp_real*>(g->data[i]), {d1, d2}));
        }
        *out = gc.release();
    });
}

int rrtmgp_api_gas_concs_destroy(void* h) { delete static_cast<Gas_concs_gpu*>(h); return RRTMGP_OK; }

int rrtmgp_api_get_col_dry(int ncol, int nlay, const rrtmgp_real* vmr_h2o, int64_t d1, int64_t d2,
                           const rrtmgp_real* plev, rrtmgp_real* col_dry)
{
    return guarded([&] {
        auto cd = view<2>(col_dry, {ncol, nlay});
        Gas_optics_rrtmgp_gpu::get_col_dry(cd, view<2>(vmr_h2o, {int(d1), int(d2)}), view<2>(plev, {ncol, nlay + 1}));
    });
}

int rrtmgp_api_gas_optics_lw(void* h, int ncol, int nlay, const rrtmgp_real* play, const rrtmgp_real* plev,
                             const rrtmgp_real* tlay, const rrtmgp_real* tsfc, const rrtmgp_real* tlev,
                             void* gas_concs, const rrtmgp_real* col_dry, rrtmgp_real* tau,
                             rrtmgp_real* lay_src, rrtmgp_real* lev_src_inc, rrtmgp_real* lev_src_dec,
                             rrtmgp_real* sfc_src, rrtmgp_real* sfc_src_jac)
{
    return guarded([&] {
        auto& k = *K_(h).k;
        const int ngpt = k.get_ngpt();
        struct View1scl : public Optical_props_arry_gpu
        {
            Array_gpu<Float,3> t;
            View1scl(const Optical_props_gpu& p, Float* ptr, int nc, int nl, int ng) :
                Optical_props_arry_gpu(p), t(ptr, {nc, nl, ng}) {}
            Array_gpu<Float,3>& get_tau() { return t; }
            Array_gpu<Float,3>& get_ssa() { throw std::runtime_error("no ssa"); }
            Array_gpu<Float,3>& get_g() { throw std::runtime_error("no g"); }
            const Array_gpu<Float,3>& get_tau() const { return t; }
            const Array_gpu<Float,3>& get_ssa() const { throw std::runtime_error("no ssa"); }
            const Array_gpu<Float,3>& get_g() const { throw std::runtime_error("no g"); }
            void delta_scale(const Array_gpu<Float,3>&) {}
            int get_ncol() const { return t.dim(1); }
            int get_nlay() const { return t.dim(2); }
        } props(k, tau, ncol, nlay, ngpt);
        // Sources are written straight into the caller's arrays through a source object whose
        // storage is replaced by views.
        Source_func_lw_gpu src(0, 0, k);
        src.get_lay_source() = Array_gpu<Float,3>(lay_src, {ncol, nlay, ngpt});
        src.get_lev_source_inc() = Array_gpu<Float,3>(lev_src_inc, {ncol, nlay, ngpt});
        src.get_lev_source_dec() = Array_gpu<Float,3>(lev_src_dec, {ncol, nlay, ngpt});
        src.get_sfc_source() = Array_gpu<Float,2>(sfc_src, {ncol, ngpt});
        src.get_sfc_source_jac() = Array_gpu<Float,2>(sfc_src_jac, {ncol, ngpt});
        k.gas_optics_with_clouds(view<2>(play, {ncol, nlay}), view<2>(plev, {ncol, nlay + 1}),
                                 view<2>(tlay, {ncol, nlay}), view<1>(tsfc, {ncol}),
                                 *static_cast<Gas_concs_gpu*>(gas_concs), props, src,
                                 view<2>(col_dry, {ncol, nlay}), view<2>(tlev, {ncol, nlay + 1}), nullptr);
    });
}

namespace
{
int gas_optics_sw(void* h, int ncol, int nlay, const rrtmgp_real* play, const rrtmgp_real* plev,
                  const rrtmgp_real* tlay, void* gas_concs, const rrtmgp_real* col_dry,
                  rrtmgp_real* tau, rrtmgp_real* ssa, rrtmgp_real* g, rrtmgp_real* toa_src, bool coalbedo)
{
    return guarded([&] {
        auto& k = *K_(h).k;
        const int ngpt = k.get_ngpt();
        struct View2str : public Optical_props_arry_gpu
        {
            Array_gpu<Float,3> t, s, gg;
            View2str(const Optical_props_gpu& p, Float* a, Float* b, Float* c, int nc, int nl, int ng) :
                Optical_props_arry_gpu(p), t(a, {nc, nl, ng}), s(b, {nc, nl, ng}), gg(c, {nc, nl, ng}) {}
            Array_gpu<Float,3>& get_tau() { return t; }
            Array_gpu<Float,3>& get_ssa() { return s; }
            Array_gpu<Float,3>& get_g() { return gg; }
            const Array_gpu<Float,3>& get_tau() const { return t; }
            const Array_gpu<Float,3>& get_ssa() const { return s; }
            const Array_gpu<Float,3>& get_g() const { return gg; }
            void delta_scale(const Array_gpu<Float,3>&) {}
            int get_ncol() const { return t.dim(1); }
            int get_nlay() const { return t.dim(2); }
        } props(k, tau, ssa, g, ncol, nlay, ngpt);
        auto toa = view<2>(toa_src, {ncol, ngpt});
        k.gas_optics_with_clouds(view<2>(play, {ncol, nlay}), view<2>(plev, {ncol, nlay + 1}),
                                 view<2>(tlay, {ncol, nlay}), *static_cast<Gas_concs_gpu*>(gas_concs), props,
                                 toa, view<2>(col_dry, {ncol, nlay}), nullptr, coalbedo);
    });
}
}

int rrtmgp_api_gas_optics_sw(void* h, int ncol, int nlay, const rrtmgp_real* play, const rrtmgp_real* plev,
                             const rrtmgp_real* tlay, void* gas_concs, const rrtmgp_real* col_dry,
                             rrtmgp_real* tau, rrtmgp_real* ssa, rrtmgp_real* g, rrtmgp_real* toa_src)
{
    return gas_optics_sw(h, ncol, nlay, play, plev, tlay, gas_concs, col_dry, tau, ssa, g, toa_src, false);
}

int rrtmgp_api_gas_optics_sw_coalbedo(void* h, int ncol, int nlay, const rrtmgp_real* play, const rrtmgp_real* plev,
                                      const rrtmgp_real* tlay, void* gas_concs, const rrtmgp_real* col_dry,
                                      rrtmgp_real* tau, rrtmgp_real* coalbedo, rrtmgp_real* g, rrtmgp_real* toa_src)
{
    return gas_optics_sw(h, ncol, nlay, play, plev, tlay, gas_concs, col_dry, tau, coalbedo, g, toa_src, true);
}

int rrtmgp_api_kdist_fill_col_gas(void* h, int ncol, int nlay, void* gas_concs, const rrtmgp_real* col_dry,
                                  rrtmgp_real* col_gas)
{
    return guarded([&] {
        auto& k = *K_(h).k;
        rrtmgp_gas_vmr v;
        k.vmr_descriptor(*static_cast<Gas_concs_gpu*>(gas_concs), ncol, nlay, v);
        rrtmgp_cpp::hip_check(K<Float>::fill_col_gas(ncol, nlay, &v, col_dry, col_gas, rrtmgp_cpp::stream()));
    });
}

int rrtmgp_api_kdist_interpolation(void* h, int ncol, int nlay, const rrtmgp_real* play, const rrtmgp_real* tlay,
                                   const rrtmgp_real* col_gas, int* jtemp, rrtmgp_real* fmajor, rrtmgp_real* fminor,
                                   rrtmgp_real* col_mix, signed char* tropo, int* jeta, int* jpress)
{
    return guarded([&] {
        auto& k = *K_(h).k;
        rrtmgp_cpp::hip_check(K<Float>::interpolation(&k.device_tables(), ncol, nlay, play, tlay, col_gas, jtemp,
                                                      fmajor, fminor, col_mix, tropo, jeta, jpress,
                                                      rrtmgp_cpp::stream()));
    });
}

int rrtmgp_api_cloud_create(const rrtmgp_cloud_lut* l, void** out)
{
    return guarded([&] {
        std::unique_ptr<Cloud> c(new Cloud);
        c->c = std::make_shared<Cloud_optics_gpu>(
            real_arr<2>(l->band_lims_wvn, {2, l->nbnd}), Float(l->radliq_lwr), Float(l->radliq_upr),
            Float(l->radliq_fac), Float(l->radice_lwr), Float(l->radice_upr), Float(l->radice_fac),
            real_arr<2>(l->lut_extliq, {l->nsize_liq, l->nbnd}), real_arr<2>(l->lut_ssaliq, {l->nsize_liq, l->nbnd}),
            real_arr<2>(l->lut_asyliq, {l->nsize_liq, l->nbnd}), real_arr<3>(l->lut_extice, {l->nsize_ice, l->nbnd, 3}),
            real_arr<3>(l->lut_ssaice, {l->nsize_ice, l->nbnd, 3}), real_arr<3>(l->lut_asyice, {l->nsize_ice, l->nbnd, 3}));
        *out = c.release();
    });
}

int rrtmgp_api_cloud_destroy(void* h) { delete static_cast<Cloud*>(h); return RRTMGP_OK; }

int rrtmgp_api_aerosol_create(const rrtmgp_aerosol_lut* l, void** out)
{
    return guarded([&] {
        if (!l || !out) throw std::invalid_argument("aerosol_create: NULL argument");
        std::unique_ptr<Aerosol> a(new Aerosol);
        a->a = std::make_shared<Aerosol_optics_gpu>(
            real_arr<2>(l->band_lims_wvn, {2, l->nbnd}), real_arr<1>(l->rh_upper, {l->nhum}),
            real_arr<2>(l->mext_phobic, {l->nbnd, l->nphobic}), real_arr<2>(l->ssa_phobic, {l->nbnd, l->nphobic}),
            real_arr<2>(l->g_phobic, {l->nbnd, l->nphobic}),
            real_arr<3>(l->mext_philic, {l->nbnd, l->nhum, l->nphilic}),
            real_arr<3>(l->ssa_philic, {l->nbnd, l->nhum, l->nphilic}),
            real_arr<3>(l->g_philic, {l->nbnd, l->nhum, l->nphilic}));
        *out = a.release();
    });
}

int rrtmgp_api_aerosol_destroy(void* h) { delete static_cast<Aerosol*>(h); return RRTMGP_OK; }

int rrtmgp_api_aerosol_optics(void* h, int ncol, int nlay, void* concs, const rrtmgp_real* rh,
                              const rrtmgp_real* plev, int delta_scale, int coalbedo, rrtmgp_real* tau,
                              rrtmgp_real* ssa, rrtmgp_real* g)
{
    return guarded([&] {
        auto& a = *static_cast<Aerosol*>(h)->a;
        const int nb = a.get_nband();
        Optical_props_2str_gpu p(0, 0, a);
        p.get_tau() = Array_gpu<Float,3>(tau, {ncol, nlay, nb});
        p.get_ssa() = Array_gpu<Float,3>(ssa, {ncol, nlay, nb});
        p.get_g() = Array_gpu<Float,3>(g, {ncol, nlay, nb});
        a.compute(*static_cast<Gas_concs_gpu*>(concs), view<2>(rh, {ncol, nlay}), view<2>(plev, {ncol, nlay + 1}),
                  delta_scale != 0, p, coalbedo != 0);
    });
}

int rrtmgp_api_cloud_optics(void* h, int ncol, int nlay, const rrtmgp_real* lwp, const rrtmgp_real* iwp,
                            const rrtmgp_real* rel, const rrtmgp_real* rei, int two_stream, int delta_scale,
                            rrtmgp_real* tau, rrtmgp_real* ssa, rrtmgp_real* g)
{
    return guarded([&] {
        auto& c = *static_cast<Cloud*>(h)->c;
        const int nb = c.get_nband();
        if (two_stream)
        {
            Optical_props_2str_gpu p(0, 0, c);
            p.get_tau() = Array_gpu<Float,3>(tau, {ncol, nlay, nb});
            p.get_ssa() = Array_gpu<Float,3>(ssa, {ncol, nlay, nb});
            p.get_g() = Array_gpu<Float,3>(g, {ncol, nlay, nb});
            if (two_stream == 2)
                c.cloud_optics_coalbedo(view<2>(lwp, {ncol, nlay}), view<2>(iwp, {ncol, nlay}),
                                        view<2>(rel, {ncol, nlay}), view<2>(rei, {ncol, nlay}), delta_scale != 0, p);
            else if (delta_scale)
                c.cloud_optics_delta_scaled(view<2>(lwp, {ncol, nlay}), view<2>(iwp, {ncol, nlay}),
                                            view<2>(rel, {ncol, nlay}), view<2>(rei, {ncol, nlay}), p);
            else
                c.cloud_optics(view<2>(lwp, {ncol, nlay}), view<2>(iwp, {ncol, nlay}),
                               view<2>(rel, {ncol, nlay}), view<2>(rei, {ncol, nlay}), p);
        }
        else
        {
            Optical_props_1scl_gpu p(0, 0, c);
            p.get_tau() = Array_gpu<Float,3>(tau, {ncol, nlay, nb});
            c.cloud_optics(view<2>(lwp, {ncol, nlay}), view<2>(iwp, {ncol, nlay}),
                           view<2>(rel, {ncol, nlay}), view<2>(rei, {ncol, nlay}), p);
        }
    });
}

int rrtmgp_api_add_to(void* h, int ncol, int nlay, int nspec_in, rrtmgp_real* tau, rrtmgp_real* ssa,
                      rrtmgp_real* g, const rrtmgp_real* tau_in, const rrtmgp_real* ssa_in,
                      const rrtmgp_real* g_in)
{
    return guarded([&] {
        auto& k = *K_(h).k;
        const int ngpt = k.get_ngpt();
        const Optical_props_gpu band_only(k.get_band_lims_wavenumber());
        const Optical_props_gpu& grid_in = (nspec_in == ngpt) ? static_cast<const Optical_props_gpu&>(k) : band_only;
        if (nspec_in != ngpt && nspec_in != k.get_nband())
            throw std::runtime_error("Cannot add optical properties with incompatible band - gpoint combination");
        if (ssa)
        {
            Optical_props_2str_gpu a(0, 0, k), b(0, 0, grid_in);
            a.get_tau() = Array_gpu<Float,3>(tau, {ncol, nlay, ngpt});
            a.get_ssa() = Array_gpu<Float,3>(ssa, {ncol, nlay, ngpt});
            a.get_g() = Array_gpu<Float,3>(g, {ncol, nlay, ngpt});
            b.get_tau() = view<3>(tau_in, {ncol, nlay, nspec_in});
            b.get_ssa() = view<3>(ssa_in, {ncol, nlay, nspec_in});
            b.get_g() = view<3>(g_in, {ncol, nlay, nspec_in});
            add_to(a, b);
        }
        else
        {
            Optical_props_1scl_gpu a(0, 0, k), b(0, 0, grid_in);
            a.get_tau() = Array_gpu<Float,3>(tau, {ncol, nlay, ngpt});
            b.get_tau() = view<3>(tau_in, {ncol, nlay, nspec_in});
            add_to(a, b);
        }
    });
}

int rrtmgp_api_add_to_coalbedo(void* h, int ncol, int nlay, rrtmgp_real* tau, rrtmgp_real* coalbedo,
                               rrtmgp_real* g, const rrtmgp_real* tau_in, const rrtmgp_real* coalbedo_in,
                               const rrtmgp_real* g_in)
{
    return guarded([&] {
        auto& k = *K_(h).k;
        const int ngpt = k.get_ngpt(), nbnd = k.get_nband();
        const Optical_props_gpu band_only(k.get_band_lims_wavenumber());
        Optical_props_2str_gpu a(0, 0, k), b(0, 0, band_only);
        a.get_tau() = Array_gpu<Float,3>(tau, {ncol, nlay, ngpt});
        a.get_ssa() = Array_gpu<Float,3>(coalbedo, {ncol, nlay, ngpt});
        a.get_g() = Array_gpu<Float,3>(g, {ncol, nlay, ngpt});
        b.get_tau() = view<3>(tau_in, {ncol, nlay, nbnd});
        b.get_ssa() = view<3>(coalbedo_in, {ncol, nlay, nbnd});
        b.get_g() = view<3>(g_in, {ncol, nlay, nbnd});
        add_to_coalbedo(a, b);
    });
}

int rrtmgp_api_delta_scale(int ncol, int nlay, int ngpt, rrtmgp_real* tau, rrtmgp_real* ssa, rrtmgp_real* g)
{
    return guarded([&] {
        rrtmgp_cpp::hip_check(sizeof(Float) == 8
            ? rrtmgp_hip_delta_scale_2str_f64(ncol, nlay, ngpt, reinterpret_cast<double*>(tau),
                                              reinterpret_cast<double*>(ssa), reinterpret_cast<double*>(g), rrtmgp_cpp::stream())
            : rrtmgp_hip_delta_scale_2str_f32(ncol, nlay, ngpt, reinterpret_cast<float*>(tau),
                                              reinterpret_cast<float*>(ssa), reinterpret_cast<float*>(g), rrtmgp_cpp::stream()));
    });
}

namespace
{
struct ViewProps : public Optical_props_arry_gpu
{
    Array_gpu<Float,3> t, s, gg;
    ViewProps(const Optical_props_gpu& p, const Float* a, const Float* b, const Float* c, int nc, int nl, int ng) :
        Optical_props_arry_gpu(p), t(const_cast<Float*>(a), {nc, nl, ng})
    {
        if (b) s = Array_gpu<Float,3>(const_cast<Float*>(b), {nc, nl, ng});
        if (c) gg = Array_gpu<Float,3>(const_cast<Float*>(c), {nc, nl, ng});
    }
    Array_gpu<Float,3>& get_tau() { return t; }
    Array_gpu<Float,3>& get_ssa() { return s; }
    Array_gpu<Float,3>& get_g() { return gg; }
    const Array_gpu<Float,3>& get_tau() const { return t; }
    const Array_gpu<Float,3>& get_ssa() const { return s; }
    const Array_gpu<Float,3>& get_g() const { return gg; }
    void delta_scale(const Array_gpu<Float,3>&) {}
    int get_ncol() const { return t.dim(1); }
    int get_nlay() const { return t.dim(2); }
};
}

int rrtmgp_api_rte_lw(void* h, int ncol, int nlay, int top_at_1, const rrtmgp_real* tau,
                      const rrtmgp_real* lay_src, const rrtmgp_real* lev_src_inc, const rrtmgp_real* lev_src_dec,
                      const rrtmgp_real* sfc_src, const rrtmgp_real* sfc_emis, const rrtmgp_real* inc_flux,
                      int n_gauss_angles, int do_broadband, rrtmgp_real* flux_up, rrtmgp_real* flux_dn)
{
    return guarded([&] {
        auto& k = *K_(h).k;
        const int ngpt = k.get_ngpt(), nbnd = k.get_nband();
        std::unique_ptr<Optical_props_arry_gpu> p(new ViewProps(k, tau, nullptr, nullptr, ncol, nlay, ngpt));
        Source_func_lw_gpu src(0, 0, k);
        src.get_lay_source() = view<3>(lay_src, {ncol, nlay, ngpt});
        src.get_lev_source_inc() = view<3>(lev_src_inc, {ncol, nlay, ngpt});
        src.get_lev_source_dec() = view<3>(lev_src_dec, {ncol, nlay, ngpt});
        src.get_sfc_source() = view<2>(sfc_src, {ncol, ngpt});
        const int ng_out = do_broadband ? 1 : ngpt;
        auto up = Array_gpu<Float,3>(flux_up, {ncol, nlay + 1, ng_out});
        auto dn = Array_gpu<Float,3>(flux_dn, {ncol, nlay + 1, ng_out});
        Rte_lw_gpu rte;
        rte.rte_lw(p, top_at_1, src, view<2>(sfc_emis, {nbnd, ncol}),
                   inc_flux ? view<2>(inc_flux, {ncol, ngpt}) : Array_gpu<Float,2>(), up, dn, n_gauss_angles);
    });
}

int rrtmgp_api_rte_sw(void* h, int ncol, int nlay, int top_at_1, const rrtmgp_real* tau, const rrtmgp_real* ssa,
                      const rrtmgp_real* g, const rrtmgp_real* mu0, const rrtmgp_real* inc_flux_dir,
                      const rrtmgp_real* sfc_alb_dir, const rrtmgp_real* sfc_alb_dif,
                      const rrtmgp_real* inc_flux_dif, int do_broadband, rrtmgp_real* flux_up,
                      rrtmgp_real* flux_dn, rrtmgp_real* flux_dir)
{
    return guarded([&] {
        auto& k = *K_(h).k;
        const int ngpt = k.get_ngpt(), nbnd = k.get_nband();
        std::unique_ptr<Optical_props_arry_gpu> p(new ViewProps(k, tau, ssa, g, ncol, nlay, ngpt));
        const int ng_out = do_broadband ? 1 : ngpt;
        auto up = Array_gpu<Float,3>(flux_up, {ncol, nlay + 1, ng_out});
        auto dn = Array_gpu<Float,3>(flux_dn, {ncol, nlay + 1, ng_out});
        auto dr = Array_gpu<Float,3>(flux_dir, {ncol, nlay + 1, ng_out});
        Rte_sw_gpu rte;
        rte.rte_sw(p, top_at_1, view<1>(mu0, {ncol}), view<2>(inc_flux_dir, {ncol, ngpt}),
                   view<2>(sfc_alb_dir, {nbnd, ncol}), view<2>(sfc_alb_dif, {nbnd, ncol}),
                   inc_flux_dif ? view<2>(inc_flux_dif, {ncol, ngpt}) : Array_gpu<Float,2>(), up, dn, dr);
    });
}

int rrtmgp_api_fluxes_reduce(void* h, int ncol, int nlev, const rrtmgp_real* gpt_up, const rrtmgp_real* gpt_dn,
                             const rrtmgp_real* gpt_dir, rrtmgp_real* flux_up, rrtmgp_real* flux_dn,
                             rrtmgp_real* flux_dir, rrtmgp_real* flux_net, rrtmgp_real* bnd_up,
                             rrtmgp_real* bnd_dn, rrtmgp_real* bnd_dir, rrtmgp_real* bnd_net)
{
    return guarded([&] {
        auto& k = *K_(h).k;
        const int ngpt = k.get_ngpt(), nbnd = k.get_nband();
        std::unique_ptr<Optical_props_arry_gpu> p(new ViewProps(k, gpt_up, nullptr, nullptr, ncol, nlev, ngpt));
        auto up = view<3>(gpt_up, {ncol, nlev, ngpt}), dn = view<3>(gpt_dn, {ncol, nlev, ngpt});
        std::unique_ptr<Fluxes_broadband_gpu> f;
        if (bnd_up) f.reset(new Fluxes_byband_gpu(ncol, nlev, nbnd));
        else f.reset(new Fluxes_broadband_gpu(ncol, nlev));
        if (gpt_dir) f->reduce(up, dn, view<3>(gpt_dir, {ncol, nlev, ngpt}), p, true);
        else f->reduce(up, dn, p, true);
        auto out2 = [&](rrtmgp_real* dst, const Array_gpu<Float,2>& a) {
            if (dst) rrtmgp_cpp::hip_check(rrtmgp_hip_memcpy_d2d(dst, a.ptr(), a.bytes(), rrtmgp_cpp::stream()));
        };
        auto out3 = [&](rrtmgp_real* dst, const Array_gpu<Float,3>& a) {
            if (dst) rrtmgp_cpp::hip_check(rrtmgp_hip_memcpy_d2d(dst, a.ptr(), a.bytes(), rrtmgp_cpp::stream()));
        };
        out2(flux_up, f->get_flux_up()); out2(flux_dn, f->get_flux_dn()); out2(flux_net, f->get_flux_net());
        if (gpt_dir) out2(flux_dir, f->get_flux_dn_dir());
        if (bnd_up)
        {
            out3(bnd_up, f->get_bnd_flux_up()); out3(bnd_dn, f->get_bnd_flux_dn()); out3(bnd_net, f->get_bnd_flux_net());
            if (gpt_dir) out3(bnd_dir, f->get_bnd_flux_dn_dir());
        }
        rrtmgp_cpp::hip_check(rrtmgp_hip_stream_synchronize(rrtmgp_cpp::stream()));
    });
}

int rrtmgp_api_heating_rate(int ncol, int nlay, const rrtmgp_real* flux_up, const rrtmgp_real* flux_dn,
                            const rrtmgp_real* plev, rrtmgp_real* hr)
{
    return guarded([&] {
        auto h = view<2>(hr, {ncol, nlay});
        compute_heating_rate(view<2>(flux_up, {ncol, nlay + 1}), view<2>(flux_dn, {ncol, nlay + 1}),
                             view<2>(plev, {ncol, nlay + 1}), h);
    });
}

int rrtmgp_api_solver_lw_create(void* kd, void* cloud, void** out)
{
    return guarded([&] {
        std::unique_ptr<Solver> s(new Solver);
        s->lw.reset(new Radiation_solver_longwave_gpu(K_(kd).k, cloud ? static_cast<Cloud*>(cloud)->c : nullptr));
        *out = s.release();
    });
}

int rrtmgp_api_solver_sw_create(void* kd, void* cloud, void** out)
{
    return guarded([&] {
        std::unique_ptr<Solver> s(new Solver);
        s->sw.reset(new Radiation_solver_shortwave_gpu(K_(kd).k, cloud ? static_cast<Cloud*>(cloud)->c : nullptr));
        *out = s.release();
    });
}

int rrtmgp_api_solver_sw_create_aerosol(void* kd, void* cloud, void* aerosol, void** out)
{
    return guarded([&] {
        std::unique_ptr<Solver> s(new Solver);
        s->sw.reset(new Radiation_solver_shortwave_gpu(K_(kd).k, cloud ? static_cast<Cloud*>(cloud)->c : nullptr,
                                                       aerosol ? static_cast<Aerosol*>(aerosol)->a : nullptr));
        *out = s.release();
    });
}

int rrtmgp_api_solver_destroy(void* h) { delete static_cast<Solver*>(h); return RRTMGP_OK; }

int rrtmgp_api_solver_set_block(void* h, int n)
{
    return guarded([&] {
        auto* s = static_cast<Solver*>(h);
        if (s->lw) s->lw->set_n_col_block(n);
        if (s->sw) s->sw->set_n_col_block(n);
    });
}

int rrtmgp_api_solve_lw(void* h, int ncol, int nlay, int cloudy, int bnd, const rrtmgp_real* play,
                        const rrtmgp_real* plev, const rrtmgp_real* tlay, const rrtmgp_real* tlev,
                        const rrtmgp_real* tsfc, void* gas_concs, const rrtmgp_real* col_dry,
                        const rrtmgp_real* emis_sfc, const rrtmgp_real* lwp, const rrtmgp_real* iwp,
                        const rrtmgp_real* rel, const rrtmgp_real* rei, rrtmgp_real* flux_up,
                        rrtmgp_real* flux_dn, rrtmgp_real* flux_net, rrtmgp_real* bnd_up,
                        rrtmgp_real* bnd_dn, rrtmgp_real* bnd_net, rrtmgp_real* heating_rate)
{
    return guarded([&] {
        auto* s = static_cast<Solver*>(h);
        if (!s->lw) throw std::invalid_argument("not a longwave solver");
        auto hr = heating_rate ? view<2>(heating_rate, {ncol, nlay}) : Array_gpu<Float,2>();
        const int nlev = nlay + 1, nbnd = s->lw->get_n_bnd();
        auto fu = view<2>(flux_up, {ncol, nlev}), fd = view<2>(flux_dn, {ncol, nlev}), fn = view<2>(flux_net, {ncol, nlev});
        auto bu = bnd ? view<3>(bnd_up, {ncol, nlev, nbnd}) : Array_gpu<Float,3>();
        auto bd = bnd ? view<3>(bnd_dn, {ncol, nlev, nbnd}) : Array_gpu<Float,3>();
        auto bn = bnd ? view<3>(bnd_net, {ncol, nlev, nbnd}) : Array_gpu<Float,3>();
        s->lw->solve_gpu(cloudy != 0, bnd != 0, *static_cast<Gas_concs_gpu*>(gas_concs),
                         view<2>(play, {ncol, nlay}), view<2>(plev, {ncol, nlev}), view<2>(tlay, {ncol, nlay}),
                         view<2>(tlev, {ncol, nlev}), col_dry ? view<2>(col_dry, {ncol, nlay}) : Array_gpu<Float,2>(),
                         view<1>(tsfc, {ncol}), view<2>(emis_sfc, {nbnd, ncol}),
                         cloudy ? view<2>(lwp, {ncol, nlay}) : Array_gpu<Float,2>(),
                         cloudy ? view<2>(iwp, {ncol, nlay}) : Array_gpu<Float,2>(),
                         cloudy ? view<2>(rel, {ncol, nlay}) : Array_gpu<Float,2>(),
                         cloudy ? view<2>(rei, {ncol, nlay}) : Array_gpu<Float,2>(),
                         fu, fd, fn, bu, bd, bn, heating_rate ? &hr : nullptr);
    });
}

int rrtmgp_api_solve_sw(void* h, int ncol, int nlay, int cloudy, int bnd, int delta_cloud,
                        const rrtmgp_real* play, const rrtmgp_real* plev, const rrtmgp_real* tlay,
                        const rrtmgp_real* tlev, void* gas_concs, const rrtmgp_real* col_dry,
                        const rrtmgp_real* sfc_alb_dir, const rrtmgp_real* sfc_alb_dif,
                        const rrtmgp_real* tsi_scaling, const rrtmgp_real* mu0, const rrtmgp_real* lwp,
                        const rrtmgp_real* iwp, const rrtmgp_real* rel, const rrtmgp_real* rei,
                        rrtmgp_real* flux_up, rrtmgp_real* flux_dn, rrtmgp_real* flux_dir, rrtmgp_real* flux_net,
                        rrtmgp_real* bnd_up, rrtmgp_real* bnd_dn, rrtmgp_real* bnd_dir, rrtmgp_real* bnd_net,
                        rrtmgp_real* heating_rate)
{
    return guarded([&] {
        auto* s = static_cast<Solver*>(h);
        if (!s->sw) throw std::invalid_argument("not a shortwave solver");
        auto hr = heating_rate ? view<2>(heating_rate, {ncol, nlay}) : Array_gpu<Float,2>();
        const int nlev = nlay + 1, nbnd = s->sw->get_n_bnd();
        auto fu = view<2>(flux_up, {ncol, nlev}), fd = view<2>(flux_dn, {ncol, nlev});
        auto fdir = view<2>(flux_dir, {ncol, nlev}), fn = view<2>(flux_net, {ncol, nlev});
        auto bu = bnd ? view<3>(bnd_up, {ncol, nlev, nbnd}) : Array_gpu<Float,3>();
        auto bd = bnd ? view<3>(bnd_dn, {ncol, nlev, nbnd}) : Array_gpu<Float,3>();
        auto bdir = bnd ? view<3>(bnd_dir, {ncol, nlev, nbnd}) : Array_gpu<Float,3>();
        auto bn = bnd ? view<3>(bnd_net, {ncol, nlev, nbnd}) : Array_gpu<Float,3>();
        s->sw->solve_gpu(cloudy != 0, bnd != 0, delta_cloud != 0, *static_cast<Gas_concs_gpu*>(gas_concs),
                         view<2>(play, {ncol, nlay}), view<2>(plev, {ncol, nlev}), view<2>(tlay, {ncol, nlay}),
                         view<2>(tlev, {ncol, nlev}), col_dry ? view<2>(col_dry, {ncol, nlay}) : Array_gpu<Float,2>(),
                         view<2>(sfc_alb_dir, {nbnd, ncol}), view<2>(sfc_alb_dif, {nbnd, ncol}),
                         view<1>(tsi_scaling, {ncol}), view<1>(mu0, {ncol}),
                         cloudy ? view<2>(lwp, {ncol, nlay}) : Array_gpu<Float,2>(),
                         cloudy ? view<2>(iwp, {ncol, nlay}) : Array_gpu<Float,2>(),
                         cloudy ? view<2>(rel, {ncol, nlay}) : Array_gpu<Float,2>(),
                         cloudy ? view<2>(rei, {ncol, nlay}) : Array_gpu<Float,2>(),
                         fu, fd, fdir, fn, bu, bd, bdir, bn, heating_rate ? &hr : nullptr);
    });
}


int rrtmgp_api_solve_sw_aerosol(void* h, int ncol, int nlay, int cloudy, int aerosol, int bnd, int delta_cloud,
                                int delta_aerosol, const rrtmgp_real* play, const rrtmgp_real* plev,
                                const rrtmgp_real* tlay, const rrtmgp_real* tlev, void* gas_concs,
                                const rrtmgp_real* col_dry, const rrtmgp_real* sfc_alb_dir,
                                const rrtmgp_real* sfc_alb_dif, const rrtmgp_real* tsi_scaling,
                                const rrtmgp_real* mu0, const rrtmgp_real* lwp, const rrtmgp_real* iwp,
                                const rrtmgp_real* rel, const rrtmgp_real* rei, const rrtmgp_real* rh,
                                void* aerosol_concs, rrtmgp_real* flux_up, rrtmgp_real* flux_dn,
                                rrtmgp_real* flux_dir, rrtmgp_real* flux_net, rrtmgp_real* bnd_up,
                                rrtmgp_real* bnd_dn, rrtmgp_real* bnd_dir, rrtmgp_real* bnd_net,
                                rrtmgp_real* heating_rate)
{
    return guarded([&] {
        auto* s = static_cast<Solver*>(h);
        if (!s->sw) throw std::invalid_argument("not a shortwave solver");
        if (aerosol && (!rh || !aerosol_concs)) throw std::invalid_argument("aerosols need rh and aerosol_concs");
        auto hr = heating_rate ? view<2>(heating_rate, {ncol, nlay}) : Array_gpu<Float,2>();
        const int nlev = nlay + 1, nbnd = s->sw->get_n_bnd();
        auto fu = view<2>(flux_up, {ncol, nlev}), fd = view<2>(flux_dn, {ncol, nlev});
        auto fdir = view<2>(flux_dir, {ncol, nlev}), fn = view<2>(flux_net, {ncol, nlev});
        auto bu = bnd ? view<3>(bnd_up, {ncol, nlev, nbnd}) : Array_gpu<Float,3>();
        auto bd = bnd ? view<3>(bnd_dn, {ncol, nlev, nbnd}) : Array_gpu<Float,3>();
        auto bdir = bnd ? view<3>(bnd_dir, {ncol, nlev, nbnd}) : Array_gpu<Float,3>();
        auto bn = bnd ? view<3>(bnd_net, {ncol, nlev, nbnd}) : Array_gpu<Float,3>();
        auto c = [&](const rrtmgp_real* p) { return cloudy ? view<2>(p, {ncol, nlay}) : Array_gpu<Float,2>(); };
        const Aerosol_concs_gpu empty;
        s->sw->solve_gpu(cloudy != 0, aerosol != 0, bnd != 0, delta_cloud != 0, delta_aerosol != 0,
                         *static_cast<Gas_concs_gpu*>(gas_concs),
                         view<2>(play, {ncol, nlay}), view<2>(plev, {ncol, nlev}), view<2>(tlay, {ncol, nlay}),
                         view<2>(tlev, {ncol, nlev}), col_dry ? view<2>(col_dry, {ncol, nlay}) : Array_gpu<Float,2>(),
                         view<2>(sfc_alb_dir, {nbnd, ncol}), view<2>(sfc_alb_dif, {nbnd, ncol}),
                         view<1>(tsi_scaling, {ncol}), view<1>(mu0, {ncol}), c(lwp), c(iwp), c(rel), c(rei),
                         aerosol ? view<2>(rh, {ncol, nlay}) : Array_gpu<Float,2>(),
                         aerosol ? *static_cast<Gas_concs_gpu*>(aerosol_concs) : empty,
                         fu, fd, fdir, fn, bu, bd, bdir, bn, heating_rate ? &hr : nullptr);
    });
}

int rrtmgp_api_adjoint_lw(void* h, int ncol, int nlay, int cloudy, const rrtmgp_real* play,
                          const rrtmgp_real* plev, const rrtmgp_real* tlay, const rrtmgp_real* tlev,
                          const rrtmgp_real* tsfc, void* gas_concs, const rrtmgp_real* emis_sfc,
                          const rrtmgp_real* lwp, const rrtmgp_real* iwp, const rrtmgp_real* rel,
                          const rrtmgp_real* rei, const rrtmgp_real* g_flux_up,
                          const rrtmgp_real* g_flux_dn, rrtmgp_real* d_tlay, rrtmgp_real* d_tlev,
                          rrtmgp_real* d_tsfc, rrtmgp_real* d_lnp_lay, rrtmgp_real* d_plev,
                          rrtmgp_real* d_vmr, rrtmgp_real* d_emis_sfc)
{
    return guarded([&] {
        auto* s = static_cast<Solver*>(h);
        if (!s || !s->lw) throw std::invalid_argument("not a longwave solver");
        const int nlev = nlay + 1, nbnd = s->lw->get_n_bnd();
        Gas_optics_rrtmgp_gpu& kd = s->lw->get_kdist();
        Lw_adjoint_out o;
        o.d_t_lay = view<2>(d_tlay, {ncol, nlay}); o.d_t_lev = view<2>(d_tlev, {ncol, nlev});
        o.d_t_sfc = view<1>(d_tsfc, {ncol}); o.d_lnp_lay = view<2>(d_lnp_lay, {ncol, nlay});
        o.d_p_lev = view<2>(d_plev, {ncol, nlev}); o.d_vmr = view<3>(d_vmr, {ncol, nlay, kd.get_ngas()});
        o.d_emis_sfc = view<2>(d_emis_sfc, {nbnd, ncol});
        auto c = [&](const rrtmgp_real* p) { return cloudy ? view<2>(p, {ncol, nlay}) : Array_gpu<Float,2>(); };
        lw_adjoint_gpu(kd, cloudy ? s->lw->get_cloud_optics() : nullptr, *static_cast<Gas_concs_gpu*>(gas_concs),
                       view<2>(play, {ncol, nlay}), view<2>(plev, {ncol, nlev}), view<2>(tlay, {ncol, nlay}),
                       view<2>(tlev, {ncol, nlev}), view<1>(tsfc, {ncol}), view<2>(emis_sfc, {nbnd, ncol}),
                       c(lwp), c(iwp), c(rel), c(rei), view<2>(g_flux_up, {ncol, nlev}),
                       view<2>(g_flux_dn, {ncol, nlev}), o);
    });
}

int rrtmgp_api_adjoint_sw(void* h, int ncol, int nlay, int cloudy, int delta_cloud,
                          const rrtmgp_real* play, const rrtmgp_real* plev, const rrtmgp_real* tlay,
                          void* gas_concs, const rrtmgp_real* sfc_alb_dir,
                          const rrtmgp_real* sfc_alb_dif, const rrtmgp_real* tsi_scaling,
                          const rrtmgp_real* mu0, const rrtmgp_real* lwp, const rrtmgp_real* iwp,
                          const rrtmgp_real* rel, const rrtmgp_real* rei,
                          const rrtmgp_real* g_flux_up, const rrtmgp_real* g_flux_dn,
                          const rrtmgp_real* g_flux_dir, rrtmgp_real* d_tlay,
                          rrtmgp_real* d_lnp_lay, rrtmgp_real* d_plev, rrtmgp_real* d_vmr,
                          rrtmgp_real* d_sfc_alb_dir, rrtmgp_real* d_sfc_alb_dif)
{
    return guarded([&] {
        auto* s = static_cast<Solver*>(h);
        if (!s || !s->sw) throw std::invalid_argument("not a shortwave solver");
        const int nlev = nlay + 1, nbnd = s->sw->get_n_bnd();
        Gas_optics_rrtmgp_gpu& kd = s->sw->get_kdist();
        Sw_adjoint_out o;
        o.d_t_lay = view<2>(d_tlay, {ncol, nlay}); o.d_lnp_lay = view<2>(d_lnp_lay, {ncol, nlay});
        o.d_p_lev = view<2>(d_plev, {ncol, nlev}); o.d_vmr = view<3>(d_vmr, {ncol, nlay, kd.get_ngas()});
        o.d_sfc_alb_dir = view<2>(d_sfc_alb_dir, {nbnd, ncol}); o.d_sfc_alb_dif = view<2>(d_sfc_alb_dif, {nbnd, ncol});
        auto c = [&](const rrtmgp_real* p) { return cloudy ? view<2>(p, {ncol, nlay}) : Array_gpu<Float,2>(); };
        sw_adjoint_gpu(kd, cloudy ? s->sw->get_cloud_optics() : nullptr, delta_cloud != 0,
                       *static_cast<Gas_concs_gpu*>(gas_concs), view<2>(play, {ncol, nlay}),
                       view<2>(plev, {ncol, nlev}), view<2>(tlay, {ncol, nlay}), view<2>(sfc_alb_dir, {nbnd, ncol}),
                       view<2>(sfc_alb_dif, {nbnd, ncol}), view<1>(tsi_scaling, {ncol}), view<1>(mu0, {ncol}),
                       c(lwp), c(iwp), c(rel), c(rei), view<2>(g_flux_up, {ncol, nlev}),
                       view<2>(g_flux_dn, {ncol, nlev}), view<2>(g_flux_dir, {ncol, nlev}), o);
    });
}

}  // extern "C"
