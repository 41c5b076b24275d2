// Gas_optics_rrtmgp_gpu: host-side initialisation of the k-distribution (restating
// src/Gas_optics_rrtmgp.cpp:40-363 helpers, :391-742 constructors/init_abs_coeffs,
// :745-760 set_solar_variability) and the fused device gas-optics pass.
#include "Gas_optics_rrtmgp.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "Gas_concs.h"
#include "Optical_props.h"
#include "Source_functions.h"
#include "hip_dispatch.h"

namespace
{
// The surrogate's entry points are FP32 only; set_surrogate rejects FP64 builds before any call.
template<typename T> struct Mlp
{
    template<typename... A> static int lw(A...) { throw std::invalid_argument("MLP surrogate: FP32 only"); }
    template<typename... A> static int lw_factored(A...) { throw std::invalid_argument("MLP surrogate: FP32 only"); }
    template<typename... A> static int sw(A...) { throw std::invalid_argument("MLP surrogate: FP32 only"); }
};
template<> struct Mlp<float>
{
    static constexpr auto lw = rrtmgp_hip_mlp_gas_optics_lw;
    static constexpr auto lw_factored = rrtmgp_hip_mlp_gas_optics_lw_factored;
    static constexpr auto sw = rrtmgp_hip_mlp_gas_optics_sw;
};

// The surrogate's input gases, in its order (h2o, o3, co2, ch4, n2o), from the gas concentrations.
rrtmgp_gas_vmr mlp_vmr_desc(const Gas_concs_gpu& gas_desc, int ncol, int nlay)
{
    static const char* names[] = {"h2o", "o3", "co2", "ch4", "n2o"};
    rrtmgp_gas_vmr out{};
    out.ngas = 5;
    for (int i = 0; i < 5; ++i)
    {
        if (!gas_desc.exists(names[i]))
            throw std::runtime_error(std::string("MLP surrogate: gas ") + names[i] + " missing from gas_desc");
        const Array_gpu<Float,2>& v = gas_desc.get_vmr(names[i]);
        if ((v.dim(1) != 1 && v.dim(1) != ncol) || (v.dim(2) != 1 && v.dim(2) != nlay))
            throw std::runtime_error(std::string("gas ") + names[i] + " has dims incompatible with (ncol, nlay)");
        out.vmr[i] = v.ptr();
        out.dim1[i] = v.dim(1);
        out.dim2[i] = v.dim(2);
    }
    return out;
}
}  // namespace

using rrtmgp_cpp::hip_check;
using rrtmgp_cpp::stream;

namespace
{
int find_index(const Array<std::string,1>& data, const std::string& value)
{
    auto it = std::find(data.v().begin(), data.v().end(), value);
    return it == data.v().end() ? -1 : int(it - data.v().begin()) + 1;
}

std::string trimmed(const std::string& s)
{
    const size_t b = s.find_first_not_of(" \t");
    if (b == std::string::npos) return "";
    return s.substr(b, s.find_last_not_of(" \t") - b + 1);
}

// Keep the minor contributors whose gas is available; compact kminor (file order
// (ncontrib, neta, ntemp)) and re-base kminor_start, then reorder to (ntemp, neta, ncontrib).
void reduce_minor_arrays(
        const Gas_concs_gpu& available_gases,
        const Array<std::string,1>& gas_minor,
        const Array<std::string,1>& identifier_minor,
        const Array<Float,3>& kminor_atm,
        const Array<std::string,1>& minor_gases_atm,
        const Array<int,2>& minor_limits_gpt_atm,
        const Array<Bool,1>& minor_scales_with_density_atm,
        const Array<std::string,1>& scaling_gas_atm,
        const Array<Bool,1>& scale_by_complement_atm,
        const Array<int,1>& kminor_start_atm,
        Array<Float,3>& kminor_red,
        Array<std::string,1>& minor_gases_red,
        Array<int,2>& limits_red,
        Array<Bool,1>& dens_red,
        Array<std::string,1>& scaling_gas_red,
        Array<Bool,1>& compl_red,
        Array<int,1>& kstart_red)
{
    const int nm = minor_gases_atm.dim(1);
    std::vector<int> keep;
    int tot_g = 0;
    for (int i = 1; i <= nm; ++i)
    {
        const int idx_mnr = find_index(identifier_minor, minor_gases_atm({i}));
        if (idx_mnr < 1) throw std::runtime_error("unknown minor gas identifier " + minor_gases_atm({i}));
        if (available_gases.exists(trimmed(gas_minor({idx_mnr}))))
        {
            keep.push_back(i);
            tot_g += minor_limits_gpt_atm({2, i}) - minor_limits_gpt_atm({1, i}) + 1;
        }
    }
    const int red_nm = int(keep.size());
    const int neta = kminor_atm.dim(2), ntemp = kminor_atm.dim(3);
    const int ncontrib = (red_nm == nm) ? kminor_atm.dim(1) : tot_g;

    minor_gases_red = Array<std::string,1>({red_nm});
    scaling_gas_red = Array<std::string,1>({red_nm});
    limits_red = Array<int,2>({2, red_nm});
    dens_red = Array<Bool,1>({red_nm});
    compl_red = Array<Bool,1>({red_nm});
    kstart_red = Array<int,1>({red_nm});
    Array<Float,3> kt({ncontrib, neta, ntemp});

    int n_elim = 0, icnt = 0, prev = 0;
    for (int i : keep)
    {
        // widths of the dropped intervals before i shift the contributor offsets
        for (int j = prev + 1; j < i; ++j)
            n_elim += minor_limits_gpt_atm({2, j}) - minor_limits_gpt_atm({1, j}) + 1;
        prev = i;
        ++icnt;
        minor_gases_red({icnt}) = minor_gases_atm({i});
        scaling_gas_red({icnt}) = scaling_gas_atm({i});
        limits_red({1, icnt}) = minor_limits_gpt_atm({1, i});
        limits_red({2, icnt}) = minor_limits_gpt_atm({2, i});
        dens_red({icnt}) = minor_scales_with_density_atm({i});
        compl_red({icnt}) = scale_by_complement_atm({i});
        kstart_red({icnt}) = (red_nm == nm) ? kminor_start_atm({i}) : kminor_start_atm({i}) - n_elim;
        const int ng = minor_limits_gpt_atm({2, i}) - minor_limits_gpt_atm({1, i}) + 1;
        for (int j = 1; j <= ng; ++j)
            for (int e = 1; e <= neta; ++e)
                for (int t = 1; t <= ntemp; ++t)
                    kt({kstart_red({icnt}) + j - 1, e, t}) = kminor_atm({kminor_start_atm({i}) + j - 1, e, t});
    }
    kminor_red = Array<Float,3>({ntemp, neta, ncontrib});
    for (int c = 1; c <= ncontrib; ++c)
        for (int e = 1; e <= neta; ++e)
            for (int t = 1; t <= ntemp; ++t)
                kminor_red({t, e, c}) = kt({c, e, t});
}

// Unique key-species pairs in band order, (0,0) rewritten to (2,2)
// (create_flavor, src/Gas_optics_rrtmgp.cpp:247-313).
Array<int,2> create_flavor(const Array<int,3>& key_species)
{
    std::vector<std::pair<int,int>> list, uniq;
    for (int b = 1; b <= key_species.dim(3); ++b)
        for (int a = 1; a <= key_species.dim(2); ++a)
        {
            int k1 = key_species({1, a, b}), k2 = key_species({2, a, b});
            if (k1 == 0 && k2 == 0) { k1 = 2; k2 = 2; }
            list.emplace_back(k1, k2);
        }
    for (const auto& p : list)
        if (std::find(uniq.begin(), uniq.end(), p) == uniq.end()) uniq.push_back(p);
    Array<int,2> flavor({2, int(uniq.size())});
    for (int i = 0; i < int(uniq.size()); ++i) { flavor({1, i + 1}) = uniq[i].first; flavor({2, i + 1}) = uniq[i].second; }
    return flavor;
}

// Flavor of every g-point in the lower/upper atmosphere (create_gpoint_flavor, :331-363).
Array<int,2> create_gpoint_flavor(const Array<int,3>& key_species, const Array<int,1>& gpt2band,
                                  const Array<int,2>& flavor)
{
    const int ngpt = gpt2band.dim(1);
    Array<int,2> gf({2, ngpt});
    for (int g = 1; g <= ngpt; ++g)
        for (int a = 1; a <= 2; ++a)
        {
            int k1 = key_species({1, a, gpt2band({g})}), k2 = key_species({2, a, gpt2band({g})});
            if (k1 == 0 && k2 == 0) { k1 = 2; k2 = 2; }
            int f = -1;
            for (int i = 1; i <= flavor.dim(2); ++i)
                if (flavor({1, i}) == k1 && flavor({2, i}) == k2) { f = i; break; }
            gf({a, g}) = f;
        }
    return gf;
}

template<typename T, int N>
void up(Array_gpu<T,N>& d, const Array<T,N>& h) { if (h.size()) d = Array_gpu<T,N>(h); }
}  // namespace


Gas_optics_rrtmgp_gpu::Gas_optics_rrtmgp_gpu(
        const Gas_concs_gpu& available_gases, const Array<std::string,1>& gas_names,
        const Array<int,3>& key_species, const Array<int,2>& band2gpt,
        const Array<Float,2>& band_lims_wavenum, const Array<Float,1>& press_ref,
        const Float press_ref_trop, const Array<Float,1>& temp_ref, const Float temp_ref_p,
        const Float temp_ref_t, const Array<Float,3>& vmr_ref, const Array<Float,4>& kmajor,
        const Array<Float,3>& kminor_lower, const Array<Float,3>& kminor_upper,
        const Array<std::string,1>& gas_minor, const Array<std::string,1>& identifier_minor,
        const Array<std::string,1>& minor_gases_lower, const Array<std::string,1>& minor_gases_upper,
        const Array<int,2>& minor_limits_gpt_lower, const Array<int,2>& minor_limits_gpt_upper,
        const Array<Bool,1>& minor_scales_with_density_lower,
        const Array<Bool,1>& minor_scales_with_density_upper,
        const Array<std::string,1>& scaling_gas_lower, const Array<std::string,1>& scaling_gas_upper,
        const Array<Bool,1>& scale_by_complement_lower, const Array<Bool,1>& scale_by_complement_upper,
        const Array<int,1>& kminor_start_lower, const Array<int,1>& kminor_start_upper,
        const Array<Float,2>& totplnk, const Array<Float,4>& planck_frac,
        const Array<Float,3>& rayl_lower, const Array<Float,3>& rayl_upper) :
    Gas_optics_gpu(band_lims_wavenum, band2gpt), totplnk(totplnk)
{
    // planck_frac (ngpt, neta, npres+1, ntemp) -> (ntemp, neta, npres+1, ngpt) (:428-433).
    this->planck_frac = Array<Float,4>({planck_frac.dim(4), planck_frac.dim(2), planck_frac.dim(3), planck_frac.dim(1)});
    for (int i4 = 1; i4 <= this->planck_frac.dim(4); ++i4)
        for (int i3 = 1; i3 <= this->planck_frac.dim(3); ++i3)
            for (int i2 = 1; i2 <= this->planck_frac.dim(2); ++i2)
                for (int i1 = 1; i1 <= this->planck_frac.dim(1); ++i1)
                    this->planck_frac({i1, i2, i3, i4}) = planck_frac({i4, i2, i3, i1});

    init_abs_coeffs(available_gases, gas_names, key_species, band2gpt, band_lims_wavenum, press_ref,
                    temp_ref, press_ref_trop, temp_ref_p, temp_ref_t, vmr_ref, kmajor, kminor_lower,
                    kminor_upper, gas_minor, identifier_minor, minor_gases_lower, minor_gases_upper,
                    minor_limits_gpt_lower, minor_limits_gpt_upper, minor_scales_with_density_lower,
                    minor_scales_with_density_upper, scaling_gas_lower, scaling_gas_upper,
                    scale_by_complement_lower, scale_by_complement_upper, kminor_start_lower,
                    kminor_start_upper, rayl_lower, rayl_upper);

    totplnk_delta = (temp_ref_max - temp_ref_min) / (totplnk.dim(1) - 1);
    upload();
}


Gas_optics_rrtmgp_gpu::Gas_optics_rrtmgp_gpu(
        const Gas_concs_gpu& available_gases, const Array<std::string,1>& gas_names,
        const Array<int,3>& key_species, const Array<int,2>& band2gpt,
        const Array<Float,2>& band_lims_wavenum, const Array<Float,1>& press_ref,
        const Float press_ref_trop, const Array<Float,1>& temp_ref, const Float temp_ref_p,
        const Float temp_ref_t, const Array<Float,3>& vmr_ref, const Array<Float,4>& kmajor,
        const Array<Float,3>& kminor_lower, const Array<Float,3>& kminor_upper,
        const Array<std::string,1>& gas_minor, const Array<std::string,1>& identifier_minor,
        const Array<std::string,1>& minor_gases_lower, const Array<std::string,1>& minor_gases_upper,
        const Array<int,2>& minor_limits_gpt_lower, const Array<int,2>& minor_limits_gpt_upper,
        const Array<Bool,1>& minor_scales_with_density_lower,
        const Array<Bool,1>& minor_scales_with_density_upper,
        const Array<std::string,1>& scaling_gas_lower, const Array<std::string,1>& scaling_gas_upper,
        const Array<Bool,1>& scale_by_complement_lower, const Array<Bool,1>& scale_by_complement_upper,
        const Array<int,1>& kminor_start_lower, const Array<int,1>& kminor_start_upper,
        const Array<Float,1>& solar_src_quiet, const Array<Float,1>& solar_src_facular,
        const Array<Float,1>& solar_src_sunspot, const Float tsi_default, const Float mg_default,
        const Float sb_default, const Array<Float,3>& rayl_lower, const Array<Float,3>& rayl_upper) :
    Gas_optics_gpu(band_lims_wavenum, band2gpt)
{
    init_abs_coeffs(available_gases, gas_names, key_species, band2gpt, band_lims_wavenum, press_ref,
                    temp_ref, press_ref_trop, temp_ref_p, temp_ref_t, vmr_ref, kmajor, kminor_lower,
                    kminor_upper, gas_minor, identifier_minor, minor_gases_lower, minor_gases_upper,
                    minor_limits_gpt_lower, minor_limits_gpt_upper, minor_scales_with_density_lower,
                    minor_scales_with_density_upper, scaling_gas_lower, scaling_gas_upper,
                    scale_by_complement_lower, scale_by_complement_upper, kminor_start_lower,
                    kminor_start_upper, rayl_lower, rayl_upper);
    solar_source_quiet = solar_src_quiet;
    solar_source_facular = solar_src_facular;
    solar_source_sunspot = solar_src_sunspot;
    solar_source = Array<Float,1>(solar_src_quiet.get_dims());
    set_solar_variability(mg_default, sb_default);
    (void)tsi_default;
    upload();
}


void Gas_optics_rrtmgp_gpu::init_abs_coeffs(
        const Gas_concs_gpu& available_gases, const Array<std::string,1>& gas_names_in,
        const Array<int,3>& key_species, const Array<int,2>& band2gpt,
        const Array<Float,2>&, const Array<Float,1>& press_ref_in, const Array<Float,1>& temp_ref_in,
        const Float press_ref_trop, const Float, const Float, const Array<Float,3>& vmr_ref_in,
        const Array<Float,4>& kmajor_in, const Array<Float,3>& kminor_lower_in,
        const Array<Float,3>& kminor_upper_in, const Array<std::string,1>& gas_minor,
        const Array<std::string,1>& identifier_minor, const Array<std::string,1>& minor_gases_lower,
        const Array<std::string,1>& minor_gases_upper, const Array<int,2>& minor_limits_gpt_lower_in,
        const Array<int,2>& minor_limits_gpt_upper_in, const Array<Bool,1>& dens_lower_in,
        const Array<Bool,1>& dens_upper_in, const Array<std::string,1>& scaling_gas_lower,
        const Array<std::string,1>& scaling_gas_upper, const Array<Bool,1>& compl_lower_in,
        const Array<Bool,1>& compl_upper_in, const Array<int,1>& kstart_lower_in,
        const Array<int,1>& kstart_upper_in, const Array<Float,3>& rayl_lower,
        const Array<Float,3>& rayl_upper)
{
    // Gases known to the k-distribution and present in the host model (:560-573).
    std::vector<std::string> use;
    for (const std::string& s : gas_names_in.v())
        if (available_gases.exists(s)) use.push_back(s);
    const int n_gas = int(use.size());
    gas_names = Array<std::string,1>(use, {n_gas});

    // vmr_ref reduced to (2, ngas+1, ntemp), gas 0 = dry air (:575-595).
    vmr_ref = Array<Float,3>({vmr_ref_in.dim(1), n_gas + 1, vmr_ref_in.dim(3)});
    for (int i1 = 1; i1 <= vmr_ref.dim(1); ++i1)
        for (int i3 = 1; i3 <= vmr_ref.dim(3); ++i3)
        {
            vmr_ref({i1, 1, i3}) = vmr_ref_in({i1, 1, i3});
            for (int i = 1; i <= n_gas; ++i)
                vmr_ref({i1, i + 1, i3}) = vmr_ref_in({i1, find_index(gas_names_in, gas_names({i})) + 1, i3});
        }

    Array<std::string,1> mg_lower_red, sg_lower_red, mg_upper_red, sg_upper_red;
    reduce_minor_arrays(available_gases, gas_minor, identifier_minor, kminor_lower_in, minor_gases_lower,
                        minor_limits_gpt_lower_in, dens_lower_in, scaling_gas_lower, compl_lower_in,
                        kstart_lower_in, kminor_lower, mg_lower_red, minor_limits_gpt_lower,
                        minor_scales_with_density_lower, sg_lower_red, scale_by_complement_lower,
                        kminor_start_lower);
    reduce_minor_arrays(available_gases, gas_minor, identifier_minor, kminor_upper_in, minor_gases_upper,
                        minor_limits_gpt_upper_in, dens_upper_in, scaling_gas_upper, compl_upper_in,
                        kstart_upper_in, kminor_upper, mg_upper_red, minor_limits_gpt_upper,
                        minor_scales_with_density_upper, sg_upper_red, scale_by_complement_upper,
                        kminor_start_upper);

    press_ref = press_ref_in;
    temp_ref = temp_ref_in;

    // kmajor (ngpt, neta, npres+1, ntemp) -> (ntemp, neta, npres+1, ngpt) (:658-663).
    kmajor = Array<Float,4>({kmajor_in.dim(4), kmajor_in.dim(2), kmajor_in.dim(3), kmajor_in.dim(1)});
    for (int i4 = 1; i4 <= kmajor.dim(4); ++i4)
        for (int i3 = 1; i3 <= kmajor.dim(3); ++i3)
            for (int i2 = 1; i2 <= kmajor.dim(2); ++i2)
                for (int i1 = 1; i1 <= kmajor.dim(1); ++i1)
                    kmajor({i1, i2, i3, i4}) = kmajor_in({i4, i2, i3, i1});
    kmajor_ntemp = kmajor.dim(1); kmajor_neta = kmajor.dim(2); kmajor_npres = kmajor.dim(3) - 1;

    // krayl (ntemp, neta, ngpt, 2) from rayl_lower/upper (ngpt, neta, ntemp) (:667-677).
    if (rayl_lower.size() > 0)
    {
        krayl = Array<Float,4>({rayl_lower.dim(3), rayl_lower.dim(2), rayl_lower.dim(1), 2});
        for (int i3 = 1; i3 <= krayl.dim(3); ++i3)
            for (int i2 = 1; i2 <= krayl.dim(2); ++i2)
                for (int i1 = 1; i1 <= krayl.dim(1); ++i1)
                {
                    krayl({i1, i2, i3, 1}) = rayl_lower({i3, i2, i1});
                    krayl({i1, i2, i3, 2}) = rayl_upper({i3, i2, i1});
                }
    }

    press_ref_log = Array<Float,1>(press_ref.get_dims());
    for (int i = 1; i <= press_ref.dim(1); ++i) press_ref_log({i}) = std::log(press_ref({i}));
    press_ref_trop_log = std::log(press_ref_trop);

    // create_idx_minor / create_idx_minor_scaling (:165-196).
    auto idx_minor = [&](const Array<std::string,1>& mg) {
        Array<int,1> o({mg.dim(1)});
        for (int i = 1; i <= mg.dim(1); ++i)
            o({i}) = find_index(gas_names, gas_minor({find_index(identifier_minor, mg({i}))}));
        return o;
    };
    idx_minor_lower = idx_minor(mg_lower_red);
    idx_minor_upper = idx_minor(mg_upper_red);
    auto idx_scal = [&](const Array<std::string,1>& sg) {
        Array<int,1> o({sg.dim(1)});
        for (int i = 1; i <= sg.dim(1); ++i) o({i}) = find_index(gas_names, sg({i}));
        return o;
    };
    idx_minor_scaling_lower = idx_scal(sg_lower_red);
    idx_minor_scaling_upper = idx_scal(sg_upper_red);

    // Key species reduced to the present gases; a missing key gas is an error (:198-245).
    Array<int,3> ks_red(key_species.get_dims());
    std::vector<int> present(gas_names_in.dim(1), 1);
    for (int64_t i = 0; i < key_species.size(); ++i)
    {
        const int ks = key_species.v()[size_t(i)];
        if (ks != 0)
        {
            const int ksr = find_index(gas_names, gas_names_in({ks}));
            ks_red.v()[size_t(i)] = ksr;
            if (ksr == -1) present[size_t(ks - 1)] = 0;
        }
        else
            ks_red.v()[size_t(i)] = 0;
    }
    for (int i = 0; i < gas_names_in.dim(1); ++i)
        if (!present[size_t(i)])
            throw std::runtime_error("Gas optics: required gas " + gas_names_in({i + 1}) + " is missing");

    flavor = create_flavor(ks_red);
    gpoint_flavor = create_gpoint_flavor(ks_red, this->get_gpoint_bands(), flavor);

    temp_ref_min = temp_ref({1});
    temp_ref_max = temp_ref({temp_ref.dim(1)});
    press_ref_min = press_ref({press_ref.dim(1)});
    press_ref_max = press_ref({1});
    press_ref_log_delta = (std::log(press_ref_min) - std::log(press_ref_max)) / (press_ref.dim(1) - 1);
    temp_ref_delta = (temp_ref_max - temp_ref_min) / (temp_ref.dim(1) - 1);

    is_key = Array<int,1>({n_gas});
    for (int j = 1; j <= flavor.dim(2); ++j)
        for (int i = 1; i <= 2; ++i)
            if (flavor({i, j}) != 0) is_key({flavor({i, j})}) = 1;

    idx_h2o = find_index(gas_names, "h2o");
    (void)band2gpt;
}


void Gas_optics_rrtmgp_gpu::set_solar_variability(const Float mg_index, const Float sb_index)
{
    constexpr Float a_offset = Float(0.1495954);
    constexpr Float b_offset = Float(0.00066696);
    for (int igpt = 1; igpt <= solar_source_quiet.dim(1); ++igpt)
        solar_source({igpt}) = solar_source_quiet({igpt})
            + (mg_index - a_offset) * solar_source_facular({igpt})
            + (sb_index - b_offset) * solar_source_sunspot({igpt});
}


Float Gas_optics_rrtmgp_gpu::get_tsi() const
{
    if (solar_source.size() != get_ngpt())
        throw std::runtime_error("get_tsi: the k-distribution has no solar source (longwave tables)");
    Float tsi = 0.;
    for (int igpt = 1; igpt <= get_ngpt(); ++igpt) tsi += solar_source({igpt});
    return tsi;
}


void Gas_optics_rrtmgp_gpu::upload()
{
    up(press_ref_log_gpu, press_ref_log); up(temp_ref_gpu, temp_ref); up(vmr_ref_gpu, vmr_ref);
    up(kmajor_gpu, kmajor); up(kminor_lower_gpu, kminor_lower); up(kminor_upper_gpu, kminor_upper);
    up(flavor_gpu, flavor); up(gpoint_flavor_gpu, gpoint_flavor);
    up(minor_limits_gpt_lower_gpu, minor_limits_gpt_lower); up(minor_limits_gpt_upper_gpu, minor_limits_gpt_upper);
    up(minor_scales_with_density_lower_gpu, minor_scales_with_density_lower);
    up(minor_scales_with_density_upper_gpu, minor_scales_with_density_upper);
    up(scale_by_complement_lower_gpu, scale_by_complement_lower);
    up(scale_by_complement_upper_gpu, scale_by_complement_upper);
    up(kminor_start_lower_gpu, kminor_start_lower); up(kminor_start_upper_gpu, kminor_start_upper);
    up(idx_minor_lower_gpu, idx_minor_lower); up(idx_minor_upper_gpu, idx_minor_upper);
    up(idx_minor_scaling_lower_gpu, idx_minor_scaling_lower);
    up(idx_minor_scaling_upper_gpu, idx_minor_scaling_upper);
    up(totplnk_gpu, totplnk); up(planck_frac_gpu, planck_frac); up(krayl_gpu, krayl);
    up(solar_source_gpu, solar_source);

    dev = rrtmgp_kdist_dev{};
    dev.ngas = get_ngas(); dev.nflav = get_nflav(); dev.neta = kmajor_neta; dev.npres = kmajor_npres;
    dev.ntemp = kmajor_ntemp; dev.ngpt = get_ngpt(); dev.nbnd = get_nband();
    dev.nplanck = totplnk.size() ? totplnk.dim(1) : 0;
    dev.idx_h2o = idx_h2o;
    dev.press_ref_log_delta = press_ref_log_delta; dev.temp_ref_min = temp_ref_min;
    dev.temp_ref_delta = temp_ref_delta; dev.press_ref_trop_log = press_ref_trop_log;
    dev.totplnk_delta = totplnk_delta;
    dev.nminor_lower = minor_limits_gpt_lower.dim(2); dev.nminor_upper = minor_limits_gpt_upper.dim(2);
    dev.flavor = flavor_gpu.ptr(); dev.gpoint_flavor = gpoint_flavor_gpu.ptr();
    dev.band_lims_gpt = get_band_lims_gpoint_gpu().ptr();
    dev.press_ref_log = press_ref_log_gpu.ptr(); dev.temp_ref = temp_ref_gpu.ptr();
    dev.vmr_ref = vmr_ref_gpu.ptr(); dev.kmajor = kmajor_gpu.ptr();
    dev.kminor_lower = kminor_lower_gpu.ptr(); dev.kminor_upper = kminor_upper_gpu.ptr();
    dev.minor_limits_gpt_lower = minor_limits_gpt_lower_gpu.ptr();
    dev.minor_limits_gpt_upper = minor_limits_gpt_upper_gpu.ptr();
    dev.minor_scales_with_density_lower = minor_scales_with_density_lower_gpu.ptr();
    dev.minor_scales_with_density_upper = minor_scales_with_density_upper_gpu.ptr();
    dev.scale_by_complement_lower = scale_by_complement_lower_gpu.ptr();
    dev.scale_by_complement_upper = scale_by_complement_upper_gpu.ptr();
    dev.idx_minor_lower = idx_minor_lower_gpu.ptr(); dev.idx_minor_upper = idx_minor_upper_gpu.ptr();
    dev.idx_minor_scaling_lower = idx_minor_scaling_lower_gpu.ptr();
    dev.idx_minor_scaling_upper = idx_minor_scaling_upper_gpu.ptr();
    dev.kminor_start_lower = kminor_start_lower_gpu.ptr(); dev.kminor_start_upper = kminor_start_upper_gpu.ptr();
    dev.totplnk = totplnk_gpu.ptr(); dev.planck_frac = planck_frac_gpu.ptr(); dev.krayl = krayl_gpu.ptr();

    dev.ncontrib_lower = kminor_lower.size() ? kminor_lower.dim(3) : 0;
    dev.ncontrib_upper = kminor_upper.size() ? kminor_upper.dim(3) : 0;

    // Band-packed tables of the fused gas-optics kernels.
    void* pk = nullptr;
    hip_check(K<Float>::kdist_pack(&dev, &pk));
    go_pack.reset(pk, [](void* q) { rrtmgp_hip_kdist_pack_free(q); });
}


void Gas_optics_rrtmgp_gpu::get_col_dry(Array_gpu<Float,2>& col_dry, const Array_gpu<Float,2>& vmr_h2o,
                                        const Array_gpu<Float,2>& plev)
{
    hip_check(K<Float>::get_col_dry(col_dry.dim(1), col_dry.dim(2), vmr_h2o.ptr(), vmr_h2o.dim(1),
                                    vmr_h2o.dim(2), plev.ptr(), col_dry.ptr(), stream()));
}


void Gas_optics_rrtmgp_gpu::make_vmr_desc(const Gas_concs_gpu& gas_desc, int ncol, int nlay,
                                          rrtmgp_gas_vmr& out) const
{
    out = rrtmgp_gas_vmr{};
    out.ngas = get_ngas();
    if (out.ngas > RRTMGP_MAX_GAS) throw std::runtime_error("too many gases");
    for (int i = 1; i <= get_ngas(); ++i)
    {
        const std::string& name = gas_names({i});
        if (!gas_desc.exists(name)) throw std::runtime_error("gas " + name + " missing from gas_desc");
        const Array_gpu<Float,2>& v = gas_desc.get_vmr(name);
        out.vmr[i - 1] = v.ptr();
        out.dim1[i - 1] = v.dim(1);
        out.dim2[i - 1] = v.dim(2);
        if ((v.dim(1) != 1 && v.dim(1) != ncol) || (v.dim(2) != 1 && v.dim(2) != nlay))
            throw std::runtime_error("gas " + name + " has dims incompatible with (ncol, nlay)");
    }
}


int Gas_optics_rrtmgp_gpu::check_inputs(const Array_gpu<Float,2>& play, const Array_gpu<Float,2>& plev,
                                        const Array_gpu<Float,2>& tlay, const Array_gpu<Float,2>& col_dry,
                                        const Array_gpu<Float,2>* tlev, const Array_gpu<Float,1>* tsfc) const
{
    const Float nan = std::nan("");
    // In the reference's order of the checks (and of its messages).
    const char* what[6] = {"play", "plev", "tlay", "tlev", "tsfc", "col_dry"};
    const Float* arrays[6] = {play.ptr(), plev.ptr(), tlay.ptr(), tlev ? tlev->ptr() : nullptr,
                              tsfc ? tsfc->ptr() : nullptr, col_dry.ptr()};
    int64_t sizes[6] = {play.size(), plev.size(), tlay.size(), tlev ? tlev->size() : 0,
                        tsfc ? tsfc->size() : 0, col_dry.size()};
    const Float lo[6] = {press_ref_min, press_ref_min, temp_ref_min, temp_ref_min, temp_ref_min, Float(0.)};
    const Float hi[6] = {press_ref_max, press_ref_max, temp_ref_max, temp_ref_max, temp_ref_max, nan};
    if (!check_ranges)
        for (auto& n : sizes) n = 0;
    int64_t counts[6] = {0, 0, 0, 0, 0, 0};
    Float ends[2] = {Float(0.), Float(0.)};
    const int ncol = play.dim(1), nlay = play.dim(2);
    hip_check(K<Float>::check_ranges(6, arrays, sizes, lo, hi, counts, play.ptr(), ncol, nlay, ends, stream()));
    for (int k = 0; k < 6; ++k)
        if (counts[k] > 0) throw std::range_error(std::string(what[k]) + " is out of range");
    // Surface layer from the first column (src/Gas_optics_rrtmgp.cpp:1333).
    return ends[0] > ends[1] ? 1 : nlay;
}


void Gas_optics_rrtmgp_gpu::gas_optics(
        const Array_gpu<Float,2>& play, const Array_gpu<Float,2>& plev, const Array_gpu<Float,2>& tlay,
        const Array_gpu<Float,1>& tsfc, const Gas_concs_gpu& gas_desc,
        std::unique_ptr<Optical_props_arry_gpu>& optical_props, Source_func_lw_gpu& sources,
        const Array_gpu<Float,2>& col_dry, const Array_gpu<Float,2>& tlev)
{
    gas_optics_with_clouds(play, plev, tlay, tsfc, gas_desc, *optical_props, sources, col_dry, tlev, nullptr);
}


void Gas_optics_rrtmgp_gpu::gas_optics(
        const Array_gpu<Float,2>& play, const Array_gpu<Float,2>& plev, const Array_gpu<Float,2>& tlay,
        const Gas_concs_gpu& gas_desc, std::unique_ptr<Optical_props_arry_gpu>& optical_props,
        Array_gpu<Float,2>& toa_src, const Array_gpu<Float,2>& col_dry)
{
    gas_optics_with_clouds(play, plev, tlay, gas_desc, *optical_props, toa_src, col_dry, nullptr);
}


void Gas_optics_rrtmgp_gpu::gas_optics_with_clouds(
        const Array_gpu<Float,2>& play, const Array_gpu<Float,2>& plev, const Array_gpu<Float,2>& tlay,
        const Array_gpu<Float,1>& tsfc, const Gas_concs_gpu& gas_desc, Optical_props_arry_gpu& props,
        Source_func_lw_gpu& sources, const Array_gpu<Float,2>& col_dry, const Array_gpu<Float,2>& tlev,
        const Optical_props_arry_gpu* cloud)
{
    if (!source_is_internal()) throw std::runtime_error("longwave gas_optics called on a shortwave k-distribution");
    const int ncol = play.dim(1), nlay = play.dim(2);
    const int sfc_lay = check_inputs(play, plev, tlay, col_dry, &tlev, &tsfc);
    rrtmgp_gas_vmr v;
    make_vmr_desc(gas_desc, ncol, nlay, v);
    if (cloud && cloud->get_ngpt() != get_nband())
        throw std::runtime_error("Cannot add optical properties with incompatible band - gpoint combination");
    if (sources.is_factored())
    {
        const Float* ctau = cloud ? cloud->get_tau().ptr() : nullptr;
        const int* tiles = nullptr;
        int64_t ntiles = 0;
        if (use_mlp)
        {
            const int need = 1 + (ncol + 63) / 64 * nlay;
            if (fallback_gpu.size() < need) fallback_gpu = Array_gpu<int,1>({need});
            const rrtmgp_gas_vmr mv = mlp_vmr_desc(gas_desc, ncol, nlay);
            hip_check(Mlp<Float>::lw_factored(
                    &mlp, ncol, nlay, get_ngpt(), get_nband(), gpt2band_gpu.ptr(), play.ptr(), tlay.ptr(),
                    tlev.ptr(), tsfc.ptr(), col_dry.ptr(), &mv, static_cast<const Float*>(dev.totplnk), dev.nplanck,
                    dev.temp_ref_min, dev.totplnk_delta, sfc_lay, ctau, props.get_tau().ptr(),
                    sources.get_planck_frac().ptr(), sources.get_band_lay_source().ptr(),
                    sources.get_band_lev_source().ptr(), sources.get_sfc_source().ptr(),
                    sources.get_sfc_source_jac().ptr(), fallback_gpu.ptr(), stream()));
            if (run_fallback_list(ncol, nlay) == 0) return;
            tiles = fallback_gpu.ptr() + 1;
            ntiles = n_fallback;
        }
        hip_check(K<Float>::gas_optics_lw_factored(
                &dev, ncol, nlay, tiles, ntiles, play.ptr(), tlay.ptr(), tlev.ptr(), tsfc.ptr(), col_dry.ptr(), &v,
                sfc_lay, ctau, props.get_tau().ptr(), sources.get_planck_frac().ptr(),
                sources.get_band_lay_source().ptr(), sources.get_band_lev_source().ptr(),
                sources.get_sfc_source().ptr(), sources.get_sfc_source_jac().ptr(), stream()));
        return;
    }
    if (use_mlp)
    {
        const int need = 1 + (ncol + 63) / 64 * nlay;
        if (fallback_gpu.size() < need) fallback_gpu = Array_gpu<int,1>({need});
        const rrtmgp_gas_vmr mv = mlp_vmr_desc(gas_desc, ncol, nlay);
        hip_check(Mlp<Float>::lw(
                &mlp, ncol, nlay, get_ngpt(), get_nband(), gpt2band_gpu.ptr(), play.ptr(), tlay.ptr(),
                tlev.ptr(), tsfc.ptr(), col_dry.ptr(), &mv, static_cast<const Float*>(dev.totplnk), dev.nplanck,
                dev.temp_ref_min, dev.totplnk_delta, sfc_lay, cloud ? cloud->get_tau().ptr() : nullptr,
                props.get_tau().ptr(), sources.get_lay_source().ptr(), sources.get_lev_source_inc().ptr(),
                sources.get_lev_source_dec().ptr(), sources.get_sfc_source().ptr(),
                sources.get_sfc_source_jac().ptr(), fallback_gpu.ptr(), stream()));
        if (run_fallback_list(ncol, nlay) > 0)
            hip_check(K<Float>::gas_optics_lw_tiles(
                    &dev, ncol, nlay, fallback_gpu.ptr() + 1, n_fallback, play.ptr(), tlay.ptr(), tlev.ptr(),
                    tsfc.ptr(), col_dry.ptr(), &v, sfc_lay, cloud ? cloud->get_tau().ptr() : nullptr,
                    props.get_tau().ptr(), sources.get_lay_source().ptr(), sources.get_lev_source_inc().ptr(),
                    sources.get_lev_source_dec().ptr(), sources.get_sfc_source().ptr(),
                    sources.get_sfc_source_jac().ptr(), stream()));
        return;
    }
    hip_check(K<Float>::gas_optics_lw(&dev, ncol, nlay, play.ptr(), tlay.ptr(), tlev.ptr(), tsfc.ptr(),
                                      col_dry.ptr(), &v, sfc_lay, cloud ? cloud->get_tau().ptr() : nullptr,
                                      props.get_tau().ptr(), sources.get_lay_source().ptr(),
                                      sources.get_lev_source_inc().ptr(), sources.get_lev_source_dec().ptr(),
                                      sources.get_sfc_source().ptr(), sources.get_sfc_source_jac().ptr(),
                                      stream()));
}


void Gas_optics_rrtmgp_gpu::gas_optics_with_clouds(
        const Array_gpu<Float,2>& play, const Array_gpu<Float,2>& plev, const Array_gpu<Float,2>& tlay,
        const Gas_concs_gpu& gas_desc, Optical_props_arry_gpu& props, Array_gpu<Float,2>& toa_src,
        const Array_gpu<Float,2>& col_dry, const Optical_props_arry_gpu* cloud, bool coalbedo)
{
    if (krayl.size() == 0) throw std::runtime_error("shortwave gas_optics called on a longwave k-distribution");
    if (coalbedo && use_mlp) throw std::invalid_argument("co-albedo gas optics with a surrogate set");
    const int ncol = play.dim(1), nlay = play.dim(2);
    (void)check_inputs(play, plev, tlay, col_dry, nullptr, nullptr);
    rrtmgp_gas_vmr v;
    make_vmr_desc(gas_desc, ncol, nlay, v);
    if (cloud && cloud->get_ngpt() != get_nband())
        throw std::runtime_error("Cannot add optical properties with incompatible band - gpoint combination");
    const Float* ct = cloud ? cloud->get_tau().ptr() : nullptr;
    const Float* cs = cloud ? cloud->get_ssa().ptr() : nullptr;
    const Float* cg = cloud ? cloud->get_g().ptr() : nullptr;
    if (use_mlp)
    {
        const int need = 1 + (ncol + 63) / 64 * nlay;
        if (fallback_gpu.size() < need) fallback_gpu = Array_gpu<int,1>({need});
        const rrtmgp_gas_vmr mv = mlp_vmr_desc(gas_desc, ncol, nlay);
        hip_check(Mlp<Float>::sw(&mlp, ncol, nlay, get_ngpt(), get_nband(), gpt2band_gpu.ptr(), play.ptr(),
                                 tlay.ptr(), col_dry.ptr(), &mv, ct, cs, cg, props.get_tau().ptr(),
                                 props.get_ssa().ptr(), props.get_g().ptr(), fallback_gpu.ptr(), stream()));
        if (run_fallback_list(ncol, nlay) > 0)
            hip_check(K<Float>::gas_optics_sw_tiles(&dev, ncol, nlay, fallback_gpu.ptr() + 1, n_fallback,
                                                    play.ptr(), tlay.ptr(), col_dry.ptr(), &v, ct, cs, cg,
                                                    props.get_tau().ptr(), props.get_ssa().ptr(),
                                                    props.get_g().ptr(), stream()));
    }
    else if (coalbedo)
        hip_check(K<Float>::gas_optics_sw_coalbedo(&dev, ncol, nlay, play.ptr(), tlay.ptr(), col_dry.ptr(), &v, ct,
                                                   cs, cg, props.get_tau().ptr(), props.get_ssa().ptr(),
                                                   props.get_g().ptr(), stream()));
    else
        hip_check(K<Float>::gas_optics_sw(&dev, ncol, nlay, play.ptr(), tlay.ptr(), col_dry.ptr(), &v, ct, cs, cg,
                                          props.get_tau().ptr(), props.get_ssa().ptr(), props.get_g().ptr(),
                                          stream()));
    hip_check(K<Float>::spread_col(ncol, get_ngpt(), solar_source_gpu.ptr(), toa_src.ptr(), stream()));
}


void Gas_optics_rrtmgp_gpu::set_surrogate(const rrtmgp_mlp_net* net)
{
    if (!net)
    {
        use_mlp = false;
        return;
    }
    if (!std::is_same<Float, float>::value)
        throw std::invalid_argument("the MLP gas-optics surrogate runs in FP32 builds only");
    if (net->n_out != 2 * int64_t(get_ngpt()))
        throw std::invalid_argument("MLP surrogate: the output count must be 2 ngpt");
    if (source_is_internal() && dev.nbnd > 16)
        throw std::invalid_argument("MLP surrogate: at most 16 longwave bands");
    // 0-based band of each g-point, from band_lims_gpt (1-based, inclusive).
    const Array<int,2>& lims = get_band_lims_gpoint();
    Array<int,1> g2b({get_ngpt()});
    for (int b = 1; b <= get_nband(); ++b)
        for (int g = lims({1, b}); g <= lims({2, b}); ++g) g2b({g}) = b - 1;
    gpt2band_gpu = Array_gpu<int,1>(g2b);
    mlp = *net;
    use_mlp = true;
}


// Reads the number of (tile, layer) units the surrogate listed for the plain kernels (a 4-byte
// device-to-host copy on the library stream) and sizes the list for the next call.
int64_t Gas_optics_rrtmgp_gpu::run_fallback_list(const int ncol, const int nlay)
{
    int count = 0;
    hip_check(rrtmgp_hip_memcpy_d2h(&count, fallback_gpu.ptr(), sizeof(int), stream()));
    hip_check(rrtmgp_hip_stream_synchronize(stream()));
    const int64_t units = int64_t((ncol + 63) / 64) * nlay;
    if (count < 0 || count > units) throw std::runtime_error("MLP surrogate: corrupt fallback list");
    n_fallback = count;
    return n_fallback;
}
