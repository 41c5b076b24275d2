// Bundle reader and coefficient loaders (include/Coefficient_bundle.h).
#include "Coefficient_bundle.h"

#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "Cloud_optics.h"
#include "Gas_concs.h"
#include "Gas_optics_rrtmgp.h"

namespace
{
size_t type_size(const std::string& t)
{
    if (t == "f8") return 8;
    if (t == "f4" || t == "i4") return 4;
    if (t == "i1" || t == "S1") return 1;
    throw std::runtime_error("bundle: unknown type " + t);
}

template<typename T, typename S>
void convert(const std::vector<char>& b, std::vector<T>& out)
{
    const size_t n = b.size() / sizeof(S);
    out.resize(n);
    for (size_t i = 0; i < n; ++i)
    {
        S v;
        std::memcpy(&v, b.data() + i * sizeof(S), sizeof(S));
        out[i] = static_cast<T>(v);
    }
}
}  // namespace

Coefficient_bundle::Coefficient_bundle(const std::string& dir_in) : dir(dir_in)
{
    std::ifstream f(dir + "/manifest.txt");
    if (!f) throw std::runtime_error("bundle: cannot open " + dir + "/manifest.txt");
    std::string line;
    while (std::getline(f, line))
    {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is(line);
        std::string kind, name;
        is >> kind >> name;
        if (kind == "dim")
        {
            int n = 0;
            is >> n;
            dims[name] = n;
        }
        else if (kind == "var")
        {
            Var v;
            int nd = 0;
            is >> v.type >> nd;
            v.dims.resize(size_t(nd));
            for (auto& d : v.dims) is >> d;
            if (!is) throw std::runtime_error("bundle: bad manifest line: " + line);
            type_size(v.type);
            vars[name] = v;
        }
        else
            throw std::runtime_error("bundle: bad manifest line: " + line);
    }
}

int Coefficient_bundle::get_dimension_size(const std::string& name) const
{
    auto it = dims.find(name);
    if (it == dims.end()) throw std::runtime_error("bundle " + dir + ": no dimension " + name);
    return it->second;
}

bool Coefficient_bundle::variable_exists(const std::string& name) const { return vars.count(name) > 0; }

std::vector<std::string> Coefficient_bundle::get_variable_dimensions(const std::string& name) const
{
    return var(name).dims;
}

const Coefficient_bundle::Var& Coefficient_bundle::var(const std::string& name) const
{
    auto it = vars.find(name);
    if (it == vars.end()) throw std::runtime_error("bundle " + dir + ": no variable " + name);
    return it->second;
}

std::vector<char> Coefficient_bundle::raw(const std::string& name, size_t elem_size, int64_t count) const
{
    std::ifstream f(dir + "/" + name + ".bin", std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("bundle " + dir + ": cannot open " + name + ".bin");
    const int64_t bytes = int64_t(f.tellg());
    if (bytes != count * int64_t(elem_size))
        throw std::runtime_error("bundle " + dir + ": " + name + ".bin has the wrong size");
    std::vector<char> b(static_cast<size_t>(bytes));
    f.seekg(0);
    f.read(b.data(), bytes);
    return b;
}

template<typename T>
std::vector<T> Coefficient_bundle::get_variable(const std::string& name, const std::vector<int>& want) const
{
    const Var& v = var(name);
    // A scalar request reads the first value of a variable of any shape of one element.
    int64_t count = 1;
    for (const auto& d : v.dims) count *= get_dimension_size(d);
    if (!want.empty())
    {
        if (want.size() != v.dims.size())
            throw std::runtime_error("bundle: " + name + " has " + std::to_string(v.dims.size()) + " dimensions");
        for (size_t k = 0; k < want.size(); ++k)
            if (want[k] != get_dimension_size(v.dims[k]))
                throw std::runtime_error("bundle: " + name + " dimension " + v.dims[k] + " has another size");
    }
    else if (count != 1)
        throw std::runtime_error("bundle: " + name + " is not a scalar");
    const std::vector<char> b = raw(name, type_size(v.type), count);
    std::vector<T> out;
    if (v.type == "f8") convert<T, double>(b, out);
    else if (v.type == "f4") convert<T, float>(b, out);
    else if (v.type == "i4") convert<T, int32_t>(b, out);
    else if (v.type == "i1") convert<T, int8_t>(b, out);
    else throw std::runtime_error("bundle: " + name + " is a character variable");
    return out;
}

template std::vector<double> Coefficient_bundle::get_variable<double>(const std::string&, const std::vector<int>&) const;
template std::vector<float> Coefficient_bundle::get_variable<float>(const std::string&, const std::vector<int>&) const;
template std::vector<int> Coefficient_bundle::get_variable<int>(const std::string&, const std::vector<int>&) const;
template std::vector<signed char> Coefficient_bundle::get_variable<signed char>(const std::string&, const std::vector<int>&) const;

std::vector<std::string> Coefficient_bundle::get_strings(const std::string& name) const
{
    const Var& v = var(name);
    if (v.type != "S1" || v.dims.size() != 2) throw std::runtime_error("bundle: " + name + " is not a string list");
    const int n = get_dimension_size(v.dims[0]), len = get_dimension_size(v.dims[1]);
    const std::vector<char> b = raw(name, 1, int64_t(n) * len);
    std::vector<std::string> out;
    for (int i = 0; i < n; ++i)
    {
        std::string s(b.begin() + int64_t(i) * len, b.begin() + int64_t(i + 1) * len);
        const size_t e = s.find_last_not_of(std::string(" \0", 2));
        s = (e == std::string::npos) ? std::string() : s.substr(0, e + 1);
        const size_t a = s.find_first_not_of(' ');
        out.push_back(a == std::string::npos ? std::string() : s.substr(a));
    }
    return out;
}

namespace
{
// A variable of NetCDF dims nc_dims (slowest first) as the reference's Array, dims reversed.
template<typename T, int N>
Array<T,N> arr(const Coefficient_bundle& c, const std::string& name, const std::array<int,N>& nc_dims)
{
    std::array<int,N> fd;
    for (int k = 0; k < N; ++k) fd[k] = nc_dims[N - 1 - k];
    return Array<T,N>(c.get_variable<T>(name, std::vector<int>(nc_dims.begin(), nc_dims.end())), fd);
}

Array<std::string,1> strs(const Coefficient_bundle& c, const std::string& name)
{
    std::vector<std::string> v = c.get_strings(name);
    const int n = int(v.size());
    return Array<std::string,1>(std::move(v), {n});
}
}  // namespace

std::unique_ptr<Gas_optics_rrtmgp_gpu> load_and_init_gas_optics(const Gas_concs_gpu& gas_concs,
                                                                const std::string& bundle_dir)
{
    const Coefficient_bundle c(bundle_dir);
    const int n_temps = c.get_dimension_size("temperature");
    const int n_press = c.get_dimension_size("pressure");
    const int n_absorbers = c.get_dimension_size("absorber");
    const int n_minorabsorbers = c.get_dimension_size("minor_absorber");
    const int n_extabsorbers = c.get_dimension_size("absorber_ext");
    const int n_mixingfracs = c.get_dimension_size("mixing_fraction");
    const int n_layers = c.get_dimension_size("atmos_layer");
    const int n_bnds = c.get_dimension_size("bnd");
    const int n_gpts = c.get_dimension_size("gpt");
    const int n_pairs = c.get_dimension_size("pair");
    const int n_lo = c.get_dimension_size("minor_absorber_intervals_lower");
    const int n_up = c.get_dimension_size("minor_absorber_intervals_upper");
    const int n_con_lo = c.get_dimension_size("contributors_lower");
    const int n_con_up = c.get_dimension_size("contributors_upper");

    const Array<std::string,1> gas_names = strs(c, "gas_names");
    if (gas_names.dim(1) != n_absorbers) throw std::runtime_error("bundle: gas_names length");
    const auto key_species = arr<int,3>(c, "key_species", {n_bnds, n_layers, 2});
    const auto band_lims = arr<Float,2>(c, "bnd_limits_wavenumber", {n_bnds, 2});
    const auto band2gpt = arr<int,2>(c, "bnd_limits_gpt", {n_bnds, 2});
    const auto press_ref = arr<Float,1>(c, "press_ref", {n_press});
    const auto temp_ref = arr<Float,1>(c, "temp_ref", {n_temps});
    const Float temp_ref_p = c.get_variable<Float>("absorption_coefficient_ref_P");
    const Float temp_ref_t = c.get_variable<Float>("absorption_coefficient_ref_T");
    const Float press_ref_trop = c.get_variable<Float>("press_ref_trop");
    const auto kminor_lower = arr<Float,3>(c, "kminor_lower", {n_temps, n_mixingfracs, n_con_lo});
    const auto kminor_upper = arr<Float,3>(c, "kminor_upper", {n_temps, n_mixingfracs, n_con_up});
    const auto gas_minor = strs(c, "gas_minor");
    const auto identifier_minor = strs(c, "identifier_minor");
    if (gas_minor.dim(1) != n_minorabsorbers) throw std::runtime_error("bundle: gas_minor length");
    const auto minor_gases_lower = strs(c, "minor_gases_lower");
    const auto minor_gases_upper = strs(c, "minor_gases_upper");
    const auto minor_limits_gpt_lower = arr<int,2>(c, "minor_limits_gpt_lower", {n_lo, n_pairs});
    const auto minor_limits_gpt_upper = arr<int,2>(c, "minor_limits_gpt_upper", {n_up, n_pairs});
    const auto minor_scales_with_density_lower = arr<Bool,1>(c, "minor_scales_with_density_lower", {n_lo});
    const auto minor_scales_with_density_upper = arr<Bool,1>(c, "minor_scales_with_density_upper", {n_up});
    const auto scale_by_complement_lower = arr<Bool,1>(c, "scale_by_complement_lower", {n_lo});
    const auto scale_by_complement_upper = arr<Bool,1>(c, "scale_by_complement_upper", {n_up});
    const auto scaling_gas_lower = strs(c, "scaling_gas_lower");
    const auto scaling_gas_upper = strs(c, "scaling_gas_upper");
    const auto kminor_start_lower = arr<int,1>(c, "kminor_start_lower", {n_lo});
    const auto kminor_start_upper = arr<int,1>(c, "kminor_start_upper", {n_up});
    const auto vmr_ref = arr<Float,3>(c, "vmr_ref", {n_temps, n_extabsorbers, n_layers});
    const auto kmajor = arr<Float,4>(c, "kmajor", {n_temps, n_press + 1, n_mixingfracs, n_gpts});

    // Rayleigh tables stay empty when absent.
    Array<Float,3> rayl_lower, rayl_upper;
    if (c.variable_exists("rayl_lower"))
    {
        rayl_lower = arr<Float,3>(c, "rayl_lower", {n_temps, n_mixingfracs, n_gpts});
        rayl_upper = arr<Float,3>(c, "rayl_upper", {n_temps, n_mixingfracs, n_gpts});
    }

    if (c.variable_exists("totplnk"))
    {
        const int n_planck = c.get_dimension_size("temperature_Planck");
        const auto totplnk = arr<Float,2>(c, "totplnk", {n_bnds, n_planck});
        const auto planck_frac = arr<Float,4>(c, "plank_fraction", {n_temps, n_press + 1, n_mixingfracs, n_gpts});
        return std::make_unique<Gas_optics_rrtmgp_gpu>(
                gas_concs, gas_names, key_species, band2gpt, band_lims, press_ref, press_ref_trop, temp_ref,
                temp_ref_p, temp_ref_t, vmr_ref, kmajor, kminor_lower, kminor_upper, gas_minor, identifier_minor,
                minor_gases_lower, minor_gases_upper, minor_limits_gpt_lower, minor_limits_gpt_upper,
                minor_scales_with_density_lower, minor_scales_with_density_upper, scaling_gas_lower,
                scaling_gas_upper, scale_by_complement_lower, scale_by_complement_upper, kminor_start_lower,
                kminor_start_upper, totplnk, planck_frac, rayl_lower, rayl_upper);
    }
    const auto solar_src_quiet = arr<Float,1>(c, "solar_source_quiet", {n_gpts});
    const auto solar_src_facular = arr<Float,1>(c, "solar_source_facular", {n_gpts});
    const auto solar_src_sunspot = arr<Float,1>(c, "solar_source_sunspot", {n_gpts});
    const Float tsi = c.get_variable<Float>("tsi_default");
    const Float mg_index = c.get_variable<Float>("mg_default");
    const Float sb_index = c.get_variable<Float>("sb_default");
    return std::make_unique<Gas_optics_rrtmgp_gpu>(
            gas_concs, gas_names, key_species, band2gpt, band_lims, press_ref, press_ref_trop, temp_ref,
            temp_ref_p, temp_ref_t, vmr_ref, kmajor, kminor_lower, kminor_upper, gas_minor, identifier_minor,
            minor_gases_lower, minor_gases_upper, minor_limits_gpt_lower, minor_limits_gpt_upper,
            minor_scales_with_density_lower, minor_scales_with_density_upper, scaling_gas_lower,
            scaling_gas_upper, scale_by_complement_lower, scale_by_complement_upper, kminor_start_lower,
            kminor_start_upper, solar_src_quiet, solar_src_facular, solar_src_sunspot, tsi, mg_index, sb_index,
            rayl_lower, rayl_upper);
}

std::unique_ptr<Cloud_optics_gpu> load_and_init_cloud_optics(const std::string& bundle_dir)
{
    const Coefficient_bundle c(bundle_dir);
    const int n_band = c.get_dimension_size("nband");
    const int n_rghice = c.get_dimension_size("nrghice");
    const int n_size_liq = c.get_dimension_size("nsize_liq");
    const int n_size_ice = c.get_dimension_size("nsize_ice");
    const auto band_lims_wvn = arr<Float,2>(c, "bnd_limits_wavenumber", {n_band, 2});
    const Float radliq_lwr = c.get_variable<Float>("radliq_lwr");
    const Float radliq_upr = c.get_variable<Float>("radliq_upr");
    const Float radliq_fac = c.get_variable<Float>("radliq_fac");
    const Float radice_lwr = c.get_variable<Float>("radice_lwr");
    const Float radice_upr = c.get_variable<Float>("radice_upr");
    const Float radice_fac = c.get_variable<Float>("radice_fac");
    const auto lut_extliq = arr<Float,2>(c, "lut_extliq", {n_band, n_size_liq});
    const auto lut_ssaliq = arr<Float,2>(c, "lut_ssaliq", {n_band, n_size_liq});
    const auto lut_asyliq = arr<Float,2>(c, "lut_asyliq", {n_band, n_size_liq});
    const auto lut_extice = arr<Float,3>(c, "lut_extice", {n_rghice, n_band, n_size_ice});
    const auto lut_ssaice = arr<Float,3>(c, "lut_ssaice", {n_rghice, n_band, n_size_ice});
    const auto lut_asyice = arr<Float,3>(c, "lut_asyice", {n_rghice, n_band, n_size_ice});
    return std::make_unique<Cloud_optics_gpu>(band_lims_wvn, radliq_lwr, radliq_upr, radliq_fac, radice_lwr,
                                              radice_upr, radice_fac, lut_extliq, lut_ssaliq, lut_asyliq,
                                              lut_extice, lut_ssaice, lut_asyice);
}
