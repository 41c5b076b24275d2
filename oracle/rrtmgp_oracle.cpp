// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement of the RTE+RRTMGP column path of rte-rrtmgp-cpp, used as the checker for
// the HIP implementation and as the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and
// the cpu_baseline leg of bench.py may load this library; the product never does.
//
// The reference's CPU path cannot be compiled here: its hot arithmetic is Fortran from the
// un-vendored rte-rrtmgp submodule (src_kernels/CMakeLists.txt:4-15) and gfortran, boost and
// netCDF are absent.  This file therefore restates
//   * host logic of src/ (Gas_optics_rrtmgp.cpp, Cloud_optics.cpp, Optical_props.cpp,
//     Rte_lw.cpp, Rte_sw.cpp, Fluxes.cpp, src_test/Radiation_solver.cpp),
//   * gas-optics arithmetic of src_kernels_cuda/gas_optics_rrtmgp_kernels.cu (the in-tree
//     restatement of rrtmgp_interpolation / compute_tau_absorption / rayleigh / Planck),
//   * solver arithmetic of the vendored src_kernels/mo_rte_solver_kernels.F90, and
//   * by-band sums of src_kernels/mo_fluxes_byband_kernels.F90.
// Every function cites the file:line it follows.  Where the reference's CPU and GPU paths
// differ the CPU semantics are kept; the decisions are listed at each site.
//
// How each part is pinned (tests/test_oracle_*.py): closed forms and invariants for every
// function -- sum rules of the interpolation weights, node reproduction of the tables, linearity,
// the Planck sum rule, transparent / non-scattering / conservative limits of the solvers, a dense
// linear solve of the two-stream equations for the adding method, by-hand sums for the
// reductions -- and, for the Planck-fraction, k-minor and Rayleigh gathers, tables linear in the
// temperature, eta and pressure nodes with per-g-point (per-contributor) coefficients and
// distinct lower/upper slopes, whose interpolation has a closed form that a swapped corner, a
// j0/j1 mix-up, a lost jpress+itropo shift or the wrong flavor or contributor breaks.  The
// oracle's outputs on two 8-column cases are kept as fixtures (tests/golden/) so that a later
// change shows.  The compiled Fortran of the submodule is the one thing nothing here can pin
// (parity against it is unpinned).
//
// Build: g++ -O3 -march=x86-64-v3 -ffp-contract=off -fPIC -shared (oracle/build_oracle.py).

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
using Bool = signed char;   // RTE_USE_CBOOL semantics (include/types.h:13-17)

thread_local std::string last_error;

std::vector<std::string> split_names(const char* s, int n)
{
    // Names are '\n'-separated; empty entries are kept (empty scaling-gas names).
    std::vector<std::string> out;
    if (n == 0) return out;
    std::string cur;
    for (const char* p = s; ; ++p)
    {
        if (*p == '\n' || *p == '\0')
        {
            out.push_back(cur);
            cur.clear();
            if (*p == '\0') break;
        }
        else
            cur.push_back(*p);
    }
    if ((int)out.size() != n)
        throw std::runtime_error("name list has " + std::to_string(out.size()) +
                                 " entries, expected " + std::to_string(n));
    return out;
}

std::string trim(const std::string& s)
{
    const size_t b = s.find_first_not_of(" \t");
    if (b == std::string::npos) return "";
    const size_t e = s.find_last_not_of(" \t");
    return s.substr(b, e - b + 1);
}

// find_index: src/Gas_optics_rrtmgp.cpp:42-50 (1-based, -1 if absent).
int find_index(const std::vector<std::string>& v, const std::string& x)
{
    auto it = std::find(v.begin(), v.end(), x);
    return it == v.end() ? -1 : int(it - v.begin()) + 1;
}
}  // namespace


extern "C" {

// Coefficient-file view, Fortran layouts as passed to the Gas_optics_rrtmgp constructors
// (include/Gas_optics_rrtmgp.h:50-121).  String lists are '\n'-joined.
struct OracleKdistFile
{
    int ngas; const char* gas_names;
    int nbnd, ngpt, ntemp, npres, neta, nplanck;   // nplanck = 0 for SW
    const int* key_species;          // (2,2,nbnd)
    const int* band2gpt;             // (2,nbnd)
    const double* band_lims_wavenum; // (2,nbnd)
    const double* press_ref;         // (npres)
    double press_ref_trop;
    const double* temp_ref;          // (ntemp)
    double temp_ref_p, temp_ref_t;
    const double* vmr_ref;           // (2, ngas+1, ntemp)
    const double* kmajor;            // (ngpt, neta, npres+1, ntemp)
    int nminor_ident; const char* gas_minor; const char* identifier_minor;
    int nminor_lower, ncontrib_lower;
    const double* kminor_lower;      // (ncontrib, neta, ntemp)
    const char* minor_gases_lower;
    const int* minor_limits_gpt_lower;             // (2, nminor)
    const Bool* minor_scales_with_density_lower;
    const char* scaling_gas_lower;
    const Bool* scale_by_complement_lower;
    const int* kminor_start_lower;
    int nminor_upper, ncontrib_upper;
    const double* kminor_upper;
    const char* minor_gases_upper;
    const int* minor_limits_gpt_upper;
    const Bool* minor_scales_with_density_upper;
    const char* scaling_gas_upper;
    const Bool* scale_by_complement_upper;
    const int* kminor_start_upper;
    const double* totplnk;           // (nplanck, nbnd)      LW
    const double* planck_frac;       // (ngpt, neta, npres+1, ntemp)  LW
    const double* solar_src_quiet;   // (ngpt)  SW
    const double* solar_src_facular;
    const double* solar_src_sunspot;
    double tsi_default, mg_default, sb_default;
    const double* rayl_lower;        // (ngpt, neta, ntemp)  SW
    const double* rayl_upper;
    int n_available; const char* available_gases;
};

// Gas concentrations (the Gas_concs name -> (ncol|1, nlay|1) map, include/Gas_concs.h:40-91).
struct OracleGasConcs
{
    int n; const char* names; const double* const* data; const int* dims;   // dims: 2*n
};

}  // extern "C"


namespace
{
struct Minor
{
    int n = 0, ncontrib = 0;
    std::vector<double> kminor;          // (ntemp, neta, ncontrib)
    std::vector<int> limits;             // (2, n)
    std::vector<Bool> scales_with_density, scale_by_complement;
    std::vector<int> kminor_start, idx_minor, idx_minor_scaling;
};

struct Kdist
{
    bool is_lw;
    int ngas, nflav, neta, npres, ntemp, ngpt, nbnd, nplanck;
    std::vector<std::string> gas_names;
    std::vector<int> flavor, gpoint_flavor, band2gpt, gpt2band, is_key;
    std::vector<double> press_ref, press_ref_log, temp_ref, vmr_ref, kmajor;
    double press_ref_min, press_ref_max, temp_ref_min, temp_ref_max;
    double press_ref_log_delta, temp_ref_delta, press_ref_trop_log, totplnk_delta = 0;
    Minor lower, upper;
    std::vector<double> totplnk, planck_frac, krayl;
    std::vector<double> solar_source_quiet, solar_source_facular, solar_source_sunspot, solar_source;
    int idx_h2o = -1;
};

// reduce_minor_arrays: src/Gas_optics_rrtmgp.cpp:53-162.  Keeps the minor contributors whose
// gas is available, compacts kminor and re-bases kminor_start, then reorders kminor from
// (ncontrib, neta, ntemp) to (ntemp, neta, ncontrib) ("new ordering in v1.5", :156-161).
void reduce_minor(const std::vector<std::string>& available,
                  const std::vector<std::string>& gas_minor,
                  const std::vector<std::string>& identifier_minor,
                  const double* kminor, int ncontrib, int neta, int ntemp,
                  const std::vector<std::string>& minor_gases, const int* limits,
                  const Bool* dens, const std::vector<std::string>& scaling_gas,
                  const Bool* compl_, const int* kstart, int nm,
                  Minor& out, std::vector<std::string>& minor_gases_red,
                  std::vector<std::string>& scaling_gas_red)
{
    std::vector<Bool> present(nm);
    int tot_g = 0;
    for (int i = 0; i < nm; ++i)
    {
        const int idx_mnr = find_index(identifier_minor, minor_gases[i]);
        if (idx_mnr < 1) throw std::runtime_error("unknown minor identifier " + minor_gases[i]);
        const std::string g = trim(gas_minor[idx_mnr - 1]);
        present[i] = std::find(available.begin(), available.end(), g) != available.end();
        if (present[i]) tot_g += limits[2*i+1] - limits[2*i] + 1;
    }
    int red_nm = 0;
    for (int i = 0; i < nm; ++i) red_nm += present[i];

    std::vector<double> kt;   // (ncontrib_red, neta, ntemp), file order
    int ncontrib_red;
    out.limits.clear(); out.scales_with_density.clear(); out.scale_by_complement.clear();
    out.kminor_start.clear(); minor_gases_red.clear(); scaling_gas_red.clear();
    if (red_nm == nm)
    {
        kt.assign(kminor, kminor + size_t(ncontrib) * neta * ntemp);
        ncontrib_red = ncontrib;
        for (int i = 0; i < nm; ++i)
        {
            out.limits.push_back(limits[2*i]); out.limits.push_back(limits[2*i+1]);
            out.scales_with_density.push_back(dens[i]); out.scale_by_complement.push_back(compl_[i]);
            out.kminor_start.push_back(kstart[i]);
            minor_gases_red.push_back(minor_gases[i]); scaling_gas_red.push_back(scaling_gas[i]);
        }
    }
    else
    {
        ncontrib_red = tot_g;
        kt.assign(size_t(tot_g) * neta * ntemp, 0.0);
        int n_elim = 0;
        for (int i = 0; i < nm; ++i)
        {
            const int ng = limits[2*i+1] - limits[2*i] + 1;
            if (present[i])
            {
                out.limits.push_back(limits[2*i]); out.limits.push_back(limits[2*i+1]);
                out.scales_with_density.push_back(dens[i]); out.scale_by_complement.push_back(compl_[i]);
                const int ks = kstart[i] - n_elim;
                out.kminor_start.push_back(ks);
                minor_gases_red.push_back(minor_gases[i]); scaling_gas_red.push_back(scaling_gas[i]);
                for (int j = 1; j <= ng; ++j)
                    for (int i2 = 0; i2 < neta; ++i2)
                        for (int i3 = 0; i3 < ntemp; ++i3)
                            kt[(ks+j-2) + size_t(ncontrib_red)*(i2 + size_t(neta)*i3)] =
                                kminor[(kstart[i]+j-2) + size_t(ncontrib)*(i2 + size_t(neta)*i3)];
            }
            else
                n_elim += ng;
        }
    }
    out.n = red_nm;
    out.ncontrib = ncontrib_red;
    out.kminor.assign(size_t(ntemp) * neta * ncontrib_red, 0.0);
    for (int c = 0; c < ncontrib_red; ++c)
        for (int e = 0; e < neta; ++e)
            for (int t = 0; t < ntemp; ++t)
                out.kminor[t + size_t(ntemp)*(e + size_t(neta)*c)] = kt[c + size_t(ncontrib_red)*(e + size_t(neta)*t)];
}

// create_flavor: src/Gas_optics_rrtmgp.cpp:247-313.
void create_flavor(const std::vector<int>& ks, int nbnd, std::vector<int>& flavor)
{
    std::vector<int> list;   // (2, 2*nbnd)
    for (int ib = 0; ib < nbnd; ++ib)
        for (int ia = 0; ia < 2; ++ia)
        {
            int a = ks[0 + 2*(ia + 2*ib)], b = ks[1 + 2*(ia + 2*ib)];
            if (a == 0 && b == 0) { a = 2; b = 2; }
            list.push_back(a); list.push_back(b);
        }
    flavor.clear();
    for (size_t i = 0; i < list.size()/2; ++i)
    {
        bool exists = false;
        for (size_t ii = 0; ii < i; ++ii)
            if (list[2*i] == list[2*ii] && list[2*i+1] == list[2*ii+1]) { exists = true; break; }
        if (!exists) { flavor.push_back(list[2*i]); flavor.push_back(list[2*i+1]); }
    }
}

// create_gpoint_flavor: src/Gas_optics_rrtmgp.cpp:331-363.
void create_gpoint_flavor(const std::vector<int>& ks, const std::vector<int>& gpt2band,
                          const std::vector<int>& flavor, std::vector<int>& gpoint_flavor)
{
    const int ngpt = gpt2band.size();
    const int nflav = flavor.size() / 2;
    gpoint_flavor.assign(2*ngpt, -1);
    for (int ig = 0; ig < ngpt; ++ig)
        for (int ia = 0; ia < 2; ++ia)
        {
            const int b = gpt2band[ig] - 1;
            int p1 = ks[0 + 2*(ia + 2*b)], p2 = ks[1 + 2*(ia + 2*b)];
            if (p1 == 0 && p2 == 0) { p1 = 2; p2 = 2; }
            for (int f = 0; f < nflav; ++f)
                if (flavor[2*f] == p1 && flavor[2*f+1] == p2) { gpoint_flavor[ia + 2*ig] = f + 1; break; }
        }
}

Kdist* make_kdist(const OracleKdistFile& f)
{
    std::unique_ptr<Kdist> kd(new Kdist);
    kd->is_lw = f.nplanck > 0;
    const auto file_gas = split_names(f.gas_names, f.ngas);
    const auto available = split_names(f.available_gases, f.n_available);
    kd->neta = f.neta; kd->npres = f.npres; kd->ntemp = f.ntemp; kd->ngpt = f.ngpt; kd->nbnd = f.nbnd;
    kd->nplanck = f.nplanck;

    // Optical_props ctor: band2gpt and gpt2band (src/Optical_props.cpp:31-47).
    kd->band2gpt.assign(f.band2gpt, f.band2gpt + 2*f.nbnd);
    kd->gpt2band.assign(f.ngpt, 0);
    for (int b = 0; b < f.nbnd; ++b)
        for (int g = f.band2gpt[2*b]; g <= f.band2gpt[2*b+1]; ++g)
            kd->gpt2band[g-1] = b + 1;

    // init_abs_coeffs: src/Gas_optics_rrtmgp.cpp:539-742.
    for (const auto& s : file_gas)
        if (std::find(available.begin(), available.end(), s) != available.end())
            kd->gas_names.push_back(s);
    const int ngas = kd->gas_names.size();
    kd->ngas = ngas;

    // vmr_ref reduced to (2, ngas+1, ntemp) with gas 0 = dry air (:575-595).
    kd->vmr_ref.assign(2 * (ngas+1) * f.ntemp, 0.0);
    auto vr_in = [&](int i1, int ig1, int it) { return f.vmr_ref[i1 + 2*(ig1 + (f.ngas+1)*it)]; };
    for (int i1 = 0; i1 < 2; ++i1)
        for (int it = 0; it < f.ntemp; ++it)
        {
            kd->vmr_ref[i1 + 2*(0 + (ngas+1)*it)] = vr_in(i1, 0, it);
            for (int i = 1; i <= ngas; ++i)
            {
                const int idx = find_index(file_gas, kd->gas_names[i-1]);
                kd->vmr_ref[i1 + 2*(i + (ngas+1)*it)] = vr_in(i1, idx, it);   // "idx+1" 1-based
            }
        }

    const auto gas_minor = split_names(f.gas_minor, f.nminor_ident);
    const auto ident = split_names(f.identifier_minor, f.nminor_ident);
    std::vector<std::string> mg_lo_red, sg_lo_red, mg_up_red, sg_up_red;
    reduce_minor(available, gas_minor, ident, f.kminor_lower, f.ncontrib_lower, f.neta, f.ntemp,
                 split_names(f.minor_gases_lower, f.nminor_lower), f.minor_limits_gpt_lower,
                 f.minor_scales_with_density_lower, split_names(f.scaling_gas_lower, f.nminor_lower),
                 f.scale_by_complement_lower, f.kminor_start_lower, f.nminor_lower,
                 kd->lower, mg_lo_red, sg_lo_red);
    reduce_minor(available, gas_minor, ident, f.kminor_upper, f.ncontrib_upper, f.neta, f.ntemp,
                 split_names(f.minor_gases_upper, f.nminor_upper), f.minor_limits_gpt_upper,
                 f.minor_scales_with_density_upper, split_names(f.scaling_gas_upper, f.nminor_upper),
                 f.scale_by_complement_upper, f.kminor_start_upper, f.nminor_upper,
                 kd->upper, mg_up_red, sg_up_red);

    kd->press_ref.assign(f.press_ref, f.press_ref + f.npres);
    kd->temp_ref.assign(f.temp_ref, f.temp_ref + f.ntemp);

    // kmajor (ngpt, neta, npres+1, ntemp) -> (ntemp, neta, npres+1, ngpt) (:658-663).
    const int np1 = f.npres + 1;
    kd->kmajor.assign(size_t(f.ntemp) * f.neta * np1 * f.ngpt, 0.0);
    for (int g = 0; g < f.ngpt; ++g)
        for (int p = 0; p < np1; ++p)
            for (int e = 0; e < f.neta; ++e)
                for (int t = 0; t < f.ntemp; ++t)
                    kd->kmajor[t + size_t(f.ntemp)*(e + size_t(f.neta)*(p + size_t(np1)*g))] =
                        f.kmajor[g + size_t(f.ngpt)*(e + size_t(f.neta)*(p + size_t(np1)*t))];

    // krayl (ntemp, neta, ngpt, 2) from rayl_lower/upper (ngpt, neta, ntemp) (:667-677).
    if (f.rayl_lower)
    {
        kd->krayl.assign(size_t(f.ntemp) * f.neta * f.ngpt * 2, 0.0);
        for (int g = 0; g < f.ngpt; ++g)
            for (int e = 0; e < f.neta; ++e)
                for (int t = 0; t < f.ntemp; ++t)
                {
                    const size_t o = t + size_t(f.ntemp)*(e + size_t(f.neta)*g);
                    const size_t i = g + size_t(f.ngpt)*(e + size_t(f.neta)*t);
                    kd->krayl[o] = f.rayl_lower[i];
                    kd->krayl[o + size_t(f.ntemp)*f.neta*f.ngpt] = f.rayl_upper[i];
                }
    }

    kd->press_ref_log.resize(f.npres);
    for (int i = 0; i < f.npres; ++i) kd->press_ref_log[i] = std::log(kd->press_ref[i]);
    kd->press_ref_trop_log = std::log(f.press_ref_trop);

    // create_idx_minor / create_idx_minor_scaling (:165-196).
    auto idx_minor = [&](const std::vector<std::string>& mg, std::vector<int>& out) {
        out.clear();
        for (const auto& m : mg)
            out.push_back(find_index(kd->gas_names, gas_minor[find_index(ident, m) - 1]));
    };
    idx_minor(mg_lo_red, kd->lower.idx_minor);
    idx_minor(mg_up_red, kd->upper.idx_minor);
    kd->lower.idx_minor_scaling.clear();
    for (const auto& s : sg_lo_red) kd->lower.idx_minor_scaling.push_back(find_index(kd->gas_names, s));
    kd->upper.idx_minor_scaling.clear();
    for (const auto& s : sg_up_red) kd->upper.idx_minor_scaling.push_back(find_index(kd->gas_names, s));

    // create_key_species_reduce + check (:198-245).
    std::vector<int> ks_red(4 * f.nbnd);
    std::vector<Bool> present_init(f.ngas, 1);
    for (int i = 0; i < 4 * f.nbnd; ++i)
    {
        const int ks = f.key_species[i];
        if (ks != 0)
        {
            const int ksr = find_index(kd->gas_names, file_gas[ks-1]);
            ks_red[i] = ksr;
            if (ksr == -1) present_init[ks-1] = 0;
        }
        else
            ks_red[i] = 0;
    }
    for (int i = 0; i < f.ngas; ++i)
        if (!present_init[i])
            throw std::runtime_error("Gas optics: required gas " + file_gas[i] + " is missing");

    create_flavor(ks_red, f.nbnd, kd->flavor);
    kd->nflav = kd->flavor.size() / 2;
    create_gpoint_flavor(ks_red, kd->gpt2band, kd->flavor, kd->gpoint_flavor);

    kd->temp_ref_min = kd->temp_ref.front();
    kd->temp_ref_max = kd->temp_ref.back();
    kd->press_ref_min = kd->press_ref.back();
    kd->press_ref_max = kd->press_ref.front();
    kd->press_ref_log_delta = (std::log(kd->press_ref_min) - std::log(kd->press_ref_max)) / (f.npres - 1);
    kd->temp_ref_delta = (kd->temp_ref_max - kd->temp_ref_min) / (f.ntemp - 1);

    kd->is_key.assign(ngas, 0);
    for (int j = 0; j < kd->nflav; ++j)
        for (int i = 0; i < 2; ++i)
            if (kd->flavor[i + 2*j] != 0) kd->is_key[kd->flavor[i + 2*j] - 1] = 1;

    if (kd->is_lw)
    {
        // LW ctor: planck_frac reorder (:428-433), totplnk_delta (:461).
        kd->totplnk.assign(f.totplnk, f.totplnk + size_t(f.nplanck) * f.nbnd);
        kd->planck_frac.assign(size_t(f.ntemp) * f.neta * np1 * f.ngpt, 0.0);
        for (int g = 0; g < f.ngpt; ++g)
            for (int p = 0; p < np1; ++p)
                for (int e = 0; e < f.neta; ++e)
                    for (int t = 0; t < f.ntemp; ++t)
                        kd->planck_frac[t + size_t(f.ntemp)*(e + size_t(f.neta)*(p + size_t(np1)*g))] =
                            f.planck_frac[g + size_t(f.ngpt)*(e + size_t(f.neta)*(p + size_t(np1)*t))];
        kd->totplnk_delta = (kd->temp_ref_max - kd->temp_ref_min) / (f.nplanck - 1);
    }
    else
    {
        // SW ctor + set_solar_variability (:524-535, :745-760).
        kd->solar_source_quiet.assign(f.solar_src_quiet, f.solar_src_quiet + f.ngpt);
        kd->solar_source_facular.assign(f.solar_src_facular, f.solar_src_facular + f.ngpt);
        kd->solar_source_sunspot.assign(f.solar_src_sunspot, f.solar_src_sunspot + f.ngpt);
        kd->solar_source.resize(f.ngpt);
        const double a_offset = 0.1495954, b_offset = 0.00066696;
        for (int g = 0; g < f.ngpt; ++g)
            kd->solar_source[g] = kd->solar_source_quiet[g]
                + (f.mg_default - a_offset) * kd->solar_source_facular[g]
                + (f.sb_default - b_offset) * kd->solar_source_sunspot[g];
    }
    kd->idx_h2o = find_index(kd->gas_names, "h2o");
    return kd.release();
}

// ---------------------------------------------------------------------------------------------
// Gas optics kernels.

struct Interp
{
    std::vector<int> jtemp, jpress, jeta;            // (ncol,nlay), (2,ncol,nlay,nflav)
    std::vector<Bool> tropo;
    std::vector<double> col_mix, fminor, fmajor;     // (2,..), (2,2,..), (2,2,2,..)
};

// interpolation_kernel: src_kernels_cuda/gas_optics_rrtmgp_kernels.cu:215-293, with
// tmin = numeric_limits<double>::min() (gas_optics_rrtmgp_kernels_launchers.cu:115).
void interpolation(const Kdist& kd, int ncol, int nlay, const double* play, const double* tlay,
                   const double* col_gas, Interp& o)
{
    const size_t ncl = size_t(ncol) * nlay;
    const int nflav = kd.nflav, neta = kd.neta, ngas = kd.ngas;
    o.jtemp.assign(ncl, 0); o.jpress.assign(ncl, 0); o.tropo.assign(ncl, 0);
    o.jeta.assign(2*ncl*nflav, 0); o.col_mix.assign(2*ncl*nflav, 0);
    o.fminor.assign(4*ncl*nflav, 0); o.fmajor.assign(8*ncl*nflav, 0);
    const double tmin = std::numeric_limits<double>::min();
    for (int iflav = 0; iflav < nflav; ++iflav)
        for (int ilay = 0; ilay < nlay; ++ilay)
            for (int icol = 0; icol < ncol; ++icol)
            {
                const size_t idx = icol + size_t(ncol)*ilay;
                int jt = int((tlay[idx] - (kd.temp_ref_min - kd.temp_ref_delta)) / kd.temp_ref_delta);
                jt = std::min(kd.ntemp - 1, std::max(1, jt));
                const double ftemp = (tlay[idx] - kd.temp_ref[jt-1]) / kd.temp_ref_delta;
                const double locpress = 1. + (std::log(play[idx]) - kd.press_ref_log[0]) / kd.press_ref_log_delta;
                const int jp = std::min(kd.npres - 1, std::max(1, int(locpress)));
                const double fpress = locpress - double(jp);
                const Bool tropo = std::log(play[idx]) > kd.press_ref_trop_log;
                const int itropo = !tropo;
                o.jtemp[idx] = jt; o.jpress[idx] = jp; o.tropo[idx] = tropo;

                const int gas1 = kd.flavor[2*iflav], gas2 = kd.flavor[2*iflav+1];
                const size_t fl = idx + ncl*iflav;
                for (int itemp = 0; itemp < 2; ++itemp)
                {
                    const size_t vb = itropo + size_t(jt+itemp-1) * (ngas+1) * 2;
                    const double ratio = kd.vmr_ref[vb + 2*gas1] / kd.vmr_ref[vb + 2*gas2];
                    const double c1 = col_gas[idx + ncl*gas1];
                    const double cm = c1 + ratio * col_gas[idx + ncl*gas2];
                    o.col_mix[itemp + 2*fl] = cm;
                    const double eta = (cm > 2.*tmin) ? c1 / cm : 0.5;
                    const double loceta = eta * double(neta-1);
                    o.jeta[itemp + 2*fl] = std::min(int(loceta) + 1, neta - 1);
                    const double feta = std::fmod(loceta, 1.0);
                    const double ftemp_term = double(1-itemp) + double(2*itemp-1) * ftemp;
                    const double f0 = (1. - feta) * ftemp_term, f1 = feta * ftemp_term;
                    o.fminor[2*(itemp + 2*fl)    ] = f0;
                    o.fminor[2*(itemp + 2*fl) + 1] = f1;
                    o.fmajor[4*(itemp + 2*fl)    ] = (1. - fpress) * f0;
                    o.fmajor[4*(itemp + 2*fl) + 1] = (1. - fpress) * f1;
                    o.fmajor[4*(itemp + 2*fl) + 2] = fpress * f0;
                    o.fmajor[4*(itemp + 2*fl) + 3] = fpress * f1;
                }
            }
}

// gas_optical_depths_major_kernel (gas_optics_rrtmgp_kernels.cu:296-341) followed by the minor
// kernel for the lower (idx_tropo = 1) then upper (idx_tropo = 0) sets (:356-476; order and
// idx_tropo from gas_optics_rrtmgp_kernels_launchers.cu:307,374).  tau must be zero on entry.
void compute_tau_absorption(const Kdist& kd, int ncol, int nlay, const double* play,
                            const double* tlay, const double* col_gas, const Interp& in, double* tau)
{
    const size_t ncl = size_t(ncol) * nlay;
    const int ntemp = kd.ntemp, neta = kd.neta, np1 = kd.npres + 1;
    for (int igpt = 0; igpt < kd.ngpt; ++igpt)
        for (int ilay = 0; ilay < nlay; ++ilay)
            for (int icol = 0; icol < ncol; ++icol)
            {
                const size_t idx = icol + size_t(ncol)*ilay;
                const int itropo = !in.tropo[idx];
                const int iflav = kd.gpoint_flavor[itropo + 2*igpt] - 1;
                const int jt = in.jtemp[idx];
                const int jpi = in.jpress[idx] + itropo;
                const size_t fl = idx + ncl*iflav;
                const double* fm = &in.fmajor[8*fl];
                double& t = tau[idx + ncl*igpt];
                for (int i = 0; i < 2; ++i)
                {
                    const int je = in.jeta[i + 2*fl];
                    auto K = [&](int e, int p) {
                        return kd.kmajor[(jt-1+i) + size_t(ntemp)*((e) + size_t(neta)*((p) + size_t(np1)*igpt))];
                    };
                    t += in.col_mix[i + 2*fl] *
                        (fm[i*4+0] * K(je-1, jpi-1) + fm[i*4+1] * K(je, jpi-1) +
                         fm[i*4+2] * K(je-1, jpi)   + fm[i*4+3] * K(je, jpi));
                }
            }

    for (int set = 0; set < 2; ++set)
    {
        const Minor& m = set == 0 ? kd.lower : kd.upper;
        const int idx_tropo = set == 0 ? 1 : 0;
        for (int ilay = 0; ilay < nlay; ++ilay)
            for (int icol = 0; icol < ncol; ++icol)
            {
                const size_t idx = icol + size_t(ncol)*ilay;
                if (in.tropo[idx] != idx_tropo) continue;
                for (int imnr = 0; imnr < m.n; ++imnr)
                {
                    double scaling = col_gas[idx + ncl*m.idx_minor[imnr]];
                    if (m.scales_with_density[imnr])
                    {
                        const double PaTohPa = 0.01;
                        scaling *= PaTohPa * play[idx] / tlay[idx];
                        if (m.idx_minor_scaling[imnr] > 0)
                        {
                            const double vmr_fact = 1. / col_gas[idx];
                            const double dry_fact = 1. / (1. + col_gas[idx + ncl*kd.idx_h2o] * vmr_fact);
                            if (m.scale_by_complement[imnr])
                                scaling *= (1. - col_gas[idx + ncl*m.idx_minor_scaling[imnr]] * vmr_fact * dry_fact);
                            else
                                scaling *= col_gas[idx + ncl*m.idx_minor_scaling[imnr]] * vmr_fact * dry_fact;
                        }
                    }
                    const int gs = m.limits[2*imnr] - 1, ge = m.limits[2*imnr+1];
                    const int gpt_offs = 1 - idx_tropo;
                    const int iflav = kd.gpoint_flavor[2*gs + gpt_offs] - 1;
                    const size_t fl = idx + ncl*iflav;
                    const double* f = &in.fminor[4*fl];
                    const int j0 = in.jeta[2*fl], j1 = in.jeta[2*fl+1];
                    const int jt = in.jtemp[idx];
                    const int goff = m.kminor_start[imnr] - 1;
                    for (int ig = 0; ig < ge - gs; ++ig)
                    {
                        auto K = [&](int t, int e) { return m.kminor[t + size_t(ntemp)*(e + size_t(neta)*(ig + goff))]; };
                        const double lt = f[0] * K(jt-1, j0-1) + f[1] * K(jt-1, j0) +
                                          f[2] * K(jt,   j1-1) + f[3] * K(jt,   j1);
                        tau[idx + ncl*(ig + gs)] += lt * scaling;
                    }
                }
            }
    }
}

// compute_tau_rayleigh_kernel: gas_optics_rrtmgp_kernels.cu:572-616.
void compute_tau_rayleigh(const Kdist& kd, int ncol, int nlay, const double* col_dry,
                          const double* col_gas, const Interp& in, double* tau_ray)
{
    const size_t ncl = size_t(ncol) * nlay;
    const int ntemp = kd.ntemp, neta = kd.neta;
    for (int igpt = 0; igpt < kd.ngpt; ++igpt)
        for (int ilay = 0; ilay < nlay; ++ilay)
            for (int icol = 0; icol < ncol; ++icol)
            {
                const size_t idx = icol + size_t(ncol)*ilay;
                const int itropo = !in.tropo[idx];
                const size_t kb = size_t(itropo) * ntemp * neta * kd.ngpt;
                const int jt = in.jtemp[idx];
                const int iflav = kd.gpoint_flavor[itropo + 2*igpt] - 1;
                const size_t fl = idx + ncl*iflav;
                const double* f = &in.fminor[4*fl];
                const int j0 = in.jeta[2*fl], j1 = in.jeta[2*fl+1];
                auto K = [&](int t, int e) { return kd.krayl[kb + t + size_t(ntemp)*(e + size_t(neta)*igpt)]; };
                const double k = f[0] * K(jt-1, j0-1) + f[1] * K(jt-1, j0) + f[2] * K(jt, j1-1) + f[3] * K(jt, j1);
                tau_ray[idx + ncl*igpt] = k * (col_gas[idx + ncl*kd.idx_h2o] + col_dry[idx]);
            }
}

// interpolate1D: gas_optics_rrtmgp_kernels.cu:1-12.
double interpolate1D(double val, double offset, double delta, int len, const double* table)
{
    const double val0 = (val - offset) / delta;
    const double frac = val0 - int(val0);
    const int idx = std::min(len - 1, std::max(1, int(val0) + 1));
    return table[idx-1] + frac * (table[idx] - table[idx-1]);
}

// Planck_source_kernel: gas_optics_rrtmgp_kernels.cu:129-212, delta_Tsurf = 1
// (gas_optics_rrtmgp_kernels_launchers.cu:468); sfc_lay from Gas_optics_rrtmgp.cpp:1333.
void compute_planck_source(const Kdist& kd, int ncol, int nlay, const double* play,
                           const double* tlay, const double* tlev, const double* tsfc,
                           const Interp& in, double* sfc_src, double* lay_src, double* lev_inc,
                           double* lev_dec, double* sfc_src_jac)
{
    const size_t ncl = size_t(ncol) * nlay;
    const int ntemp = kd.ntemp, neta = kd.neta, np1 = kd.npres + 1;
    const int sfc_lay = play[0] > play[size_t(ncol)*(nlay-1)] ? 1 : nlay;
    const double dT = 1.;
    for (int igpt = 0; igpt < kd.ngpt; ++igpt)
        for (int ilay = 0; ilay < nlay; ++ilay)
            for (int icol = 0; icol < ncol; ++icol)
            {
                const size_t idx = icol + size_t(ncol)*ilay;
                const int ibnd = kd.gpt2band[igpt] - 1;
                const int itropo = !in.tropo[idx];
                const int iflav = kd.gpoint_flavor[itropo + 2*igpt] - 1;
                const size_t fl = idx + ncl*iflav;
                const double* f = &in.fmajor[8*fl];
                const int j0 = in.jeta[2*fl], j1 = in.jeta[2*fl+1];
                const int jt = in.jtemp[idx];
                const int jp = in.jpress[idx] + itropo;
                const double* tab = &kd.totplnk[size_t(ibnd) * kd.nplanck];
                auto P = [&](int t, int e, int p) { return kd.planck_frac[t + size_t(ntemp)*(e + size_t(neta)*(p + size_t(np1)*igpt))]; };
                const double pfrac =
                      (f[0] * P(jt-1, j0-1, jp-1) + f[1] * P(jt-1, j0, jp-1) + f[2] * P(jt-1, j0-1, jp) + f[3] * P(jt-1, j0, jp))
                    + (f[4] * P(jt,   j1-1, jp-1) + f[5] * P(jt,   j1, jp-1) + f[6] * P(jt,   j1-1, jp) + f[7] * P(jt,   j1, jp));
                const double b_lay = interpolate1D(tlay[idx], kd.temp_ref_min, kd.totplnk_delta, kd.nplanck, tab);
                const double b_lev1 = interpolate1D(tlev[icol + size_t(ncol)*(ilay+1)], kd.temp_ref_min, kd.totplnk_delta, kd.nplanck, tab);
                const double b_lev2 = interpolate1D(tlev[icol + size_t(ncol)*ilay], kd.temp_ref_min, kd.totplnk_delta, kd.nplanck, tab);
                lay_src[idx + ncl*igpt] = pfrac * b_lay;
                lev_inc[idx + ncl*igpt] = pfrac * b_lev1;
                lev_dec[idx + ncl*igpt] = pfrac * b_lev2;
                if (ilay == sfc_lay - 1)
                {
                    const double s1 = interpolate1D(tsfc[icol], kd.temp_ref_min, kd.totplnk_delta, kd.nplanck, tab);
                    const double s2 = interpolate1D(tsfc[icol] + dT, kd.temp_ref_min, kd.totplnk_delta, kd.nplanck, tab);
                    sfc_src[icol + size_t(ncol)*igpt] = pfrac * s1;
                    sfc_src_jac[icol + size_t(ncol)*igpt] = pfrac * (s2 - s1);
                }
            }
}

bool any_outside(const double* a, size_t n, double lo, double hi)
{
    for (size_t i = 0; i < n; ++i) if (a[i] < lo || a[i] > hi) return true;
    return false;
}

// Gas_optics_rrtmgp::gas_optics LW (:796-849) / SW (:852-898) -> compute_gas_taus (:1090-1275).
// Range checks follow :813-826 (CPU semantics; the GPU reference skips them).
void gas_optics(const Kdist& kd, int ncol, int nlay, const double* play, const double* plev,
                const double* tlay, const double* tsfc, const double* tlev,
                const OracleGasConcs& gc, const double* col_dry,
                double* tau, double* ssa, double* g, double* lay_src, double* lev_inc,
                double* lev_dec, double* sfc_src, double* sfc_src_jac, double* toa_src,
                Interp* interp_out)
{
    const size_t ncl = size_t(ncol) * nlay;
    const int nlev = nlay + 1;
    if (any_outside(play, ncl, kd.press_ref_min, kd.press_ref_max)) throw std::range_error("play is out of range");
    if (any_outside(plev, size_t(ncol)*nlev, kd.press_ref_min, kd.press_ref_max)) throw std::range_error("plev is out of range");
    if (any_outside(tlay, ncl, kd.temp_ref_min, kd.temp_ref_max)) throw std::range_error("tlay is out of range");
    if (kd.is_lw)
    {
        if (any_outside(tlev, size_t(ncol)*nlev, kd.temp_ref_min, kd.temp_ref_max)) throw std::range_error("tlev is out of range");
        if (any_outside(tsfc, ncol, kd.temp_ref_min, kd.temp_ref_max)) throw std::range_error("tsfc is out of range");
    }
    for (size_t i = 0; i < ncl; ++i) if (col_dry[i] < 0.) throw std::range_error("col_dry is out of range");

    // vmr fill and col_gas (:1121-1160); col_gas(:,:,0) = col_dry.
    const auto names = split_names(gc.names, gc.n);
    std::vector<double> col_gas(ncl * (kd.ngas + 1));
    for (size_t i = 0; i < ncl; ++i) col_gas[i] = col_dry[i];
    for (int ig = 1; ig <= kd.ngas; ++ig)
    {
        const int k = find_index(names, kd.gas_names[ig-1]);
        if (k < 1) throw std::runtime_error("gas " + kd.gas_names[ig-1] + " missing from gas_desc");
        const double* v = gc.data[k-1];
        const int d1 = gc.dims[2*(k-1)], d2 = gc.dims[2*(k-1)+1];
        for (int ilay = 0; ilay < nlay; ++ilay)
            for (int icol = 0; icol < ncol; ++icol)
            {
                double vmr;
                if (d1 == 1 && d2 == 1) vmr = v[0];
                else if (d1 == 1) vmr = v[ilay];
                else vmr = v[icol + size_t(ncol)*ilay];
                const size_t idx = icol + size_t(ncol)*ilay;
                col_gas[idx + ncl*ig] = vmr * col_dry[idx];
            }
    }
    if (kd.idx_h2o < 1) throw std::runtime_error("idx_h2o cannot be found");

    Interp in;
    interpolation(kd, ncol, nlay, play, tlay, col_gas.data(), in);

    const size_t ng = ncl * kd.ngpt;
    if (!kd.krayl.empty())
    {
        std::vector<double> tau_abs(ng, 0.0), tau_ray(ng, 0.0);
        compute_tau_absorption(kd, ncol, nlay, play, tlay, col_gas.data(), in, tau_abs.data());
        compute_tau_rayleigh(kd, ncol, nlay, col_dry, col_gas.data(), in, tau_ray.data());
        // combine_abs_and_rayleigh (CPU, :365-385): threshold 2*epsilon, g = 0.
        const double eps = std::numeric_limits<double>::epsilon();
        for (size_t i = 0; i < ng; ++i)
        {
            const double t = tau_abs[i] + tau_ray[i];
            ssa[i] = (t > 2. * eps) ? tau_ray[i] / t : 0.;
            tau[i] = t;
            g[i] = 0.;
        }
        if (toa_src)
            for (int igpt = 0; igpt < kd.ngpt; ++igpt)
                for (int icol = 0; icol < ncol; ++icol)
                    toa_src[icol + size_t(ncol)*igpt] = kd.solar_source[igpt];
    }
    else
    {
        std::fill(tau, tau + ng, 0.0);
        compute_tau_absorption(kd, ncol, nlay, play, tlay, col_gas.data(), in, tau);
        if (lay_src)
            compute_planck_source(kd, ncol, nlay, play, tlay, tlev, tsfc, in, sfc_src, lay_src,
                                  lev_inc, lev_dec, sfc_src_jac);
    }
    if (interp_out) *interp_out = std::move(in);
}

// ---------------------------------------------------------------------------------------------
// Cloud optics: src/Cloud_optics.cpp.

struct Cloud
{
    int nbnd, liq_nsteps, ice_nsteps;
    double liq_step_size, ice_step_size, radliq_lwr, radliq_upr, radice_lwr, radice_upr;
    std::vector<double> extliq, ssaliq, asyliq, extice, ssaice, asyice;   // (nsize, nbnd)
};

// compute_all_from_table: src/Cloud_optics.cpp:72-107.
void compute_all_from_table(int ncol, int nlay, int nbnd, const double* cwp, const double* re,
                            int nsteps, double step_size, double offset, const double* tt,
                            const double* st, const double* at, double* tau, double* taussa,
                            double* taussag)
{
    const size_t ncl = size_t(ncol) * nlay;
    for (int ibnd = 0; ibnd < nbnd; ++ibnd)
        for (size_t i = 0; i < ncl; ++i)
        {
            const size_t o = i + ncl*ibnd;
            if (cwp[i] > 0.)   // mask: cwp > mask_min_value = 0 (:127-134)
            {
                const int index = std::min(int((re[i] - offset) / step_size) + 1, nsteps - 1);
                const double fint = (re[i] - offset) / step_size - (index - 1);
                auto L = [&](const double* t) {
                    const double a = t[(index-1) + size_t(nsteps)*ibnd], b = t[index + size_t(nsteps)*ibnd];
                    return a + fint * (b - a);
                };
                const double tl = cwp[i] * L(tt);
                const double tsl = tl * L(st);
                tau[o] = tl; taussa[o] = tsl; taussag[o] = tsl * L(at);
            }
            else
            {
                tau[o] = 0.; taussa[o] = 0.; taussag[o] = 0.;
            }
        }
}

// Cloud_optics::cloud_optics 2str (:111-172) and 1scl (:176-232).
void cloud_optics(const Cloud& c, int ncol, int nlay, const double* lwp, const double* iwp,
                  const double* rel, const double* rei, bool two_stream, double* tau,
                  double* ssa, double* g)
{
    const size_t n = size_t(ncol) * nlay * c.nbnd;
    std::vector<double> lt(n), ls(n), lg(n), it(n), is(n), ig(n);
    compute_all_from_table(ncol, nlay, c.nbnd, lwp, rel, c.liq_nsteps, c.liq_step_size, c.radliq_lwr,
                           c.extliq.data(), c.ssaliq.data(), c.asyliq.data(), lt.data(), ls.data(), lg.data());
    compute_all_from_table(ncol, nlay, c.nbnd, iwp, rei, c.ice_nsteps, c.ice_step_size, c.radice_lwr,
                           c.extice.data(), c.ssaice.data(), c.asyice.data(), it.data(), is.data(), ig.data());
    const double eps = std::numeric_limits<double>::epsilon();
    for (size_t i = 0; i < n; ++i)
    {
        if (two_stream)
        {
            const double t = lt[i] + it[i], ts = ls[i] + is[i], tsg = lg[i] + ig[i];
            tau[i] = t;
            ssa[i] = ts / std::max(t, eps);
            g[i] = tsg / std::max(ts, eps);
        }
        else
            tau[i] = (lt[i] - ls[i]) + (it[i] - is[i]);
    }
}

// ---------------------------------------------------------------------------------------------
// Aerosol optics: src/Aerosol_optics.cpp:26-224.
//
// rh_class (:26-36): the first 1-based class whose upper bound is >= rh.  The reference reads
// past the table when rh exceeds every bound; here the last class is taken.
int aerosol_rh_class(double rh, const double* rh_upper, int nhum)
{
    int ihum = 1;
    while (ihum < nhum && rh_upper[ihum - 1] < rh) ++ihum;
    return ihum;
}

// Species of compute_all_from_table (:49-146), in its summation order: the mass-mixing-ratio
// field aermrNN, whether the species is hydrophilic (table indexed by humidity class), and its
// 1-based column of the hydrophobic or hydrophilic table.
struct AerosolSpecies { int field; bool philic; int column; };
const AerosolSpecies aerosol_species[11] = {
    {1, true, 1},    // SS1
    {2, true, 2},    // SS2
    {3, true, 3},    // SS3
    {4, false, 1},   // DU1
    {5, false, 8},   // DU2
    {6, false, 6},   // DU3
    {8, false, 10},  // OM1
    {7, true, 4},    // OM2
    {9, false, 11},  // BC1
    {10, false, 11}, // BC2
    {11, true, 5},   // SU
};

// compute_all_from_table (:38-150) and Aerosol_optics::aerosol_optics (:168-224) for (ncol, nlay)
// mixing ratios aermr[0..10], rh (ncol, nlay) and plev (ncol, nlay+1); tables mext/ssa/g_phobic
// (nbnd, nphobic) and _philic (nbnd, nhum, nphilic); outputs (ncol, nlay, nbnd).  The layer mass
// is |p(ilay) - p(ilay+1)| / 9.81: the CPU path takes the signed difference (:55), which is
// the same for levels ordered from the surface up and negative the other way round; the GPU path
// takes its magnitude (src_cuda/Aerosol_optics.cu:67), which is kept here.
void aerosol_optics(int ncol, int nlay, int nbnd, int nhum, int nphobic, int nphilic, const double* const* aermr,
                    const double* rh, const double* plev, const double* rh_upper, const double* mext_phobic,
                    const double* ssa_phobic, const double* g_phobic, const double* mext_philic,
                    const double* ssa_philic, const double* g_philic, double* tau, double* ssa, double* g)
{
    for (const auto& s : aerosol_species)
        if (s.column > (s.philic ? nphilic : nphobic))
            throw std::runtime_error("aerosol_optics: table has too few species");
    const double eps = std::numeric_limits<double>::epsilon();
    const size_t ncl = size_t(ncol) * nlay;
    for (int ibnd = 0; ibnd < nbnd; ++ibnd)
        for (int ilay = 0; ilay < nlay; ++ilay)
            for (int icol = 0; icol < ncol; ++icol)
            {
                const size_t i2 = icol + size_t(ncol) * ilay;
                const double dpg = std::abs(plev[i2] - plev[i2 + ncol]) / 9.81;
                const int ihum = aerosol_rh_class(rh[i2], rh_upper, nhum);
                double t = 0., ts = 0., tsg = 0.;
                for (const auto& s : aerosol_species)
                {
                    const size_t it = s.philic
                        ? ibnd + size_t(nbnd) * ((ihum - 1) + size_t(nhum) * (s.column - 1))
                        : ibnd + size_t(nbnd) * (s.column - 1);
                    const double mext = s.philic ? mext_philic[it] : mext_phobic[it];
                    const double sa = s.philic ? ssa_philic[it] : ssa_phobic[it];
                    const double ga = s.philic ? g_philic[it] : g_phobic[it];
                    const double od = aermr[s.field - 1][i2] * dpg * mext;
                    t += od;
                    ts += od * sa;
                    tsg += od * sa * ga;
                }
                const size_t o = i2 + ncl * ibnd;
                tau[o] = t;
                ssa[o] = ts / std::max(t, eps);
                g[o] = tsg / std::max(ts, eps);
            }
}

// ---------------------------------------------------------------------------------------------
// Optical-property arithmetic: src_kernels_cuda/optical_props_kernels.cu, eps = 3*DBL_MIN
// (optical_props_kernels_launchers.cu:80,123,172).

const double op_eps = 3. * std::numeric_limits<double>::min();

// delta_scale_2str_k_kernel (:139-160).
void delta_scale_2str(size_t n, double* tau, double* ssa, double* g)
{
    for (size_t i = 0; i < n; ++i)
    {
        const double gv = g[i], sv = ssa[i];
        const double f = gv * gv, wf = sv * f;
        tau[i] *= (1. - wf);
        ssa[i] = (sv - wf) / std::max(op_eps, 1. - wf);
        g[i] = (gv - f) / std::max(op_eps, 1. - f);
    }
}

// inc_1scalar_by_1scalar_bybnd_kernel (:74-96) and inc_2stream_by_2stream_bybnd_kernel (:99-136).
void inc_bybnd(int ncol, int nlay, int ngpt, double* tau1, double* ssa1, double* g1,
               const double* tau2, const double* ssa2, const double* g2, int nbnd,
               const int* band_lims_gpt)
{
    const size_t ncl = size_t(ncol) * nlay;
    for (int ibnd = 0; ibnd < nbnd; ++ibnd)
        for (int igpt = band_lims_gpt[2*ibnd] - 1; igpt < band_lims_gpt[2*ibnd+1]; ++igpt)
            for (size_t i = 0; i < ncl; ++i)
            {
                const size_t a = i + ncl*igpt, b = i + ncl*ibnd;
                if (!ssa1)
                    tau1[a] = tau1[a] + tau2[b];
                else
                {
                    const double t1 = tau1[a], t2 = tau2[b], s1 = ssa1[a], s2 = ssa2[b];
                    const double tau12 = t1 + t2;
                    const double tauscat12 = (t1 * s1) + (t2 * s2);
                    g1[a] = ((t1 * s1 * g1[a]) + (t2 * s2 * g2[b])) / std::max(tauscat12, op_eps);
                    ssa1[a] = tauscat12 / std::max(op_eps, tau12);
                    tau1[a] = tau12;
                }
            }
}

// increment_{1scalar,2stream}_by_{1scalar,2stream}_kernel (:30-71), same spectral grid.
void increment(size_t n, double* tau1, double* ssa1, double* g1, const double* tau2,
               const double* ssa2, const double* g2)
{
    for (size_t i = 0; i < n; ++i)
    {
        if (!ssa1) { tau1[i] = tau1[i] + tau2[i]; continue; }
        const double t1 = tau1[i], t2 = tau2[i];
        const double tau12 = t1 + t2;
        const double tauscat12 = (t1 * ssa1[i]) + (t2 * ssa2[i]);
        g1[i] = ((t1 * ssa1[i] * g1[i]) + (t2 * ssa2[i] * g2[i])) / std::max(tauscat12, op_eps);
        ssa1[i] = tauscat12 / std::max(op_eps, tau12);
        tau1[i] = tau12;
    }
}

// ---------------------------------------------------------------------------------------------
// Longwave solver: vendored src_kernels/mo_rte_solver_kernels.F90.

const double pi = std::acos(-1.0);

// lw_transport_1rescl (F90:731-822): the up transport with the adjustment of Tang et al. (2018)
// Eq. 18b/20 for the linear-in-tau source, then a second down transport with that of Eq. 19/21,
// overwriting the first.  The up pass reads the first down pass's radiance at the top of each
// layer in both orientations; the second down pass reads the up radiance at level ilev of layer
// ilev, which is the layer's top when top_at_1 and its bottom otherwise (F90:773-776, 806-809),
// as the reference does.  The Jacobian is transported without adjustment.
void lw_transport_1rescl(int ncol, int nlay, bool top_at_1, const double* trans, const double* src_dn,
                         const double* src_up, double* up, double* dn, const double* An, const double* Cn,
                         bool do_jac, double* jac)
{
    auto L = [ncol](int i, int k) { return i + size_t(ncol)*k; };
    if (top_at_1)
    {
        for (int il = nlay - 1; il >= 0; --il)
        {
            for (int i = 0; i < ncol; ++i)
            {
                const size_t l = L(i, il);
                const double adj = Cn[l] * (An[l] * dn[l] - trans[l] * src_dn[l] - src_up[l]);
                up[l] = trans[l] * up[L(i, il+1)] + src_up[l] + adj;
            }
            if (do_jac) for (int i = 0; i < ncol; ++i) jac[L(i, il)] = trans[L(i, il)] * jac[L(i, il+1)];
        }
        for (int il = 0; il < nlay; ++il)
            for (int i = 0; i < ncol; ++i)
            {
                const size_t l = L(i, il);
                const double adj = Cn[l] * (An[l] * up[l] - trans[l] * src_up[l] - src_dn[l]);
                dn[L(i, il+1)] = trans[l] * dn[l] + src_dn[l] + adj;
            }
    }
    else
    {
        for (int il = 0; il < nlay; ++il)
        {
            for (int i = 0; i < ncol; ++i)
            {
                const size_t l = L(i, il);
                const double adj = Cn[l] * (An[l] * dn[L(i, il+1)] - trans[l] * src_dn[l] - src_up[l]);
                up[L(i, il+1)] = trans[l] * up[l] + src_up[l] + adj;
            }
            if (do_jac) for (int i = 0; i < ncol; ++i) jac[L(i, il+1)] = trans[L(i, il)] * jac[L(i, il)];
        }
        for (int il = nlay - 1; il >= 0; --il)
            for (int i = 0; i < ncol; ++i)
            {
                const size_t l = L(i, il);
                const double adj = Cn[l] * (An[l] * up[l] - trans[l] * src_up[l] - src_dn[l]);
                dn[l] = trans[l] * dn[L(i, il+1)] + src_dn[l] + adj;
            }
    }
}

// lw_source_noscat (F90:616-653), transport dn (:659-689) and up (:691-723), and the g-point
// loop of lw_solver_noscat (F90:58-260).  Fluxes are (ncol, nlay+1, ngpt) or broadband
// (ncol, nlay+1) when do_broadband.  Jacobian is the broadband sfc-T Jacobian (F90:236-256).
// do_rescaling (F90:168-200, Tang et al. 2018): tau_loc = tau D (1 - w + w (1 - g)/2), the
// adjustment coefficients Cn = 0.4 w (1 - g)/2 / scaleTau and An = 1 - trans^2, and the up and
// second down transport of lw_transport_1rescl (F90:731-822) in place of the up transport.
void lw_solver_noscat(int ncol, int nlay, int ngpt, bool top_at_1, const double* D, double weight,
                      const double* tau, const double* lay_source, const double* lev_source_inc,
                      const double* lev_source_dec, const double* sfc_emis, const double* sfc_src,
                      const double* inc_flux, double* flux_up, double* flux_dn, bool do_broadband,
                      double* bb_up, double* bb_dn, bool do_jac, const double* sfc_src_jac,
                      double* flux_up_jac, bool do_rescaling = false, const double* ssa = nullptr,
                      const double* g = nullptr)
{
    const int nlev = nlay + 1;
    const int top_level = top_at_1 ? 0 : nlay;
    const int sfc_level = top_at_1 ? nlay : 0;
    const double* lev_up = top_at_1 ? lev_source_dec : lev_source_inc;
    const double* lev_dn = top_at_1 ? lev_source_inc : lev_source_dec;
    const double tau_thresh = std::sqrt(std::numeric_limits<double>::epsilon());
    const size_t ncl = size_t(ncol) * nlay, ncv = size_t(ncol) * nlev;

    std::vector<double> tau_loc(ncl), trans(ncl), src_dn(ncl), src_up(ncl), An, Cn;
    std::vector<double> loc_up(ncv), loc_dn(ncv), jac(ncv);
    if (do_rescaling)
    {
        if (!ssa || !g) throw std::runtime_error("lw_solver_noscat: rescaling needs ssa and g");
        An.resize(ncl);
        Cn.resize(ncl);
    }
    if (do_broadband) { std::fill(bb_up, bb_up + ncv, 0.); std::fill(bb_dn, bb_dn + ncv, 0.); }
    if (do_jac) std::fill(flux_up_jac, flux_up_jac + ncv, 0.);

    for (int igpt = 0; igpt < ngpt; ++igpt)
    {
        double* gup = do_broadband ? loc_up.data() : flux_up + ncv*igpt;
        double* gdn = do_broadband ? loc_dn.data() : flux_dn + ncv*igpt;
        for (int icol = 0; icol < ncol; ++icol)
            gdn[icol + size_t(ncol)*top_level] = inc_flux[icol + size_t(ncol)*igpt] / (2. * pi * weight);
        for (size_t i = 0; i < ncl; ++i)
        {
            const size_t icol = i % ncol;
            tau_loc[i] = tau[i + ncl*igpt] * D[icol + size_t(ncol)*igpt];
            if (do_rescaling)
            {
                const double ssal = ssa[i + ncl*igpt];
                const double wb = ssal * (1. - g[i + ncl*igpt]) * 0.5;
                const double scale_tau = 1. - ssal + wb;
                Cn[i] = 0.4 * wb / scale_tau;
                tau_loc[i] *= scale_tau;
            }
            trans[i] = std::exp(-tau_loc[i]);
            if (do_rescaling) An[i] = 1. - trans[i] * trans[i];
        }
        for (size_t i = 0; i < ncl; ++i)
        {
            const double t = tau_loc[i];
            const double fact = (t > tau_thresh) ? (1. - trans[i]) / t - trans[i]
                                                 : t * (0.5 - 1./3.*t);
            const double ls = lay_source[i + ncl*igpt];
            const double ld = lev_dn[i + ncl*igpt], lu = lev_up[i + ncl*igpt];
            src_dn[i] = (1. - trans[i]) * ld + 2. * fact * (ls - ld);
            src_up[i] = (1. - trans[i]) * lu + 2. * fact * (ls - lu);
        }
        if (top_at_1)
            for (int ilev = 1; ilev <= nlay; ++ilev)
                for (int icol = 0; icol < ncol; ++icol)
                    gdn[icol + size_t(ncol)*ilev] = trans[icol + size_t(ncol)*(ilev-1)] * gdn[icol + size_t(ncol)*(ilev-1)]
                                                  + src_dn[icol + size_t(ncol)*(ilev-1)];
        else
            for (int ilev = nlay - 1; ilev >= 0; --ilev)
                for (int icol = 0; icol < ncol; ++icol)
                    gdn[icol + size_t(ncol)*ilev] = trans[icol + size_t(ncol)*ilev] * gdn[icol + size_t(ncol)*(ilev+1)]
                                                  + src_dn[icol + size_t(ncol)*ilev];
        for (int icol = 0; icol < ncol; ++icol)
        {
            const double e = sfc_emis[icol + size_t(ncol)*igpt];
            const size_t s = icol + size_t(ncol)*sfc_level;
            gup[s] = gdn[s] * (1. - e) + e * sfc_src[icol + size_t(ncol)*igpt];
            if (do_jac) jac[s] = e * sfc_src_jac[icol + size_t(ncol)*igpt];
        }
        if (do_rescaling)
            lw_transport_1rescl(ncol, nlay, top_at_1, trans.data(), src_dn.data(), src_up.data(), gup, gdn,
                                An.data(), Cn.data(), do_jac, jac.data());
        else if (top_at_1)
            for (int ilev = nlay - 1; ilev >= 0; --ilev)
                for (int icol = 0; icol < ncol; ++icol)
                {
                    const size_t l = icol + size_t(ncol)*ilev;
                    gup[l] = trans[l] * gup[l + ncol] + src_up[l];
                    if (do_jac) jac[l] = trans[l] * jac[l + ncol];
                }
        else
            for (int ilev = 1; ilev <= nlay; ++ilev)
                for (int icol = 0; icol < ncol; ++icol)
                {
                    const size_t l = icol + size_t(ncol)*ilev, lm = l - ncol;
                    gup[l] = trans[lm] * gup[lm] + src_up[lm];
                    if (do_jac) jac[l] = trans[lm] * jac[lm];
                }
        if (do_broadband)
            for (size_t i = 0; i < ncv; ++i) { bb_up[i] += gup[i]; bb_dn[i] += gdn[i]; }
        else
            for (size_t i = 0; i < ncv; ++i) { gdn[i] = 2. * pi * weight * gdn[i]; gup[i] = 2. * pi * weight * gup[i]; }
        if (do_jac)
            for (size_t i = 0; i < ncv; ++i) flux_up_jac[i] += jac[i];
    }
    if (do_broadband)
        for (size_t i = 0; i < ncv; ++i) { bb_up[i] = 2. * pi * weight * bb_up[i]; bb_dn[i] = 2. * pi * weight * bb_dn[i]; }
    if (do_jac)
        for (size_t i = 0; i < ncv; ++i) flux_up_jac[i] = 2. * pi * weight * flux_up_jac[i];
}

// Gauss tables of Rte_lw::rte_lw (src/Rte_lw.cpp:100-113).
const double gauss_Ds[4][4] = {{1.66, 0., 0., 0.},
                               {1.18350343, 2.81649655, 0., 0.},
                               {1.09719858, 1.69338507, 4.70941630, 0.},
                               {1.06056257, 1.38282560, 2.40148179, 7.15513024}};
const double gauss_wts[4][4] = {{0.5, 0., 0., 0.},
                                {0.3180413817, 0.1819586183, 0., 0.},
                                {0.2009319137, 0.2292411064, 0.0698269799, 0.},
                                {0.1355069134, 0.2034645680, 0.1298475476, 0.0311809710}};

// Rte_lw::rte_lw (src/Rte_lw.cpp:90-162) + lw_solver_noscat_GaussQuad (F90:268-376).
void rte_lw(int ncol, int nlay, int ngpt, int nbnd, const int* band_lims, bool top_at_1,
            const double* tau, const double* lay_src, const double* lev_inc, const double* lev_dec,
            const double* sfc_src, const double* sfc_emis_bnd, const double* inc_flux, int nmus,
            bool do_broadband, double* flux_up, double* flux_dn, const double* sfc_src_jac,
            double* flux_up_jac, bool do_rescaling = false, const double* ssa = nullptr,
            const double* g = nullptr)
{
    if (nmus < 1 || nmus > 4) throw std::runtime_error("n_gauss_angles must be 1..4");
    const int nlev = nlay + 1;
    const size_t ncv = size_t(ncol) * nlev;
    // expand_and_transpose (src/Rte_lw.cpp:165-178): (nbnd, ncol) -> (ncol, ngpt).
    std::vector<double> emis(size_t(ncol) * ngpt);
    for (int ib = 0; ib < nbnd; ++ib)
        for (int icol = 0; icol < ncol; ++icol)
            for (int ig = band_lims[2*ib] - 1; ig < band_lims[2*ib+1]; ++ig)
                emis[icol + size_t(ncol)*ig] = sfc_emis_bnd[ib + size_t(nbnd)*icol];
    std::vector<double> zero_inc;
    if (!inc_flux) { zero_inc.assign(size_t(ncol) * ngpt, 0.); inc_flux = zero_inc.data(); }
    const bool do_jac = flux_up_jac != nullptr;
    std::vector<double> D(size_t(ncol) * ngpt);
    const size_t nout = do_broadband ? ncv : ncv * ngpt;
    std::vector<double> tup, tdn, tjac;
    for (int imu = 0; imu < nmus; ++imu)
    {
        std::fill(D.begin(), D.end(), gauss_Ds[nmus-1][imu]);
        const double w = gauss_wts[nmus-1][imu];
        if (imu == 0)
            lw_solver_noscat(ncol, nlay, ngpt, top_at_1, D.data(), w, tau, lay_src, lev_inc, lev_dec,
                             emis.data(), sfc_src, inc_flux, flux_up, flux_dn, do_broadband,
                             flux_up, flux_dn, do_jac, sfc_src_jac, flux_up_jac, do_rescaling, ssa, g);
        else
        {
            tup.assign(nout, 0.); tdn.assign(nout, 0.); tjac.assign(ncv, 0.);
            lw_solver_noscat(ncol, nlay, ngpt, top_at_1, D.data(), w, tau, lay_src, lev_inc, lev_dec,
                             emis.data(), sfc_src, inc_flux, tup.data(), tdn.data(), do_broadband,
                             tup.data(), tdn.data(), do_jac, sfc_src_jac, tjac.data(), do_rescaling, ssa, g);
            for (size_t i = 0; i < nout; ++i) { flux_up[i] += tup[i]; flux_dn[i] += tdn[i]; }
            if (do_jac) for (size_t i = 0; i < ncv; ++i) flux_up_jac[i] += tjac[i];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Shortwave solver.

// sw_dif_and_source (F90:996-1109) for one g-point.  Decisions:
//  * k_min = 1e-12 (FP64, F90:40-44);
//  * the 1-k^2 mu0^2 guard uses ">= epsilon" (F90:1080-1082; CUDA uses ">", :570);
//  * the Rdir/Tdir clamp of the in-tree CUDA restatement (rte_solver_kernels.cu:581-583) is
//    applied; the vendored F90 predates it.
void sw_dif_and_source(int ncol, int nlay, bool top_at_1, const double* mu0, const double* sfc_alb,
                       const double* tau, const double* w0, const double* g, double* Rdif,
                       double* Tdif, double* source_dn, double* source_up, double* source_sfc,
                       double* flux_dn_dir)
{
    const double k_min = 1.e-12;
    const double eps = std::numeric_limits<double>::epsilon();
    int last_trans_lev = 0;
    for (int j = 0; j < nlay; ++j)
    {
        const int lay = top_at_1 ? j : nlay - 1 - j;
        const int lev_inc = top_at_1 ? lay : lay + 1;
        const int lev_trans = top_at_1 ? lay + 1 : lay;
        last_trans_lev = lev_trans;
        for (int i = 0; i < ncol; ++i)
        {
            const size_t l = i + size_t(ncol)*lay;
            const double tau_s = tau[l], w0_s = w0[l], g_s = g[l], mu0_s = mu0[i];
            const double gamma1 = (8. - w0_s * (5. + 3. * g_s)) * .25;
            const double gamma2 = 3. * (w0_s * (1. - g_s)) * .25;
            const double gamma3 = (2. - 3. * mu0_s * g_s) * .25;
            const double gamma4 = 1. - gamma3;
            const double alpha1 = gamma1 * gamma4 + gamma2 * gamma3;
            const double alpha2 = gamma1 * gamma3 + gamma2 * gamma4;
            const double k = std::sqrt(std::max((gamma1 - gamma2) * (gamma1 + gamma2), k_min));
            const double k_mu = k * mu0_s, k_gamma3 = k * gamma3, k_gamma4 = k * gamma4;
            const double exp_minusktau = std::exp(-tau_s * k);
            const double exp_minus2ktau = exp_minusktau * exp_minusktau;
            double RT_term = 1. / (k * (1. + exp_minus2ktau) + gamma1 * (1. - exp_minus2ktau));
            Rdif[l] = RT_term * gamma2 * (1. - exp_minus2ktau);
            Tdif[l] = RT_term * 2. * k * exp_minusktau;
            const double denom = (std::abs(1. - k_mu*k_mu) >= eps) ? 1. - k_mu*k_mu : eps;
            RT_term = w0_s * RT_term / denom;
            const double Tnoscat = std::exp(-tau_s / mu0_s);
            double Rdir = RT_term * ((1. - k_mu) * (alpha2 + k_gamma3) -
                                     (1. + k_mu) * (alpha2 - k_gamma3) * exp_minus2ktau -
                                     2.0 * (k_gamma3 - alpha2 * k_mu) * exp_minusktau * Tnoscat);
            double Tdir = -RT_term * ((1. + k_mu) * (alpha1 + k_gamma4) * Tnoscat -
                                      (1. - k_mu) * (alpha1 - k_gamma4) * exp_minus2ktau * Tnoscat -
                                      2.0 * (k_gamma4 + alpha1 * k_mu) * exp_minusktau);
            Rdir = std::max(eps, std::min(Rdir, 1.0 - Tnoscat));
            Tdir = std::max(eps, std::min(Tdir, 1.0 - Tnoscat - Rdir));
            const double inc = flux_dn_dir[i + size_t(ncol)*lev_inc];
            source_up[l] = Rdir * inc;
            source_dn[l] = Tdir * inc;
            flux_dn_dir[i + size_t(ncol)*lev_trans] = Tnoscat * inc;
        }
    }
    for (int i = 0; i < ncol; ++i)
        source_sfc[i] = flux_dn_dir[i + size_t(ncol)*last_trans_lev] * sfc_alb[i];
}

// adding (F90:1117-1227), Shonk and Hogan 2008.
void adding(int ncol, int nlay, bool top_at_1, const double* albedo_sfc, const double* rdif,
            const double* tdif, const double* src_dn, const double* src_up, const double* src_sfc,
            double* flux_up, double* flux_dn)
{
    const int nlev = nlay + 1;
    std::vector<double> albedo(size_t(ncol) * nlev), src(size_t(ncol) * nlev), denom(size_t(ncol) * nlay);
    auto L = [ncol](int i, int k) { return i + size_t(ncol)*k; };
    if (top_at_1)
    {
        for (int i = 0; i < ncol; ++i) { albedo[L(i, nlay)] = albedo_sfc[i]; src[L(i, nlay)] = src_sfc[i]; }
        for (int il = nlay - 1; il >= 0; --il)
            for (int i = 0; i < ncol; ++i)
            {
                denom[L(i, il)] = 1. / (1. - rdif[L(i, il)] * albedo[L(i, il+1)]);
                albedo[L(i, il)] = rdif[L(i, il)] + tdif[L(i, il)] * tdif[L(i, il)] * albedo[L(i, il+1)] * denom[L(i, il)];
                src[L(i, il)] = src_up[L(i, il)] + tdif[L(i, il)] * denom[L(i, il)] *
                                (src[L(i, il+1)] + albedo[L(i, il+1)] * src_dn[L(i, il)]);
            }
        for (int i = 0; i < ncol; ++i) flux_up[L(i, 0)] = flux_dn[L(i, 0)] * albedo[L(i, 0)] + src[L(i, 0)];
        for (int il = 1; il <= nlay; ++il)
            for (int i = 0; i < ncol; ++i)
            {
                flux_dn[L(i, il)] = (tdif[L(i, il-1)] * flux_dn[L(i, il-1)] + rdif[L(i, il-1)] * src[L(i, il)] +
                                     src_dn[L(i, il-1)]) * denom[L(i, il-1)];
                flux_up[L(i, il)] = flux_dn[L(i, il)] * albedo[L(i, il)] + src[L(i, il)];
            }
    }
    else
    {
        for (int i = 0; i < ncol; ++i) { albedo[L(i, 0)] = albedo_sfc[i]; src[L(i, 0)] = src_sfc[i]; }
        for (int il = 0; il < nlay; ++il)
            for (int i = 0; i < ncol; ++i)
            {
                denom[L(i, il)] = 1. / (1. - rdif[L(i, il)] * albedo[L(i, il)]);
                albedo[L(i, il+1)] = rdif[L(i, il)] + tdif[L(i, il)] * tdif[L(i, il)] * albedo[L(i, il)] * denom[L(i, il)];
                src[L(i, il+1)] = src_up[L(i, il)] + tdif[L(i, il)] * denom[L(i, il)] *
                                  (src[L(i, il)] + albedo[L(i, il)] * src_dn[L(i, il)]);
            }
        for (int i = 0; i < ncol; ++i) flux_up[L(i, nlay)] = flux_dn[L(i, nlay)] * albedo[L(i, nlay)] + src[L(i, nlay)];
        for (int il = nlay - 1; il >= 0; --il)
            for (int i = 0; i < ncol; ++i)
            {
                flux_dn[L(i, il)] = (tdif[L(i, il)] * flux_dn[L(i, il+1)] + rdif[L(i, il)] * src[L(i, il)] +
                                     src_dn[L(i, il)]) * denom[L(i, il)];
                flux_up[L(i, il)] = flux_dn[L(i, il)] * albedo[L(i, il)] + src[L(i, il)];
            }
    }
}

// lw_solver_2stream (F90:386-453): per g-point, lw_combine_sources (F90:896-920), lw_two_stream
// (F90:832-887, Fu et al. 1997 coefficients with the diffusivity secant 1.66), lw_source_2str
// (F90:928-978, Toon et al. 1989 Eqs. 26-27; layer sources zero where tau <= 1e-8), then the
// adding method with surface albedo 1 - sfc_emis and flux_dn(top) = inc_flux.  sfc_emis, sfc_src
// and inc_flux are (ncol, ngpt); fluxes (ncol, nlay+1, ngpt).  lay_source is an argument of the
// reference but no part of its arithmetic.
void lw_solver_2stream(int ncol, int nlay, int ngpt, bool top_at_1, const double* tau, const double* ssa,
                       const double* g, const double* lev_source_inc, const double* lev_source_dec,
                       const double* sfc_emis, const double* sfc_src, const double* inc_flux, double* flux_up,
                       double* flux_dn)
{
    const double k_min = 1.e-12;
    const double lw_diff_sec = 1.66;
    const int nlev = nlay + 1;
    const size_t ncl = size_t(ncol) * nlay, ncv = size_t(ncol) * nlev;
    auto L = [ncol](int i, int k) { return i + size_t(ncol)*k; };
    std::vector<double> lev(ncv), rdif(ncl), tdif(ncl), g1(ncl), g2(ncl), sdn(ncl), sup(ncl), ssfc(ncol), alb(ncol);
    for (int igpt = 0; igpt < ngpt; ++igpt)
    {
        const double* inc = lev_source_inc + ncl*igpt;
        const double* dec = lev_source_dec + ncl*igpt;
        const double* tg = tau + ncl*igpt;
        for (int i = 0; i < ncol; ++i)
        {
            lev[L(i, 0)] = dec[L(i, 0)];
            lev[L(i, nlay)] = inc[L(i, nlay-1)];
        }
        for (int il = 1; il < nlay; ++il)
            for (int i = 0; i < ncol; ++i)
                lev[L(i, il)] = std::sqrt(dec[L(i, il)] * inc[L(i, il-1)]);
        for (size_t j = 0; j < ncl; ++j)
        {
            const double w0 = ssa[j + ncl*igpt], gg = g[j + ncl*igpt];
            g1[j] = lw_diff_sec * (1. - 0.5 * w0 * (1. + gg));
            g2[j] = lw_diff_sec * 0.5 * w0 * (1. - gg);
            const double k = std::sqrt(std::max((g1[j] - g2[j]) * (g1[j] + g2[j]), k_min));
            const double e = std::exp(-tg[j] * k);
            const double e2 = e * e;
            const double rt = 1. / (k * (1. + e2) + g1[j] * (1. - e2));
            rdif[j] = rt * g2[j] * (1. - e2);
            tdif[j] = rt * 2. * k * e;
        }
        for (int il = 0; il < nlay; ++il)
            for (int i = 0; i < ncol; ++i)
            {
                const size_t j = L(i, il);
                const double top = top_at_1 ? lev[L(i, il)] : lev[L(i, il+1)];
                const double bot = top_at_1 ? lev[L(i, il+1)] : lev[L(i, il)];
                if (tg[j] > 1.e-8)
                {
                    const double Z = (bot - top) / (tg[j] * (g1[j] + g2[j]));
                    const double zup_top = Z + top, zup_bot = Z + bot, zdn_top = -Z + top, zdn_bot = -Z + bot;
                    sup[j] = pi * (zup_top - rdif[j] * zdn_top - tdif[j] * zup_bot);
                    sdn[j] = pi * (zdn_bot - rdif[j] * zup_bot - tdif[j] * zdn_top);
                }
                else
                {
                    sup[j] = 0.;
                    sdn[j] = 0.;
                }
            }
        for (int i = 0; i < ncol; ++i)
        {
            ssfc[i] = pi * sfc_emis[i + size_t(ncol)*igpt] * sfc_src[i + size_t(ncol)*igpt];
            alb[i] = 1. - sfc_emis[i + size_t(ncol)*igpt];
        }
        double* up = flux_up + ncv*igpt;
        double* dn = flux_dn + ncv*igpt;
        const int top_level = top_at_1 ? 0 : nlay;
        for (int i = 0; i < ncol; ++i) dn[L(i, top_level)] = inc_flux ? inc_flux[i + size_t(ncol)*igpt] : 0.;
        adding(ncol, nlay, top_at_1, alb.data(), rdif.data(), tdif.data(), sdn.data(), sup.data(), ssfc.data(),
               up, dn);
    }
}

// Rte_sw::rte_sw (src/Rte_sw.cpp:97-145) + sw_solver_2stream (F90:512-605).
void rte_sw(int ncol, int nlay, int ngpt, int nbnd, const int* band_lims, bool top_at_1,
            const double* tau, const double* ssa, const double* g, const double* mu0,
            const double* inc_flux_dir, const double* alb_dir_bnd, const double* alb_dif_bnd,
            const double* inc_flux_dif, bool do_broadband, double* flux_up, double* flux_dn,
            double* flux_dir)
{
    const int nlev = nlay + 1;
    const size_t ncl = size_t(ncol) * nlay, ncv = size_t(ncol) * nlev;
    std::vector<double> adir(size_t(ncol) * ngpt), adif(size_t(ncol) * ngpt);
    for (int ib = 0; ib < nbnd; ++ib)
        for (int icol = 0; icol < ncol; ++icol)
            for (int ig = band_lims[2*ib] - 1; ig < band_lims[2*ib+1]; ++ig)
            {
                adir[icol + size_t(ncol)*ig] = alb_dir_bnd[ib + size_t(nbnd)*icol];
                adif[icol + size_t(ncol)*ig] = alb_dif_bnd[ib + size_t(nbnd)*icol];
            }
    const int top_level = top_at_1 ? 0 : nlay;
    std::vector<double> Rdif(ncl), Tdif(ncl), sup(ncl), sdn(ncl), ssfc(ncol);
    std::vector<double> lup(ncv), ldn(ncv), ldir(ncv);
    if (do_broadband)
    {
        std::fill(flux_up, flux_up + ncv, 0.); std::fill(flux_dn, flux_dn + ncv, 0.);
        std::fill(flux_dir, flux_dir + ncv, 0.);
    }
    for (int igpt = 0; igpt < ngpt; ++igpt)
    {
        double* gup = do_broadband ? lup.data() : flux_up + ncv*igpt;
        double* gdn = do_broadband ? ldn.data() : flux_dn + ncv*igpt;
        double* gdir = do_broadband ? ldir.data() : flux_dir + ncv*igpt;
        for (int i = 0; i < ncol; ++i)
        {
            gdir[i + size_t(ncol)*top_level] = inc_flux_dir[i + size_t(ncol)*igpt] * mu0[i];
            gdn[i + size_t(ncol)*top_level] = inc_flux_dif ? inc_flux_dif[i + size_t(ncol)*igpt] : 0.;
        }
        sw_dif_and_source(ncol, nlay, top_at_1, mu0, &adir[size_t(ncol)*igpt], tau + ncl*igpt,
                          ssa + ncl*igpt, g + ncl*igpt, Rdif.data(), Tdif.data(), sdn.data(),
                          sup.data(), ssfc.data(), gdir);
        adding(ncol, nlay, top_at_1, &adif[size_t(ncol)*igpt], Rdif.data(), Tdif.data(), sdn.data(),
               sup.data(), ssfc.data(), gup, gdn);
        if (do_broadband)
            for (size_t i = 0; i < ncv; ++i)
            {
                flux_up[i] = flux_up[i] + gup[i];
                flux_dn[i] = flux_dn[i] + gdn[i] + gdir[i];
                flux_dir[i] = flux_dir[i] + gdir[i];
            }
        else
            for (size_t i = 0; i < ncv; ++i) gdn[i] = gdn[i] + gdir[i];
    }
}

// ---------------------------------------------------------------------------------------------
// Fluxes.

// rte_sum_broadband (called at src/Fluxes.cpp:39-45): sum over g-points in order.
void sum_broadband(int ncol, int nlev, int ngpt, const double* gf, double* f)
{
    const size_t ncv = size_t(ncol) * nlev;
    for (size_t i = 0; i < ncv; ++i)
    {
        double s = 0.;
        for (int g = 0; g < ngpt; ++g) s += gf[i + ncv*g];
        f[i] = s;
    }
}

// sum_byband (mo_fluxes_byband_kernels.F90:22-41), 1-based inclusive band_lims.
void sum_byband(int ncol, int nlev, int nbnd, const int* band_lims, const double* gf, double* bf)
{
    const size_t ncv = size_t(ncol) * nlev;
    for (int b = 0; b < nbnd; ++b)
        for (size_t i = 0; i < ncv; ++i)
        {
            double s = gf[i + ncv*(band_lims[2*b] - 1)];
            for (int g = band_lims[2*b]; g < band_lims[2*b+1]; ++g) s += gf[i + ncv*g];
            bf[i + ncv*b] = s;
        }
}

// Heating rate (no reference counterpart): dT/dt = g0/cp * (Fnet(k) - Fnet(k+1)) / (p(k+1) - p(k))
// in K/day, with net = dn - up; valid for either vertical orientation.
void heating_rate(int ncol, int nlay, const double* flux_net, const double* plev, double* hr)
{
    const double g0 = 9.80665, cp = 1004.64, sec_per_day = 86400.;
    for (int il = 0; il < nlay; ++il)
        for (int i = 0; i < ncol; ++i)
        {
            const size_t a = i + size_t(ncol)*il, b = a + ncol;
            hr[a] = g0 / cp * (flux_net[a] - flux_net[b]) / (plev[b] - plev[a]) * sec_per_day;
        }
}

template<typename F>
int guarded(F&& f)
{
    try { f(); return 0; }
    catch (const std::range_error& e) { last_error = e.what(); return 2; }
    catch (const std::exception& e) { last_error = e.what(); return 1; }
}

// Threads of the column-block loops of oracle_solve_lw/sw.  1 (the default) runs the blocks in
// order on the calling thread, as the reference's driver does; more spreads independent blocks
// over OpenMP threads (a parallel variant of the CPU baseline, not a reference behaviour).
int g_threads = 1;

template<typename F>
void for_col_blocks(int ncol, int n_col_block, F&& body)
{
    const int nblocks = (ncol + n_col_block - 1) / n_col_block;
    if (g_threads <= 1)
    {
        for (int b = 0; b < nblocks; ++b) body(b * n_col_block);
        return;
    }
    std::exception_ptr first;
    #pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
    for (int b = 0; b < nblocks; ++b)
    {
        try { body(b * n_col_block); }
        catch (...)
        {
            #pragma omp critical(oracle_first_error)
            { if (!first) first = std::current_exception(); }
        }
    }
    if (first) std::rethrow_exception(first);
}
}  // namespace


// =============================================================================================
// C ABI (ctypes).

extern "C" {

const char* oracle_last_error() { return last_error.c_str(); }

// Sets the thread count of oracle_solve_lw/sw (values < 1 mean 1); returns the previous one.
int oracle_set_threads(int n)
{
    const int prev = g_threads;
    g_threads = n < 1 ? 1 : n;
    return prev;
}

void* oracle_kdist_create(const OracleKdistFile* f)
{
    try { return make_kdist(*f); }
    catch (const std::exception& e) { last_error = e.what(); return nullptr; }
}

void oracle_kdist_destroy(void* h) { delete static_cast<Kdist*>(h); }

// info: ngas, nflav, neta, npres, ntemp, ngpt, nbnd, nplanck, nminor_lower, nminor_upper,
//       ncontrib_lower, ncontrib_upper, is_lw, idx_h2o
void oracle_kdist_info(void* h, int* info)
{
    const Kdist& k = *static_cast<Kdist*>(h);
    const int v[] = {k.ngas, k.nflav, k.neta, k.npres, k.ntemp, k.ngpt, k.nbnd, k.nplanck,
                     k.lower.n, k.upper.n, k.lower.ncontrib, k.upper.ncontrib, k.is_lw, k.idx_h2o};
    std::copy(v, v + 14, info);
}

// Derived integer tables, for cross-checking the product's constructor.
int oracle_kdist_get_int(void* h, const char* name, int* out)
{
    const Kdist& k = *static_cast<Kdist*>(h);
    const std::string n(name);
    const std::vector<int>* v = nullptr;
    std::vector<int> tmp;
    if (n == "flavor") v = &k.flavor;
    else if (n == "gpoint_flavor") v = &k.gpoint_flavor;
    else if (n == "gpt2band") v = &k.gpt2band;
    else if (n == "is_key") v = &k.is_key;
    else if (n == "minor_limits_gpt_lower") v = &k.lower.limits;
    else if (n == "minor_limits_gpt_upper") v = &k.upper.limits;
    else if (n == "kminor_start_lower") v = &k.lower.kminor_start;
    else if (n == "kminor_start_upper") v = &k.upper.kminor_start;
    else if (n == "idx_minor_lower") v = &k.lower.idx_minor;
    else if (n == "idx_minor_upper") v = &k.upper.idx_minor;
    else if (n == "idx_minor_scaling_lower") v = &k.lower.idx_minor_scaling;
    else if (n == "idx_minor_scaling_upper") v = &k.upper.idx_minor_scaling;
    else if (n == "minor_scales_with_density_lower") { tmp.assign(k.lower.scales_with_density.begin(), k.lower.scales_with_density.end()); v = &tmp; }
    else if (n == "minor_scales_with_density_upper") { tmp.assign(k.upper.scales_with_density.begin(), k.upper.scales_with_density.end()); v = &tmp; }
    else if (n == "scale_by_complement_lower") { tmp.assign(k.lower.scale_by_complement.begin(), k.lower.scale_by_complement.end()); v = &tmp; }
    else if (n == "scale_by_complement_upper") { tmp.assign(k.upper.scale_by_complement.begin(), k.upper.scale_by_complement.end()); v = &tmp; }
    if (!v) return -1;
    std::copy(v->begin(), v->end(), out);
    return int(v->size());
}

int oracle_kdist_get_real(void* h, const char* name, double* out)
{
    const Kdist& k = *static_cast<Kdist*>(h);
    const std::string n(name);
    const std::vector<double>* v = nullptr;
    if (n == "kmajor") v = &k.kmajor;
    else if (n == "kminor_lower") v = &k.lower.kminor;
    else if (n == "kminor_upper") v = &k.upper.kminor;
    else if (n == "vmr_ref") v = &k.vmr_ref;
    else if (n == "planck_frac") v = &k.planck_frac;
    else if (n == "totplnk") v = &k.totplnk;
    else if (n == "krayl") v = &k.krayl;
    else if (n == "solar_source") v = &k.solar_source;
    else if (n == "press_ref_log") v = &k.press_ref_log;
    else if (n == "temp_ref") v = &k.temp_ref;
    if (v)
    {
        std::copy(v->begin(), v->end(), out);
        return int(v->size());
    }
    double s;
    if (n == "press_ref_log_delta") s = k.press_ref_log_delta;
    else if (n == "temp_ref_delta") s = k.temp_ref_delta;
    else if (n == "temp_ref_min") s = k.temp_ref_min;
    else if (n == "temp_ref_max") s = k.temp_ref_max;
    else if (n == "press_ref_min") s = k.press_ref_min;
    else if (n == "press_ref_max") s = k.press_ref_max;
    else if (n == "press_ref_trop_log") s = k.press_ref_trop_log;
    else if (n == "totplnk_delta") s = k.totplnk_delta;
    else if (n == "tsi")
    {
        s = 0.;
        for (double x : k.solar_source) s += x;   // get_tsi (:1347-1356)
    }
    else return -1;
    out[0] = s;
    return 1;
}

// Gas names of the reduced k-distribution, '\n'-joined into buf.
int oracle_kdist_gas_names(void* h, char* buf, int len)
{
    const Kdist& k = *static_cast<Kdist*>(h);
    std::string s;
    for (size_t i = 0; i < k.gas_names.size(); ++i) { if (i) s += "\n"; s += k.gas_names[i]; }
    if (int(s.size()) + 1 > len) return -1;
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return int(k.gas_names.size());
}

// Gas_optics_rrtmgp::get_col_dry (src/Gas_optics_rrtmgp.cpp:764-792).
void oracle_get_col_dry(int ncol, int nlay, const double* vmr_h2o, const double* plev, double* col_dry)
{
    const double g0 = 9.80665, avogad = 6.02214076e23, m_dry = 0.028964, m_h2o = 0.018016;
    for (int il = 0; il < nlay; ++il)
        for (int i = 0; i < ncol; ++i)
        {
            const size_t a = i + size_t(ncol)*il;
            const double dp = std::abs(plev[a] - plev[a + ncol]);
            const double m_air = (m_dry + m_h2o * vmr_h2o[a]) / (1. + vmr_h2o[a]);
            double c = 10. * dp * avogad / (1000. * m_air * 100. * g0);
            c /= (1. + vmr_h2o[a]);
            col_dry[a] = c;
        }
}

// Full gas optics.  LW: tau + sources (ssa, g, toa_src unused).  SW: tau, ssa, g, toa_src.
// Interpolation outputs are optional (NULL to skip).
int oracle_gas_optics(void* h, int ncol, int nlay, const double* play, const double* plev,
                      const double* tlay, const double* tsfc, const double* tlev,
                      const OracleGasConcs* gc, const double* col_dry, double* tau, double* ssa,
                      double* g, double* lay_src, double* lev_inc, double* lev_dec, double* sfc_src,
                      double* sfc_src_jac, double* toa_src, int* jtemp, int* jpress, Bool* tropo,
                      int* jeta, double* col_mix, double* fmajor, double* fminor)
{
    const Kdist& k = *static_cast<Kdist*>(h);
    return guarded([&] {
        Interp in;
        gas_optics(k, ncol, nlay, play, plev, tlay, tsfc, tlev, *gc, col_dry, tau, ssa, g, lay_src,
                   lev_inc, lev_dec, sfc_src, sfc_src_jac, toa_src, jtemp ? &in : nullptr);
        if (jtemp)
        {
            std::copy(in.jtemp.begin(), in.jtemp.end(), jtemp);
            std::copy(in.jpress.begin(), in.jpress.end(), jpress);
            std::copy(in.tropo.begin(), in.tropo.end(), tropo);
            std::copy(in.jeta.begin(), in.jeta.end(), jeta);
            std::copy(in.col_mix.begin(), in.col_mix.end(), col_mix);
            std::copy(in.fmajor.begin(), in.fmajor.end(), fmajor);
            std::copy(in.fminor.begin(), in.fminor.end(), fminor);
        }
    });
}

void* oracle_cloud_create(int nbnd, double radliq_lwr, double radliq_upr, double radice_lwr,
                          double radice_upr, int nsize_liq, const double* extliq,
                          const double* ssaliq, const double* asyliq, int nsize_ice,
                          const double* extice, const double* ssaice, const double* asyice)
{
    // Cloud_optics ctor (src/Cloud_optics.cpp:29-69); ice roughness icergh = 2 (:61-68).
    Cloud* c = new Cloud;
    c->nbnd = nbnd;
    c->liq_nsteps = nsize_liq; c->ice_nsteps = nsize_ice;
    c->liq_step_size = (radliq_upr - radliq_lwr) / (nsize_liq - 1.);
    c->ice_step_size = (radice_upr - radice_lwr) / (nsize_ice - 1.);
    c->radliq_lwr = radliq_lwr; c->radliq_upr = radliq_upr;
    c->radice_lwr = radice_lwr; c->radice_upr = radice_upr;
    const size_t nl = size_t(nsize_liq) * nbnd, ni = size_t(nsize_ice) * nbnd;
    c->extliq.assign(extliq, extliq + nl); c->ssaliq.assign(ssaliq, ssaliq + nl); c->asyliq.assign(asyliq, asyliq + nl);
    const int icergh = 2;
    const size_t off = size_t(icergh - 1) * ni;
    c->extice.assign(extice + off, extice + off + ni);
    c->ssaice.assign(ssaice + off, ssaice + off + ni);
    c->asyice.assign(asyice + off, asyice + off + ni);
    return c;
}

void oracle_cloud_destroy(void* h) { delete static_cast<Cloud*>(h); }

int oracle_cloud_optics(void* h, int ncol, int nlay, const double* lwp, const double* iwp,
                        const double* rel, const double* rei, int two_stream, double* tau,
                        double* ssa, double* g)
{
    return guarded([&] { cloud_optics(*static_cast<Cloud*>(h), ncol, nlay, lwp, iwp, rel, rei,
                                      two_stream != 0, tau, ssa, g); });
}

// Aerosol_optics::aerosol_optics (src/Aerosol_optics.cpp:168-224); aermr holds the 11 (ncol, nlay)
// fields aermr01..aermr11 back to back.
int oracle_aerosol_optics(int ncol, int nlay, int nbnd, int nhum, int nphobic, int nphilic, const double* aermr,
                          const double* rh, const double* plev, const double* rh_upper,
                          const double* mext_phobic, const double* ssa_phobic, const double* g_phobic,
                          const double* mext_philic, const double* ssa_philic, const double* g_philic,
                          double* tau, double* ssa, double* g)
{
    return guarded([&] {
        const double* fields[11];
        for (int k = 0; k < 11; ++k) fields[k] = aermr + size_t(k) * ncol * nlay;
        aerosol_optics(ncol, nlay, nbnd, nhum, nphobic, nphilic, fields, rh, plev, rh_upper, mext_phobic,
                       ssa_phobic, g_phobic, mext_philic, ssa_philic, g_philic, tau, ssa, g);
    });
}

void oracle_delta_scale_2str(int ncol, int nlay, int ngpt, double* tau, double* ssa, double* g)
{
    delta_scale_2str(size_t(ncol) * nlay * ngpt, tau, ssa, g);
}

// ssa1 == NULL selects the 1scl variants.
void oracle_inc_bybnd(int ncol, int nlay, int ngpt, double* tau1, double* ssa1, double* g1,
                      const double* tau2, const double* ssa2, const double* g2, int nbnd,
                      const int* band_lims_gpt)
{
    inc_bybnd(ncol, nlay, ngpt, tau1, ssa1, g1, tau2, ssa2, g2, nbnd, band_lims_gpt);
}

void oracle_increment(int ncol, int nlay, int ngpt, double* tau1, double* ssa1, double* g1,
                      const double* tau2, const double* ssa2, const double* g2)
{
    increment(size_t(ncol) * nlay * ngpt, tau1, ssa1, g1, tau2, ssa2, g2);
}

int oracle_rte_lw(int ncol, int nlay, int ngpt, int nbnd, const int* band_lims, int top_at_1,
                  const double* tau, const double* lay_src, const double* lev_inc,
                  const double* lev_dec, const double* sfc_src, const double* sfc_emis,
                  const double* inc_flux, int n_gauss_angles, int do_broadband, double* flux_up,
                  double* flux_dn, const double* sfc_src_jac, double* flux_up_jac)
{
    return guarded([&] {
        rte_lw(ncol, nlay, ngpt, nbnd, band_lims, top_at_1 != 0, tau, lay_src, lev_inc, lev_dec,
               sfc_src, sfc_emis, inc_flux, n_gauss_angles, do_broadband != 0, flux_up, flux_dn,
               sfc_src_jac, flux_up_jac);
    });
}

// rte_lw_solver_noscat (include/rrtmgp_kernels.h:150-158) with do_rescaling = true: Gauss
// quadrature over n_gauss_angles as oracle_rte_lw, with the Tang et al. rescaling by ssa, g.
int oracle_rte_lw_rescaled(int ncol, int nlay, int ngpt, int nbnd, const int* band_lims, int top_at_1,
                           const double* tau, const double* ssa, const double* g, const double* lay_src,
                           const double* lev_inc, const double* lev_dec, const double* sfc_src,
                           const double* sfc_emis, const double* inc_flux, int n_gauss_angles, int do_broadband,
                           double* flux_up, double* flux_dn, const double* sfc_src_jac, double* flux_up_jac)
{
    return guarded([&] {
        rte_lw(ncol, nlay, ngpt, nbnd, band_lims, top_at_1 != 0, tau, lay_src, lev_inc, lev_dec,
               sfc_src, sfc_emis, inc_flux, n_gauss_angles, do_broadband != 0, flux_up, flux_dn,
               sfc_src_jac, flux_up_jac, true, ssa, g);
    });
}

// lw_solver_2stream (F90:386-453), spectral fluxes (ncol, nlay+1, ngpt); sfc_emis, sfc_src and
// inc_flux (NULL = 0) per (ncol, ngpt) as the Fortran kernel takes them.
int oracle_lw_solver_2stream(int ncol, int nlay, int ngpt, int top_at_1, const double* tau, const double* ssa,
                             const double* g, const double* lev_inc, const double* lev_dec,
                             const double* sfc_emis, const double* sfc_src, const double* inc_flux,
                             double* flux_up, double* flux_dn)
{
    return guarded([&] {
        lw_solver_2stream(ncol, nlay, ngpt, top_at_1 != 0, tau, ssa, g, lev_inc, lev_dec, sfc_emis, sfc_src,
                          inc_flux, flux_up, flux_dn);
    });
}

int oracle_rte_sw(int ncol, int nlay, int ngpt, int nbnd, const int* band_lims, int top_at_1,
                  const double* tau, const double* ssa, const double* g, const double* mu0,
                  const double* inc_flux_dir, const double* sfc_alb_dir, const double* sfc_alb_dif,
                  const double* inc_flux_dif, int do_broadband, double* flux_up, double* flux_dn,
                  double* flux_dir)
{
    return guarded([&] {
        rte_sw(ncol, nlay, ngpt, nbnd, band_lims, top_at_1 != 0, tau, ssa, g, mu0, inc_flux_dir,
               sfc_alb_dir, sfc_alb_dif, inc_flux_dif, do_broadband != 0, flux_up, flux_dn, flux_dir);
    });
}

// Two-stream layer properties of sw_dif_and_source (one layer per element), for pinning tests.
void oracle_sw_two_stream(int n, const double* tau, const double* w0, const double* g,
                          const double* mu0, double* Rdif, double* Tdif, double* Rdir,
                          double* Tdir, double* Tnoscat)
{
    for (int i = 0; i < n; ++i)
    {
        double dir[2] = {1.0, 0.0}, sup, sdn, ssfc, alb = 0.;
        sw_dif_and_source(1, 1, true, &mu0[i], &alb, &tau[i], &w0[i], &g[i], &Rdif[i], &Tdif[i],
                          &sdn, &sup, &ssfc, dir);
        Rdir[i] = sup; Tdir[i] = sdn; Tnoscat[i] = dir[1];
    }
}

void oracle_sum_broadband(int ncol, int nlev, int ngpt, const double* gf, double* f)
{
    sum_broadband(ncol, nlev, ngpt, gf, f);
}

void oracle_sum_byband(int ncol, int nlev, int nbnd, const int* band_lims, const double* gf, double* bf)
{
    sum_byband(ncol, nlev, nbnd, band_lims, gf, bf);
}

void oracle_heating_rate(int ncol, int nlay, const double* flux_net, const double* plev, double* hr)
{
    heating_rate(ncol, nlay, flux_net, plev, hr);
}

// Radiation_solver_longwave::solve (src_test/Radiation_solver.cpp:384-617): 12-column blocks,
// every input subset-copied per block, gas optics -> cloud optics 1scl -> add_to -> rte_lw with
// one angle and zero incident flux; broadband fluxes are accumulated inside the solver unless
// band fluxes are requested.  cloud may be NULL (clear sky).  bnd_* may be NULL.
int oracle_solve_lw(void* kd_h, void* cloud_h, int ncol, int nlay, const double* play,
                    const double* plev, const double* tlay, const double* tlev,
                    const double* tsfc, const OracleGasConcs* gc, const double* col_dry_in,
                    const double* emis_sfc, const double* lwp, const double* iwp,
                    const double* rel, const double* rei, double* flux_up, double* flux_dn,
                    double* flux_net, double* bnd_up, double* bnd_dn, double* bnd_net,
                    int n_col_block)
{
    const Kdist& kd = *static_cast<Kdist*>(kd_h);
    const Cloud* cl = static_cast<Cloud*>(cloud_h);
    return guarded([&] {
        const int nlev = nlay + 1, ngpt = kd.ngpt, nbnd = kd.nbnd;
        const bool top_at_1 = play[0] < play[size_t(ncol)*(nlay-1)];
        const bool out_bnd = bnd_up != nullptr;
        const auto gnames = split_names(gc->names, gc->n);
        if (n_col_block <= 0) n_col_block = 12;
        for_col_blocks(ncol, n_col_block, [&](const int cs)
        {
            const int n = std::min(n_col_block, ncol - cs);
            auto sub = [&](const double* a, int nz) {
                std::vector<double> o(size_t(n) * nz);
                for (int k = 0; k < nz; ++k)
                    for (int i = 0; i < n; ++i) o[i + size_t(n)*k] = a[cs + i + size_t(ncol)*k];
                return o;
            };
            // Gas_concs subset ctor (src/Gas_concs.cpp:29-42).
            std::vector<std::vector<double>> vstore;
            std::vector<const double*> vptr;
            std::vector<int> vdims;
            for (int i = 0; i < gc->n; ++i)
            {
                const int d1 = gc->dims[2*i], d2 = gc->dims[2*i+1];
                if (d1 == 1) vstore.emplace_back(gc->data[i], gc->data[i] + d2);
                else vstore.push_back(sub(gc->data[i], d2));
                vdims.push_back(d1 == 1 ? 1 : n); vdims.push_back(d2);
            }
            for (auto& v : vstore) vptr.push_back(v.data());
            OracleGasConcs gsub{gc->n, gc->names, vptr.data(), vdims.data()};

            auto p_lay = sub(play, nlay), p_lev = sub(plev, nlev), t_lay = sub(tlay, nlay), t_lev = sub(tlev, nlev);
            std::vector<double> t_sfc(tsfc + cs, tsfc + cs + n);
            std::vector<double> col_dry(size_t(n) * nlay);
            if (col_dry_in) col_dry = sub(col_dry_in, nlay);
            else
            {
                const int k = find_index(gnames, "h2o");
                const int d1 = gc->dims[2*(k-1)];
                std::vector<double> h2o(size_t(n) * nlay);
                for (int il = 0; il < nlay; ++il)
                    for (int i = 0; i < n; ++i)
                        h2o[i + size_t(n)*il] = d1 == 1 ? vstore[k-1][gc->dims[2*(k-1)+1] == 1 ? 0 : il] : vstore[k-1][i + size_t(n)*il];
                oracle_get_col_dry(n, nlay, h2o.data(), p_lev.data(), col_dry.data());
            }
            const size_t ng = size_t(n) * nlay * ngpt;
            std::vector<double> tau(ng), lay(ng), linc(ng), ldec(ng), ssrc(size_t(n)*ngpt), sjac(size_t(n)*ngpt);
            gas_optics(kd, n, nlay, p_lay.data(), p_lev.data(), t_lay.data(), t_sfc.data(), t_lev.data(),
                       gsub, col_dry.data(), tau.data(), nullptr, nullptr, lay.data(), linc.data(), ldec.data(),
                       ssrc.data(), sjac.data(), nullptr, nullptr);
            if (cl)
            {
                std::vector<double> ctau(size_t(n) * nlay * cl->nbnd);
                cloud_optics(*cl, n, nlay, sub(lwp, nlay).data(), sub(iwp, nlay).data(), sub(rel, nlay).data(),
                             sub(rei, nlay).data(), false, ctau.data(), nullptr, nullptr);
                inc_bybnd(n, nlay, ngpt, tau.data(), nullptr, nullptr, ctau.data(), nullptr, nullptr, nbnd, kd.band2gpt.data());
            }
            std::vector<double> emis(size_t(nbnd) * n);
            for (int i = 0; i < n; ++i)
                for (int b = 0; b < nbnd; ++b) emis[b + size_t(nbnd)*i] = emis_sfc[b + size_t(nbnd)*(cs + i)];
            const size_t ncv = size_t(n) * nlev;
            std::vector<double> gup(out_bnd ? ncv*ngpt : ncv), gdn(out_bnd ? ncv*ngpt : ncv);
            rte_lw(n, nlay, ngpt, nbnd, kd.band2gpt.data(), top_at_1, tau.data(), lay.data(), linc.data(),
                   ldec.data(), ssrc.data(), emis.data(), nullptr, 1, !out_bnd, gup.data(), gdn.data(),
                   nullptr, nullptr);
            std::vector<double> fu(ncv), fd(ncv);
            if (out_bnd)
            {
                sum_broadband(n, nlev, ngpt, gup.data(), fu.data());
                sum_broadband(n, nlev, ngpt, gdn.data(), fd.data());
                std::vector<double> bu(ncv*nbnd), bd(ncv*nbnd);
                sum_byband(n, nlev, nbnd, kd.band2gpt.data(), gup.data(), bu.data());
                sum_byband(n, nlev, nbnd, kd.band2gpt.data(), gdn.data(), bd.data());
                for (int b = 0; b < nbnd; ++b)
                    for (int k = 0; k < nlev; ++k)
                        for (int i = 0; i < n; ++i)
                        {
                            const size_t o = cs + i + size_t(ncol)*(k + size_t(nlev)*b), s = i + size_t(n)*(k + size_t(nlev)*b);
                            bnd_up[o] = bu[s]; bnd_dn[o] = bd[s]; bnd_net[o] = bd[s] - bu[s];
                        }
            }
            else { fu = gup; fd = gdn; }
            for (int k = 0; k < nlev; ++k)
                for (int i = 0; i < n; ++i)
                {
                    const size_t o = cs + i + size_t(ncol)*k, s = i + size_t(n)*k;
                    flux_up[o] = fu[s]; flux_dn[o] = fd[s]; flux_net[o] = fd[s] - fu[s];
                }
        });
    });
}

// Radiation_solver_shortwave::solve (src_test/Radiation_solver.cpp:642-913): 12-column blocks,
// gas optics -> toa_src * tsi_scaling -> cloud optics 2str -> delta_scale (if delta_cloud) ->
// add_to -> rte_sw (no diffuse BC) -> broadband (or band) fluxes.
int oracle_solve_sw(void* kd_h, void* cloud_h, int delta_cloud, int ncol, int nlay,
                    const double* play, const double* plev, const double* tlay,
                    const OracleGasConcs* gc, const double* col_dry_in, const double* sfc_alb_dir,
                    const double* sfc_alb_dif, const double* tsi_scaling, const double* mu0,
                    const double* lwp, const double* iwp, const double* rel, const double* rei,
                    double* flux_up, double* flux_dn, double* flux_dn_dir, double* flux_net,
                    double* bnd_up, double* bnd_dn, double* bnd_dn_dir, double* bnd_net,
                    int n_col_block)
{
    const Kdist& kd = *static_cast<Kdist*>(kd_h);
    const Cloud* cl = static_cast<Cloud*>(cloud_h);
    return guarded([&] {
        const int nlev = nlay + 1, ngpt = kd.ngpt, nbnd = kd.nbnd;
        const bool top_at_1 = play[0] < play[size_t(ncol)*(nlay-1)];
        const bool out_bnd = bnd_up != nullptr;
        const auto gnames = split_names(gc->names, gc->n);
        if (n_col_block <= 0) n_col_block = 12;
        for_col_blocks(ncol, n_col_block, [&](const int cs)
        {
            const int n = std::min(n_col_block, ncol - cs);
            auto sub = [&](const double* a, int nz) {
                std::vector<double> o(size_t(n) * nz);
                for (int k = 0; k < nz; ++k)
                    for (int i = 0; i < n; ++i) o[i + size_t(n)*k] = a[cs + i + size_t(ncol)*k];
                return o;
            };
            std::vector<std::vector<double>> vstore;
            std::vector<const double*> vptr;
            std::vector<int> vdims;
            for (int i = 0; i < gc->n; ++i)
            {
                const int d1 = gc->dims[2*i], d2 = gc->dims[2*i+1];
                if (d1 == 1) vstore.emplace_back(gc->data[i], gc->data[i] + d2);
                else vstore.push_back(sub(gc->data[i], d2));
                vdims.push_back(d1 == 1 ? 1 : n); vdims.push_back(d2);
            }
            for (auto& v : vstore) vptr.push_back(v.data());
            OracleGasConcs gsub{gc->n, gc->names, vptr.data(), vdims.data()};
            auto p_lay = sub(play, nlay), p_lev = sub(plev, nlev), t_lay = sub(tlay, nlay);
            std::vector<double> col_dry(size_t(n) * nlay);
            if (col_dry_in) col_dry = sub(col_dry_in, nlay);
            else
            {
                const int k = find_index(gnames, "h2o");
                const int d1 = gc->dims[2*(k-1)];
                std::vector<double> h2o(size_t(n) * nlay);
                for (int il = 0; il < nlay; ++il)
                    for (int i = 0; i < n; ++i)
                        h2o[i + size_t(n)*il] = d1 == 1 ? vstore[k-1][gc->dims[2*(k-1)+1] == 1 ? 0 : il] : vstore[k-1][i + size_t(n)*il];
                oracle_get_col_dry(n, nlay, h2o.data(), p_lev.data(), col_dry.data());
            }
            const size_t ng = size_t(n) * nlay * ngpt;
            std::vector<double> tau(ng), ssa(ng), g(ng), toa(size_t(n) * ngpt);
            gas_optics(kd, n, nlay, p_lay.data(), p_lev.data(), t_lay.data(), nullptr, nullptr, gsub,
                       col_dry.data(), tau.data(), ssa.data(), g.data(), nullptr, nullptr, nullptr,
                       nullptr, nullptr, toa.data(), nullptr);
            for (int ig = 0; ig < ngpt; ++ig)
                for (int i = 0; i < n; ++i) toa[i + size_t(n)*ig] *= tsi_scaling[cs + i];
            if (cl)
            {
                const size_t nb = size_t(n) * nlay * cl->nbnd;
                std::vector<double> ct(nb), cs_(nb), cg(nb);
                cloud_optics(*cl, n, nlay, sub(lwp, nlay).data(), sub(iwp, nlay).data(), sub(rel, nlay).data(),
                             sub(rei, nlay).data(), true, ct.data(), cs_.data(), cg.data());
                if (delta_cloud) delta_scale_2str(nb, ct.data(), cs_.data(), cg.data());
                inc_bybnd(n, nlay, ngpt, tau.data(), ssa.data(), g.data(), ct.data(), cs_.data(), cg.data(),
                          nbnd, kd.band2gpt.data());
            }
            std::vector<double> adir(size_t(nbnd) * n), adif(size_t(nbnd) * n), m0(mu0 + cs, mu0 + cs + n);
            for (int i = 0; i < n; ++i)
                for (int b = 0; b < nbnd; ++b)
                {
                    adir[b + size_t(nbnd)*i] = sfc_alb_dir[b + size_t(nbnd)*(cs + i)];
                    adif[b + size_t(nbnd)*i] = sfc_alb_dif[b + size_t(nbnd)*(cs + i)];
                }
            const size_t ncv = size_t(n) * nlev;
            const size_t no = out_bnd ? ncv*ngpt : ncv;
            std::vector<double> gup(no), gdn(no), gdir(no);
            rte_sw(n, nlay, ngpt, nbnd, kd.band2gpt.data(), top_at_1, tau.data(), ssa.data(), g.data(),
                   m0.data(), toa.data(), adir.data(), adif.data(), nullptr, !out_bnd, gup.data(),
                   gdn.data(), gdir.data());
            std::vector<double> fu(ncv), fd(ncv), fdir(ncv);
            if (out_bnd)
            {
                sum_broadband(n, nlev, ngpt, gup.data(), fu.data());
                sum_broadband(n, nlev, ngpt, gdn.data(), fd.data());
                sum_broadband(n, nlev, ngpt, gdir.data(), fdir.data());
                std::vector<double> bu(ncv*nbnd), bd(ncv*nbnd), bdir(ncv*nbnd);
                sum_byband(n, nlev, nbnd, kd.band2gpt.data(), gup.data(), bu.data());
                sum_byband(n, nlev, nbnd, kd.band2gpt.data(), gdn.data(), bd.data());
                sum_byband(n, nlev, nbnd, kd.band2gpt.data(), gdir.data(), bdir.data());
                for (int b = 0; b < nbnd; ++b)
                    for (int k = 0; k < nlev; ++k)
                        for (int i = 0; i < n; ++i)
                        {
                            const size_t o = cs + i + size_t(ncol)*(k + size_t(nlev)*b), s = i + size_t(n)*(k + size_t(nlev)*b);
                            bnd_up[o] = bu[s]; bnd_dn[o] = bd[s]; bnd_dn_dir[o] = bdir[s]; bnd_net[o] = bd[s] - bu[s];
                        }
            }
            else { fu = gup; fd = gdn; fdir = gdir; }
            for (int k = 0; k < nlev; ++k)
                for (int i = 0; i < n; ++i)
                {
                    const size_t o = cs + i + size_t(ncol)*k, s = i + size_t(n)*k;
                    flux_up[o] = fu[s]; flux_dn[o] = fd[s]; flux_dn_dir[o] = fdir[s]; flux_net[o] = fd[s] - fu[s];
                }
        });
    });
}

}  // extern "C"
