// Gas optics kernels for gfx950.
//
// Arithmetic follows the reference restatement in src_kernels_cuda/gas_optics_rrtmgp_kernels.cu
// (interpolation :215-293, major :296-341, minor :356-476, Rayleigh :572-616, Planck :129-212,
// interpolate1D :1-12) with the CPU combine rule of src/Gas_optics_rrtmgp.cpp:365-385.
//
// Design.  In the fused kernels a work-group (4 wavefronts) computes one layer of 64 columns
// for one chunk of at most 16 g-points of a band.  Lane l of a wavefront takes g-point l % 16 of
// pair l / 16 of its current four (column, layer) pairs.  The k-tables are gathered from copies
// with the g-point index fastest, so the 16 lanes of a pair read one contiguous 64-byte segment
// per table corner (4 segments per wavefront gather).  A pair's interpolation state (indices and
// weights of the band's flavor) is computed by its 16 lanes together, one wavefront instruction
// per four pairs, and never written to memory.  Results pass through LDS (rows padded to 68
// floats: conflict-free for this lane layout) and leave as 256-byte column rows of the reference
// layout (ncol, nlay, ngpt).  Work-groups of one chunk are consecutive and kept on one XCD, so
// that XCD's L2 holds the chunk's table slice.
#include <algorithm>
#include <vector>

#include "common.h"

namespace rrtmgp_hip
{
constexpr int GO_THREADS = 256;   // work-group: 4 wavefronts
constexpr int GO_COLS = 64;       // columns of a work-group's layer
constexpr int GO_GPTS = 16;       // g-points per chunk: the lanes of one pair
constexpr int GO_ROW = 68;        // LDS row length (64 columns + padding)

template<typename Real>
struct Kd
{
    int ngas, nflav, neta, npres, ntemp, ngpt, nbnd, nplanck, idx_h2o;
    Real press_ref_log_delta, temp_ref_min, temp_ref_delta, press_ref_trop_log, totplnk_delta;
    int nminor[2];                 // [0] lower (tropo), [1] upper: indexed by itropo = !tropo
    const int* flavor;
    const int* gpoint_flavor;
    const int* band_lims_gpt;
    const Real* press_ref_log;
    const Real* temp_ref;
    const Real* vmr_ref;
    const Real* kmajor;
    const Real* kminor[2];
    const int* minor_limits[2];
    const Bool* dens[2];
    const Bool* compl_[2];
    const int* idx_minor[2];
    const int* idx_scal[2];
    const int* kstart[2];
    const int* band_start[2];
    const int* band_list[2];
    const Real* totplnk;
    const Real* planck_frac;
    const Real* krayl;
    // g-point-fastest copies and their strides
    const Real* kmajor_g;          // (ngpt, npres+1, neta, ntemp)
    const Real* planck_frac_g;
    const Real* kminor_g[2];       // (ncontrib, neta, ntemp)
    const Real* krayl_g;           // (ngpt, neta, ntemp, 2)
    int64_t ncontrib[2];
    int max_band_gpts;
};

template<typename Real>
Kd<Real> make_kd(const rrtmgp_kdist_dev& d)
{
    Kd<Real> k;
    k.ngas = int(d.ngas); k.nflav = int(d.nflav); k.neta = int(d.neta); k.npres = int(d.npres);
    k.ntemp = int(d.ntemp); k.ngpt = int(d.ngpt); k.nbnd = int(d.nbnd); k.nplanck = int(d.nplanck);
    k.idx_h2o = int(d.idx_h2o);
    k.press_ref_log_delta = Real(d.press_ref_log_delta);
    k.temp_ref_min = Real(d.temp_ref_min);
    k.temp_ref_delta = Real(d.temp_ref_delta);
    k.press_ref_trop_log = Real(d.press_ref_trop_log);
    k.totplnk_delta = Real(d.totplnk_delta);
    k.nminor[0] = int(d.nminor_lower); k.nminor[1] = int(d.nminor_upper);
    k.flavor = d.flavor; k.gpoint_flavor = d.gpoint_flavor; k.band_lims_gpt = d.band_lims_gpt;
    k.press_ref_log = static_cast<const Real*>(d.press_ref_log);
    k.temp_ref = static_cast<const Real*>(d.temp_ref);
    k.vmr_ref = static_cast<const Real*>(d.vmr_ref);
    k.kmajor = static_cast<const Real*>(d.kmajor);
    k.kminor[0] = static_cast<const Real*>(d.kminor_lower);
    k.kminor[1] = static_cast<const Real*>(d.kminor_upper);
    k.minor_limits[0] = d.minor_limits_gpt_lower; k.minor_limits[1] = d.minor_limits_gpt_upper;
    k.dens[0] = d.minor_scales_with_density_lower; k.dens[1] = d.minor_scales_with_density_upper;
    k.compl_[0] = d.scale_by_complement_lower; k.compl_[1] = d.scale_by_complement_upper;
    k.idx_minor[0] = d.idx_minor_lower; k.idx_minor[1] = d.idx_minor_upper;
    k.idx_scal[0] = d.idx_minor_scaling_lower; k.idx_scal[1] = d.idx_minor_scaling_upper;
    k.kstart[0] = d.kminor_start_lower; k.kstart[1] = d.kminor_start_upper;
    k.band_start[0] = d.minor_band_start_lower; k.band_start[1] = d.minor_band_start_upper;
    k.band_list[0] = d.minor_band_list_lower; k.band_list[1] = d.minor_band_list_upper;
    k.totplnk = static_cast<const Real*>(d.totplnk);
    k.planck_frac = static_cast<const Real*>(d.planck_frac);
    k.krayl = static_cast<const Real*>(d.krayl);
    k.kmajor_g = static_cast<const Real*>(d.kmajor_g);
    k.planck_frac_g = static_cast<const Real*>(d.planck_frac_g);
    k.kminor_g[0] = static_cast<const Real*>(d.kminor_lower_g);
    k.kminor_g[1] = static_cast<const Real*>(d.kminor_upper_g);
    k.krayl_g = static_cast<const Real*>(d.krayl_g);
    k.ncontrib[0] = d.ncontrib_lower; k.ncontrib[1] = d.ncontrib_upper;
    k.max_band_gpts = int(d.max_band_gpts);
    return k;
}

template<typename Real>
struct Gases
{
    const Real* p[RRTMGP_MAX_GAS];
    int64_t scol[RRTMGP_MAX_GAS];
    int64_t slay[RRTMGP_MAX_GAS];
};

template<typename Real>
int make_gases(const rrtmgp_gas_vmr* g, int64_t ngas_kd, int64_t ncol, int64_t nlay, Gases<Real>& out)
{
    if (!g) return arg_error("gas concentrations are NULL");
    if (g->ngas != ngas_kd)
        return arg_error("gas concentrations hold " + std::to_string(g->ngas) +
                         " gases, the k-distribution expects " + std::to_string(ngas_kd));
    for (int i = 0; i < g->ngas; ++i)
    {
        const int64_t d1 = g->dim1[i], d2 = g->dim2[i];
        if (!g->vmr[i]) return arg_error("vmr pointer of gas " + std::to_string(i + 1) + " is NULL");
        if (!((d1 == 1 || d1 == ncol) && (d2 == 1 || d2 == nlay)) || (d1 == ncol && ncol > 1 && d2 == 1))
            return arg_error("vmr of gas " + std::to_string(i + 1) + " has dims (" + std::to_string(d1) +
                             ", " + std::to_string(d2) + "); expected (1|ncol, 1|nlay)");
        out.p[i] = static_cast<const Real*>(g->vmr[i]);
        out.scol[i] = (d1 == 1) ? 0 : 1;
        out.slay[i] = (d2 == 1) ? 0 : d1;
    }
    return RRTMGP_OK;
}

// col_gas of gas index `gas` (0 = dry air): vmr * col_dry (compute_gas_taus, :1150-1154).
template<typename Real>
__device__ __forceinline__ Real col_of(const Gases<Real>& g, int gas, int64_t icol, int64_t ilay, Real col_dry)
{
    if (gas == 0) return col_dry;
    const int i = gas - 1;
    return g.p[i][icol * g.scol[i] + ilay * g.slay[i]] * col_dry;
}

template<typename Real>
struct CL { int jtemp, jpress, itropo; Real ftemp, fpress; };

template<typename Real>
struct FL { int jeta[2]; Real col_mix[2]; Real fminor[2][2]; Real fmajor[2][4]; };

// Per (col, lay) part of interpolation_kernel (.cu:236-249).  FMA contraction is off so the
// FP64 indices and weights reproduce a contraction-free CPU evaluation bit for bit.
template<typename Real>
__device__ __forceinline__ CL<Real> interp_cl(const Kd<Real>& k, Real p, Real t)
{
#pragma clang fp contract(off)
    CL<Real> c;
    int jt = int((t - (k.temp_ref_min - k.temp_ref_delta)) / k.temp_ref_delta);
    jt = min(k.ntemp - 1, max(1, jt));
    c.jtemp = jt;
    c.ftemp = (t - k.temp_ref[jt - 1]) / k.temp_ref_delta;
    const Real lp = log(p);
    const Real locpress = Real(1) + (lp - k.press_ref_log[0]) / k.press_ref_log_delta;
    c.jpress = min(k.npres - 1, max(1, int(locpress)));
    c.fpress = locpress - Real(c.jpress);
    c.itropo = (lp > k.press_ref_trop_log) ? 0 : 1;
    return c;
}

// Per-flavor part of interpolation_kernel (.cu:254-291).
template<typename Real>
__device__ __forceinline__ void interp_flav(const Kd<Real>& k, const CL<Real>& c, int gas1, int gas2,
                                            Real col1, Real col2, FL<Real>& f)
{
#pragma clang fp contract(off)
#pragma unroll
    for (int itemp = 0; itemp < 2; ++itemp)
    {
        const int64_t vb = c.itropo + int64_t(c.jtemp + itemp - 1) * (k.ngas + 1) * 2;
        const Real ratio = k.vmr_ref[vb + 2 * gas1] / k.vmr_ref[vb + 2 * gas2];
        const Real cm = col1 + ratio * col2;
        f.col_mix[itemp] = cm;
        const Real eta = (cm > Real(2) * Limits<Real>::tiny) ? col1 / cm : Real(0.5);
        const Real loceta = eta * Real(k.neta - 1);
        // The lower clamp only matters for non-finite inputs; it keeps every gather in bounds.
        f.jeta[itemp] = max(1, min(int(loceta) + 1, k.neta - 1));
        const Real feta = fmod(loceta, Real(1));
        const Real ftemp_term = Real(1 - itemp) + Real(2 * itemp - 1) * c.ftemp;
        const Real m0 = (Real(1) - feta) * ftemp_term;
        const Real m1 = feta * ftemp_term;
        f.fminor[itemp][0] = m0;
        f.fminor[itemp][1] = m1;
        f.fmajor[itemp][0] = (Real(1) - c.fpress) * m0;
        f.fmajor[itemp][1] = (Real(1) - c.fpress) * m1;
        f.fmajor[itemp][2] = c.fpress * m0;
        f.fmajor[itemp][3] = c.fpress * m1;
    }
}

// Major absorption of one g-point (.cu:317-339); kmajor is (ntemp, neta, npres+1, ngpt).
template<typename Real>
__device__ __forceinline__ Real major_tau(const Kd<Real>& k, const CL<Real>& c, const FL<Real>& f, int igpt)
{
    const int jpi = c.jpress + c.itropo;
    const int64_t se = k.ntemp, sp = int64_t(k.ntemp) * k.neta;
    Real tau = Real(0);
#pragma unroll
    for (int i = 0; i < 2; ++i)
    {
        const int64_t b = (c.jtemp - 1 + i) + se * (f.jeta[i] - 1) + sp * ((jpi - 1) + int64_t(k.npres + 1) * igpt);
        const Real* km = k.kmajor + b;
        tau += f.col_mix[i] * (f.fmajor[i][0] * km[0] + f.fmajor[i][1] * km[se] +
                               f.fmajor[i][2] * km[sp] + f.fmajor[i][3] * km[sp + se]);
    }
    return tau;
}

// Planck fraction of one g-point (.cu:170-180); planck_frac has the kmajor layout.
template<typename Real>
__device__ __forceinline__ Real planck_fraction(const Kd<Real>& k, const CL<Real>& c, const FL<Real>& f, int igpt)
{
    const int jp = c.jpress + c.itropo;
    const int64_t se = k.ntemp, sp = int64_t(k.ntemp) * k.neta;
    const Real* p0 = k.planck_frac + (c.jtemp - 1) + se * (f.jeta[0] - 1) + sp * ((jp - 1) + int64_t(k.npres + 1) * igpt);
    const Real* p1 = k.planck_frac + (c.jtemp) + se * (f.jeta[1] - 1) + sp * ((jp - 1) + int64_t(k.npres + 1) * igpt);
    return (f.fmajor[0][0] * p0[0] + f.fmajor[0][1] * p0[se] + f.fmajor[0][2] * p0[sp] + f.fmajor[0][3] * p0[sp + se])
         + (f.fmajor[1][0] * p1[0] + f.fmajor[1][1] * p1[se] + f.fmajor[1][2] * p1[sp] + f.fmajor[1][3] * p1[sp + se]);
}

// interpolate1D (.cu:1-12).
template<typename Real>
__device__ __forceinline__ Real interpolate1D(Real val, Real offset, Real delta, int len, const Real* table)
{
    const Real val0 = (val - offset) / delta;
    const Real frac = val0 - Real(int(val0));
    const int idx = min(len - 1, max(1, int(val0) + 1));
    return table[idx - 1] + frac * (table[idx] - table[idx - 1]);
}

// Scaling of minor interval m of set s at one (col, lay) (.cu:388-420).
template<typename Real>
__device__ __forceinline__ Real minor_scaling(const Kd<Real>& k, const Gases<Real>& g, int s, int m,
                                              int64_t icol, int64_t ilay, Real p, Real t, Real col_dry)
{
    Real scaling = col_of(g, k.idx_minor[s][m], icol, ilay, col_dry);
    if (k.dens[s][m])
    {
        const Real PaTohPa = Real(0.01);
        scaling *= PaTohPa * p / t;
        const int is = k.idx_scal[s][m];
        if (is > 0)
        {
            const Real vmr_fact = Real(1) / col_dry;
            const Real dry_fact = Real(1) / (Real(1) + col_of(g, k.idx_h2o, icol, ilay, col_dry) * vmr_fact);
            const Real cs = col_of(g, is, icol, ilay, col_dry);
            if (k.compl_[s][m])
                scaling *= (Real(1) - cs * vmr_fact * dry_fact);
            else
                scaling *= cs * vmr_fact * dry_fact;
        }
    }
    return scaling;
}

// ---------------------------------------------------------------------------------------------
// Gathers from the g-fastest tables for g-point gg (same arithmetic order as major_tau,
// planck_fraction and the minor / Rayleigh sums of the reference-layout kernels).
// The g-fastest tables hold ntemp * neta * (npres+1) * ngpt < 2^31 entries (checked at launch),
// so table offsets are 32-bit.
template<typename Real>
__device__ __forceinline__ Real major_tau_g(const Kd<Real>& k, const CL<Real>& c, const FL<Real>& f, int gg)
{
    const int sP = k.ngpt, sE = (k.npres + 1) * k.ngpt, sT = sE * k.neta;
    const int jp = c.jpress + c.itropo - 1;
    const Real* km = k.kmajor_g;
    Real tau = Real(0);
#pragma unroll
    for (int i = 0; i < 2; ++i)
    {
        const int o = (c.jtemp - 1 + i) * sT + (f.jeta[i] - 1) * sE + jp * sP + gg;
        tau += f.col_mix[i] * (f.fmajor[i][0] * km[o] + f.fmajor[i][1] * km[o + sE] +
                               f.fmajor[i][2] * km[o + sP] + f.fmajor[i][3] * km[o + sP + sE]);
    }
    return tau;
}

template<typename Real>
__device__ __forceinline__ Real planck_fraction_g(const Kd<Real>& k, const CL<Real>& c, const FL<Real>& f, int gg)
{
    const int sP = k.ngpt, sE = (k.npres + 1) * k.ngpt, sT = sE * k.neta;
    const int jp = c.jpress + c.itropo - 1;
    const Real* pf = k.planck_frac_g;
    const int o0 = (c.jtemp - 1) * sT + (f.jeta[0] - 1) * sE + jp * sP + gg;
    const int o1 = (c.jtemp) * sT + (f.jeta[1] - 1) * sE + jp * sP + gg;
    return (f.fmajor[0][0] * pf[o0] + f.fmajor[0][1] * pf[o0 + sE] + f.fmajor[0][2] * pf[o0 + sP] + f.fmajor[0][3] * pf[o0 + sP + sE])
         + (f.fmajor[1][0] * pf[o1] + f.fmajor[1][1] * pf[o1 + sE] + f.fmajor[1][2] * pf[o1 + sP] + f.fmajor[1][3] * pf[o1 + sP + sE]);
}

// Grid of the fused kernels: units (chunk, layer, 64-column tile), chunk = (band, 16-g-point
// part of the band), ordered so that consecutive units share a chunk.
struct GoGrid
{
    int64_t ncoltiles, nlay, cpb, nunits;
};

inline GoGrid make_go_grid(int64_t ncol, int64_t nlay, int64_t nbnd, int64_t max_band_gpts)
{
    GoGrid t;
    t.ncoltiles = (ncol + GO_COLS - 1) / GO_COLS;
    t.nlay = nlay;
    t.cpb = std::max<int64_t>(1, (max_band_gpts + GO_GPTS - 1) / GO_GPTS);
    t.nunits = t.ncoltiles * nlay * nbnd * t.cpb;
    return t;
}

struct GoUnit
{
    int ibnd, gs, ge;
    int64_t ilay, col0;
};

template<typename Real>
__device__ __forceinline__ GoUnit go_unit(const Kd<Real>& k, const GoGrid& t, int64_t u)
{
    GoUnit r;
    const int64_t rest = u / t.ncoltiles;
    r.col0 = (u % t.ncoltiles) * GO_COLS;
    r.ilay = rest % t.nlay;
    const int64_t chunk = rest / t.nlay;
    r.ibnd = int(chunk / t.cpb);
    const int gs_b = k.band_lims_gpt[2 * r.ibnd] - 1, ge_b = k.band_lims_gpt[2 * r.ibnd + 1];
    r.gs = gs_b + GO_GPTS * int(chunk % t.cpb);
    r.ge = min(ge_b, r.gs + GO_GPTS);
    return r;
}

constexpr int GO_MAXM = 8;   // minor intervals of one band per pair held in LDS

// State of one (column, layer) pair for one band, computed once per pair (phase 1 of the fused
// kernels) and read by the 16 g-point lanes of the pair from LDS (phase 2).
template<typename Real>
struct PairState
{
    int om[2];           // kmajor_g / planck_frac_g offsets for itemp 0/1, without the g-point
    int mn[2];           // kminor_g offsets for itemp 0/1 (table of half itropo), without contributor
    int ry[2];           // krayl_g offsets for itemp 0/1, without the g-point
    int itropo;
    int flags;           // bit 0: valid column
    Real col_mix[2];
    Real fmajor[2][4];
    Real fminor[2][2];
    Real scaling[GO_MAXM];
    Real aux[5];         // LW: B(T_lay), B(T_lev above), B(T_lev below), B(T_sfc), B(T_sfc + 1); SW: col_ray
};

// Phase 1 for one pair.
template<typename Real, bool LW>
__device__ __forceinline__ void pair_state(const Kd<Real>& k, const Gases<Real>& g, int ibnd, int band_g0,
                                           int64_t icol, int64_t ilay, int64_t ncol, const Real* play,
                                           const Real* tlay, const Real* tlev, const Real* tsfc,
                                           const Real* col_dry_a, bool at_sfc, PairState<Real>& st)
{
    const int64_t idx = icol + ncol * ilay;
    const Real pv = play[idx], tv = tlay[idx], cd = col_dry_a[idx];
    const CL<Real> c = interp_cl(k, pv, tv);
    const int bflav = k.gpoint_flavor[2 * band_g0 + c.itropo] - 1;
    const int gas1 = k.flavor[2 * bflav], gas2 = k.flavor[2 * bflav + 1];
    FL<Real> f;
    interp_flav(k, c, gas1, gas2, col_of(g, gas1, icol, ilay, cd), col_of(g, gas2, icol, ilay, cd), f);
    const int sP = k.ngpt, sE = (k.npres + 1) * k.ngpt, sT = sE * k.neta;
    const int jp = c.jpress + c.itropo - 1;
    const int s = c.itropo;
    const int mE = int(k.ncontrib[s]), mT = mE * k.neta;
    const int rE = k.ngpt, rT = k.neta * k.ngpt, rS = rT * k.ntemp;
    int flags = 1;
#pragma unroll
    for (int i = 0; i < 2; ++i)
    {
        st.om[i] = (c.jtemp - 1 + i) * sT + (f.jeta[i] - 1) * sE + jp * sP;
        st.mn[i] = (c.jtemp - 1 + i) * mT + (f.jeta[i] - 1) * mE;
        st.ry[i] = s * rS + (c.jtemp - 1 + i) * rT + (f.jeta[i] - 1) * rE;
        st.col_mix[i] = f.col_mix[i];
#pragma unroll
        for (int q = 0; q < 4; ++q) st.fmajor[i][q] = f.fmajor[i][q];
        st.fminor[i][0] = f.fminor[i][0];
        st.fminor[i][1] = f.fminor[i][1];
    }
    st.itropo = s;
    // The launch checks that a band has at most GO_MAXM minor intervals, all of its flavor.
    const int q0 = k.band_start[s][ibnd], q1 = k.band_start[s][ibnd + 1];
    for (int q = q0; q < q1; ++q)
        st.scaling[q - q0] = minor_scaling(k, g, s, k.band_list[s][q], icol, ilay, pv, tv, cd);
    st.flags = flags;
    if constexpr (LW)
    {
        const Real* tab = k.totplnk + int64_t(ibnd) * k.nplanck;
        st.aux[0] = interpolate1D(tv, k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
        st.aux[1] = interpolate1D(tlev[icol + ncol * (ilay + 1)], k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
        st.aux[2] = interpolate1D(tlev[idx], k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
        if (at_sfc)
        {
            st.aux[3] = interpolate1D(tsfc[icol], k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
            st.aux[4] = interpolate1D(tsfc[icol] + Real(1), k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
        }
    }
    else
        st.aux[0] = col_of(g, k.idx_h2o, icol, ilay, cd) + cd;
}

// Phase 2: major + minor absorption of g-point gg from a pair's state (same arithmetic order as
// major_tau and the minor sum of tau_absorption_kernel).
template<typename Real>
__device__ __forceinline__ Real absorption_from_state(const Kd<Real>& k, const PairState<Real>& st, int ibnd, int gg)
{
    const int sP = k.ngpt, sE = (k.npres + 1) * k.ngpt;
    const Real* km = k.kmajor_g;
    Real tau = Real(0);
#pragma unroll
    for (int i = 0; i < 2; ++i)
    {
        const int o = st.om[i] + gg;
        tau += st.col_mix[i] * (st.fmajor[i][0] * km[o] + st.fmajor[i][1] * km[o + sE] +
                                st.fmajor[i][2] * km[o + sP] + st.fmajor[i][3] * km[o + sP + sE]);
    }
    const int s = st.itropo;
    const int q0 = k.band_start[s][ibnd], q1 = k.band_start[s][ibnd + 1];
    const int mE = int(k.ncontrib[s]);
    const Real* kn = k.kminor_g[s];
    Real minor = Real(0);
    for (int q = q0; q < q1; ++q)
    {
        const int m = k.band_list[s][q];
        const int gs = k.minor_limits[s][2 * m] - 1, ge = k.minor_limits[s][2 * m + 1];
        if (gg >= gs && gg < ge)
        {
            const int ic = k.kstart[s][m] - 1 + (gg - gs);
            const int o0 = st.mn[0] + ic, o1 = st.mn[1] + ic;
            const Real lt = st.fminor[0][0] * kn[o0] + st.fminor[0][1] * kn[o0 + mE] +
                            st.fminor[1][0] * kn[o1] + st.fminor[1][1] * kn[o1 + mE];
            minor += lt * st.scaling[q - q0];
        }
    }
    return tau + minor;
}

// ---------------------------------------------------------------------------------------------
// Fused longwave gas optics: tau (+ cloud absorption per band) and Planck sources.
//   phase 1 (64 threads, lane = column): pair states -> LDS;
//   phase 2 (all lanes, lane = (pair, g-point)): absorption and Planck fraction -> LDS;
//   phase 3 (lane = column, wavefront = g-point row): sources, cloud add_to, coalesced stores.
template<typename Real>
__global__ __launch_bounds__(GO_THREADS)
void gas_optics_lw_kernel(Kd<Real> k, Gases<Real> g, int64_t ncol, int64_t nlay, GoGrid grid,
                          const Real* __restrict__ play, const Real* __restrict__ tlay,
                          const Real* __restrict__ tlev, const Real* __restrict__ tsfc,
                          const Real* __restrict__ col_dry_a, int sfc_lay,
                          const Real* __restrict__ cloud_tau, Real* __restrict__ tau,
                          Real* __restrict__ lay_src, Real* __restrict__ lev_inc,
                          Real* __restrict__ lev_dec, Real* __restrict__ sfc_src,
                          Real* __restrict__ sfc_jac)
{
    __shared__ PairState<Real> sst[GO_COLS];
    __shared__ Real s_tau[GO_GPTS][GO_ROW];
    __shared__ Real s_pf[GO_GPTS][GO_ROW];
    const int64_t u = xcd_remap(blockIdx.x, grid.nunits);
    if (u >= grid.nunits) return;                       // uniform over the work-group
    const GoUnit un = go_unit(k, grid, u);
    const int ng = un.ge - un.gs;
    if (ng <= 0) return;                                // uniform: an empty chunk of a short band
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int64_t ilay = un.ilay, ncl = ncol * nlay;
    const bool at_sfc = (ilay == sfc_lay - 1);

    if (w == 0)
    {
        const int64_t icol = un.col0 + lane;
        if (icol < ncol)
            pair_state<Real, true>(k, g, un.ibnd, un.gs, icol, ilay, ncol, play, tlay, tlev, tsfc, col_dry_a,
                                   at_sfc, sst[lane]);
        else
            sst[lane].flags = 0;
    }
    __syncthreads();

    const int p = lane / GO_GPTS, gl = lane % GO_GPTS;
    const int gg = un.gs + gl;
    for (int it = 0; it < GO_COLS / 16; ++it)
    {
        const int pi = it * 16 + w * 4 + p;
        const PairState<Real>& st = sst[pi];
        Real t = Real(0), pf = Real(0);
        if ((st.flags & 1) && gl < ng)
        {
            t = absorption_from_state(k, st, un.ibnd, gg);
            const int sP = k.ngpt, sE = (k.npres + 1) * k.ngpt;
            const Real* pfr = k.planck_frac_g;
            const int o0 = st.om[0] + gg, o1 = st.om[1] + gg;
            pf = (st.fmajor[0][0] * pfr[o0] + st.fmajor[0][1] * pfr[o0 + sE] + st.fmajor[0][2] * pfr[o0 + sP] +
                  st.fmajor[0][3] * pfr[o0 + sP + sE])
               + (st.fmajor[1][0] * pfr[o1] + st.fmajor[1][1] * pfr[o1 + sE] + st.fmajor[1][2] * pfr[o1 + sP] +
                  st.fmajor[1][3] * pfr[o1 + sP + sE]);
        }
        s_tau[gl][pi] = t;
        s_pf[gl][pi] = pf;
    }
    __syncthreads();

    const int64_t icol = un.col0 + lane;
    if (icol >= ncol) return;
    const PairState<Real>& st = sst[lane];
    const int64_t idx = icol + ncol * ilay;
    const Real ctau = cloud_tau ? cloud_tau[idx + ncl * un.ibnd] : Real(0);
    for (int row = w; row < ng; row += GO_THREADS / 64)
    {
        const int gr = un.gs + row;
        const int64_t o = idx + ncl * gr;
        const Real pf = s_pf[row][lane];
        tau[o] = cloud_tau ? s_tau[row][lane] + ctau : s_tau[row][lane];
        lay_src[o] = pf * st.aux[0];
        lev_inc[o] = pf * st.aux[1];
        lev_dec[o] = pf * st.aux[2];
        if (at_sfc)
        {
            sfc_src[icol + ncol * gr] = pf * st.aux[3];
            sfc_jac[icol + ncol * gr] = pf * (st.aux[4] - st.aux[3]);
        }
    }
}

// Fused shortwave gas optics: absorption + Rayleigh + combine (+ band cloud add_to), in the
// same three phases.
template<typename Real>
__global__ __launch_bounds__(GO_THREADS)
void gas_optics_sw_kernel(Kd<Real> k, Gases<Real> g, int64_t ncol, int64_t nlay, GoGrid grid,
                          const Real* __restrict__ play, const Real* __restrict__ tlay,
                          const Real* __restrict__ col_dry_a, const Real* __restrict__ ctau_a,
                          const Real* __restrict__ cssa_a, const Real* __restrict__ cg_a,
                          Real* __restrict__ tau, Real* __restrict__ ssa, Real* __restrict__ gout)
{
    __shared__ PairState<Real> sst[GO_COLS];
    __shared__ Real s_abs[GO_GPTS][GO_ROW];
    __shared__ Real s_ray[GO_GPTS][GO_ROW];
    const int64_t u = xcd_remap(blockIdx.x, grid.nunits);
    if (u >= grid.nunits) return;
    const GoUnit un = go_unit(k, grid, u);
    const int ng = un.ge - un.gs;
    if (ng <= 0) return;
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int64_t ilay = un.ilay, ncl = ncol * nlay;

    if (w == 0)
    {
        const int64_t icol = un.col0 + lane;
        if (icol < ncol)
            pair_state<Real, false>(k, g, un.ibnd, un.gs, icol, ilay, ncol, play, tlay, nullptr, nullptr,
                                    col_dry_a, false, sst[lane]);
        else
            sst[lane].flags = 0;
    }
    __syncthreads();

    const int p = lane / GO_GPTS, gl = lane % GO_GPTS;
    const int gg = un.gs + gl;
    const int rE = k.ngpt;
    for (int it = 0; it < GO_COLS / 16; ++it)
    {
        const int pi = it * 16 + w * 4 + p;
        const PairState<Real>& st = sst[pi];
        Real tg = Real(0), tray = Real(0);
        if ((st.flags & 1) && gl < ng)
        {
            tg = absorption_from_state(k, st, un.ibnd, gg);
            // Rayleigh (.cu:572-616).
            const Real* kr = k.krayl_g;
            const int o0 = st.ry[0] + gg, o1 = st.ry[1] + gg;
            const Real kloc = st.fminor[0][0] * kr[o0] + st.fminor[0][1] * kr[o0 + rE] +
                              st.fminor[1][0] * kr[o1] + st.fminor[1][1] * kr[o1 + rE];
            tray = kloc * st.aux[0];
        }
        s_abs[gl][pi] = tg;
        s_ray[gl][pi] = tray;
    }
    __syncthreads();

    const int64_t icol = un.col0 + lane;
    if (icol >= ncol) return;
    const int64_t idx = icol + ncol * ilay;
    const Real eps2 = Real(2) * Limits<Real>::eps;
    const Real op_eps = Real(3) * Limits<Real>::tiny;
    Real t2 = Real(0), s2 = Real(0), g2 = Real(0);
    if (ctau_a)
    {
        const int64_t ib = idx + ncl * un.ibnd;
        t2 = ctau_a[ib]; s2 = cssa_a[ib]; g2 = cg_a[ib];
    }
    for (int row = w; row < ng; row += GO_THREADS / 64)
    {
        const Real tray = s_ray[row][lane];
        const Real tt = s_abs[row][lane] + tray;
        // The CPU combine rule (Gas_optics_rrtmgp.cpp:365-385).
        Real tt_out = tt;
        Real ss = (tt > eps2) ? tray / tt : Real(0);
        Real gg_out = Real(0);
        if (ctau_a)
        {
            // inc_2stream_by_2stream_bybnd (optical_props_kernels.cu:99-136).
            const Real tau12 = tt + t2;
            const Real tauscat12 = (tt * ss) + (t2 * s2);
            gg_out = ((tt * ss * Real(0)) + (t2 * s2 * g2)) / max(tauscat12, op_eps);
            ss = tauscat12 / max(op_eps, tau12);
            tt_out = tau12;
        }
        const int64_t o = idx + ncl * (un.gs + row);
        tau[o] = tt_out;
        ssa[o] = ss;
        gout[o] = gg_out;
    }
}

// ---------------------------------------------------------------------------------------------
// API-parity kernels that materialize the interpolation state (reference launch structure).

template<typename Real>
__global__ void interpolation_kernel(Kd<Real> k, int64_t ncol, int64_t nlay,
                                     const Real* __restrict__ play, const Real* __restrict__ tlay,
                                     const Real* __restrict__ col_gas, int* __restrict__ jtemp,
                                     Real* __restrict__ fmajor, Real* __restrict__ fminor,
                                     Real* __restrict__ col_mix, Bool* __restrict__ tropo,
                                     int* __restrict__ jeta, int* __restrict__ jpress)
{
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t ncl = ncol * nlay;
    if (idx >= ncl) return;
    const CL<Real> c = interp_cl(k, play[idx], tlay[idx]);
    jtemp[idx] = c.jtemp;
    jpress[idx] = c.jpress;
    tropo[idx] = Bool(c.itropo == 0);
    for (int iflav = 0; iflav < k.nflav; ++iflav)
    {
        const int gas1 = k.flavor[2 * iflav], gas2 = k.flavor[2 * iflav + 1];
        FL<Real> f;
        interp_flav(k, c, gas1, gas2, col_gas[idx + ncl * gas1], col_gas[idx + ncl * gas2], f);
        const int64_t fl = idx + ncl * iflav;
        for (int i = 0; i < 2; ++i)
        {
            jeta[i + 2 * fl] = f.jeta[i];
            col_mix[i + 2 * fl] = f.col_mix[i];
            fminor[2 * (i + 2 * fl)] = f.fminor[i][0];
            fminor[2 * (i + 2 * fl) + 1] = f.fminor[i][1];
            for (int q = 0; q < 4; ++q) fmajor[4 * (i + 2 * fl) + q] = f.fmajor[i][q];
        }
    }
}

template<typename Real>
__device__ __forceinline__ void load_state(const Kd<Real>& k, int64_t idx, int64_t ncl, int iflav,
                                           const Bool* tropo, const int* jtemp, const int* jpress,
                                           const int* jeta, const Real* col_mix, const Real* fminor,
                                           const Real* fmajor, CL<Real>& c, FL<Real>& f)
{
    c.itropo = !tropo[idx];
    c.jtemp = jtemp[idx];
    c.jpress = jpress ? jpress[idx] : 1;
    const int64_t fl = idx + ncl * iflav;
    for (int i = 0; i < 2; ++i)
    {
        f.jeta[i] = jeta[i + 2 * fl];
        f.col_mix[i] = col_mix ? col_mix[i + 2 * fl] : Real(0);
        f.fminor[i][0] = fminor ? fminor[2 * (i + 2 * fl)] : Real(0);
        f.fminor[i][1] = fminor ? fminor[2 * (i + 2 * fl) + 1] : Real(0);
        for (int q = 0; q < 4; ++q) f.fmajor[i][q] = fmajor ? fmajor[4 * (i + 2 * fl) + q] : Real(0);
    }
}

// compute_tau_absorption from materialized state: thread per (col, lay, band).
template<typename Real>
__global__ void tau_absorption_kernel(Kd<Real> k, int64_t ncol, int64_t nlay,
                                      const Bool* __restrict__ tropo, const Real* __restrict__ col_mix,
                                      const Real* __restrict__ fmajor, const Real* __restrict__ fminor,
                                      const Real* __restrict__ play, const Real* __restrict__ tlay,
                                      const Real* __restrict__ col_gas, const int* __restrict__ jeta,
                                      const int* __restrict__ jtemp, const int* __restrict__ jpress,
                                      Real* __restrict__ tau)
{
    const int64_t ncl = ncol * nlay;
    const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (tid >= ncl * k.nbnd) return;
    const int64_t idx = tid % ncl;
    const int ibnd = int(tid / ncl);
    const int64_t icol = idx % ncol, ilay = idx / ncol;
    const int itropo = !tropo[idx];
    const int gs_b = k.band_lims_gpt[2 * ibnd] - 1, ge_b = k.band_lims_gpt[2 * ibnd + 1];
    const Real p = play[idx], t = tlay[idx];
    for (int gg = gs_b; gg < ge_b; ++gg)
    {
        const int iflav = k.gpoint_flavor[itropo + 2 * gg] - 1;
        CL<Real> c; FL<Real> f;
        load_state(k, idx, ncl, iflav, tropo, jtemp, jpress, jeta, col_mix, fminor, fmajor, c, f);
        Real tt = tau[idx + ncl * gg] + major_tau(k, c, f, gg);
        // minor intervals of the active set covering this g-point, in interval order
        const int s = itropo;
        for (int m = 0; m < k.nminor[s]; ++m)
        {
            const int gs = k.minor_limits[s][2 * m] - 1, ge = k.minor_limits[s][2 * m + 1];
            if (gg < gs || gg >= ge) continue;
            Real scaling = col_gas[idx + ncl * k.idx_minor[s][m]];
            if (k.dens[s][m])
            {
                scaling *= Real(0.01) * p / t;
                const int is = k.idx_scal[s][m];
                if (is > 0)
                {
                    const Real vmr_fact = Real(1) / col_gas[idx];
                    const Real dry_fact = Real(1) / (Real(1) + col_gas[idx + ncl * k.idx_h2o] * vmr_fact);
                    if (k.compl_[s][m])
                        scaling *= (Real(1) - col_gas[idx + ncl * is] * vmr_fact * dry_fact);
                    else
                        scaling *= col_gas[idx + ncl * is] * vmr_fact * dry_fact;
                }
            }
            const int mflav = k.gpoint_flavor[2 * gs + s] - 1;
            CL<Real> cm; FL<Real> fm;
            load_state(k, idx, ncl, mflav, tropo, jtemp, jpress, jeta, col_mix, fminor, fmajor, cm, fm);
            const int64_t se = k.ntemp, sc = int64_t(k.ntemp) * k.neta;
            const int64_t off = sc * (gg - gs + k.kstart[s][m] - 1);
            const Real* kb0 = k.kminor[s] + (cm.jtemp - 1) + se * (fm.jeta[0] - 1) + off;
            const Real* kb1 = k.kminor[s] + (cm.jtemp) + se * (fm.jeta[1] - 1) + off;
            const Real lt = fm.fminor[0][0] * kb0[0] + fm.fminor[0][1] * kb0[se] +
                            fm.fminor[1][0] * kb1[0] + fm.fminor[1][1] * kb1[se];
            tt += lt * scaling;
        }
        tau[idx + ncl * gg] = tt;
        (void)icol; (void)ilay;
    }
}

template<typename Real>
__global__ void tau_rayleigh_kernel(Kd<Real> k, int64_t ncol, int64_t nlay, const Real* __restrict__ col_dry,
                                    const Real* __restrict__ col_gas, const Real* __restrict__ fminor,
                                    const int* __restrict__ jeta, const Bool* __restrict__ tropo,
                                    const int* __restrict__ jtemp, Real* __restrict__ tau_ray)
{
    const int64_t ncl = ncol * nlay;
    const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (tid >= ncl * k.ngpt) return;
    const int64_t idx = tid % ncl;
    const int gg = int(tid / ncl);
    const int itropo = !tropo[idx];
    const int iflav = k.gpoint_flavor[itropo + 2 * gg] - 1;
    CL<Real> c; FL<Real> f;
    load_state<Real>(k, idx, ncl, iflav, tropo, jtemp, nullptr, jeta, nullptr, fminor, nullptr, c, f);
    const int64_t se = k.ntemp;
    const Real* kr = k.krayl + int64_t(itropo) * k.ntemp * k.neta * k.ngpt + int64_t(k.ntemp) * k.neta * gg;
    const Real* k0 = kr + (c.jtemp - 1) + se * (f.jeta[0] - 1);
    const Real* k1 = kr + (c.jtemp) + se * (f.jeta[1] - 1);
    const Real kloc = f.fminor[0][0] * k0[0] + f.fminor[0][1] * k0[se] + f.fminor[1][0] * k1[0] + f.fminor[1][1] * k1[se];
    tau_ray[idx + ncl * gg] = kloc * (col_gas[idx + ncl * k.idx_h2o] + col_dry[idx]);
}

template<typename Real>
__global__ void combine_kernel(int64_t n, const Real* __restrict__ tau_abs, const Real* __restrict__ tau_ray,
                               Real* __restrict__ tau, Real* __restrict__ ssa, Real* __restrict__ g)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Real tr = tau_ray[i];
    const Real tt = tau_abs[i] + tr;
    ssa[i] = (tt > Real(2) * Limits<Real>::eps) ? tr / tt : Real(0);
    tau[i] = tt;
    g[i] = Real(0);
}

template<typename Real>
__global__ void planck_kernel(Kd<Real> k, int64_t ncol, int64_t nlay, const Real* __restrict__ tlay,
                              const Real* __restrict__ tlev, const Real* __restrict__ tsfc, int sfc_lay,
                              const Real* __restrict__ fmajor, const int* __restrict__ jeta,
                              const Bool* __restrict__ tropo, const int* __restrict__ jtemp,
                              const int* __restrict__ jpress, Real* __restrict__ sfc_src,
                              Real* __restrict__ lay_src, Real* __restrict__ lev_inc,
                              Real* __restrict__ lev_dec, Real* __restrict__ sfc_jac)
{
    const int64_t ncl = ncol * nlay;
    const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (tid >= ncl * k.ngpt) return;
    const int64_t idx = tid % ncl;
    const int gg = int(tid / ncl);
    const int64_t icol = idx % ncol, ilay = idx / ncol;
    const int itropo = !tropo[idx];
    const int iflav = k.gpoint_flavor[itropo + 2 * gg] - 1;
    CL<Real> c; FL<Real> f;
    load_state<Real>(k, idx, ncl, iflav, tropo, jtemp, jpress, jeta, nullptr, nullptr, fmajor, c, f);
    int ibnd = 0;
    while (gg >= k.band_lims_gpt[2 * ibnd + 1]) ++ibnd;
    const Real* tab = k.totplnk + int64_t(ibnd) * k.nplanck;
    const Real pf = planck_fraction(k, c, f, gg);
    lay_src[tid] = pf * interpolate1D(tlay[idx], k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
    lev_inc[tid] = pf * interpolate1D(tlev[icol + ncol * (ilay + 1)], k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
    lev_dec[tid] = pf * interpolate1D(tlev[idx], k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
    if (ilay == sfc_lay - 1)
    {
        const Real s1 = interpolate1D(tsfc[icol], k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
        const Real s2 = interpolate1D(tsfc[icol] + Real(1), k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
        sfc_src[icol + ncol * gg] = pf * s1;
        sfc_jac[icol + ncol * gg] = pf * (s2 - s1);
    }
}

// get_col_dry (src/Gas_optics_rrtmgp.cpp:764-792).
template<typename Real>
__global__ void col_dry_kernel(int64_t ncol, int64_t nlay, const Real* __restrict__ h2o, int64_t scol,
                               int64_t slay, const Real* __restrict__ plev, Real* __restrict__ col_dry)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncol * nlay) return;
    const int64_t icol = i % ncol, ilay = i / ncol;
    const Real g0 = Real(9.80665), avogad = Real(6.02214076e23), m_dry = Real(0.028964), m_h2o = Real(0.018016);
    const Real v = h2o[icol * scol + ilay * slay];
    const Real dp = fabs(plev[i] - plev[i + ncol]);
    const Real m_air = (m_dry + m_h2o * v) / (Real(1) + v);
    Real cdry = Real(10) * dp * avogad / (Real(1000) * m_air * Real(100) * g0);
    cdry /= (Real(1) + v);
    col_dry[i] = cdry;
}

template<typename Real>
__global__ void fill_col_gas_kernel(int64_t ncol, int64_t nlay, int ngas, Gases<Real> g,
                                    const Real* __restrict__ col_dry, Real* __restrict__ col_gas)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t ncl = ncol * nlay;
    if (i >= ncl) return;
    const int64_t icol = i % ncol, ilay = i / ncol;
    const Real cd = col_dry[i];
    col_gas[i] = cd;
    for (int ig = 1; ig <= ngas; ++ig) col_gas[i + ncl * ig] = col_of(g, ig, icol, ilay, cd);
}

// ---------------------------------------------------------------------------------------------
// Host launchers.

int check_kd(const rrtmgp_kdist_dev* kd, bool need_planck, bool need_rayleigh)
{
    if (!kd) return arg_error("k-distribution is NULL");
    if (kd->ngas < 1 || kd->ngas > RRTMGP_MAX_GAS) return arg_error("ngas out of range");
    if (!kd->kmajor || !kd->flavor || !kd->gpoint_flavor || !kd->band_lims_gpt)
        return arg_error("k-distribution tables are missing");
    if (need_planck && (!kd->totplnk || !kd->planck_frac || kd->nplanck < 2))
        return arg_error("k-distribution has no Planck tables (longwave entry on a shortwave kdist)");
    if (need_rayleigh && !kd->krayl)
        return arg_error("k-distribution has no Rayleigh tables (shortwave entry on a longwave kdist)");
    if (kd->idx_h2o < 1) return arg_error("idx_h2o cannot be found");
    return RRTMGP_OK;
}

// The fused kernels also read the g-fastest table copies.
int check_kd_fused(const rrtmgp_kdist_dev* kd, bool lw)
{
    RRTMGP_TRY(check_kd(kd, lw, !lw));
    if (!kd->kmajor_g || (lw && !kd->planck_frac_g) || (!lw && !kd->krayl_g))
        return arg_error("k-distribution has no g-point-fastest tables");
    if ((kd->nminor_lower > 0 && !kd->kminor_lower_g) || (kd->nminor_upper > 0 && !kd->kminor_upper_g))
        return arg_error("k-distribution has no g-point-fastest minor tables");
    if (kd->max_band_gpts < 1 || kd->max_band_gpts > kd->ngpt) return arg_error("max_band_gpts out of range");
    if (kd->max_band_minors > GO_MAXM || !kd->minor_flavor_uniform)
        return arg_error("fused gas optics: a band has more than " + std::to_string(GO_MAXM) +
                         " minor intervals or an interval crossing into another band's flavor");
    const int64_t nmaj = kd->ntemp * kd->neta * (kd->npres + 1) * kd->ngpt;
    const int64_t nmin = kd->ntemp * kd->neta * std::max(kd->ncontrib_lower, kd->ncontrib_upper);
    if (nmaj >= (int64_t(1) << 31) || nmin >= (int64_t(1) << 31)) return arg_error("k-tables too large for 32-bit offsets");
    return RRTMGP_OK;
}

template<typename Real>
int gas_optics_lw(const rrtmgp_kdist_dev* kd, int64_t ncol, int64_t nlay, const Real* play,
                  const Real* tlay, const Real* tlev, const Real* tsfc, const Real* col_dry,
                  const rrtmgp_gas_vmr* gases, int sfc_lay, const Real* cloud_tau, Real* tau,
                  Real* lay_src, Real* lev_inc, Real* lev_dec, Real* sfc_src, Real* sfc_jac,
                  void* stream)
{
    RRTMGP_TRY(check_kd_fused(kd, true));
    if (ncol <= 0 || nlay <= 0) return arg_error("ncol and nlay must be positive");
    if (sfc_lay != 1 && sfc_lay != nlay) return arg_error("sfc_lay must be 1 or nlay");
    if (!play || !tlay || !tlev || !tsfc || !col_dry || !tau || !lay_src || !lev_inc || !lev_dec || !sfc_src || !sfc_jac)
        return arg_error("gas_optics_lw: NULL array");
    Gases<Real> g;
    RRTMGP_TRY(make_gases(gases, kd->ngas, ncol, nlay, g));
    const Kd<Real> k = make_kd<Real>(*kd);
    const GoGrid grid = make_go_grid(ncol, nlay, kd->nbnd, kd->max_band_gpts);
    RRTMGP_TIMED("gas_optics_lw_kernel", as_stream(stream));
    hipLaunchKernelGGL(gas_optics_lw_kernel<Real>, dim3(xcd_grid(grid.nunits)), dim3(GO_THREADS), 0,
                       as_stream(stream), k, g, ncol, nlay, grid, play, tlay, tlev, tsfc, col_dry,
                       sfc_lay, cloud_tau, tau, lay_src, lev_inc, lev_dec, sfc_src, sfc_jac);
    RRTMGP_LAUNCHED("gas_optics_lw_kernel");
    return RRTMGP_OK;
}

template<typename Real>
int gas_optics_sw(const rrtmgp_kdist_dev* kd, int64_t ncol, int64_t nlay, const Real* play,
                  const Real* tlay, const Real* col_dry, const rrtmgp_gas_vmr* gases,
                  const Real* ctau, const Real* cssa, const Real* cg, Real* tau, Real* ssa, Real* g,
                  void* stream)
{
    RRTMGP_TRY(check_kd_fused(kd, false));
    if (ncol <= 0 || nlay <= 0) return arg_error("ncol and nlay must be positive");
    if (!play || !tlay || !col_dry || !tau || !ssa || !g) return arg_error("gas_optics_sw: NULL array");
    if (ctau && (!cssa || !cg)) return arg_error("cloud ssa/g missing");
    Gases<Real> gs;
    RRTMGP_TRY(make_gases(gases, kd->ngas, ncol, nlay, gs));
    const Kd<Real> k = make_kd<Real>(*kd);
    const GoGrid grid = make_go_grid(ncol, nlay, kd->nbnd, kd->max_band_gpts);
    RRTMGP_TIMED("gas_optics_sw_kernel", as_stream(stream));
    hipLaunchKernelGGL(gas_optics_sw_kernel<Real>, dim3(xcd_grid(grid.nunits)), dim3(GO_THREADS), 0,
                       as_stream(stream), k, gs, ncol, nlay, grid, play, tlay, col_dry, ctau, cssa, cg,
                       tau, ssa, g);
    RRTMGP_LAUNCHED("gas_optics_sw_kernel");
    return RRTMGP_OK;
}
}  // namespace rrtmgp_hip

using namespace rrtmgp_hip;

#define RRTMGP_DEF_GAS_OPTICS(S, T)                                                               \
extern "C" int rrtmgp_hip_gas_optics_lw_##S(const rrtmgp_kdist_dev* kd, int64_t ncol, int64_t nlay, \
    const T* play, const T* tlay, const T* tlev, const T* tsfc, const T* col_dry,                   \
    const rrtmgp_gas_vmr* gases, int sfc_lay, const T* cloud_tau, T* tau, T* lay_src,              \
    T* lev_src_inc, T* lev_src_dec, T* sfc_src, T* sfc_src_jac, void* stream)                      \
{                                                                                                  \
    return gas_optics_lw<T>(kd, ncol, nlay, play, tlay, tlev, tsfc, col_dry, gases, sfc_lay,      \
                            cloud_tau, tau, lay_src, lev_src_inc, lev_src_dec, sfc_src,            \
                            sfc_src_jac, stream);                                                  \
}                                                                                                  \
extern "C" int rrtmgp_hip_gas_optics_sw_##S(const rrtmgp_kdist_dev* kd, int64_t ncol, int64_t nlay, \
    const T* play, const T* tlay, const T* col_dry, const rrtmgp_gas_vmr* gases,                   \
    const T* cloud_tau, const T* cloud_ssa, const T* cloud_g, T* tau, T* ssa, T* g, void* stream)  \
{                                                                                                  \
    return gas_optics_sw<T>(kd, ncol, nlay, play, tlay, col_dry, gases, cloud_tau, cloud_ssa,     \
                            cloud_g, tau, ssa, g, stream);                                         \
}                                                                                                  \
extern "C" int rrtmgp_hip_interpolation_##S(const rrtmgp_kdist_dev* kd, int64_t ncol, int64_t nlay, \
    const T* play, const T* tlay, const T* col_gas, int* jtemp, T* fmajor, T* fminor, T* col_mix,  \
    rrtmgp_bool* tropo, int* jeta, int* jpress, void* stream)                                      \
{                                                                                                  \
    RRTMGP_TRY(check_kd(kd, false, false));                                                       \
    const Kd<T> k = make_kd<T>(*kd);                                                               \
    const int64_t n = ncol * nlay;                                                                 \
    RRTMGP_TIMED("interpolation_kernel", as_stream(stream));                                       \
    hipLaunchKernelGGL(interpolation_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0,                \
                       as_stream(stream), k, ncol, nlay, play, tlay, col_gas, jtemp, fmajor,       \
                       fminor, col_mix, tropo, jeta, jpress);                                      \
    RRTMGP_LAUNCHED("interpolation_kernel");                                                      \
    return RRTMGP_OK;                                                                              \
}                                                                                                  \
extern "C" int rrtmgp_hip_compute_tau_absorption_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,   \
    int64_t nlay, const rrtmgp_bool* tropo, const T* col_mix, const T* fmajor, const T* fminor,    \
    const T* play, const T* tlay, const T* col_gas, const int* jeta, const int* jtemp,             \
    const int* jpress, T* tau, void* stream)                                                       \
{                                                                                                  \
    RRTMGP_TRY(check_kd(kd, false, false));                                                       \
    const Kd<T> k = make_kd<T>(*kd);                                                               \
    const int64_t n = ncol * nlay * kd->nbnd;                                                      \
    RRTMGP_TIMED("tau_absorption_kernel", as_stream(stream));                                      \
    hipLaunchKernelGGL(tau_absorption_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0,               \
                       as_stream(stream), k, ncol, nlay, tropo, col_mix, fmajor, fminor, play,     \
                       tlay, col_gas, jeta, jtemp, jpress, tau);                                   \
    RRTMGP_LAUNCHED("tau_absorption_kernel");                                                     \
    return RRTMGP_OK;                                                                              \
}                                                                                                  \
extern "C" int rrtmgp_hip_compute_tau_rayleigh_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,     \
    int64_t nlay, const T* col_dry, const T* col_gas, const T* fminor, const int* jeta,            \
    const rrtmgp_bool* tropo, const int* jtemp, T* tau_rayleigh, void* stream)                     \
{                                                                                                  \
    RRTMGP_TRY(check_kd(kd, false, true));                                                        \
    const Kd<T> k = make_kd<T>(*kd);                                                               \
    const int64_t n = ncol * nlay * kd->ngpt;                                                      \
    RRTMGP_TIMED("tau_rayleigh_kernel", as_stream(stream));                                        \
    hipLaunchKernelGGL(tau_rayleigh_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0,                 \
                       as_stream(stream), k, ncol, nlay, col_dry, col_gas, fminor, jeta, tropo,    \
                       jtemp, tau_rayleigh);                                                       \
    RRTMGP_LAUNCHED("tau_rayleigh_kernel");                                                       \
    return RRTMGP_OK;                                                                              \
}                                                                                                  \
extern "C" int rrtmgp_hip_combine_abs_and_rayleigh_##S(int64_t ncol, int64_t nlay, int64_t ngpt,  \
    const T* tau_abs, const T* tau_rayleigh, T* tau, T* ssa, T* g, void* stream)                   \
{                                                                                                  \
    const int64_t n = ncol * nlay * ngpt;                                                          \
    RRTMGP_TIMED("combine_kernel", as_stream(stream));                                             \
    hipLaunchKernelGGL(combine_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),   \
                       n, tau_abs, tau_rayleigh, tau, ssa, g);                                     \
    RRTMGP_LAUNCHED("combine_kernel");                                                            \
    return RRTMGP_OK;                                                                              \
}                                                                                                  \
extern "C" int rrtmgp_hip_compute_planck_source_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,    \
    int64_t nlay, const T* tlay, const T* tlev, const T* tsfc, int sfc_lay, const T* fmajor,       \
    const int* jeta, const rrtmgp_bool* tropo, const int* jtemp, const int* jpress, T* sfc_src,    \
    T* lay_src, T* lev_src_inc, T* lev_src_dec, T* sfc_src_jac, void* stream)                      \
{                                                                                                  \
    RRTMGP_TRY(check_kd(kd, true, false));                                                        \
    const Kd<T> k = make_kd<T>(*kd);                                                               \
    const int64_t n = ncol * nlay * kd->ngpt;                                                      \
    RRTMGP_TIMED("planck_kernel", as_stream(stream));                                              \
    hipLaunchKernelGGL(planck_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),    \
                       k, ncol, nlay, tlay, tlev, tsfc, sfc_lay, fmajor, jeta, tropo, jtemp,       \
                       jpress, sfc_src, lay_src, lev_src_inc, lev_src_dec, sfc_src_jac);           \
    RRTMGP_LAUNCHED("planck_kernel");                                                             \
    return RRTMGP_OK;                                                                              \
}                                                                                                  \
extern "C" int rrtmgp_hip_get_col_dry_##S(int64_t ncol, int64_t nlay, const T* vmr_h2o,          \
    int64_t h2o_dim1, int64_t h2o_dim2, const T* plev, T* col_dry, void* stream)                   \
{                                                                                                  \
    const int64_t n = ncol * nlay;                                                                 \
    const int64_t scol = h2o_dim1 == 1 ? 0 : 1, slay = h2o_dim2 == 1 ? 0 : h2o_dim1;              \
    RRTMGP_TIMED("col_dry_kernel", as_stream(stream));                                             \
    hipLaunchKernelGGL(col_dry_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),   \
                       ncol, nlay, vmr_h2o, scol, slay, plev, col_dry);                            \
    RRTMGP_LAUNCHED("col_dry_kernel");                                                            \
    return RRTMGP_OK;                                                                              \
}                                                                                                  \
extern "C" int rrtmgp_hip_fill_col_gas_##S(int64_t ncol, int64_t nlay, const rrtmgp_gas_vmr* gases,\
    const T* col_dry, T* col_gas, void* stream)                                                    \
{                                                                                                  \
    Gases<T> g;                                                                                    \
    RRTMGP_TRY(make_gases(gases, gases ? gases->ngas : 0, ncol, nlay, g));                        \
    const int64_t n = ncol * nlay;                                                                 \
    RRTMGP_TIMED("fill_col_gas_kernel", as_stream(stream));                                        \
    hipLaunchKernelGGL(fill_col_gas_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0,                 \
                       as_stream(stream), ncol, nlay, int(gases->ngas), g, col_dry, col_gas);      \
    RRTMGP_LAUNCHED("fill_col_gas_kernel");                                                       \
    return RRTMGP_OK;                                                                              \
}

RRTMGP_DEF_GAS_OPTICS(f32, float)
RRTMGP_DEF_GAS_OPTICS(f64, double)
