// Cloud optics, optical-property arithmetic, flux reductions and heating rates for gfx950.
//
// Element-wise kernels: one thread per element of the reference layout, columns fastest, so
// every access is coalesced.  Arithmetic follows src/Cloud_optics.cpp:72-232,
// src_kernels_cuda/optical_props_kernels.cu:30-160 (eps = 3*tiny, launchers :80,123,172),
// src_kernels_cuda/fluxes_kernels.cu:26-62 and src_kernels/mo_fluxes_byband_kernels.F90:22-79.
#include "common.h"

namespace rrtmgp_hip
{
// Table lookup of compute_all_from_table (src/Cloud_optics.cpp:80-101) for one element.
// taua, when given, receives the absorption optical depth tau (1 - ssa), with 1 - ssa interpolated
// from the table's own 1 - ssa (exact in either precision for ssa >= 1/2) instead of formed as
// tau - taussa.
template<typename Real>
__device__ __forceinline__ void cloud_lookup(Real cwp, Real re, int64_t nsteps, Real step, Real offset,
                                             const Real* et, const Real* st, const Real* at, int64_t ibnd,
                                             Real& tau, Real& taussa, Real& taussag, Real* taua = nullptr)
{
    if (cwp > Real(0))
    {
        // Clamped below as well: an effective radius under the table's lower bound would
        // otherwise index before the table (the reference reads out of bounds there).
        const int index = max(1, min(int((re - offset) / step) + 1, int(nsteps) - 1));
        const Real fint = (re - offset) / step - Real(index - 1);
        const int64_t b = nsteps * ibnd;
        const Real e0 = et[b + index - 1], s0 = st[b + index - 1], a0 = at[b + index - 1];
        const Real ds = st[b + index] - s0;
        tau = cwp * (e0 + fint * (et[b + index] - e0));
        taussa = tau * (s0 + fint * ds);
        taussag = taussa * (a0 + fint * (at[b + index] - a0));
        if (taua) *taua = tau * ((Real(1) - s0) - fint * ds);
    }
    else
    {
        tau = Real(0); taussa = Real(0); taussag = Real(0);
        if (taua) *taua = Real(0);
    }
}

template<typename Real>
__global__ void cloud_optics_kernel(int64_t ncol, int64_t nlay, int64_t nbnd,
                                    int64_t liq_nsteps, Real liq_step, Real liq_lwr,
                                    const Real* __restrict__ extl, const Real* __restrict__ ssal,
                                    const Real* __restrict__ asyl, int64_t ice_nsteps, Real ice_step,
                                    Real ice_lwr, const Real* __restrict__ exti, const Real* __restrict__ ssai,
                                    const Real* __restrict__ asyi, const Real* __restrict__ clwp,
                                    const Real* __restrict__ ciwp, const Real* __restrict__ rel,
                                    const Real* __restrict__ rei, int two_stream, int delta_scale,
                                    Real* __restrict__ tau, Real* __restrict__ ssa, Real* __restrict__ g)
{
    const int64_t ncl = ncol * nlay;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncl) return;
    const Real lw = clwp[i], iw = ciwp[i], rl = rel[i], ri = rei[i];
    const Real eps = Real(prop_eps);
    const Real op_eps = Real(3) * Limits<Real>::tiny;
    for (int64_t b = 0; b < nbnd; ++b)
    {
        Real lt, ls, lg, it, is, ig, la = Real(0), ia = Real(0);
        cloud_lookup(lw, rl, liq_nsteps, liq_step, liq_lwr, extl, ssal, asyl, b, lt, ls, lg,
                     two_stream == 2 ? &la : nullptr);
        cloud_lookup(iw, ri, ice_nsteps, ice_step, ice_lwr, exti, ssai, asyi, b, it, is, ig,
                     two_stream == 2 ? &ia : nullptr);
        const int64_t o = i + ncl * b;
        if (two_stream == 2)
        {
            // As below, with the co-albedo 1 - ssa written in place of ssa, formed from the
            // absorption optical depth (tau - tau ssa per phase, each from 1 - the table's ssa):
            // it keeps the precision that 1 - ssa in FP32 loses near conservative scattering.
            // Delta-scaling in the same terms: 1 - ssa' = (1 - ssa) / (1 - f + f (1 - ssa)).
            Real t = lt + it;
            const Real ts = ls + is, tsg = lg + ig;
            const Real ta = la + ia;
            Real co = (t > eps) ? ta / t : Real(1) - ts / eps;
            Real gg = tsg / max(ts, eps);
            if (delta_scale)
            {
                const Real f = gg * gg;
                const Real d = (Real(1) - f) + f * co;         // 1 - ssa f
                t *= d;
                co = co / max(op_eps, d);
                gg = (gg - f) / max(op_eps, Real(1) - f);
            }
            tau[o] = t; ssa[o] = co; g[o] = gg;
        }
        else if (two_stream)
        {
            // Cloud_optics::cloud_optics 2str (src/Cloud_optics.cpp:159-171).
            Real t = lt + it;
            const Real ts = ls + is, tsg = lg + ig;
            Real s = ts / max(t, eps);
            Real gg = tsg / max(ts, eps);
            if (delta_scale)
            {
                // delta_scale_2str_k_kernel (optical_props_kernels.cu:139-160).
                const Real f = gg * gg, wf = s * f;
                t *= (Real(1) - wf);
                s = (s - wf) / max(op_eps, Real(1) - wf);
                gg = (gg - f) / max(op_eps, Real(1) - f);
            }
            tau[o] = t; ssa[o] = s; g[o] = gg;
        }
        else
            tau[o] = (lt - ls) + (it - is);   // 1scl (:219-230)
    }
}

// Aerosol optics (src/Aerosol_optics.cpp:38-224): per layer, 11 species with mass mixing
// ratios aermr01..aermr11, each with its column of the hydrophobic table (nbnd, nphobic) or, for
// hydrophilic species, of the humidity class's slice of (nbnd, nhum, nphilic).  Species in the
// summation order of compute_all_from_table (:49-146): field (0-based), hydrophilic, column
// (0-based).
struct AerosolSpeciesRef { int field, philic, column; };
constexpr AerosolSpeciesRef aerosol_species_order[11] = {
    {0, 1, 0},    // SS1: aermr01, hydrophilic 1
    {1, 1, 1},    // SS2: aermr02, hydrophilic 2
    {2, 1, 2},    // SS3: aermr03, hydrophilic 3
    {3, 0, 0},    // DU1: aermr04, hydrophobic 1
    {4, 0, 7},    // DU2: aermr05, hydrophobic 8
    {5, 0, 5},    // DU3: aermr06, hydrophobic 6
    {7, 0, 9},    // OM1: aermr08, hydrophobic 10
    {6, 1, 3},    // OM2: aermr07, hydrophilic 4
    {8, 0, 10},   // BC1: aermr09, hydrophobic 11
    {9, 0, 10},   // BC2: aermr10, hydrophobic 11
    {10, 1, 4},   // SU:  aermr11, hydrophilic 5
};
constexpr int AEROSOL_NSPECIES = 11;

template<typename Real>
struct AerosolMix
{
    const Real* p[AEROSOL_NSPECIES];
    int64_t scol[AEROSOL_NSPECIES], slay[AEROSOL_NSPECIES];
};

// One thread per (column, layer); the band loop writes (ncol, nlay, nbnd) with columns fastest.
// The layer mass is |p(ilay) - p(ilay+1)| / 9.81 (src_cuda/Aerosol_optics.cu:67; the CPU path's
// signed difference agrees for levels ordered from the surface up).  The humidity class is the
// first whose upper bound is >= rh, the last when rh exceeds them all.  delta_scale != 0 applies
// delta_scale_2str_k (optical_props_kernels.cu:139-160) to the result.
template<typename Real>
__global__ void aerosol_optics_kernel(int64_t ncol, int64_t nlay, int64_t nbnd, int64_t nhum, AerosolMix<Real> mix,
                                      const Real* __restrict__ rh, const Real* __restrict__ plev,
                                      const Real* __restrict__ rh_upper, const Real* __restrict__ mext_phobic,
                                      const Real* __restrict__ ssa_phobic, const Real* __restrict__ g_phobic,
                                      const Real* __restrict__ mext_philic, const Real* __restrict__ ssa_philic,
                                      const Real* __restrict__ g_philic, int delta_scale, int coalb,
                                      Real* __restrict__ tau, Real* __restrict__ ssa, Real* __restrict__ g)
{
    const int64_t ncl = ncol * nlay;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncl) return;
    const int64_t icol = i % ncol, ilay = i / ncol;
    const Real dpg = fabs(plev[i] - plev[i + ncol]) / Real(9.81);
    const Real r = rh[i];
    int64_t ihum = 0;
    while (ihum + 1 < nhum && rh_upper[ihum] < r) ++ihum;
    Real m[AEROSOL_NSPECIES];
#pragma unroll
    for (int k = 0; k < AEROSOL_NSPECIES; ++k) m[k] = mix.p[k][icol * mix.scol[k] + ilay * mix.slay[k]];
    const Real eps = Real(prop_eps);
    const Real op_eps = Real(3) * Limits<Real>::tiny;
    for (int64_t b = 0; b < nbnd; ++b)
    {
        Real t = Real(0), ts = Real(0), tsg = Real(0), ta = Real(0);
#pragma unroll
        for (int k = 0; k < AEROSOL_NSPECIES; ++k)
        {
            const AerosolSpeciesRef sp = aerosol_species_order[k];
            const int64_t it = sp.philic ? b + nbnd * (ihum + nhum * sp.column) : b + nbnd * sp.column;
            const Real mext = sp.philic ? mext_philic[it] : mext_phobic[it];
            const Real sa = sp.philic ? ssa_philic[it] : ssa_phobic[it];
            const Real ga = sp.philic ? g_philic[it] : g_phobic[it];
            const Real od = m[sp.field] * dpg * mext;
            t += od;
            ts += od * sa;
            tsg += od * sa * ga;
            ta += od * (Real(1) - sa);
        }
        Real gg = tsg / max(ts, eps);
        if (coalb)
        {
            // The co-albedo 1 - ssa from the absorption optical depth (see cloud_optics_kernel).
            Real co = (t > eps) ? ta / t : Real(1) - ts / eps;
            if (delta_scale)
            {
                const Real f = gg * gg;
                const Real d = (Real(1) - f) + f * co;
                t *= d;
                co = co / max(op_eps, d);
                gg = (gg - f) / max(op_eps, Real(1) - f);
            }
            const int64_t o = i + ncl * b;
            tau[o] = t; ssa[o] = co; g[o] = gg;
            continue;
        }
        Real s = ts / max(t, eps);
        if (delta_scale)
        {
            const Real f = gg * gg, wf = s * f;
            t *= (Real(1) - wf);
            s = (s - wf) / max(op_eps, Real(1) - wf);
            gg = (gg - f) / max(op_eps, Real(1) - f);
        }
        const int64_t o = i + ncl * b;
        tau[o] = t; ssa[o] = s; g[o] = gg;
    }
}

template<typename Real>
__global__ void increment_kernel(int64_t n, Real* __restrict__ tau1, Real* __restrict__ ssa1,
                                 Real* __restrict__ g1, const Real* __restrict__ tau2,
                                 const Real* __restrict__ ssa2, const Real* __restrict__ g2)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!ssa1) { tau1[i] = tau1[i] + tau2[i]; return; }
    const Real eps = Real(3) * Limits<Real>::tiny;
    const Real t1 = tau1[i], t2 = tau2[i], s1 = ssa1[i], s2 = ssa2[i];
    const Real tau12 = t1 + t2;
    const Real tauscat12 = (t1 * s1) + (t2 * s2);
    g1[i] = ((t1 * s1 * g1[i]) + (t2 * s2 * g2[i])) / max(tauscat12, eps);
    ssa1[i] = tauscat12 / max(eps, tau12);
    tau1[i] = tau12;
}

// inc_*_bybnd: thread per (col, lay, gpt); the band of a g-point is found by scanning the
// (few) band limits, as the reference kernel does (optical_props_kernels.cu:86-95).
template<typename Real>
__global__ void inc_bybnd_kernel(int64_t ncol, int64_t nlay, int64_t ngpt, Real* __restrict__ tau1,
                                 Real* __restrict__ ssa1, Real* __restrict__ g1,
                                 const Real* __restrict__ tau2, const Real* __restrict__ ssa2,
                                 const Real* __restrict__ g2, int64_t nbnd, const int* __restrict__ band_lims,
                                 int coalb = 0)
{
    const int64_t ncl = ncol * nlay;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncl * ngpt) return;
    const int64_t cl = i % ncl;
    const int gpt1 = int(i / ncl) + 1;
    const Real eps = Real(3) * Limits<Real>::tiny;
    for (int64_t b = 0; b < nbnd; ++b)
    {
        if (gpt1 >= band_lims[2 * b] && gpt1 <= band_lims[2 * b + 1])
        {
            const int64_t ib = cl + ncl * b;
            if (!ssa1) { tau1[i] = tau1[i] + tau2[ib]; continue; }
            if (coalb)
            {
                // ssa1, ssa2 hold co-albedos 1 - ssa: the same increment in those terms.
                const Real t1 = tau1[i], t2 = tau2[ib], c1 = ssa1[i], c2 = ssa2[ib];
                const Real tau12 = t1 + t2;
                const Real tauscat12 = (t1 * (Real(1) - c1)) + (t2 * (Real(1) - c2));
                g1[i] = ((t1 * (Real(1) - c1) * g1[i]) + (t2 * (Real(1) - c2) * g2[ib])) / max(tauscat12, eps);
                ssa1[i] = (tau12 > eps) ? ((t1 * c1) + (t2 * c2)) / tau12 : Real(1) - tauscat12 / eps;
                tau1[i] = tau12;
                continue;
            }
            const Real t1 = tau1[i], t2 = tau2[ib], s1 = ssa1[i], s2 = ssa2[ib];
            const Real tau12 = t1 + t2;
            const Real tauscat12 = (t1 * s1) + (t2 * s2);
            g1[i] = ((t1 * s1 * g1[i]) + (t2 * s2 * g2[ib])) / max(tauscat12, eps);
            ssa1[i] = tauscat12 / max(eps, tau12);
            tau1[i] = tau12;
        }
    }
}

template<typename Real>
__global__ void delta_scale_kernel(int64_t n, Real* __restrict__ tau, Real* __restrict__ ssa, Real* __restrict__ g)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Real eps = Real(3) * Limits<Real>::tiny;
    const Real gv = g[i], sv = ssa[i];
    const Real f = gv * gv, wf = sv * f;
    tau[i] *= (Real(1) - wf);
    ssa[i] = (sv - wf) / max(eps, Real(1) - wf);
    g[i] = (gv - f) / max(eps, Real(1) - f);
}

// Broadband sum over g-points in order (fluxes_kernels.cu:26-43).
template<typename Real>
__global__ void sum_broadband_kernel(int64_t ncv, int64_t ngpt, const Real* __restrict__ f, Real* __restrict__ out)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncv) return;
    Real s = Real(0);
    for (int64_t g = 0; g < ngpt; ++g) s += f[i + ncv * g];
    out[i] = s;
}

template<typename Real>
__global__ void net_kernel(int64_t n, const Real* __restrict__ dn, const Real* __restrict__ up, Real* __restrict__ net)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    net[i] = dn[i] - up[i];
}

// sum_byband (mo_fluxes_byband_kernels.F90:22-41).
template<typename Real>
__global__ void sum_byband_kernel(int64_t ncv, int64_t nbnd, const int* __restrict__ bl,
                                  const Real* __restrict__ f, Real* __restrict__ out)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncv * nbnd) return;
    const int64_t cv = i % ncv, b = i / ncv;
    const int gs = bl[2 * b], ge = bl[2 * b + 1];
    Real s = f[cv + ncv * (gs - 1)];
    for (int g = gs; g < ge; ++g) s += f[cv + ncv * g];
    out[i] = s;
}

// net_byband_full (mo_fluxes_byband_kernels.F90:46-69): (net + dn) - up in g-point order.
template<typename Real>
__global__ void net_byband_full_kernel(int64_t ncv, int64_t nbnd, const int* __restrict__ bl,
                                       const Real* __restrict__ dn, const Real* __restrict__ up,
                                       Real* __restrict__ out)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncv * nbnd) return;
    const int64_t cv = i % ncv, b = i / ncv;
    const int gs = bl[2 * b], ge = bl[2 * b + 1];
    Real s = dn[cv + ncv * (gs - 1)] - up[cv + ncv * (gs - 1)];
    for (int g = gs; g < ge; ++g) s = s + dn[cv + ncv * g] - up[cv + ncv * g];
    out[i] = s;
}

// Heating rate with the net-flux and pressure differences in double precision.
template<typename Real>
__global__ void heating_rate_kernel(int64_t ncol, int64_t nlay, const Real* __restrict__ up,
                                    const Real* __restrict__ dn, const Real* __restrict__ plev,
                                    Real* __restrict__ hr)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncol * nlay) return;
    const int64_t j = i + ncol;
    const double net_a = double(dn[i]) - double(up[i]);
    const double net_b = double(dn[j]) - double(up[j]);
    const double g0 = 9.80665, cp = 1004.64, sec_per_day = 86400.0;
    hr[i] = Real(g0 / cp * (net_a - net_b) / (double(plev[j]) - double(plev[i])) * sec_per_day);
}

// Adjoint of heating_rate_kernel: hr_i = c (net_i - net_{i+1}) / (p_{i+1} - p_i), net = dn - up.
// Adds into g_up, g_dn (ncol, nlev) and d_plev (ncol, nlev); one thread per column walks its
// layers, so neighbouring layers' updates of a shared level do not race.
template<typename Real>
__global__ void heating_rate_adjoint_kernel(int64_t ncol, int64_t nlay, const Real* __restrict__ up,
                                            const Real* __restrict__ dn, const Real* __restrict__ plev,
                                            const Real* __restrict__ g_hr, Real* __restrict__ g_up,
                                            Real* __restrict__ g_dn, Real* __restrict__ d_plev)
{
    const int64_t icol = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (icol >= ncol) return;
    const double c = 9.80665 / 1004.64 * 86400.0;
    for (int64_t ilay = 0; ilay < nlay; ++ilay)
    {
        const int64_t i = icol + ncol * ilay, j = i + ncol;
        const double net_a = double(dn[i]) - double(up[i]);
        const double net_b = double(dn[j]) - double(up[j]);
        const double dp = double(plev[j]) - double(plev[i]);
        const double g = double(g_hr[i]);
        const double gn = c * g / dp;                               // d hr / d net_a = -d hr / d net_b
        const double gp = c * g * (net_a - net_b) / (dp * dp);      // -d hr / d p_j = d hr / d p_i
        g_dn[i] += Real(gn); g_up[i] -= Real(gn);
        g_dn[j] -= Real(gn); g_up[j] += Real(gn);
        d_plev[i] += Real(gp);
        d_plev[j] -= Real(gp);
    }
}

// expand_and_transpose (src_cuda/Rte_lw.cu:36-56): (nbnd, ncol) -> (ncol, ngpt).
template<typename Real>
__global__ void expand_kernel(int64_t ncol, int64_t nbnd, int64_t ngpt, const int* __restrict__ bl,
                              const Real* __restrict__ in, Real* __restrict__ out)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncol * ngpt) return;
    const int64_t icol = i % ncol;
    const int gpt1 = int(i / ncol) + 1;
    for (int64_t b = 0; b < nbnd; ++b)
        if (gpt1 >= bl[2 * b] && gpt1 <= bl[2 * b + 1]) out[i] = in[b + nbnd * icol];
}

// apply_BC variants (rte_solver_kernels.cu:342-379): 0, inc_flux, inc_flux*factor(icol).
template<typename Real>
__global__ void apply_bc_kernel(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1,
                                const Real* __restrict__ inc, const Real* __restrict__ factor,
                                Real* __restrict__ flux_dn)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncol * ngpt) return;
    const int64_t icol = i % ncol, igpt = i / ncol;
    const int64_t o = icol + (top_at_1 ? 0 : nlay * ncol) + igpt * ncol * (nlay + 1);
    Real v = inc ? inc[i] : Real(0);
    if (inc && factor) v = inc[i] * factor[icol];
    flux_dn[o] = v;
}
template<typename Real>
__global__ void spread_col_kernel(int64_t ncol, int64_t ngpt, const Real* __restrict__ src, Real* __restrict__ dst)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncol * ngpt) return;
    dst[i] = src[i / ncol];
}

template<typename Real>
__global__ void scale_by_col_kernel(int64_t ncol, int64_t n2, Real* __restrict__ arr, const Real* __restrict__ f)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncol * n2) return;
    arr[i] *= f[i % ncol];
}

template<typename Real>
__global__ void copy_cols_kernel(const Real* __restrict__ src, int64_t src_ncol, int64_t src_col0,
                                 Real* __restrict__ dst, int64_t dst_ncol, int64_t dst_col0,
                                 int64_t ncopy, int64_t nrows)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncopy * nrows) return;
    const int64_t c = i % ncopy, r = i / ncopy;
    dst[dst_col0 + c + dst_ncol * r] = src[src_col0 + c + src_ncol * r];
}
}  // namespace rrtmgp_hip

using namespace rrtmgp_hip;

#define RRTMGP_DEF_OPTPROPS(S, T)                                                                   \
extern "C" int rrtmgp_hip_cloud_optics_##S(int64_t ncol, int64_t nlay, int64_t nbnd,              \
    int64_t liq_nsteps, T liq_step_size, T radliq_lwr, const T* lut_extliq, const T* lut_ssaliq,    \
    const T* lut_asyliq, int64_t ice_nsteps, T ice_step_size, T radice_lwr, const T* lut_extice,    \
    const T* lut_ssaice, const T* lut_asyice, const T* clwp, const T* ciwp, const T* reliq,         \
    const T* reice, int two_stream, int delta_scale, T* tau, T* ssa, T* g, void* stream)            \
{                                                                                                    \
    if (ncol <= 0 || nlay <= 0 || nbnd <= 0) return arg_error("cloud_optics: empty extents");      \
    if (liq_nsteps < 2 || ice_nsteps < 2) return arg_error("cloud_optics: LUT needs >= 2 sizes");  \
    if (two_stream && (!ssa || !g)) return arg_error("cloud_optics: ssa/g required for 2str");     \
    const int64_t n = ncol * nlay;                                                                   \
    RRTMGP_TIMED("cloud_optics_kernel", as_stream(stream));                                          \
    hipLaunchKernelGGL(cloud_optics_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),\
                       ncol, nlay, nbnd, liq_nsteps, liq_step_size, radliq_lwr, lut_extliq,         \
                       lut_ssaliq, lut_asyliq, ice_nsteps, ice_step_size, radice_lwr, lut_extice,   \
                       lut_ssaice, lut_asyice, clwp, ciwp, reliq, reice, two_stream, delta_scale,   \
                       tau, ssa, g);                                                                \
    RRTMGP_LAUNCHED("cloud_optics_kernel");                                                         \
    return RRTMGP_OK;                                                                                \
}                                                                                                    \
extern "C" int rrtmgp_hip_aerosol_optics_##S(int64_t ncol, int64_t nlay, int64_t nbnd,            \
    int64_t nhum, int64_t nphobic, int64_t nphilic, const rrtmgp_gas_vmr* aermr, const T* rh,       \
    const T* plev, const T* rh_upper, const T* mext_phobic, const T* ssa_phobic, const T* g_phobic, \
    const T* mext_philic, const T* ssa_philic, const T* g_philic, int delta_scale, int coalbedo,    \
    T* tau, T* ssa, T* g, void* stream)                                                              \
{                                                                                                    \
    if (ncol <= 0 || nlay <= 0 || nbnd <= 0 || nhum <= 0) return arg_error("aerosol_optics: empty extents"); \
    if (nphobic < 11 || nphilic < 5)                                                                 \
        return arg_error("aerosol_optics: tables need 11 hydrophobic and 5 hydrophilic species");  \
    if (!aermr || aermr->ngas != 11) return arg_error("aerosol_optics: 11 mixing ratios (aermr01..aermr11) expected"); \
    if (!rh || !plev || !rh_upper || !mext_phobic || !ssa_phobic || !g_phobic || !mext_philic ||    \
        !ssa_philic || !g_philic || !tau || !ssa || !g)                                              \
        return arg_error("aerosol_optics: NULL array");                                             \
    AerosolMix<T> mix;                                                                               \
    for (int k = 0; k < AEROSOL_NSPECIES; ++k)                                                       \
    {                                                                                                \
        const int64_t d1 = aermr->dim1[k], d2 = aermr->dim2[k];                                      \
        if (!aermr->vmr[k] || !((d1 == 1 || d1 == ncol) && (d2 == 1 || d2 == nlay)))                \
            return arg_error("aerosol_optics: aermr" + std::to_string(k + 1) + " is NULL or not (1|ncol, 1|nlay)"); \
        mix.p[k] = static_cast<const T*>(aermr->vmr[k]);                                             \
        mix.scol[k] = d1 == 1 ? 0 : 1;                                                               \
        mix.slay[k] = d2 == 1 ? 0 : d1;                                                              \
    }                                                                                                \
    const int64_t n = ncol * nlay;                                                                   \
    RRTMGP_TIMED("aerosol_optics_kernel", as_stream(stream));                                        \
    hipLaunchKernelGGL(aerosol_optics_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream), \
                       ncol, nlay, nbnd, nhum, mix, rh, plev, rh_upper, mext_phobic, ssa_phobic,     \
                       g_phobic, mext_philic, ssa_philic, g_philic, delta_scale, coalbedo, tau, ssa, \
                       g);                                                                           \
    RRTMGP_LAUNCHED("aerosol_optics_kernel");                                                       \
    return RRTMGP_OK;                                                                                \
}                                                                                                    \
extern "C" int rrtmgp_hip_increment_##S(int64_t ncol, int64_t nlay, int64_t ngpt, T* tau_inout,   \
    T* ssa_inout, T* g_inout, const T* tau_in, const T* ssa_in, const T* g_in, void* stream)        \
{                                                                                                    \
    const int64_t n = ncol * nlay * ngpt;                                                            \
    RRTMGP_TIMED("increment_kernel", as_stream(stream));                                             \
    hipLaunchKernelGGL(increment_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),   \
                       n, tau_inout, ssa_inout, g_inout, tau_in, ssa_in, g_in);                     \
    RRTMGP_LAUNCHED("increment_kernel");                                                            \
    return RRTMGP_OK;                                                                                \
}                                                                                                    \
extern "C" int rrtmgp_hip_inc_bybnd_##S(int64_t ncol, int64_t nlay, int64_t ngpt, T* tau_inout,   \
    T* ssa_inout, T* g_inout, const T* tau_in, const T* ssa_in, const T* g_in, int64_t nbnd,        \
    const int* band_lims_gpt, void* stream)                                                          \
{                                                                                                    \
    const int64_t n = ncol * nlay * ngpt;                                                            \
    RRTMGP_TIMED("inc_bybnd_kernel", as_stream(stream));                                             \
    hipLaunchKernelGGL(inc_bybnd_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),   \
                       ncol, nlay, ngpt, tau_inout, ssa_inout, g_inout, tau_in, ssa_in, g_in, nbnd, \
                       band_lims_gpt);                                                               \
    RRTMGP_LAUNCHED("inc_bybnd_kernel");                                                            \
    return RRTMGP_OK;                                                                                \
}                                                                                                    \
extern "C" int rrtmgp_hip_inc_bybnd_coalbedo_##S(int64_t ncol, int64_t nlay, int64_t ngpt,       \
    T* tau_inout, T* coalbedo_inout, T* g_inout, const T* tau_in, const T* coalbedo_in,             \
    const T* g_in, int64_t nbnd, const int* band_lims_gpt, void* stream)                             \
{                                                                                                    \
    if (!coalbedo_inout || !coalbedo_in || !g_inout || !g_in)                                        \
        return arg_error("inc_bybnd_coalbedo: two-stream arrays required");                         \
    const int64_t n = ncol * nlay * ngpt;                                                            \
    RRTMGP_TIMED("inc_bybnd_kernel", as_stream(stream));                                             \
    hipLaunchKernelGGL(inc_bybnd_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),   \
                       ncol, nlay, ngpt, tau_inout, coalbedo_inout, g_inout, tau_in, coalbedo_in,   \
                       g_in, nbnd, band_lims_gpt, 1);                                                \
    RRTMGP_LAUNCHED("inc_bybnd_kernel");                                                            \
    return RRTMGP_OK;                                                                                \
}                                                                                                    \
extern "C" int rrtmgp_hip_delta_scale_2str_##S(int64_t ncol, int64_t nlay, int64_t ngpt, T* tau,  \
    T* ssa, T* g, void* stream)                                                                      \
{                                                                                                    \
    const int64_t n = ncol * nlay * ngpt;                                                            \
    RRTMGP_TIMED("delta_scale_kernel", as_stream(stream));                                           \
    hipLaunchKernelGGL(delta_scale_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream), \
                       n, tau, ssa, g);                                                              \
    RRTMGP_LAUNCHED("delta_scale_kernel");                                                          \
    return RRTMGP_OK;                                                                                \
}                                                                                                    \
extern "C" int rrtmgp_hip_sum_broadband_##S(int64_t ncol, int64_t nlev, int64_t ngpt,             \
    const T* gpt_flux, T* flux, void* stream)                                                        \
{                                                                                                    \
    const int64_t n = ncol * nlev;                                                                   \
    RRTMGP_TIMED("sum_broadband_kernel", as_stream(stream));                                         \
    hipLaunchKernelGGL(sum_broadband_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0,                  \
                       as_stream(stream), n, ngpt, gpt_flux, flux);                                 \
    RRTMGP_LAUNCHED("sum_broadband_kernel");                                                        \
    return RRTMGP_OK;                                                                                \
}                                                                                                    \
extern "C" int rrtmgp_hip_net_broadband_##S(int64_t ncol, int64_t nlev, const T* flux_dn,         \
    const T* flux_up, T* flux_net, void* stream)                                                     \
{                                                                                                    \
    const int64_t n = ncol * nlev;                                                                   \
    RRTMGP_TIMED("net_kernel", as_stream(stream));                                                   \
    hipLaunchKernelGGL(net_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream), n,      \
                       flux_dn, flux_up, flux_net);                                                  \
    RRTMGP_LAUNCHED("net_kernel");                                                                  \
    return RRTMGP_OK;                                                                                \
}                                                                                                    \
extern "C" int rrtmgp_hip_sum_byband_##S(int64_t ncol, int64_t nlev, int64_t ngpt, int64_t nbnd,  \
    const int* band_lims, const T* gpt_flux, T* bnd_flux, void* stream)                             \
{                                                                                                    \
    const int64_t n = ncol * nlev * nbnd;                                                            \
    RRTMGP_TIMED("sum_byband_kernel", as_stream(stream));                                            \
    hipLaunchKernelGGL(sum_byband_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),  \
                       ncol * nlev, nbnd, band_lims, gpt_flux, bnd_flux);                           \
    RRTMGP_LAUNCHED("sum_byband_kernel");                                                           \
    (void)ngpt;                                                                                      \
    return RRTMGP_OK;                                                                                \
}                                                                                                    \
extern "C" int rrtmgp_hip_net_byband_full_##S(int64_t ncol, int64_t nlev, int64_t ngpt,           \
    int64_t nbnd, const int* band_lims, const T* gpt_flux_dn, const T* gpt_flux_up,                 \
    T* bnd_flux_net, void* stream)                                                                   \
{                                                                                                    \
    const int64_t n = ncol * nlev * nbnd;                                                            \
    RRTMGP_TIMED("net_byband_full_kernel", as_stream(stream));                                       \
    hipLaunchKernelGGL(net_byband_full_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0,                \
                       as_stream(stream), ncol * nlev, nbnd, band_lims, gpt_flux_dn, gpt_flux_up,   \
                       bnd_flux_net);                                                                \
    RRTMGP_LAUNCHED("net_byband_full_kernel");                                                      \
    (void)ngpt;                                                                                      \
    return RRTMGP_OK;                                                                                \
}                                                                                                    \
extern "C" int rrtmgp_hip_heating_rate_##S(int64_t ncol, int64_t nlay, const T* flux_up,          \
    const T* flux_dn, const T* plev, T* heating_rate, void* stream)                                 \
{                                                                                                    \
    const int64_t n = ncol * nlay;                                                                   \
    RRTMGP_TIMED("heating_rate_kernel", as_stream(stream));                                          \
    hipLaunchKernelGGL(heating_rate_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0,                   \
                       as_stream(stream), ncol, nlay, flux_up, flux_dn, plev, heating_rate);        \
    RRTMGP_LAUNCHED("heating_rate_kernel");                                                         \
    return RRTMGP_OK;                                                                                \
}                                                                                                    \
extern "C" int rrtmgp_hip_heating_rate_adjoint_##S(int64_t ncol, int64_t nlay, const T* flux_up,  \
    const T* flux_dn, const T* plev, const T* g_heating_rate, T* g_flux_up, T* g_flux_dn, T* d_plev, \
    void* stream)                                                                                    \
{                                                                                                    \
    if (ncol <= 0 || nlay <= 0 || !flux_up || !flux_dn || !plev || !g_heating_rate || !g_flux_up ||  \
        !g_flux_dn || !d_plev)                                                                       \
        return arg_error("heating_rate_adjoint: bad arguments");                                    \
    RRTMGP_TIMED("heating_rate_adjoint_kernel", as_stream(stream));                                  \
    hipLaunchKernelGGL(heating_rate_adjoint_kernel<T>, dim3(grid1(ncol, 256)), dim3(256), 0,        \
                       as_stream(stream), ncol, nlay, flux_up, flux_dn, plev, g_heating_rate,       \
                       g_flux_up, g_flux_dn, d_plev);                                               \
    RRTMGP_LAUNCHED("heating_rate_adjoint_kernel");                                                 \
    return RRTMGP_OK;                                                                                \
}                                                                                                    \
extern "C" int rrtmgp_hip_expand_and_transpose_##S(int64_t ncol, int64_t nbnd, int64_t ngpt,      \
    const int* band_lims_gpt, const T* arr_in, T* arr_out, void* stream)                            \
{                                                                                                    \
    const int64_t n = ncol * ngpt;                                                                   \
    RRTMGP_TIMED("expand_kernel", as_stream(stream));                                                \
    hipLaunchKernelGGL(expand_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),      \
                       ncol, nbnd, ngpt, band_lims_gpt, arr_in, arr_out);                           \
    RRTMGP_LAUNCHED("expand_kernel");                                                               \
    return RRTMGP_OK;                                                                                \
}                                                                                                    \
extern "C" int rrtmgp_hip_apply_bc_##S(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1,    \
    const T* inc_flux, const T* factor, T* flux_dn, void* stream)                                   \
{                                                                                                    \
    const int64_t n = ncol * ngpt;                                                                   \
    RRTMGP_TIMED("apply_bc_kernel", as_stream(stream));                                              \
    hipLaunchKernelGGL(apply_bc_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),    \
                       ncol, nlay, ngpt, top_at_1, inc_flux, factor, flux_dn);                      \
    RRTMGP_LAUNCHED("apply_bc_kernel");                                                             \
    return RRTMGP_OK;                                                                                \
}                                                                                                    \
extern "C" int rrtmgp_hip_spread_col_##S(int64_t ncol, int64_t ngpt, const T* src, T* dst,      \
    void* stream)                                                                                    \
{                                                                                                    \
    const int64_t n = ncol * ngpt;                                                                   \
    RRTMGP_TIMED("spread_col_kernel", as_stream(stream));                                            \
    hipLaunchKernelGGL(spread_col_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),  \
                       ncol, ngpt, src, dst);                                                        \
    RRTMGP_LAUNCHED("spread_col_kernel");                                                           \
    return RRTMGP_OK;                                                                                \
}                                                                                                    \
extern "C" int rrtmgp_hip_scale_by_col_##S(int64_t ncol, int64_t n2, T* arr, const T* factor,     \
    void* stream)                                                                                    \
{                                                                                                    \
    const int64_t n = ncol * n2;                                                                     \
    RRTMGP_TIMED("scale_by_col_kernel", as_stream(stream));                                          \
    hipLaunchKernelGGL(scale_by_col_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0,                   \
                       as_stream(stream), ncol, n2, arr, factor);                                   \
    RRTMGP_LAUNCHED("scale_by_col_kernel");                                                         \
    return RRTMGP_OK;                                                                                \
}                                                                                                    \
extern "C" int rrtmgp_hip_copy_cols_##S(const T* src, int64_t src_ncol, int64_t src_col0, T* dst, \
    int64_t dst_ncol, int64_t dst_col0, int64_t ncopy, int64_t nrows, void* stream)                \
{                                                                                                    \
    const int64_t n = ncopy * nrows;                                                                 \
    if (n == 0) return RRTMGP_OK;                                                                    \
    RRTMGP_TIMED("copy_cols_kernel", as_stream(stream));                                             \
    hipLaunchKernelGGL(copy_cols_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),   \
                       src, src_ncol, src_col0, dst, dst_ncol, dst_col0, ncopy, nrows);             \
    RRTMGP_LAUNCHED("copy_cols_kernel");                                                            \
    return RRTMGP_OK;                                                                                \
}

RRTMGP_DEF_OPTPROPS(f32, float)
RRTMGP_DEF_OPTPROPS(f64, double)
