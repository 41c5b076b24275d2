// Longwave (no-scattering, Gaussian quadrature) and shortwave (two-stream + adding) solvers.
//
// Arithmetic follows src_kernels/mo_rte_solver_kernels.F90 (lw_solver_noscat :58-260,
// lw_source_noscat :616-653, transport :659-723, sw_solver_2stream :512-605,
// sw_dif_and_source :996-1109, adding :1117-1227) with the Rdir/Tdir clamp of
// src_kernels_cuda/rte_solver_kernels.cu:581-583.
//
// Layout of the work.  A work-group is 64 columns (the lanes of each wavefront) x up to 16
// g-points of one band (one wavefront each).  Each thread runs the layer recurrences of one
// (column, g-point); every layer step of a wavefront reads one contiguous 256-byte row of the
// (ncol, nlay, ngpt) inputs.  Fluxes are summed over the g-points of the work-group in LDS
// every K levels (a fixed order, so results are reproducible), giving per-band partial fluxes
// without materialising (ncol, nlev, ngpt) g-point fluxes; a small kernel then forms band and
// broadband fluxes.  Callers that ask for g-point fluxes get them written directly instead.
//
// Shortwave adding.  The reference runs the adding method as an albedo/source sweep up from
// the surface followed by a flux sweep down, keeping per-layer R, T and sources in memory.
// Here the same layer equations are solved two-sided: a sweep down from the top carries the
// reflectance R_up and diffuse source D of the column above each level (and the direct
// beam), storing them only every K levels; the sweep up from the surface carries the albedo
// A and source S of the column below, replaying each K-layer segment from its checkpoint in
// registers, and at every level the fluxes follow from both sides:
//     dn = (D + R_up S) / (1 - R_up A),   up = A dn + S.
// Per (column, g-point) the memory traffic is the inputs read twice plus a checkpoint per K
// levels, instead of four layer arrays written and read back.
//
// Precision.  The adding denominators 1/(1 - R A) and the carried adding states (A, S, R_up, D,
// the direct beam) are in the build's precision; the FP32 build forms the direct beam from the
// optical depth summed in FP64.  The two-stream layer coefficients are FP64 in the FP64 build;
// the FP32 build evaluates them in a rearranged FP32 form (sw_layer_fp32), with the FP64
// k_min = 1e-12 and epsilon, and keeps FP64 for near-conservative layers when given ssa (whose
// FP32 value cannot resolve 1 - ssa below 6e-8), not when given the co-albedo.  Inputs, outputs
// and checkpoints keep the build's precision.
#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "common.h"
#include "solver_math.h"
#include "two_stream.h"

namespace rrtmgp_hip
{
constexpr int SB_COLS = 64;     // columns per work-group: the lanes of each wavefront
constexpr int SB_GPTS = 16;     // g-points per work-group: one wavefront each, within one band
constexpr int MAX_GROUPS = 256;

// Segment lengths: levels per LDS reduction round (LW, SW) and the SW checkpoint spacing.  The
// LW segment inputs are double-buffered in registers (K1 for one angle, K4 for up to four), sized
// for 64 VGPRs in FP32 (two 1024-thread work-groups per CU) without spills; the SW segment
// registers fit the 128 VGPRs of a 1024-thread work-group.
template<typename Real> struct LwSeg;
template<> struct LwSeg<float> { static constexpr int K1 = 6, K4 = 4; };
template<> struct LwSeg<double> { static constexpr int K1 = 4, K4 = 4; };
// The SW kernel's wide variant (FP32, co-albedo input, no heating rates) fits 64 VGPRs with
// K = 3, so two 1024-thread work-groups share a CU; the others keep K = 4 at up to 128 VGPRs.
constexpr bool sw_wide(bool fp32, bool co, bool hr) { return fp32 && co && !hr; }
constexpr int sw_seg_k(bool fp32, bool co, bool hr) { return sw_wide(fp32, co, hr) ? 3 : 4; }

// G-point groups: consecutive g-points of one band, at most SB_GPTS each.
struct Groups
{
    int n;
    int start[MAX_GROUPS];
    int count[MAX_GROUPS];
    int band[MAX_GROUPS];
};

// Groups from band limits (2, nbnd), 1-based inclusive; without band limits the g-points are cut
// into consecutive chunks (enough for broadband sums).
int make_groups(int64_t ngpt, int64_t nbnd, const int* band_lims, Groups& g)
{
    g.n = 0;
    auto add = [&](int s, int e, int b) {
        for (int x = s; x < e; x += SB_GPTS)
        {
            if (g.n == MAX_GROUPS) return false;
            g.start[g.n] = x;
            g.count[g.n] = std::min(SB_GPTS, e - x);
            g.band[g.n] = b;
            ++g.n;
        }
        return true;
    };
    if (band_lims)
    {
        int expect = 1;
        for (int64_t b = 0; b < nbnd; ++b)
        {
            const int s = band_lims[2 * b], e = band_lims[2 * b + 1];
            if (s != expect || e < s || e > ngpt) return arg_error("solver: band limits do not tile the g-points");
            if (!add(s - 1, e, int(b))) return arg_error("solver: too many g-point groups");
            expect = e + 1;
        }
        if (expect != ngpt + 1) return arg_error("solver: band limits do not cover all g-points");
    }
    else if (!add(0, int(ngpt), 0))
        return arg_error("solver: too many g-point groups");
    return RRTMGP_OK;
}

// Writes the LDS slots kk < nslots, components c < NC, summed over the group's g-points, to
// out[c] at level lev_of(kk) (skipped when negative).
// red layout: ((kk * NC + c) * SB_GPTS + w) * SB_COLS + lane.
template<typename Real, int NC, int STRIDE = NC, typename LevOf>
__device__ __forceinline__ void write_group_sums(const Real* red, int nslots, int gcount, int64_t col0,
                                                 int64_t ncol, int64_t nlev, int gi, Real* const (&out)[NC],
                                                 LevOf lev_of)
{
    const int tid = threadIdx.y * SB_COLS + threadIdx.x;
    const int items = nslots * NC * SB_COLS;
    for (int it = tid; it < items; it += SB_COLS * SB_GPTS)
    {
        const int lane = it % SB_COLS;
        const int rest = it / SB_COLS;
        const int c = rest % NC, kk = rest / NC;
        const int64_t icol = col0 + lane;
        const int64_t lev = lev_of(kk);
        if (icol >= ncol || lev < 0) continue;
        const Real* r = red + (int64_t(kk) * STRIDE + c) * SB_GPTS * SB_COLS + lane;
        Real s = r[0];
        for (int w = 1; w < gcount; ++w) s += r[w * SB_COLS];
        out[c][icol + ncol * (lev + nlev * gi)] = s;
    }
}

// write_group_sums for component 0 only of an LDS layout with NC components per slot.
template<typename Real, int NC, typename LevOf>
__device__ __forceinline__ void write_group_sums_stride(const Real* red, int nslots, int gcount, int64_t col0,
                                                        int64_t ncol, int64_t nlev, int gi,
                                                        Real* const (&out)[1], LevOf lev_of)
{
    const int tid = threadIdx.y * SB_COLS + threadIdx.x;
    const int items = nslots * SB_COLS;
    for (int it = tid; it < items; it += SB_COLS * SB_GPTS)
    {
        const int lane = it % SB_COLS;
        const int kk = it / SB_COLS;
        const int64_t icol = col0 + lane;
        const int64_t lev = lev_of(kk);
        if (icol >= ncol || lev < 0) continue;
        const Real* r = red + int64_t(kk) * NC * SB_GPTS * SB_COLS + lane;
        Real s = r[0];
        for (int w = 1; w < gcount; ++w) s += r[w * SB_COLS];
        out[0][icol + ncol * (lev + nlev * gi)] = s;
    }
}

// Heating-rate island: each layer's absorbed flux (net-flux convergence) summed over the group's
// g-points in FP64, for the LDS slots kk < nslots of component comp (NCH components per slot),
// written to out[icol + ncol (layer + nlay gi)] for layer lay_of(kk) (skipped when negative).
template<typename Real, int NCH, typename LayOf>
__device__ __forceinline__ void write_group_layer_sums(const Real* red, int comp, int nslots, int gcount,
                                                       int64_t col0, int64_t ncol, int64_t nlay, int gi,
                                                       double* out, LayOf lay_of)
{
    const int tid = threadIdx.y * SB_COLS + threadIdx.x;
    const int items = nslots * SB_COLS;
    for (int it = tid; it < items; it += SB_COLS * SB_GPTS)
    {
        const int lane = it % SB_COLS;
        const int kk = it / SB_COLS;
        const int64_t icol = col0 + lane;
        const int64_t lay = lay_of(kk);
        if (icol >= ncol || lay < 0) continue;
        const Real* r = red + (int64_t(kk) * NCH + comp) * SB_GPTS * SB_COLS + lane;
        double s = double(r[0]);
        for (int w = 1; w < gcount; ++w) s += double(r[w * SB_COLS]);
        out[icol + ncol * (lay + nlay * gi)] = s;
    }
}

// Heating rate from the per-group absorbed fluxes of nparts sweeps, (ncol, nlay, ngroups) each:
// hr = g0 / cp * (sum over parts and groups, in order) / |p(lev+1) - p(lev)|, in FP64.  Equal to
// g0 / cp (net_i - net_i+1) / (p_i+1 - p_i) with net = dn - up, in either orientation.
template<typename Real>
__global__ void absorbed_heating_rate_kernel(int64_t ncol, int64_t nlay, int ngroups, int nparts,
                                             const double* __restrict__ parts, const Real* __restrict__ plev,
                                             Real* __restrict__ hr)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncol * nlay) return;
    const int64_t ncl = ncol * nlay;
    double s = 0.;
    for (int g = 0; g < ngroups; ++g)
        for (int p = 0; p < nparts; ++p)
            s += parts[i + ncl * (g + int64_t(ngroups) * p)];
    const double dp = fabs(double(plev[i + ncol]) - double(plev[i]));
    const double g0 = 9.80665, cp = 1004.64, sec_per_day = 86400.0;
    hr[i] = Real(g0 / cp * s / dp * sec_per_day);
}

// ---------------------------------------------------------------------------------------------
// Longwave: no scattering, linear-in-tau source, nmus <= NMU Gaussian angles.  The layer inputs
// of the next K-level segment are read while the current segment is computed.
template<typename Real, bool JAC, int NMU, bool FACT, bool HR = false>
__global__ __launch_bounds__(SB_COLS * SB_GPTS) __attribute__((amdgpu_waves_per_eu(sizeof(Real) == 4 ? 8 : 4)))
void lw_solver_kernel(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1, int nmus,
                      const Real* __restrict__ secants, const Real* __restrict__ weights,
                      const Real* __restrict__ tau, const Real* __restrict__ lay_src,
                      const Real* __restrict__ lev_inc, const Real* __restrict__ lev_dec,
                      const Real* __restrict__ emis, const Real* __restrict__ sfc_src,
                      const Real* __restrict__ inc_flux, Real* __restrict__ flux_up,
                      Real* __restrict__ flux_dn, const Real* __restrict__ sfc_src_jac,
                      Groups grp, Real* __restrict__ grp_up, Real* __restrict__ grp_dn,
                      Real* __restrict__ grp_jac, const Real* __restrict__ bnd_lay,
                      const Real* __restrict__ bnd_lev, int64_t emis_nbnd, int secants_per_angle,
                      double* __restrict__ absorbed)
{
    constexpr int K = NMU == 1 ? LwSeg<Real>::K1 : LwSeg<Real>::K4;
    constexpr int NC = JAC ? 2 : 1;
    // With HR, one more LDS component per slot: the layer's absorbed flux of this g-point,
    // -sum_mu fac (S - (1 - T) I_in), from the sweep's own terms (no difference of level fluxes).
    constexpr int NCH = NC + (HR ? 1 : 0);
    __shared__ Real red[K * NCH * SB_GPTS * SB_COLS];

    const int lane = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.y));   // wavefront = g-point
    const int gi = blockIdx.y;
    const int gcount = grp.count[gi];
    const int64_t col0 = int64_t(blockIdx.x) * SB_COLS;
    const int64_t icol = col0 + lane;
    const int64_t igpt = grp.start[gi] + w;
    const bool active = icol < ncol && w < gcount;
    const bool reduce = grp_dn != nullptr;
    const int64_t nlev = nlay + 1;
    const int64_t c2 = icol + ncol * igpt;
    // Per-g-point slabs at wavefront-uniform bases; lane offsets icol + ncol * level (32-bit).
    const int64_t lslab = ncol * nlay * igpt, vslab = ncol * nlev * igpt;
    const Real* tau_g = tau + lslab;
    const Real* lay_g = lay_src + lslab;
    const Real* lev_dn_g = FACT ? nullptr : (top_at_1 ? lev_inc : lev_dec) + lslab;
    const Real* lev_up_g = FACT ? nullptr : (top_at_1 ? lev_dec : lev_inc) + lslab;
    // Factored sources (FACT): lay_src holds Planck fractions; the group's band Planck values per
    // layer and level come from bnd_lay (ncol, nlay, nbnd) and bnd_lev (ncol, nlev, nbnd).
    const Real* blay_g = FACT ? bnd_lay + ncol * nlay * grp.band[gi] : nullptr;
    const Real* blev_g = FACT ? bnd_lev + ncol * nlev * grp.band[gi] : nullptr;
    Real* flux_up_g = flux_up ? flux_up + vslab : nullptr;
    Real* flux_dn_g = flux_dn ? flux_dn + vslab : nullptr;
    const uint32_t ucol = uint32_t(lane) + uint32_t(col0), uncol = uint32_t(ncol);
    const Real pi = Real(3.14159265358979323846);

    Real D[NMU], fac[NMU], I[NMU], Ij[NMU];
#pragma unroll
    for (int mu = 0; mu < NMU; ++mu)
    {
        D[mu] = Real(0); fac[mu] = Real(0); I[mu] = Real(0); Ij[mu] = Real(0);
        if (active && mu < nmus)
        {
            // Secants per (column, g-point, angle) as the reference's lw_secants_array fills them,
            // or one secant per angle (the only form the class API can express).
            D[mu] = secants_per_angle ? secants[mu] : secants[c2 + ncol * ngpt * mu];
            fac[mu] = Real(2) * pi * weights[mu];
            I[mu] = (inc_flux ? inc_flux[c2] : Real(0)) / fac[mu];
        }
    }

    // Inputs of levels s0 .. s0+K-1 of a sweep (level j > 0 ends layer layer_of(j)).  With factored
    // sources the level source of layer L is at level L + lsh (lev_src_inc: L + 1, lev_src_dec: L),
    // and the products are the gas optics' own (no contraction), so the sources are bit-identical.
    // The up sweep (nt) reads tau and the Planck fractions for the last time and streams them; the
    // down sweep's reads stay cacheable, so the bottom layers it read last are partly still in the
    // last-level cache when the up sweep starts there.
    Real bt[K], bl[K], bx[K];
    auto fetch = [&](int64_t s0, const Real* lev_g, int lsh, auto layer_of, bool nt) {
#pragma unroll
        for (int kk = 0; kk < K; ++kk)
        {
            const int64_t j = s0 + kk;
            bt[kk] = Real(0); bl[kk] = Real(0); bx[kk] = Real(0);
            if (active && j >= 1 && j <= nlay)
            {
                const uint32_t L = uint32_t(layer_of(j));
                const uint32_t l = ucol + uncol * L;
                bt[kk] = nt ? ld_nt(tau_g, l) : at32(tau_g, l);
                if constexpr (FACT)
                {
#pragma clang fp contract(off)
                    const Real pf = nt ? ld_nt(lay_g, l) : at32(lay_g, l);
                    bl[kk] = pf * at32(blay_g, l);
                    bx[kk] = pf * at32(blev_g, ucol + uncol * (L + uint32_t(lsh)));
                }
                else
                {
                    bl[kk] = at32(lay_g, l); bx[kk] = at32(lev_g, l);
                }
            }
        }
    };
    // One sweep over levels 0..nlay; lev_out(j) is the output level of sweep level j.
    auto sweep = [&](const Real* lev_g, int lsh, auto layer_of, auto lev_out, Real* flux_g, Real* grp_out, bool up) {
        fetch(0, lev_g, lsh, layer_of, up);
        for (int64_t s0 = 0; s0 <= nlay; s0 += K)
        {
            Real ct[K], cl[K], cx[K];
#pragma unroll
            for (int kk = 0; kk < K; ++kk) { ct[kk] = bt[kk]; cl[kk] = bl[kk]; cx[kk] = bx[kk]; }
            if (s0 + K <= nlay) fetch(s0 + K, lev_g, lsh, layer_of, up);
#pragma unroll
            for (int kk = 0; kk < K; ++kk)
            {
                const int64_t j = s0 + kk;
                Real f = Real(0), fj = Real(0), ab = Real(0);
                if (active && j <= nlay)
                {
                    if (j > 0)
                    {
#pragma unroll
                        for (int mu = 0; mu < NMU; ++mu)
                            if (mu < nmus)
                            {
                                Real T, fact;
                                const Real t = ct[kk] * D[mu];
                                lw_layer(t, T, fact);
                                if constexpr (HR)
                                    ab -= fac[mu] * (one_minus_T(t, T) * (cx[kk] - I[mu]) +
                                                     Real(2) * fact * (cl[kk] - cx[kk]));
                                I[mu] = T * I[mu] + ((Real(1) - T) * cx[kk] + Real(2) * fact * (cl[kk] - cx[kk]));
                                if (JAC && up) Ij[mu] = T * Ij[mu];
                            }
                    }
#pragma unroll
                    for (int mu = 0; mu < NMU; ++mu)
                        if (mu < nmus)
                        {
                            f = (mu == 0) ? fac[0] * I[0] : f + fac[mu] * I[mu];
                            if (JAC && up) fj = (mu == 0) ? fac[0] * Ij[0] : fj + fac[mu] * Ij[mu];
                        }
                    if (flux_g) at32(flux_g, ucol + uncol * uint32_t(lev_out(j))) = f;
                }
                if (reduce)
                {
                    red[(kk * NCH) * SB_GPTS * SB_COLS + w * SB_COLS + lane] = f;
                    if (JAC && up) red[(kk * NCH + 1) * SB_GPTS * SB_COLS + w * SB_COLS + lane] = fj;
                    if constexpr (HR) red[(kk * NCH + NC) * SB_GPTS * SB_COLS + w * SB_COLS + lane] = ab;
                }
            }
            if (reduce)
            {
                __syncthreads();
                const int nsl = int(std::min<int64_t>(K, nlay + 1 - s0));
                auto lev_of = [&](int kk) -> int64_t { return lev_out(s0 + kk); };
                if (JAC && up)
                {
                    Real* const outs[2] = {grp_out, grp_jac};
                    write_group_sums<Real, 2, NCH>(red, nsl, gcount, col0, ncol, nlev, gi, outs, lev_of);
                }
                else
                {
                    // NCH keeps the component stride of the Jacobian / heating-rate layouts.
                    Real* const outs[1] = {grp_out};
                    write_group_sums_stride<Real, NCH>(red, nsl, gcount, col0, ncol, nlev, gi, outs, lev_of);
                }
                if constexpr (HR)
                {
                    // Sweep level j > 0 closes layer layer_of(j); the down sweep's parts go to
                    // absorbed[.., 0], the up sweep's to absorbed[.., 1].
                    double* part = absorbed + (up ? ncol * nlay * int64_t(grp.n) : int64_t(0));
                    write_group_layer_sums<Real, NCH>(red, NC, nsl, gcount, col0, ncol, nlay, gi, part,
                                                      [&](int kk) -> int64_t {
                                                          const int64_t jj = s0 + kk;
                                                          return jj >= 1 ? int64_t(layer_of(jj)) : int64_t(-1);
                                                      });
                }
                __syncthreads();
            }
        }
    };

    // Down sweep: level j counted from the top (0 = top boundary).
    sweep(lev_dn_g, top_at_1 ? 1 : 0,
          [&](int64_t j) -> int64_t { return top_at_1 ? j - 1 : nlay - j; },
          [&](int64_t j) -> int64_t { return top_at_1 ? j : nlay - j; }, flux_dn_g, grp_dn, false);

    // Surface: reflection of the downward intensity plus emission; then the up sweep with
    // level j counted from the surface (0 = surface).
    {
        // Surface emissivity per (column, g-point), or per (band, column) read directly.
        const Real e = !active ? Real(0) : (emis_nbnd > 0 ? emis[grp.band[gi] + emis_nbnd * icol] : emis[c2]);
        const Real ss = active ? sfc_src[c2] : Real(0);
        const Real sj = (active && JAC) ? sfc_src_jac[c2] : Real(0);
#pragma unroll
        for (int mu = 0; mu < NMU; ++mu)
        {
            I[mu] = I[mu] * (Real(1) - e) + e * ss;
            Ij[mu] = e * sj;
        }
    }
    sweep(lev_up_g, top_at_1 ? 0 : 1,
          [&](int64_t j) -> int64_t { return top_at_1 ? nlay - j : j - 1; },
          [&](int64_t j) -> int64_t { return top_at_1 ? nlay - j : j; }, flux_up_g, grp_up, true);
}

// ---------------------------------------------------------------------------------------------
// Shortwave two-stream layer properties in FP64 (sw_two_stream_t of two_stream.h).
__device__ __forceinline__ void sw_two_stream(double tau, double w0, double g, double mu0, double& Rdif,
                                              double& Tdif, double& Rdir, double& Tdir, double& Tnoscat)
{
    sw_two_stream_t<double>(tau, w0, g, mu0, Rdif, Tdif, Rdir, Tdir, Tnoscat);
}

// One layer of the top-down sweep: diffuse R/T, the direct-beam sources and transmittance, in
// the precision St of the adding states.  ab = 1 - R - T (diffuse absorptance) and omTn = 1 - Tn
// are only read by the heating-rate variant (dead code otherwise).
template<typename St>
struct LayerSW { St R, T, SU, SD, Tn, ab, omTn; };

// The layer properties in FP32, rearranged so that FP32 keeps its accuracy:
//   k^2 = 2 (1 - w0) (gamma1 + gamma2)          (gamma1 - gamma2 = 2 (1 - w0) exactly),
//   1 - e^{-2k tau} = a (2 - a) with a = 1 - e^{-k tau},
// and the direct-beam solution written without the removable singularity at k mu0 = 1.  With
// G = (e^{-k tau} - e^{-tau/mu0}) / (1 - k mu0) the reference's expressions (F90:1079-1104)
// become, identically,
//   Rdir = w0 RT [(alpha2 + k gamma3)(2a - a^2) + 2 (k gamma3 - alpha2 k mu0) e^{-k tau} G] / (1 + k mu0)
//   Tdir = -w0 RT [(alpha1 - k gamma4)(2a - a^2) e^{-tau/mu0} - 2k (gamma4 + mu0 alpha1) G] / (1 + k mu0)
// where every O(1) term that cancels in the reference form has been cancelled analytically.  G is
// e^{-tau/mu0} (tau/mu0) (e^x - 1)/x with x = tau (1 - k mu0) / mu0, from a series when |x| <= 1/4
// and as the difference quotient otherwise.  a keeps its series for small k tau: where k is small
// RT ~ 1/(2k) amplifies the absolute error of 1 - e^{-k tau}.  1 - e^{-tau/mu0}, the bound of
// Rdir + Tdir, keeps its series too (see the clamp).  The transmittance e^{-tau/mu0} of this layer enters G
// and Tdir only; the direct beam itself is carried by the caller (Beam<float>).  Returns false
// where FP64 is kept: near conservative scattering (1 - w0 < 1e-6).
template<bool ISLAND, typename St>
__device__ __forceinline__ bool sw_layer_fp32(float tau, float w0, float one_m_w0, float g, float mu0, float inv_mu0,
                                              St dir, LayerSW<St>& L)
{
    if (ISLAND && one_m_w0 < 1e-6f) return false;
    const float gamma1 = (8.f - w0 * (5.f + 3.f * g)) * .25f;
    const float gamma2 = 3.f * (w0 * (1.f - g)) * .25f;
    const float k = __builtin_amdgcn_sqrtf(fmaxf(2.f * one_m_w0 * (gamma1 + gamma2), 1e-12f));
    const float k_mu = k * mu0;
    const float gamma3 = (2.f - 3.f * mu0 * g) * .25f;
    const float gamma4 = 1.f - gamma3;
    const float alpha1 = gamma1 * gamma4 + gamma2 * gamma3;
    const float alpha2 = gamma1 * gamma3 + gamma2 * gamma4;
    const float kt = k * tau;
    const float e1 = exp_neg_f32(kt);
    const float a = one_minus_exp_neg_f32(kt, e1);
    const float c2 = a * (2.f - a);
    const float y = tau * inv_mu0;
    const float Tn = exp_neg_f32(y);
    const float ek = 1.f - k_mu;
    const float x = y * ek;
    // (e^x - 1) / x below |x| = 1/4 (truncation x^6/7! < 5e-8); above, the difference quotient
    // loses at most 2^-23 / |x| < 5e-7 relative.
    float phi = 1.f / 720.f;
    phi = 1.f / 120.f + x * phi;
    phi = 1.f / 24.f + x * phi;
    phi = 1.f / 6.f + x * phi;
    phi = .5f + x * phi;
    phi = 1.f + x * phi;
    const float G = (fabsf(x) <= .25f) ? Tn * y * phi : (e1 - Tn) * rcp_f32(ek);
    // One reciprocal for RT = 1 / (k (2 - c2) + gamma1 c2) and 1 / (1 + k mu0).
    const float opk = 1.f + k_mu;
    const float q = rcp_f32((k * (2.f - c2) + gamma1 * c2) * opk);
    const float RT = q * opk;
    const float s = w0 * q;
    const float rdf = s * ((alpha2 + k * gamma3) * c2 + 2.f * (k * gamma3 - alpha2 * k_mu) * e1 * G);
    const float tdf = -s * ((alpha1 - k * gamma4) * c2 * Tn - 2.f * k * (gamma4 + mu0 * alpha1) * G);
    // Clamps of rte_solver_kernels.cu:581-583 with the FP64 epsilon of the island they replace
    // (representable in FP32).  The bound 1 - e^{-tau/mu0} is taken from its series for small
    // tau/mu0: near conservative scattering Rdir + Tdir reaches it, so the clamp is active, and
    // 1 - Tn in FP32 would cut Tdir of a thin layer by up to the FP32 spacing near 1 (6e-8).
    const float b = one_minus_exp_neg_f32(y, Tn);
    const float eps = float(Limits<double>::eps);
    const float rd = fmaxf(eps, fminf(rdf, b));
    const float td = fmaxf(eps, fminf(tdf, b - rd));
    // 1 - R - T = RT [k (2 - c2) + gamma1 c2 - gamma2 c2 - 2 k e1] = RT [k a^2 + (gamma1 - gamma2) c2] with
    // gamma1 - gamma2 = 2 (1 - w0): a sum of non-negative terms, free of the cancellation of 1 - R - T.
    const float ab = RT * (k * a * a + 2.f * one_m_w0 * c2);
    L = {St(RT * gamma2 * c2), St(RT * 2.f * k * e1), St(rd) * dir, St(td) * dir, St(Tn), St(ab), St(b)};
    return true;
}

// FP32 builds take the FP32 form where it is accurate and the FP64 island elsewhere; FP64
// builds evaluate the reference expressions in FP64.  With CO the second input is the
// co-albedo 1 - w0, exact in the caller's precision, where w0 near 1 is not (FP32 resolves
// 1 - w0 only to 6e-8).  Given it exactly, the FP32 form keeps its accuracy down to conservative
// scattering: against the two-stream in 80-bit arithmetic over tau in [1e-10, 3e3],
// 1 - w0 in [0, 1e-5], g in [0, 0.95], mu0 in [0.02, 1], its largest errors in R, T, Rdir and
// Tdir are 2-4e-6 for 1 - w0 below 1e-6 as above it, so the CO form has no FP64 island.
template<typename Real, bool CO, typename St>
__device__ __forceinline__ LayerSW<St> sw_layer(Real tau, Real w_in, Real g, Real mu0, Real inv_mu0, St dir)
{
    LayerSW<St> L;
    const Real w0 = CO ? Real(1) - w_in : w_in;
    if constexpr (std::is_same<Real, float>::value)
    {
        if constexpr (CO)
        {
            sw_layer_fp32<false>(tau, w0, w_in, g, mu0, inv_mu0, dir, L);
            return L;
        }
        else if (sw_layer_fp32<true>(tau, w0, 1.f - w_in, g, mu0, inv_mu0, dir, L))
            return L;
    }
    const double w0d = CO ? 1. - double(w_in) : double(w_in);
    double Rdif, Tdif, Rdir, Tdir, Tn;
    sw_two_stream(double(tau), w0d, double(g), double(mu0), Rdif, Tdif, Rdir, Tdir, Tn);
    return {St(Rdif), St(Tdif), St(Rdir * double(dir)), St(Tdir * double(dir)), St(Tn), St(1. - Rdif - Tdif),
            St(-expm1(-double(tau) / double(mu0)))};
}

// Reciprocal of an adding denominator 1 - R A in (0, 1].
__device__ __forceinline__ double rcp_den(double x) { return rcp_f64(x); }
__device__ __forceinline__ float rcp_den(float x) { return rcp_f32(x); }

// Column above the next level from the column above this one and the layer between.
template<typename St>
__device__ __forceinline__ void above_update(const LayerSW<St>& L, St& R_up, St& D)
{
    const St den = rcp_den(St(1) - L.R * R_up);
    D = L.SD + L.T * (D + R_up * L.SU) * den;
    R_up = L.R + L.T * L.T * R_up * den;
}

// The direct beam below each layer of the top-down sweeps.  FP64 builds multiply the layer
// transmittances as the reference does (F90 sw_solver_2stream).  FP32 builds evaluate
// dir0 e^{-sum tau/mu0} from the optical depth summed in FP64 at every checkpoint segment start
// and multiply the layer transmittances within a segment, so the rounding of the hardware
// exponential does not compound down the column.  The checkpoint value is the beam (FP64
// builds) or the summed optical depth (FP32 builds).
template<typename Real, typename St> struct Beam;
template<> struct Beam<double, double>
{
    double dir;
    __device__ void init(double dir0) { dir = dir0; }
    __device__ void step(const LayerSW<double>& L, double, double) { dir *= L.Tn; }
    __device__ void refresh() {}
    __device__ double saved() const { return dir; }
    __device__ void restore(double v) { dir = v; }
};
template<typename St> struct Beam<float, St>
{
    double ycum;
    St dir0, dir;
    __device__ void init(St d0) { dir0 = d0; ycum = 0.; dir = d0; }
    // Within a checkpoint segment the beam is carried by the layers' own transmittances (at most
    // K - 1 roundings); at each segment start it is re-formed from the summed optical depth, the
    // value the replay of that segment restores.
    __device__ void step(const LayerSW<St>& L, float tau, float inv_mu0)
    {
        ycum += double(tau * inv_mu0);
        dir *= L.Tn;
    }
    __device__ void refresh() { dir = dir0 * St(exp_neg_f32(float(ycum))); }
    __device__ float saved() const { return float(ycum); }
    __device__ void restore(float v)
    {
        ycum = double(v);
        dir = dir0 * St(exp_neg_f32(v));
    }
};

// CO: the ssa argument holds the co-albedo 1 - ssa (see sw_layer).  The adding states and
// denominators are carried in the build's precision (with FP32 states the FP32 build stays within
// 2e-3 W m-2 of the FP64 build, near-conservative cloud layers included).
template<typename Real, bool HR = false, bool CO = false>
__global__ __launch_bounds__(SB_COLS * SB_GPTS)
__attribute__((amdgpu_waves_per_eu(sw_wide(sizeof(Real) == 4, CO, HR) ? 8 : 4)))
void sw_solver_kernel(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1,
                      const Real* __restrict__ tau, const Real* __restrict__ ssa,
                      const Real* __restrict__ gasy, const Real* __restrict__ mu0,
                      const Real* __restrict__ alb_dir, const Real* __restrict__ alb_dif,
                      const Real* __restrict__ inc_dir, const Real* __restrict__ inc_dif,
                      Real* __restrict__ flux_up, Real* __restrict__ flux_dn, Real* __restrict__ flux_dir,
                      Real* __restrict__ ckpt, Groups grp, Real* __restrict__ grp_up,
                      Real* __restrict__ grp_dn, Real* __restrict__ grp_dir, int64_t alb_nbnd,
                      double* __restrict__ absorbed)
{
    using St = Real;
    constexpr int K = sw_seg_k(sizeof(Real) == 4, CO, HR);
    constexpr int NC = 3;
    // With HR a fourth component per slot: the absorbed flux of the layer below the slot's level,
    // (dn_dif(top) + up(bottom)) (1 - R - T) + dir(top) (1 - Tn) - SU - SD for this g-point, from the
    // layer's own terms instead of a difference of level fluxes.
    constexpr int NCH = NC + (HR ? 1 : 0);
    // Slots 0..K-1: the levels of a segment; slot K: the surface level.
    __shared__ Real red[(K + 1) * NCH * SB_GPTS * SB_COLS];

    const int lane = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.y));   // wavefront = g-point
    const int gi = blockIdx.y;
    const int gcount = grp.count[gi];
    const int64_t col0 = int64_t(blockIdx.x) * SB_COLS;
    const int64_t icol = col0 + lane;
    const int64_t igpt = grp.start[gi] + w;
    const bool active = icol < ncol && w < gcount;
    const bool reduce = grp_dn != nullptr;
    const int64_t nlev = nlay + 1;
    const int64_t nseg = (nlay + K - 1) / K;
    const int64_t c2 = icol + ncol * igpt;
    // Per-g-point slabs at wavefront-uniform bases; lane offsets icol + ncol * level (32-bit).
    const int64_t lslab = ncol * nlay * igpt, vslab = ncol * nlev * igpt;
    const Real* tau_g = tau + lslab;
    const Real* ssa_g = ssa + lslab;
    const Real* gasy_g = gasy + lslab;
    Real* up_g = flux_up ? flux_up + vslab : nullptr;
    Real* dn_g = flux_up ? flux_dn + vslab : nullptr;
    Real* dir_g = flux_up ? flux_dir + vslab : nullptr;
    const uint32_t ucol = uint32_t(lane) + uint32_t(col0), uncol = uint32_t(ncol);
    // Checkpoint q of segment s at ck_g[(s * 3 + q) * ncol + icol], one slab per g-point.
    Real* ck_g = ckpt + ncol * nseg * 3 * igpt;
    auto ck = [&](int64_t s, int q) -> Real& { return at32(ck_g, ucol + uncol * uint32_t(s * 3 + q)); };
    auto lay_of = [&](int64_t j) -> uint32_t { return uint32_t(top_at_1 ? j : nlay - 1 - j); };   // layer j from the top
    auto lev_of = [&](int64_t j) -> uint32_t { return uint32_t(top_at_1 ? j : nlay - j); };       // level j from the top

    const Real m0r = active ? mu0[icol] : Real(1);
    const Real inv_m0r = Real(1) / m0r;
    const St m0 = St(m0r);
    St R_up = St(0), D = St(0);
    Beam<Real, St> beam;
    beam.init(active ? St(inc_dir[c2]) * m0 : St(0));
    if (active)
    {
        D = inc_dif ? St(inc_dif[c2]) : St(0);
        // Sweep down: the column above every level; checkpoints at the top of each segment.  The
        // inputs of a segment are all requested before its dependent chain starts.
        for (int64_t s = 0; s < nseg; ++s)
        {
            const int64_t j0 = s * K;
            const int nk = int(std::min<int64_t>(K, nlay - j0));
            beam.refresh();
            ck(s, 0) = Real(R_up); ck(s, 1) = Real(D); ck(s, 2) = Real(beam.saved());
            Real tv[K], wv[K], gv[K];
#pragma unroll
            for (int kk = 0; kk < K; ++kk)
                if (kk < nk)
                {
                    const uint32_t l = ucol + uncol * lay_of(j0 + kk);
                    tv[kk] = ld_nt(tau_g, l); wv[kk] = ld_nt(ssa_g, l); gv[kk] = ld_nt(gasy_g, l);
                }
#pragma unroll
            for (int kk = 0; kk < K; ++kk)
                if (kk < nk)
                {
                    const LayerSW<St> L = sw_layer<Real, CO>(tv[kk], wv[kk], gv[kk], m0r, inv_m0r, beam.dir);
                    above_update(L, R_up, D);
                    beam.step(L, tv[kk], inv_m0r);
                }
        }
    }

    // Upward flux at the bottom of the layer being closed (heating-rate variant).
    St up_below = St(0);
    // Surface: albedos per (column, g-point), or per (band, column) read directly.
    const int64_t ia = alb_nbnd > 0 ? grp.band[gi] + alb_nbnd * icol : c2;
    St A = active ? St(alb_dif[ia]) : St(0);
    St S = beam.dir * (active ? St(alb_dir[ia]) : St(0));
    {
        const St dir = beam.dir;
        const St dn = (D + R_up * S) * rcp_den(St(1) - R_up * A);
        const St up = A * dn + S;
        if (active && up_g)
        {
            const uint32_t o = ucol + uncol * lev_of(nlay);
            at32(up_g, o) = Real(up); at32(dn_g, o) = Real(dn + dir); at32(dir_g, o) = Real(dir);
        }
        if (reduce)
        {
            red[((K * NCH + 0) * SB_GPTS + w) * SB_COLS + lane] = Real(up);
            red[((K * NCH + 1) * SB_GPTS + w) * SB_COLS + lane] = Real(dn + dir);
            red[((K * NCH + 2) * SB_GPTS + w) * SB_COLS + lane] = Real(dir);
        }
        up_below = up;
    }

    // Sweep up, one segment at a time from the bottom.
    for (int64_t s = nseg - 1; s >= 0; --s)
    {
        const int64_t j0 = s * K;
        const int nk = int(std::min<int64_t>(K, nlay - j0));
        if (active)
        {
            Real sR[K], sT[K], sSU[K], sSD[K], sRup[K], sD[K], sDir[K];
            Real sAb[HR ? K : 1], sOmTn[HR ? K : 1];
            St r_up = St(ck(s, 0)), d = St(ck(s, 1));
            Beam<Real, St> bm = beam;
            bm.restore(ck(s, 2));
            Real tv[K], wv[K], gv[K];
#pragma unroll
            for (int kk = 0; kk < K; ++kk)
                if (kk < nk)
                {
                    const uint32_t l = ucol + uncol * lay_of(j0 + kk);
                    tv[kk] = ld_nt(tau_g, l); wv[kk] = ld_nt(ssa_g, l); gv[kk] = ld_nt(gasy_g, l);
                }
#pragma unroll
            for (int kk = 0; kk < K; ++kk)
                if (kk < nk)
                {
                    sRup[kk] = Real(r_up); sD[kk] = Real(d); sDir[kk] = Real(bm.dir);
                    const LayerSW<St> L = sw_layer<Real, CO>(tv[kk], wv[kk], gv[kk], m0r, inv_m0r, bm.dir);
                    sR[kk] = Real(L.R); sT[kk] = Real(L.T); sSU[kk] = Real(L.SU); sSD[kk] = Real(L.SD);
                    if constexpr (HR) { sAb[kk] = Real(L.ab); sOmTn[kk] = Real(L.omTn); }
                    above_update(L, r_up, d);
                    bm.step(L, tv[kk], inv_m0r);
                }
#pragma unroll
            for (int kk = K - 1; kk >= 0; --kk)
                if (kk < nk)
                {
                    const St R = sR[kk], T = sT[kk];
                    const St den = rcp_den(St(1) - R * A);
                    S = St(sSU[kk]) + T * (S + A * St(sSD[kk])) * den;
                    A = R + T * T * A * den;
                    const St ru = sRup[kk];
                    const St dn = (St(sD[kk]) + ru * S) * rcp_den(St(1) - ru * A);
                    const St up = A * dn + S;
                    const St dd = sDir[kk];
                    if (up_g)
                    {
                        const uint32_t o = ucol + uncol * lev_of(j0 + kk);
                        at32(up_g, o) = Real(up); at32(dn_g, o) = Real(dn + dd); at32(dir_g, o) = Real(dd);
                    }
                    if (reduce)
                    {
                        red[((kk * NCH + 0) * SB_GPTS + w) * SB_COLS + lane] = Real(up);
                        red[((kk * NCH + 1) * SB_GPTS + w) * SB_COLS + lane] = Real(dn + dd);
                        red[((kk * NCH + 2) * SB_GPTS + w) * SB_COLS + lane] = Real(dd);
                    }
                    if constexpr (HR)
                    {
                        const St ab = (dn + up_below) * St(sAb[kk]) + dd * St(sOmTn[kk]) - St(sSU[kk]) -
                                      St(sSD[kk]);
                        red[((kk * NCH + NC) * SB_GPTS + w) * SB_COLS + lane] = Real(ab);
                        up_below = up;
                    }
                }
        }
        if (reduce)
        {
            __syncthreads();
            Real* const outs[3] = {grp_up, grp_dn, grp_dir};
            // The surface slot K goes out with the first (bottom) segment.
            const int nsl = (s == nseg - 1) ? K + 1 : nk;
            write_group_sums<Real, 3, NCH>(red, nsl, gcount, col0, ncol, nlev, gi, outs,
                                                              [&](int kk) -> int64_t {
                if (kk == K) return lev_of(nlay);
                return kk < nk ? int64_t(lev_of(j0 + kk)) : int64_t(-1);
            });
            if constexpr (HR)
                write_group_layer_sums<Real, NCH>(red, NC, nk, gcount, col0, ncol, nlay, gi, absorbed,
                                                  [&](int kk) -> int64_t { return int64_t(lay_of(j0 + kk)); });
            __syncthreads();
        }
    }
}

// Band sums (when a band spans several groups) and broadband sums of group fluxes
// (ncol, nlev, ngroups): per (column, level), groups in order, then bands in order.
template<typename Real>
__global__ void group_sums_kernel(int64_t ncv, Groups grp, const Real* __restrict__ gf,
                                  Real* __restrict__ bnd, Real* __restrict__ bb)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncv) return;
    Real total = Real(0), band = Real(0);
    for (int g = 0; g < grp.n; ++g)
    {
        const Real v = gf[i + ncv * g];
        band = (g > 0 && grp.band[g] == grp.band[g - 1]) ? band + v : v;
        const bool last = (g + 1 == grp.n) || grp.band[g + 1] != grp.band[g];
        if (last)
        {
            if (bnd) bnd[i + ncv * grp.band[g]] = band;
            total = (grp.band[g] == grp.band[0]) ? band : total + band;
        }
    }
    if (bb) bb[i] = total;
}

template<typename Real>
int group_sums(int64_t ncv, const Groups& grp, const Real* gf, Real* bnd, Real* bb, hipStream_t s)
{
    if (!bnd && !bb) return RRTMGP_OK;
    RRTMGP_TIMED("group_sums_kernel", s);
    hipLaunchKernelGGL(group_sums_kernel<Real>, dim3(grid1(ncv, 256)), dim3(256), 0, s, ncv, grp, gf, bnd, bb);
    RRTMGP_LAUNCHED("group_sums_kernel");
    return RRTMGP_OK;
}

template<typename Real>
int absorbed_heating_rate(int64_t ncol, int64_t nlay, int ngroups, int nparts, const double* parts, const Real* plev,
                          Real* hr, hipStream_t s)
{
    RRTMGP_TIMED("absorbed_heating_rate_kernel", s);
    hipLaunchKernelGGL(absorbed_heating_rate_kernel<Real>, dim3(grid1(ncol * nlay, 256)), dim3(256), 0, s, ncol,
                       nlay, ngroups, nparts, parts, plev, hr);
    RRTMGP_LAUNCHED("absorbed_heating_rate_kernel");
    return RRTMGP_OK;
}

template<typename Real>
__global__ void lw_secants_kernel(int64_t n, int64_t ncg, int64_t n_quad, int64_t max_pts,
                                  const Real* __restrict__ gauss_Ds, Real* __restrict__ secants)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t imu = i / ncg;
    secants[i] = gauss_Ds[imu + (n_quad - 1) * max_pts];
}

template<typename Real>
int lw_secants(int64_t ncol, int64_t ngpt, int64_t n_quad, int64_t max_pts, const Real* gauss_Ds,
               Real* secants, void* stream)
{
    if (ncol <= 0 || ngpt <= 0) return arg_error("lw_secants_array: empty extents");
    if (n_quad < 1 || n_quad > max_pts) return arg_error("lw_secants_array: n_gauss_quad out of range");
    if (!gauss_Ds || !secants) return arg_error("lw_secants_array: NULL array");
    const int64_t n = ncol * ngpt * n_quad;
    RRTMGP_TIMED("lw_secants_kernel", as_stream(stream));
    hipLaunchKernelGGL(lw_secants_kernel<Real>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),
                       n, ncol * ngpt, n_quad, max_pts, gauss_Ds, secants);
    RRTMGP_LAUNCHED("lw_secants_kernel");
    return RRTMGP_OK;
}

// Outputs are either g-point fluxes (flux_*) or band (bnd_*, (ncol, nlev, nbnd)) and/or broadband
// (bb_*, (ncol, nlev)) sums.  Group sums go straight into the band outputs when every band is a
// single group, else through scratch.
template<typename Real>
int lw_solver(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1, int64_t nmus, const Real* secants,
              const Real* weights, const Real* tau, const Real* lay, const Real* linc, const Real* ldec,
              const Real* emis, const Real* sfc_src, const Real* inc_flux, Real* flux_up, Real* flux_dn,
              int64_t nbnd, const int* band_lims, Real* bnd_up, Real* bnd_dn, Real* bb_up, Real* bb_dn,
              int do_jac, const Real* sfc_src_jac, Real* flux_up_jac, void* stream,
              const Real* bnd_lay = nullptr, const Real* bnd_lev = nullptr, bool by_band = false,
              const Real* plev = nullptr, Real* heating_rate = nullptr)
{
    if (ncol <= 0 || nlay <= 0 || ngpt <= 0) return arg_error("lw_solver: empty extents");
    if (heating_rate && (!plev || flux_up || flux_dn || do_jac))
        return arg_error("lw_solver: heating rates need plev and summed (not g-point) flux outputs");
    const bool factored = bnd_lay != nullptr;
    if (factored && (!bnd_lev || !band_lims || flux_up || flux_dn || do_jac))
        return arg_error("lw_solver: factored sources need both band Planck arrays, band limits and summed outputs");
    if (nmus < 1 || nmus > 4) return arg_error("lw_solver: n_gauss_angles must be 1..4");
    if (ncol * (nlay + 1) * int64_t(sizeof(Real)) >= (int64_t(1) << 32))
        return arg_error("lw_solver: ncol * (nlay + 1) values of one g-point must stay below 4 GiB");
    if (!secants || !weights || !tau || !lay || (!factored && (!linc || !ldec)) || !emis || !sfc_src)
        return arg_error("lw_solver: NULL input");
    const bool spectral = flux_up || flux_dn;
    if (spectral && (!flux_up || !flux_dn)) return arg_error("lw_solver: flux_up and flux_dn go together");
    if (!spectral && !((bnd_up && bnd_dn) || (bb_up && bb_dn))) return arg_error("lw_solver: no output");
    if (spectral && (bnd_up || bnd_dn || bb_up || bb_dn)) return arg_error("lw_solver: g-point and summed outputs are exclusive");
    if ((bnd_up || bnd_dn) && (!bnd_up || !bnd_dn || !band_lims)) return arg_error("lw_solver: band fluxes need both arrays and band limits");
    if (do_jac && (!sfc_src_jac || !flux_up_jac || spectral)) return arg_error("lw_solver: the Jacobian needs broadband output");
    if (by_band && (!band_lims || nbnd < 1)) return arg_error("lw_solver: band surface inputs need band limits");
    hipStream_t s = as_stream(stream);
    Groups grp;
    RRTMGP_TRY(make_groups(ngpt, band_lims ? nbnd : 1, band_lims, grp));
    const int64_t ncv = ncol * (nlay + 1);
    const bool direct = bnd_up != nullptr && grp.n == int(nbnd);
    const dim3 grid(grid1(ncol, SB_COLS), unsigned(grp.n));
    const dim3 block(SB_COLS, SB_GPTS);

    Real *g_up = nullptr, *g_dn = nullptr, *g_jac = nullptr, *scratch = nullptr;
    double* absorbed = nullptr;
    if (!spectral)
    {
        const int64_t nbuf = (direct ? 0 : 2) + (do_jac ? 1 : 0);
        RRTMGP_TRY(scratch_alloc(&scratch, size_t(ncv * grp.n * nbuf), s));
        g_up = direct ? bnd_up : scratch;
        g_dn = direct ? bnd_dn : scratch + ncv * grp.n;
        if (do_jac) g_jac = scratch + ncv * grp.n * (direct ? 0 : 2);
    }
    if (heating_rate)
    {
        // Per-group absorbed fluxes of the down and the up sweep, FP64, every element written.
        const int rc0 = scratch_alloc(&absorbed, size_t(ncol * nlay * grp.n * 2), s);
        if (rc0 != RRTMGP_OK) { scratch_free(scratch, s); return rc0; }
    }
    int rc = RRTMGP_OK;
    {
        RRTMGP_TIMED("lw_solver_kernel", s);
        auto launch = [&](auto kernel, const Real* jac_src, Real* jac_out) {
            hipLaunchKernelGGL(kernel, grid, block, 0, s, ncol, nlay, ngpt, top_at_1, int(nmus), secants,
                               weights, tau, lay, linc, ldec, emis, sfc_src, inc_flux, flux_up, flux_dn,
                               jac_src, grp, g_up, g_dn, jac_out, bnd_lay, bnd_lev, by_band ? nbnd : int64_t(0),
                               by_band ? 1 : 0, absorbed);
        };
        if (do_jac)
            launch(nmus == 1 ? lw_solver_kernel<Real, true, 1, false> : lw_solver_kernel<Real, true, 4, false>,
                   sfc_src_jac, g_jac);
        else if (heating_rate && factored)
            launch(nmus == 1 ? lw_solver_kernel<Real, false, 1, true, true> : lw_solver_kernel<Real, false, 4, true, true>,
                   nullptr, nullptr);
        else if (heating_rate)
            launch(nmus == 1 ? lw_solver_kernel<Real, false, 1, false, true> : lw_solver_kernel<Real, false, 4, false, true>,
                   nullptr, nullptr);
        else if (factored)
            launch(nmus == 1 ? lw_solver_kernel<Real, false, 1, true> : lw_solver_kernel<Real, false, 4, true>,
                   nullptr, nullptr);
        else
            launch(nmus == 1 ? lw_solver_kernel<Real, false, 1, false> : lw_solver_kernel<Real, false, 4, false>,
                   nullptr, nullptr);
        rc = check_hip(hipGetLastError(), "lw_solver_kernel");
    }
    if (rc == RRTMGP_OK && !spectral)
    {
        rc = group_sums<Real>(ncv, grp, g_up, direct ? nullptr : bnd_up, bb_up, s);
        if (rc == RRTMGP_OK) rc = group_sums<Real>(ncv, grp, g_dn, direct ? nullptr : bnd_dn, bb_dn, s);
        if (rc == RRTMGP_OK && do_jac) rc = group_sums<Real>(ncv, grp, g_jac, nullptr, flux_up_jac, s);
    }
    if (rc == RRTMGP_OK && heating_rate)
        rc = absorbed_heating_rate<Real>(ncol, nlay, grp.n, 2, absorbed, plev, heating_rate, s);
    scratch_free(absorbed, s);
    scratch_free(scratch, s);
    return rc;
}

// Scratch of the entry points without allocation (rrtmgp_hip_sw_solver_2stream_bnd_noalloc_*):
// checkpoints, group partial fluxes and heating-rate parts of one call, carved from this array
// in 256-byte aligned pieces.  Each device holds its own copy for the life of the process; its
// address on a device comes from hipGetSymbolAddress (noalloc_scratch).
constexpr size_t SW_NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char sw_scratch[SW_NOALLOC_SCRATCH_BYTES];

// Device address of sw_scratch on the current device, looked up once per device.
int noalloc_scratch(void** out)
{
    constexpr int max_dev = 64;
    static std::mutex m;
    static void* addr[max_dev] = {};
    int dev = 0;
    RRTMGP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= max_dev) return arg_error("sw_solver (no allocation): device index out of range");
    std::lock_guard<std::mutex> lock(m);
    if (!addr[dev])
    {
        void* p = nullptr;
        RRTMGP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(sw_scratch)));
        addr[dev] = p;
    }
    *out = addr[dev];
    return RRTMGP_OK;
}

// With workspace != nullptr the solver's scratch is carved from it (at most workspace_bytes) and
// nothing is allocated; otherwise it comes from the library's caching allocator.
template<typename Real>
int sw_solver(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1, const Real* tau, const Real* ssa,
              const Real* g, const Real* mu0, const Real* alb_dir, const Real* alb_dif, const Real* inc_dir,
              Real* flux_up, Real* flux_dn, Real* flux_dir, int has_dif_bc, const Real* inc_dif,
              int64_t nbnd, const int* band_lims, Real* bnd_up, Real* bnd_dn, Real* bnd_dir,
              Real* bb_up, Real* bb_dn, Real* bb_dir, void* stream, bool by_band = false,
              const Real* plev = nullptr, Real* heating_rate = nullptr, unsigned char* workspace = nullptr,
              size_t workspace_bytes = 0, bool coalbedo = false)
{
    if (ncol <= 0 || nlay <= 0 || ngpt <= 0) return arg_error("sw_solver: empty extents");
    if (heating_rate && (!plev || flux_up || flux_dn || flux_dir))
        return arg_error("sw_solver: heating rates need plev and summed (not g-point) flux outputs");
    if (by_band && (!band_lims || nbnd < 1)) return arg_error("sw_solver: band surface inputs need band limits");
    // 32-bit lane offsets within one g-point's slab of levels or of checkpoints.
    const int K = sw_seg_k(sizeof(Real) == 4, coalbedo, heating_rate != nullptr);
    if (ncol * std::max<int64_t>(nlay + 1, 3 * ((nlay + K - 1) / K)) * int64_t(sizeof(Real)) >= (int64_t(1) << 32))
        return arg_error("sw_solver: ncol * (nlay + 1) values of one g-point must stay below 4 GiB");
    if (!tau || !ssa || !g || !mu0 || !alb_dir || !alb_dif || !inc_dir) return arg_error("sw_solver: NULL input");
    const bool spectral = flux_up || flux_dn || flux_dir;
    const bool want_bnd = bnd_up || bnd_dn || bnd_dir, want_bb = bb_up || bb_dn || bb_dir;
    if (spectral && (!flux_up || !flux_dn || !flux_dir)) return arg_error("sw_solver: flux_up, flux_dn and flux_dir go together");
    if (spectral && (want_bnd || want_bb)) return arg_error("sw_solver: g-point and summed outputs are exclusive");
    if (want_bnd && (!bnd_up || !bnd_dn || !bnd_dir || !band_lims)) return arg_error("sw_solver: band fluxes need all three arrays and band limits");
    if (want_bb && (!bb_up || !bb_dn || !bb_dir)) return arg_error("sw_solver: broadband fluxes need all three arrays");
    if (!spectral && !want_bnd && !want_bb) return arg_error("sw_solver: no output");
    if (has_dif_bc && !inc_dif) return arg_error("sw_solver: has_dif_bc without inc_flux_dif");
    hipStream_t s = as_stream(stream);
    Groups grp;
    RRTMGP_TRY(make_groups(ngpt, band_lims ? nbnd : 1, band_lims, grp));
    const int64_t ncv = ncol * (nlay + 1), nseg = (nlay + K - 1) / K;
    const bool direct = want_bnd && grp.n == int(nbnd);
    const dim3 grid(grid1(ncol, SB_COLS), unsigned(grp.n));
    const dim3 block(SB_COLS, SB_GPTS);

    // Every element of the checkpoints, group partials and heating-rate parts is written by the
    // kernel before it is read, so none of them needs clearing.
    Real *ckpt = nullptr, *scratch = nullptr;
    double* absorbed = nullptr;
    const size_t n_ck = size_t(nseg * 3 * ngpt * ncol);
    const size_t n_part = (!spectral && !direct) ? size_t(ncv * grp.n * 3) : 0;
    const size_t n_abs = heating_rate ? size_t(ncol * nlay * grp.n) : 0;
    if (workspace)
    {
        auto round256 = [](size_t b) { return (b + 255) & ~size_t(255); };
        const size_t need = round256(n_ck * sizeof(Real)) + round256(n_part * sizeof(Real)) + round256(n_abs * sizeof(double));
        if (need > workspace_bytes)
            return arg_error("sw_solver (no allocation): this call needs " + std::to_string(need) +
                             " bytes of scratch; the entry point without allocation holds " +
                             std::to_string(workspace_bytes) + " (use fewer columns per call)");
        ckpt = reinterpret_cast<Real*>(workspace);
        scratch = n_part ? reinterpret_cast<Real*>(workspace + round256(n_ck * sizeof(Real))) : nullptr;
        absorbed = n_abs ? reinterpret_cast<double*>(workspace + round256(n_ck * sizeof(Real)) +
                                                     round256(n_part * sizeof(Real)))
                         : nullptr;
    }
    else
    {
        RRTMGP_TRY(scratch_alloc(&ckpt, n_ck, s));
        if (n_part)
        {
            const int rc0 = scratch_alloc(&scratch, n_part, s);
            if (rc0 != RRTMGP_OK) { scratch_free(ckpt, s); return rc0; }
        }
        if (n_abs)
        {
            const int rc0 = scratch_alloc(&absorbed, n_abs, s);
            if (rc0 != RRTMGP_OK) { scratch_free(ckpt, s); scratch_free(scratch, s); return rc0; }
        }
    }
    Real* gf[3] = {nullptr, nullptr, nullptr};
    Real* const bo[3] = {bnd_up, bnd_dn, bnd_dir};
    Real* const bbo[3] = {bb_up, bb_dn, bb_dir};
    if (!spectral)
        for (int c = 0; c < 3; ++c) gf[c] = direct ? bo[c] : scratch + ncv * grp.n * c;
    int rc = RRTMGP_OK;
    {
        RRTMGP_TIMED("sw_solver_kernel", s);
        auto kernel = heating_rate ? (coalbedo ? sw_solver_kernel<Real, true, true> : sw_solver_kernel<Real, true>)
                                   : (coalbedo ? sw_solver_kernel<Real, false, true> : sw_solver_kernel<Real, false>);
        hipLaunchKernelGGL(kernel, grid, block, 0, s, ncol, nlay, ngpt, top_at_1, tau, ssa, g, mu0, alb_dir, alb_dif, inc_dir,
                           has_dif_bc ? inc_dif : nullptr, flux_up, flux_dn, flux_dir, ckpt, grp, gf[0], gf[1], gf[2],
                           by_band ? nbnd : int64_t(0), absorbed);
        rc = check_hip(hipGetLastError(), "sw_solver_kernel");
    }
    for (int c = 0; c < 3 && rc == RRTMGP_OK && !spectral; ++c)
        rc = group_sums<Real>(ncv, grp, gf[c], direct ? nullptr : bo[c], bbo[c], s);
    if (rc == RRTMGP_OK && heating_rate)
        rc = absorbed_heating_rate<Real>(ncol, nlay, grp.n, 1, absorbed, plev, heating_rate, s);
    if (!workspace)
    {
        scratch_free(absorbed, s);
        scratch_free(ckpt, s);
        scratch_free(scratch, s);
    }
    return rc;
}

template<typename Real>
int sw_solver_noalloc(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1, const Real* tau, const Real* ssa,
                      const Real* g, const Real* mu0, const Real* alb_dir, const Real* alb_dif, const Real* inc_dir,
                      int has_dif_bc, const Real* inc_dif, int64_t nbnd, const int* band_lims, Real* bnd_up,
                      Real* bnd_dn, Real* bnd_dir, Real* bb_up, Real* bb_dn, Real* bb_dir, const Real* plev,
                      Real* heating_rate, void* stream)
{
    void* ws = nullptr;
    RRTMGP_TRY(noalloc_scratch(&ws));
    // Launch timing would create events: it is held off for this call.
    const TimingSuspend no_timing;
    return sw_solver<Real>(ncol, nlay, ngpt, top_at_1, tau, ssa, g, mu0, alb_dir, alb_dif, inc_dir, nullptr,
                           nullptr, nullptr, has_dif_bc, inc_dif, nbnd, band_lims, bnd_up, bnd_dn, bnd_dir, bb_up,
                           bb_dn, bb_dir, stream, true, plev, heating_rate, static_cast<unsigned char*>(ws),
                           SW_NOALLOC_SCRATCH_BYTES);
}
}  // namespace rrtmgp_hip

using namespace rrtmgp_hip;

#define RRTMGP_DEF_SOLVERS(S, T)                                                                    \
extern "C" int rrtmgp_hip_lw_solver_noscat_##S(int64_t ncol, int64_t nlay, int64_t ngpt,          \
    int top_at_1, int64_t nmus, const T* secants, const T* weights, const T* tau,                   \
    const T* lay_source, const T* lev_source_inc, const T* lev_source_dec, const T* sfc_emis,       \
    const T* sfc_src, const T* inc_flux, T* flux_up, T* flux_dn, int do_broadband,                  \
    T* broadband_up, T* broadband_dn, int do_jacobians, const T* sfc_src_jac, T* flux_up_jac,       \
    void* stream)                                                                                    \
{                                                                                                    \
    return lw_solver<T>(ncol, nlay, ngpt, top_at_1, nmus, secants, weights, tau, lay_source,        \
                        lev_source_inc, lev_source_dec, sfc_emis, sfc_src, inc_flux,                 \
                        do_broadband ? nullptr : flux_up, do_broadband ? nullptr : flux_dn,          \
                        0, nullptr, nullptr, nullptr, do_broadband ? broadband_up : nullptr,         \
                        do_broadband ? broadband_dn : nullptr, do_jacobians, sfc_src_jac,            \
                        flux_up_jac, stream);                                                        \
}                                                                                                    \
extern "C" int rrtmgp_hip_lw_solver_noscat_bnd_##S(int64_t ncol, int64_t nlay, int64_t ngpt,      \
    int top_at_1, int64_t nmus, const T* secants, const T* weights, const T* tau,                   \
    const T* lay_source, const T* lev_source_inc, const T* lev_source_dec, const T* sfc_emis,       \
    const T* sfc_src, const T* inc_flux, int64_t nbnd, const int* band_lims_gpt, T* bnd_flux_up,    \
    T* bnd_flux_dn, T* broadband_up, T* broadband_dn, const T* plev, T* heating_rate, void* stream)  \
{                                                                                                    \
    return lw_solver<T>(ncol, nlay, ngpt, top_at_1, nmus, secants, weights, tau, lay_source,        \
                        lev_source_inc, lev_source_dec, sfc_emis, sfc_src, inc_flux, nullptr,        \
                        nullptr, nbnd, band_lims_gpt, bnd_flux_up, bnd_flux_dn, broadband_up,        \
                        broadband_dn, 0, nullptr, nullptr, stream, nullptr, nullptr, true, plev,     \
                        heating_rate);                                                               \
}                                                                                                    \
extern "C" int rrtmgp_hip_lw_solver_noscat_bnd_factored_##S(int64_t ncol, int64_t nlay,           \
    int64_t ngpt, int top_at_1, int64_t nmus, const T* secants, const T* weights, const T* tau,     \
    const T* planck_frac, const T* band_lay_src, const T* band_lev_src, const T* sfc_emis,          \
    const T* sfc_src, const T* inc_flux, int64_t nbnd, const int* band_lims_gpt, T* bnd_flux_up,    \
    T* bnd_flux_dn, T* broadband_up, T* broadband_dn, const T* plev, T* heating_rate, void* stream)  \
{                                                                                                    \
    if (!band_lay_src || !band_lev_src) return arg_error("lw_solver_factored: NULL band array");    \
    return lw_solver<T>(ncol, nlay, ngpt, top_at_1, nmus, secants, weights, tau, planck_frac,       \
                        nullptr, nullptr, sfc_emis, sfc_src, inc_flux, nullptr, nullptr, nbnd,       \
                        band_lims_gpt, bnd_flux_up, bnd_flux_dn, broadband_up, broadband_dn, 0,      \
                        nullptr, nullptr, stream, band_lay_src, band_lev_src, true, plev,            \
                        heating_rate);                                                               \
}                                                                                                    \
extern "C" int rrtmgp_hip_sw_solver_2stream_##S(int64_t ncol, int64_t nlay, int64_t ngpt,         \
    int top_at_1, const T* tau, const T* ssa, const T* g, const T* mu0, const T* sfc_alb_dir,       \
    const T* sfc_alb_dif, const T* inc_flux_dir, T* flux_up, T* flux_dn, T* flux_dir,               \
    int has_dif_bc, const T* inc_flux_dif, int do_broadband, T* broadband_up, T* broadband_dn,      \
    T* broadband_dir, void* stream)                                                                  \
{                                                                                                    \
    return sw_solver<T>(ncol, nlay, ngpt, top_at_1, tau, ssa, g, mu0, sfc_alb_dir, sfc_alb_dif,     \
                        inc_flux_dir, do_broadband ? nullptr : flux_up,                              \
                        do_broadband ? nullptr : flux_dn, do_broadband ? nullptr : flux_dir,         \
                        has_dif_bc, inc_flux_dif, 0, nullptr, nullptr, nullptr, nullptr,             \
                        do_broadband ? broadband_up : nullptr, do_broadband ? broadband_dn : nullptr,\
                        do_broadband ? broadband_dir : nullptr, stream);                             \
}                                                                                                    \
extern "C" int rrtmgp_hip_sw_solver_2stream_bnd_##S(int64_t ncol, int64_t nlay, int64_t ngpt,     \
    int top_at_1, const T* tau, const T* ssa, const T* g, const T* mu0, const T* sfc_alb_dir,       \
    const T* sfc_alb_dif, const T* inc_flux_dir, int has_dif_bc, const T* inc_flux_dif,             \
    int64_t nbnd, const int* band_lims_gpt, T* bnd_flux_up, T* bnd_flux_dn, T* bnd_flux_dir,        \
    T* broadband_up, T* broadband_dn, T* broadband_dir, const T* plev, T* heating_rate,             \
    void* stream)                                                                                    \
{                                                                                                    \
    return sw_solver<T>(ncol, nlay, ngpt, top_at_1, tau, ssa, g, mu0, sfc_alb_dir, sfc_alb_dif,     \
                        inc_flux_dir, nullptr, nullptr, nullptr, has_dif_bc, inc_flux_dif, nbnd,     \
                        band_lims_gpt, bnd_flux_up, bnd_flux_dn, bnd_flux_dir, broadband_up,         \
                        broadband_dn, broadband_dir, stream, true, plev, heating_rate);              \
}                                                                                                    \
extern "C" int rrtmgp_hip_sw_solver_2stream_bnd_coalbedo_##S(int64_t ncol, int64_t nlay,           \
    int64_t ngpt, int top_at_1, const T* tau, const T* coalbedo, const T* g, const T* mu0,          \
    const T* sfc_alb_dir, const T* sfc_alb_dif, const T* inc_flux_dir, int has_dif_bc,             \
    const T* inc_flux_dif, int64_t nbnd, const int* band_lims_gpt, T* bnd_flux_up, T* bnd_flux_dn,  \
    T* bnd_flux_dir, T* broadband_up, T* broadband_dn, T* broadband_dir, const T* plev,             \
    T* heating_rate, void* stream)                                                                   \
{                                                                                                    \
    return sw_solver<T>(ncol, nlay, ngpt, top_at_1, tau, coalbedo, g, mu0, sfc_alb_dir,            \
                        sfc_alb_dif, inc_flux_dir, nullptr, nullptr, nullptr, has_dif_bc,            \
                        inc_flux_dif, nbnd, band_lims_gpt, bnd_flux_up, bnd_flux_dn, bnd_flux_dir,   \
                        broadband_up, broadband_dn, broadband_dir, stream, true, plev, heating_rate, \
                        nullptr, 0, true);                                                           \
}

#define RRTMGP_DEF_SW_NOALLOC(S, T)                                                                 \
extern "C" int rrtmgp_hip_sw_solver_2stream_bnd_noalloc_##S(int64_t ncol, int64_t nlay,            \
    int64_t ngpt, int top_at_1, const T* tau, const T* ssa, const T* g, const T* mu0,               \
    const T* sfc_alb_dir, const T* sfc_alb_dif, const T* inc_flux_dir, int has_dif_bc,             \
    const T* inc_flux_dif, int64_t nbnd, const int* band_lims_gpt, T* bnd_flux_up, T* bnd_flux_dn,  \
    T* bnd_flux_dir, T* broadband_up, T* broadband_dn, T* broadband_dir, const T* plev,             \
    T* heating_rate, void* stream)                                                                   \
{                                                                                                    \
    return sw_solver_noalloc<T>(ncol, nlay, ngpt, top_at_1, tau, ssa, g, mu0, sfc_alb_dir,          \
                                sfc_alb_dif, inc_flux_dir, has_dif_bc, inc_flux_dif, nbnd,           \
                                band_lims_gpt, bnd_flux_up, bnd_flux_dn, bnd_flux_dir, broadband_up, \
                                broadband_dn, broadband_dir, plev, heating_rate, stream);            \
}

RRTMGP_DEF_SW_NOALLOC(f32, float)
RRTMGP_DEF_SW_NOALLOC(f64, double)

#define RRTMGP_DEF_SECANTS(S, T)                                                                    \
extern "C" int rrtmgp_hip_lw_secants_array_##S(int64_t ncol, int64_t ngpt, int64_t n_gauss_quad,  \
    int64_t max_gauss_pts, const T* gauss_Ds, T* secants, void* stream)                             \
{                                                                                                    \
    return lw_secants<T>(ncol, ngpt, n_gauss_quad, max_gauss_pts, gauss_Ds, secants, stream);       \
}

RRTMGP_DEF_SECANTS(f32, float)
RRTMGP_DEF_SECANTS(f64, double)
RRTMGP_DEF_SOLVERS(f32, float)
RRTMGP_DEF_SOLVERS(f64, double)
