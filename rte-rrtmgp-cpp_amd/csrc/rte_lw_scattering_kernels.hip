// Longwave solvers with scattering: the no-scattering solver with the scattering rescaling of
// Tang et al. (2018), and the two-stream solver.
//
// Arithmetic follows src_kernels/mo_rte_solver_kernels.F90: lw_solver_noscat with do_rescaling
// (:168-200) and lw_transport_1rescl (:731-822); lw_solver_2stream (:386-453) with
// lw_combine_sources (:896-920), lw_two_stream (:832-887), lw_source_2str (:928-978) and adding
// (:1117-1227).  The reference has no GPU form of either: Rte_lw::rte_lw (src/Rte_lw.cpp:144-145)
// calls the solver with do_rescaling = false, and the two-stream solver has no C binding in
// include/rrtmgp_kernels.h.
//
// Layout of the work: one thread per (column, g-point), columns fastest, so that every layer
// step of a wavefront reads 64 consecutive columns of the (ncol, nlay, ngpt) inputs.
//
// Rescaled solver: per Gauss angle three sweeps -- down (the plain transport), up with the
// adjustment of Eq. 18b/20 (reading the first down radiance at the top of each layer), and down
// again with that of Eq. 19/21 (reading the up radiance at level ilay of layer ilay, which is the
// layer's top when top_at_1 and its bottom otherwise, as the reference does).  The first down and
// the up radiances of the angle are kept per g-point in scratch; the layer terms are recomputed
// in each sweep (one exponential).
//
// Two-stream solver: the adding method as two sweeps per (column, g-point).  The sweep up from the
// surface carries the albedo A and source S of the column below each level in registers and
// stores them in the flux outputs (flux_up holds A, flux_dn holds S); the sweep down reads them
// back level by level, recomputes the layer, and overwrites them with the fluxes.  Nothing else is
// written.  The adding denominators 1/(1 - R A) are FP64 in both builds; the FP32 build evaluates
// the layer coefficients in a rearranged FP32 form and keeps FP64 near conservative scattering.
#include <algorithm>

#include "common.h"
#include "solver_math.h"

namespace rrtmgp_hip
{
constexpr int LWS_BLOCK = 256;
constexpr double LW_PI = 3.14159265358979323846;
constexpr double LW_DIFF_SEC = 1.66;     // diffusivity secant of lw_two_stream (F90:853)

// ---------------------------------------------------------------------------------------------
// Rescaled no-scattering solver.

// Terms of one layer at one angle: transmittance, sources of the two directions, and the
// rescaling coefficients An = 1 - T^2, Cn = 0.4 w (1 - g)/2 / scale (F90:168-200).
template<typename Real>
struct RsLayer { Real T, sdn, sup, An, Cn; };

template<typename Real>
__device__ __forceinline__ RsLayer<Real> rescaled_layer(Real tau, Real ssa, Real g, Real D, Real lay, Real ldn,
                                                        Real lup)
{
    const Real wb = ssa * (Real(1) - g) * Real(.5);
    const Real scale = Real(1) - ssa + wb;
    RsLayer<Real> r;
    r.Cn = Real(.4) * wb / scale;
    const Real t = tau * D * scale;
    Real fact;
    lw_layer(t, r.T, fact);
    const Real omT = one_minus_T(t, r.T);
    r.An = omT * (Real(1) + r.T);                      // 1 - T^2 without cancellation
    r.sdn = omT * ldn + Real(2) * fact * (lay - ldn);
    r.sup = omT * lup + Real(2) * fact * (lay - lup);
    return r;
}

template<typename Real, bool JAC>
__global__ __launch_bounds__(LWS_BLOCK)
void lw_rescaled_kernel(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1, int nmus,
                        const Real* __restrict__ secants, const Real* __restrict__ weights,
                        const Real* __restrict__ tau, const Real* __restrict__ ssa, const Real* __restrict__ gasy,
                        const Real* __restrict__ lay_src, const Real* __restrict__ lev_inc,
                        const Real* __restrict__ lev_dec, const Real* __restrict__ emis,
                        const Real* __restrict__ sfc_src, const Real* __restrict__ inc_flux,
                        const Real* __restrict__ sfc_src_jac, Real* __restrict__ flux_up,
                        Real* __restrict__ flux_dn, Real* __restrict__ jac, Real* __restrict__ dn1,
                        Real* __restrict__ upr)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncol * ngpt) return;
    const int64_t icol = i % ncol, igpt = i / ncol;
    const int64_t nlev = nlay + 1;
    const int64_t ls = ncol * nlay * igpt, vs = ncol * nlev * igpt;
    const Real* ldn_a = top_at_1 ? lev_inc : lev_dec;   // level source of the layer's lower edge...
    const Real* lup_a = top_at_1 ? lev_dec : lev_inc;   // ...for the down and up directions
    auto L = [&](int64_t lay) { return icol + ncol * lay + ls; };
    auto V = [&](int64_t lev) { return icol + ncol * lev + vs; };
    auto layer = [&](int64_t il, Real D) {
        const int64_t l = L(il);
        return rescaled_layer<Real>(tau[l], ssa[l], gasy[l], D, lay_src[l], ldn_a[l], lup_a[l]);
    };
    const int64_t top_lev = top_at_1 ? 0 : nlay, sfc_lev = top_at_1 ? nlay : 0;
    const Real e = emis[i], bsfc = sfc_src[i];
    const Real inc = inc_flux ? inc_flux[i] : Real(0);

    for (int mu = 0; mu < nmus; ++mu)
    {
        const Real D = secants[i + ncol * ngpt * mu];
        const Real fac = Real(2 * LW_PI) * weights[mu];
        auto out = [&](Real* f, int64_t lev, Real I) {
            const int64_t v = V(lev);
            f[v] = (mu == 0) ? fac * I : f[v] + fac * I;
        };
        // Down: the plain transport; dn1 keeps the radiance of every level.
        Real I = inc / fac;
        dn1[V(top_lev)] = I;
        for (int64_t j = 0; j < nlay; ++j)
        {
            const int64_t il = top_at_1 ? j : nlay - 1 - j;
            const RsLayer<Real> r = layer(il, D);
            I = r.T * I + r.sdn;
            dn1[V(top_at_1 ? il + 1 : il)] = I;
        }
        // Surface.
        Real Iu = I * (Real(1) - e) + e * bsfc;
        Real Ij = JAC ? e * sfc_src_jac[i] : Real(0);
        upr[V(sfc_lev)] = Iu;
        out(flux_up, sfc_lev, Iu);
        if (JAC) out(jac, sfc_lev, Ij);
        // Up with the adjustment from the first down radiance at the top of the layer.
        for (int64_t j = 0; j < nlay; ++j)
        {
            const int64_t il = top_at_1 ? nlay - 1 - j : j;
            const int64_t lev_top = top_at_1 ? il : il + 1;
            const RsLayer<Real> r = layer(il, D);
            const Real adj = r.Cn * (r.An * dn1[V(lev_top)] - r.T * r.sdn - r.sup);
            Iu = r.T * Iu + r.sup + adj;
            upr[V(lev_top)] = Iu;
            out(flux_up, lev_top, Iu);
            if (JAC)
            {
                Ij = r.T * Ij;
                out(jac, lev_top, Ij);
            }
        }
        // Down again with the adjustment from the up radiance at level il.
        Real I2 = inc / fac;
        out(flux_dn, top_lev, I2);
        for (int64_t j = 0; j < nlay; ++j)
        {
            const int64_t il = top_at_1 ? j : nlay - 1 - j;
            const RsLayer<Real> r = layer(il, D);
            const Real adj = r.Cn * (r.An * upr[V(il)] - r.T * r.sup - r.sdn);
            I2 = r.T * I2 + r.sdn + adj;
            out(flux_dn, top_at_1 ? il + 1 : il, I2);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Two-stream solver.

struct Lw2sLayer { double R, T, sup, sdn; };

// The reference expressions of lw_two_stream and lw_source_2str in FP64.
__device__ __forceinline__ Lw2sLayer lw_two_stream_f64(double tau, double w0, double g, double top, double bot)
{
    const double g1 = LW_DIFF_SEC * (1. - .5 * w0 * (1. + g));
    const double g2 = LW_DIFF_SEC * .5 * w0 * (1. - g);
    const double k = sqrt(fmax((g1 - g2) * (g1 + g2), Limits<double>::k_min));
    const double e = exp(-tau * k);
    const double e2 = e * e;
    const double rt = 1. / (k * (1. + e2) + g1 * (1. - e2));
    Lw2sLayer r;
    r.R = rt * g2 * (1. - e2);
    r.T = rt * 2. * k * e;
    if (tau > 1e-8)
    {
        const double Z = (bot - top) / (tau * (g1 + g2));
        const double zup_top = Z + top, zup_bot = Z + bot, zdn_top = -Z + top, zdn_bot = -Z + bot;
        r.sup = LW_PI * (zup_top - r.R * zdn_top - r.T * zup_bot);
        r.sdn = LW_PI * (zdn_bot - r.R * zup_bot - r.T * zdn_top);
    }
    else
    {
        r.sup = 0.;
        r.sdn = 0.;
    }
    return r;
}

// The same layer in FP32, rearranged so that FP32 keeps its accuracy:
//   k^2 = D (1 - w0) (g1 + g2)                    (g1 - g2 = D (1 - w0) exactly),
//   1 - e^{-2 k tau} = a (2 - a) with a = 1 - e^{-k tau} (series for small k tau),
// and the source term Z (1 + R - T) of the reference, whose two factors are large and small for
// thin layers, taken as (bot - top) RT (a / tau) (k a / (g1 + g2) + 2 - a), with a / tau from the
// series k (1 - e^{-y})/y for small y = k tau.  Returns false near conservative scattering
// (1 - w0 < 1e-6), where the FP64 form is kept.
__device__ __forceinline__ bool lw_two_stream_f32(float tau, float w0, float g, float top, float bot, Lw2sLayer& out)
{
    const float one_m_w0 = 1.f - w0;
    if (one_m_w0 < 1e-6f) return false;
    const float D = float(LW_DIFF_SEC);
    const float g1 = D * (1.f - .5f * w0 * (1.f + g));
    const float g2 = D * .5f * w0 * (1.f - g);
    const float gs = g1 + g2;
    const float k = __builtin_amdgcn_sqrtf(fmaxf(D * one_m_w0 * gs, 1e-12f));
    const float kt = k * tau;
    const float e = exp_neg_f32(kt);
    const float a = one_minus_exp_neg_f32(kt, e);
    const float c2 = a * (2.f - a);
    const float rt = rcp_f32(k * (2.f - c2) + g1 * c2);
    const float R = rt * g2 * c2;
    const float T = rt * 2.f * k * e;
    float sup = 0.f, sdn = 0.f;
    if (tau > 1e-8f)
    {
        // a / tau = k (1 - e^{-y}) / y, y = k tau.
        float phi = 1.f / 5040.f;
        phi = 1.f / 720.f - kt * phi;
        phi = 1.f / 120.f - kt * phi;
        phi = 1.f / 24.f - kt * phi;
        phi = 1.f / 6.f - kt * phi;
        phi = .5f - kt * phi;
        phi = 1.f - kt * phi;
        const float a_tau = (kt < .25f) ? k * phi : a / tau;
        const float ZF = (bot - top) * rt * a_tau * (k * a / gs + 2.f - a);
        const float pi = float(LW_PI);
        sup = pi * (ZF + top * (1.f - R) - T * bot);
        sdn = pi * (-ZF + bot * (1.f - R) - T * top);
    }
    out = {double(R), double(T), double(sup), double(sdn)};
    return true;
}

template<typename Real>
__device__ __forceinline__ Lw2sLayer lw_two_stream_layer(Real tau, Real w0, Real g, Real top, Real bot)
{
    if constexpr (std::is_same<Real, float>::value)
    {
        Lw2sLayer r;
        if (lw_two_stream_f32(tau, w0, g, top, bot, r)) return r;
    }
    return lw_two_stream_f64(double(tau), double(w0), double(g), double(top), double(bot));
}

template<typename Real>
__global__ __launch_bounds__(LWS_BLOCK)
void lw_2stream_kernel(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1, const Real* __restrict__ tau,
                       const Real* __restrict__ ssa, const Real* __restrict__ gasy,
                       const Real* __restrict__ lev_inc, const Real* __restrict__ lev_dec,
                       const Real* __restrict__ emis, const Real* __restrict__ sfc_src,
                       const Real* __restrict__ inc_flux, Real* __restrict__ flux_up, Real* __restrict__ flux_dn)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncol * ngpt) return;
    const int64_t icol = i % ncol, igpt = i / ncol;
    const int64_t nlev = nlay + 1;
    const int64_t ls = ncol * nlay * igpt, vs = ncol * nlev * igpt;
    auto L = [&](int64_t lay) { return icol + ncol * lay + ls; };
    auto V = [&](int64_t lev) { return icol + ncol * lev + vs; };
    // Level source of lw_combine_sources: the outer levels from their one layer, the inner ones
    // the geometric mean of the two layers' values.
    auto lev_src = [&](int64_t lev) -> Real {
        if (lev == 0) return lev_dec[L(0)];
        if (lev == nlay) return lev_inc[L(nlay - 1)];
        return sqrt(lev_dec[L(lev)] * lev_inc[L(lev - 1)]);
    };
    auto layer = [&](int64_t il) {
        const int64_t l = L(il);
        const Real top = lev_src(top_at_1 ? il : il + 1), bot = lev_src(top_at_1 ? il + 1 : il);
        return lw_two_stream_layer<Real>(tau[l], ssa[l], gasy[l], top, bot);
    };
    const int64_t top_lev = top_at_1 ? 0 : nlay, sfc_lev = top_at_1 ? nlay : 0;

    // Sweep up: albedo and source of the column below each level, stored in the flux outputs.
    const double e = double(emis[i]);
    double A = 1. - e;
    double S = LW_PI * e * double(sfc_src[i]);
    flux_up[V(sfc_lev)] = Real(A);
    flux_dn[V(sfc_lev)] = Real(S);
    for (int64_t j = 0; j < nlay; ++j)
    {
        const int64_t il = top_at_1 ? nlay - 1 - j : j;
        const Lw2sLayer r = layer(il);
        const double den = rcp_f64(1. - r.R * A);
        S = r.sup + r.T * den * (S + A * r.sdn);
        A = r.R + r.T * r.T * A * den;
        const int64_t lev = top_at_1 ? il : il + 1;
        flux_up[V(lev)] = Real(A);
        flux_dn[V(lev)] = Real(S);
    }
    // Top boundary, then the sweep down.
    double dn = inc_flux ? double(inc_flux[i]) : 0.;
    flux_dn[V(top_lev)] = Real(dn);
    flux_up[V(top_lev)] = Real(dn * A + S);
    for (int64_t j = 0; j < nlay; ++j)
    {
        const int64_t il = top_at_1 ? j : nlay - 1 - j;
        const int64_t lev = top_at_1 ? il + 1 : il;
        const Lw2sLayer r = layer(il);
        const double Ab = double(flux_up[V(lev)]), Sb = double(flux_dn[V(lev)]);
        dn = (r.T * dn + r.R * Sb + r.sdn) * rcp_f64(1. - r.R * Ab);
        flux_dn[V(lev)] = Real(dn);
        flux_up[V(lev)] = Real(dn * Ab + Sb);
    }
}

// ---------------------------------------------------------------------------------------------
// Launchers.

template<typename Real>
int lw_rescaled(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1, int64_t nmus, const Real* secants,
                const Real* weights, const Real* tau, const Real* ssa, const Real* g, const Real* lay,
                const Real* linc, const Real* ldec, const Real* emis, const Real* sfc_src, const Real* inc_flux,
                Real* flux_up, Real* flux_dn, int do_broadband, Real* bb_up, Real* bb_dn, int do_jac,
                const Real* sfc_src_jac, Real* flux_up_jac, void* stream);

template<typename Real>
int lw_2stream(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1, const Real* tau, const Real* ssa,
               const Real* g, const Real* linc, const Real* ldec, const Real* emis, const Real* sfc_src,
               const Real* inc_flux, Real* flux_up, Real* flux_dn, void* stream)
{
    if (ncol <= 0 || nlay <= 0 || ngpt <= 0) return arg_error("lw_solver_2stream: empty extents");
    if (!tau || !ssa || !g || !linc || !ldec || !emis || !sfc_src || !flux_up || !flux_dn)
        return arg_error("lw_solver_2stream: NULL array");
    hipStream_t s = as_stream(stream);
    RRTMGP_TIMED("lw_2stream_kernel", s);
    hipLaunchKernelGGL(lw_2stream_kernel<Real>, dim3(grid1(ncol * ngpt, LWS_BLOCK)), dim3(LWS_BLOCK), 0, s, ncol,
                       nlay, ngpt, top_at_1, tau, ssa, g, linc, ldec, emis, sfc_src, inc_flux, flux_up, flux_dn);
    RRTMGP_LAUNCHED("lw_2stream_kernel");
    return RRTMGP_OK;
}
}  // namespace rrtmgp_hip

using namespace rrtmgp_hip;

// Broadband sums of rrtmgp_hip_sum_broadband_* (optical_props_kernels.hip).
extern "C" int rrtmgp_hip_sum_broadband_f32(int64_t, int64_t, int64_t, const float*, float*, void*);
extern "C" int rrtmgp_hip_sum_broadband_f64(int64_t, int64_t, int64_t, const double*, double*, void*);

namespace rrtmgp_hip
{
inline int sum_bb(int64_t ncol, int64_t nlev, int64_t ngpt, const float* f, float* out, void* s)
{
    return rrtmgp_hip_sum_broadband_f32(ncol, nlev, ngpt, f, out, s);
}
inline int sum_bb(int64_t ncol, int64_t nlev, int64_t ngpt, const double* f, double* out, void* s)
{
    return rrtmgp_hip_sum_broadband_f64(ncol, nlev, ngpt, f, out, s);
}

template<typename Real>
int lw_rescaled(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1, int64_t nmus, const Real* secants,
                const Real* weights, const Real* tau, const Real* ssa, const Real* g, const Real* lay,
                const Real* linc, const Real* ldec, const Real* emis, const Real* sfc_src, const Real* inc_flux,
                Real* flux_up, Real* flux_dn, int do_broadband, Real* bb_up, Real* bb_dn, int do_jac,
                const Real* sfc_src_jac, Real* flux_up_jac, void* stream)
{
    if (ncol <= 0 || nlay <= 0 || ngpt <= 0) return arg_error("lw_solver_noscat (rescaled): empty extents");
    if (nmus < 1 || nmus > 4) return arg_error("lw_solver_noscat (rescaled): n_gauss_angles must be 1..4");
    if (!secants || !weights || !tau || !lay || !linc || !ldec || !emis || !sfc_src)
        return arg_error("lw_solver_noscat (rescaled): NULL input");
    if (!ssa || !g) return arg_error("lw_solver_noscat (rescaled): rescaling needs ssa and g");
    if (do_broadband ? (!bb_up || !bb_dn) : (!flux_up || !flux_dn))
        return arg_error("lw_solver_noscat (rescaled): NULL output");
    if (do_jac && (!do_broadband || !sfc_src_jac || !flux_up_jac))
        return arg_error("lw_solver_noscat (rescaled): the Jacobian needs broadband output");
    hipStream_t s = as_stream(stream);
    const int64_t ncvg = ncol * (nlay + 1) * ngpt;
    // Scratch: the first down and the up radiance of the current angle, and (broadband output)
    // the g-point fluxes and Jacobian before they are summed.
    const int64_t nbuf = 2 + (do_broadband ? 2 : 0) + (do_jac ? 1 : 0);
    Real* scratch = nullptr;
    RRTMGP_TRY(scratch_alloc(&scratch, size_t(ncvg * nbuf), s));
    Real* dn1 = scratch;
    Real* upr = scratch + ncvg;
    Real* gup = do_broadband ? scratch + 2 * ncvg : flux_up;
    Real* gdn = do_broadband ? scratch + 3 * ncvg : flux_dn;
    Real* gjac = do_jac ? scratch + 4 * ncvg : nullptr;
    int rc = RRTMGP_OK;
    {
        RRTMGP_TIMED("lw_rescaled_kernel", s);
        auto kernel = do_jac ? lw_rescaled_kernel<Real, true> : lw_rescaled_kernel<Real, false>;
        hipLaunchKernelGGL(kernel, dim3(grid1(ncol * ngpt, LWS_BLOCK)), dim3(LWS_BLOCK), 0, s, ncol, nlay, ngpt,
                           top_at_1, int(nmus), secants, weights, tau, ssa, g, lay, linc, ldec, emis, sfc_src,
                           inc_flux, sfc_src_jac, gup, gdn, gjac, dn1, upr);
        rc = check_hip(hipGetLastError(), "lw_rescaled_kernel");
    }
    if (rc == RRTMGP_OK && do_broadband)
    {
        rc = sum_bb(ncol, nlay + 1, ngpt, gup, bb_up, stream);
        if (rc == RRTMGP_OK) rc = sum_bb(ncol, nlay + 1, ngpt, gdn, bb_dn, stream);
        if (rc == RRTMGP_OK && do_jac) rc = sum_bb(ncol, nlay + 1, ngpt, gjac, flux_up_jac, stream);
    }
    scratch_free(scratch, s);
    return rc;
}
}  // namespace rrtmgp_hip

#define RRTMGP_DEF_LW_SCATTERING(S, T)                                                              \
extern "C" int rrtmgp_hip_lw_solver_noscat_rescaled_##S(int64_t ncol, int64_t nlay, int64_t ngpt, \
    int top_at_1, int64_t nmus, const T* secants, const T* weights, const T* tau, const T* ssa,     \
    const T* g, const T* lay_source, const T* lev_source_inc, const T* lev_source_dec,              \
    const T* sfc_emis, const T* sfc_src, const T* inc_flux, T* flux_up, T* flux_dn,                 \
    int do_broadband, T* broadband_up, T* broadband_dn, int do_jacobians, const T* sfc_src_jac,     \
    T* flux_up_jac, void* stream)                                                                    \
{                                                                                                    \
    return lw_rescaled<T>(ncol, nlay, ngpt, top_at_1, nmus, secants, weights, tau, ssa, g,          \
                          lay_source, lev_source_inc, lev_source_dec, sfc_emis, sfc_src, inc_flux,   \
                          flux_up, flux_dn, do_broadband, broadband_up, broadband_dn, do_jacobians,  \
                          sfc_src_jac, flux_up_jac, stream);                                         \
}                                                                                                    \
extern "C" int rrtmgp_hip_lw_solver_2stream_##S(int64_t ncol, int64_t nlay, int64_t ngpt,         \
    int top_at_1, const T* tau, const T* ssa, const T* g, const T* lay_source,                      \
    const T* lev_source_inc, const T* lev_source_dec, const T* sfc_emis, const T* sfc_src,          \
    const T* inc_flux, T* flux_up, T* flux_dn, void* stream)                                         \
{                                                                                                    \
    (void)lay_source;                                                                                \
    return lw_2stream<T>(ncol, nlay, ngpt, top_at_1, tau, ssa, g, lev_source_inc, lev_source_dec,   \
                         sfc_emis, sfc_src, inc_flux, flux_up, flux_dn, stream);                     \
}

RRTMGP_DEF_LW_SCATTERING(f32, float)
RRTMGP_DEF_LW_SCATTERING(f64, double)
