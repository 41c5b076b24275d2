// Longwave (no-scattering, Gaussian quadrature) and shortwave (two-stream + adding) solvers.
//
// One thread owns one (column, g-point) and runs the layer recurrences of that profile
// sequentially; lanes of a wavefront are consecutive columns, so each layer step of the
// wavefront reads and writes one contiguous 256-byte row of the (ncol, nlay|nlev, ngpt)
// arrays.  Arithmetic follows src_kernels/mo_rte_solver_kernels.F90 (lw_solver_noscat :58-260,
// lw_source_noscat :616-653, transport :659-723, sw_solver_2stream :512-605,
// sw_dif_and_source :996-1109, adding :1117-1227) with the Rdir/Tdir clamp of
// src_kernels_cuda/rte_solver_kernels.cu:581-583.
#include "common.h"

namespace rrtmgp_hip
{
constexpr int SV_BX = 64;   // columns per work-group
constexpr int SV_BY = 4;    // g-points per work-group

template<typename Real>
__global__ __launch_bounds__(SV_BX * SV_BY)
void lw_solver_kernel(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1, int nmus,
                      const Real* __restrict__ secants, const Real* __restrict__ weights,
                      const Real* __restrict__ tau, const Real* __restrict__ lay_src,
                      const Real* __restrict__ lev_inc, const Real* __restrict__ lev_dec,
                      const Real* __restrict__ emis, const Real* __restrict__ sfc_src,
                      const Real* __restrict__ inc_flux, Real* __restrict__ flux_up,
                      Real* __restrict__ flux_dn, const Real* __restrict__ sfc_src_jac,
                      Real* __restrict__ jac)
{
    const int64_t icol = int64_t(blockIdx.x) * SV_BX + threadIdx.x;
    const int64_t igpt = int64_t(blockIdx.y) * SV_BY + threadIdx.y;
    if (icol >= ncol || igpt >= ngpt) return;
    const int64_t nlev = nlay + 1;
    const int64_t c2 = icol + ncol * igpt;
    const int64_t lb = icol + ncol * nlay * igpt;     // layer (ilay) offset: + ncol*ilay
    const int64_t vb = icol + ncol * nlev * igpt;     // level offset
    const Real* lev_dn = top_at_1 ? lev_inc : lev_dec;
    const Real* lev_up = top_at_1 ? lev_dec : lev_inc;
    const Real tau_thresh = sqrt(Limits<Real>::eps);
    const Real e = emis[c2];
    const Real ssrc = sfc_src[c2];
    const Real pi = Real(3.14159265358979323846);
    const int64_t top_lev = top_at_1 ? 0 : nlay;
    const int64_t sfc_lev = top_at_1 ? nlay : 0;

    for (int imu = 0; imu < nmus; ++imu)
    {
        const Real D = secants[c2 + ncol * ngpt * imu];
        const Real fac = Real(2) * pi * weights[imu];
        auto put = [&](Real* arr, int64_t lev, Real v) {
            const int64_t o = vb + ncol * lev;
            arr[o] = (imu == 0) ? v : arr[o] + v;
        };
        Real I = (inc_flux ? inc_flux[c2] : Real(0)) / fac;
        put(flux_dn, top_lev, fac * I);
        for (int64_t j = 0; j < nlay; ++j)
        {
            const int64_t ilay = top_at_1 ? j : nlay - 1 - j;
            const int64_t l = lb + ncol * ilay;
            const Real t = tau[l] * D;
            const Real T = exp(-t);
            const Real fact = (t > tau_thresh) ? (Real(1) - T) / t - T : t * (Real(0.5) - Real(1. / 3.) * t);
            const Real ld = lev_dn[l];
            const Real src = (Real(1) - T) * ld + Real(2) * fact * (lay_src[l] - ld);
            I = T * I + src;
            put(flux_dn, top_at_1 ? ilay + 1 : ilay, fac * I);
        }
        Real Iu = I * (Real(1) - e) + e * ssrc;
        put(flux_up, sfc_lev, fac * Iu);
        Real Ij = Real(0);
        if (jac)
        {
            Ij = e * sfc_src_jac[c2];
            put(jac, sfc_lev, fac * Ij);
        }
        for (int64_t j = 0; j < nlay; ++j)
        {
            const int64_t ilay = top_at_1 ? nlay - 1 - j : j;
            const int64_t l = lb + ncol * ilay;
            const Real t = tau[l] * D;
            const Real T = exp(-t);
            const Real fact = (t > tau_thresh) ? (Real(1) - T) / t - T : t * (Real(0.5) - Real(1. / 3.) * t);
            const Real lu = lev_up[l];
            const Real src = (Real(1) - T) * lu + Real(2) * fact * (lay_src[l] - lu);
            Iu = T * Iu + src;
            const int64_t lev = top_at_1 ? ilay : ilay + 1;
            put(flux_up, lev, fac * Iu);
            if (jac)
            {
                Ij = T * Ij;
                put(jac, lev, fac * Ij);
            }
        }
    }
}

// Two-stream layer properties of sw_dif_and_source (F90:1043-1104) for one layer.
template<typename Real>
__device__ __forceinline__ void sw_two_stream(Real tau, Real w0, Real g, Real mu0, Real& Rdif, Real& Tdif,
                                              Real& Rdir, Real& Tdir, Real& Tnoscat)
{
    const Real eps = Limits<Real>::eps;
    const Real gamma1 = (Real(8) - w0 * (Real(5) + Real(3) * g)) * Real(.25);
    const Real gamma2 = Real(3) * (w0 * (Real(1) - g)) * Real(.25);
    const Real gamma3 = (Real(2) - Real(3) * mu0 * g) * Real(.25);
    const Real gamma4 = Real(1) - gamma3;
    const Real alpha1 = gamma1 * gamma4 + gamma2 * gamma3;
    const Real alpha2 = gamma1 * gamma3 + gamma2 * gamma4;
    const Real k = sqrt(max((gamma1 - gamma2) * (gamma1 + gamma2), Limits<Real>::k_min));
    const Real k_mu = k * mu0, k_gamma3 = k * gamma3, k_gamma4 = k * gamma4;
    const Real e1 = exp(-tau * k);
    const Real e2 = e1 * e1;
    const Real RT = Real(1) / (k * (Real(1) + e2) + gamma1 * (Real(1) - e2));
    Rdif = RT * gamma2 * (Real(1) - e2);
    Tdif = RT * Real(2) * k * e1;
    const Real d = Real(1) - k_mu * k_mu;
    const Real denom = (fabs(d) >= eps) ? d : eps;
    const Real RT2 = w0 * RT / denom;
    Tnoscat = exp(-tau / mu0);
    Real rd = RT2 * ((Real(1) - k_mu) * (alpha2 + k_gamma3) - (Real(1) + k_mu) * (alpha2 - k_gamma3) * e2 -
                     Real(2) * (k_gamma3 - alpha2 * k_mu) * e1 * Tnoscat);
    Real td = -RT2 * ((Real(1) + k_mu) * (alpha1 + k_gamma4) * Tnoscat -
                      (Real(1) - k_mu) * (alpha1 - k_gamma4) * e2 * Tnoscat -
                      Real(2) * (k_gamma4 + alpha1 * k_mu) * e1);
    rd = max(eps, min(rd, Real(1) - Tnoscat));
    td = max(eps, min(td, Real(1) - Tnoscat - rd));
    Rdir = rd;
    Tdir = td;
}

// Shortwave: direct beam + layer properties (top -> surface), adding up-sweep (surface -> top),
// flux down-sweep (top -> surface).  R, T, SU, SD are per-layer scratch arrays; the level
// albedo and source of the up-sweep are parked in flux_up / flux_dn until the down-sweep
// overwrites them with the fluxes.
template<typename Real>
__global__ __launch_bounds__(SV_BX * SV_BY)
void sw_solver_kernel(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1,
                      const Real* __restrict__ tau, const Real* __restrict__ ssa,
                      const Real* __restrict__ g, const Real* __restrict__ mu0,
                      const Real* __restrict__ alb_dir, const Real* __restrict__ alb_dif,
                      const Real* __restrict__ inc_dir, const Real* __restrict__ inc_dif,
                      Real* __restrict__ flux_up, Real* __restrict__ flux_dn, Real* __restrict__ flux_dir,
                      Real* __restrict__ R, Real* __restrict__ T, Real* __restrict__ SU, Real* __restrict__ SD)
{
    const int64_t icol = int64_t(blockIdx.x) * SV_BX + threadIdx.x;
    const int64_t igpt = int64_t(blockIdx.y) * SV_BY + threadIdx.y;
    if (icol >= ncol || igpt >= ngpt) return;
    const int64_t nlev = nlay + 1;
    const int64_t c2 = icol + ncol * igpt;
    const int64_t lb = icol + ncol * nlay * igpt;
    const int64_t vb = icol + ncol * nlev * igpt;
    const Real m0 = mu0[icol];
    const int64_t top_lev = top_at_1 ? 0 : nlay;
    const int64_t sfc_lev = top_at_1 ? nlay : 0;

    Real dir = inc_dir[c2] * m0;
    flux_dir[vb + ncol * top_lev] = dir;
    for (int64_t j = 0; j < nlay; ++j)
    {
        const int64_t ilay = top_at_1 ? j : nlay - 1 - j;
        const int64_t l = lb + ncol * ilay;
        Real rd, td, rb, tb, tn;
        sw_two_stream(tau[l], ssa[l], g[l], m0, rd, td, rb, tb, tn);
        R[l] = rd; T[l] = td;
        SU[l] = rb * dir;
        SD[l] = tb * dir;
        dir = tn * dir;
        flux_dir[vb + ncol * (top_at_1 ? ilay + 1 : ilay)] = dir;
    }
    // adding (SH08): albedo and source of the column below each level
    Real alb = alb_dif[c2];
    Real src = dir * alb_dir[c2];
    flux_up[vb + ncol * sfc_lev] = alb;
    flux_dn[vb + ncol * sfc_lev] = src;
    for (int64_t j = 0; j < nlay; ++j)
    {
        const int64_t ilay = top_at_1 ? nlay - 1 - j : j;
        const int64_t l = lb + ncol * ilay;
        const Real rd = R[l], td = T[l];
        const Real denom = Real(1) / (Real(1) - rd * alb);
        const Real alb_new = rd + td * td * alb * denom;
        src = SU[l] + td * denom * (src + alb * SD[l]);
        alb = alb_new;
        SU[l] = denom;
        const int64_t lev = top_at_1 ? ilay : ilay + 1;
        flux_up[vb + ncol * lev] = alb;
        flux_dn[vb + ncol * lev] = src;
    }
    Real dn = inc_dif ? inc_dif[c2] : Real(0);
    {
        const int64_t o = vb + ncol * top_lev;
        const Real up = dn * flux_up[o] + flux_dn[o];
        flux_up[o] = up;
        flux_dn[o] = dn + flux_dir[o];
    }
    for (int64_t j = 0; j < nlay; ++j)
    {
        const int64_t ilay = top_at_1 ? j : nlay - 1 - j;
        const int64_t l = lb + ncol * ilay;
        const int64_t o = vb + ncol * (top_at_1 ? ilay + 1 : ilay);
        const Real a_b = flux_up[o], s_b = flux_dn[o];
        dn = (T[l] * dn + R[l] * s_b + SD[l]) * SU[l];
        flux_up[o] = dn * a_b + s_b;
        flux_dn[o] = dn + flux_dir[o];
    }
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

template<typename Real>
__global__ void sum_gpt_kernel(int64_t ncv, int64_t ngpt, const Real* __restrict__ f, Real* __restrict__ out)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncv) return;
    Real s = Real(0);
    for (int64_t g = 0; g < ngpt; ++g) s += f[i + ncv * g];
    out[i] = s;
}

template<typename Real>
int lw_solver(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1, int64_t nmus, const Real* secants,
              const Real* weights, const Real* tau, const Real* lay, const Real* linc, const Real* ldec,
              const Real* emis, const Real* sfc_src, const Real* inc_flux, Real* flux_up, Real* flux_dn,
              int do_broadband, Real* bb_up, Real* bb_dn, int do_jac, const Real* sfc_src_jac, Real* flux_up_jac,
              void* stream)
{
    if (ncol <= 0 || nlay <= 0 || ngpt <= 0) return arg_error("lw_solver: empty extents");
    if (nmus < 1 || nmus > 4) return arg_error("lw_solver: n_gauss_angles must be 1..4");
    if (!secants || !weights || !tau || !lay || !linc || !ldec || !emis || !sfc_src)
        return arg_error("lw_solver: NULL input");
    if (do_broadband ? (!bb_up || !bb_dn) : (!flux_up || !flux_dn)) return arg_error("lw_solver: NULL output");
    if (do_jac && (!sfc_src_jac || !flux_up_jac)) return arg_error("lw_solver: NULL Jacobian array");
    hipStream_t s = as_stream(stream);
    const int64_t ncv = ncol * (nlay + 1);
    Real *up = flux_up, *dn = flux_dn, *jac = nullptr;
    if (do_broadband)
    {
        RRTMGP_TRY(scratch_alloc(&up, size_t(ncv * ngpt), s));
        RRTMGP_TRY(scratch_alloc(&dn, size_t(ncv * ngpt), s));
    }
    if (do_jac) RRTMGP_TRY(scratch_alloc(&jac, size_t(ncv * ngpt), s));
    const dim3 grid(grid1(ncol, SV_BX), grid1(ngpt, SV_BY));
    {
        RRTMGP_TIMED("lw_solver_kernel", s);
        hipLaunchKernelGGL(lw_solver_kernel<Real>, grid, dim3(SV_BX, SV_BY), 0, s, ncol, nlay, ngpt, top_at_1,
                           int(nmus), secants, weights, tau, lay, linc, ldec, emis, sfc_src, inc_flux, up, dn,
                           sfc_src_jac, jac);
    }
    int rc = check_hip(hipGetLastError(), "lw_solver_kernel");
    if (rc == RRTMGP_OK && do_broadband)
    {
        for (int i = 0; i < 2; ++i)
        {
            RRTMGP_TIMED("sum_gpt_kernel", s);
            hipLaunchKernelGGL(sum_gpt_kernel<Real>, dim3(grid1(ncv, 256)), dim3(256), 0, s, ncv, ngpt,
                               i == 0 ? up : dn, i == 0 ? bb_up : bb_dn);
        }
        rc = check_hip(hipGetLastError(), "sum_gpt_kernel");
    }
    if (rc == RRTMGP_OK && do_jac)
    {
        RRTMGP_TIMED("sum_gpt_kernel", s);
        hipLaunchKernelGGL(sum_gpt_kernel<Real>, dim3(grid1(ncv, 256)), dim3(256), 0, s, ncv, ngpt, jac, flux_up_jac);
        rc = check_hip(hipGetLastError(), "sum_gpt_kernel");
    }
    if (do_broadband) { scratch_free(up, s); scratch_free(dn, s); }
    scratch_free(jac, s);
    return rc;
}

template<typename Real>
int sw_solver(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1, const Real* tau, const Real* ssa,
              const Real* g, const Real* mu0, const Real* alb_dir, const Real* alb_dif, const Real* inc_dir,
              Real* flux_up, Real* flux_dn, Real* flux_dir, int has_dif_bc, const Real* inc_dif,
              int do_broadband, Real* bb_up, Real* bb_dn, Real* bb_dir, void* stream)
{
    if (ncol <= 0 || nlay <= 0 || ngpt <= 0) return arg_error("sw_solver: empty extents");
    if (!tau || !ssa || !g || !mu0 || !alb_dir || !alb_dif || !inc_dir) return arg_error("sw_solver: NULL input");
    if (do_broadband ? (!bb_up || !bb_dn || !bb_dir) : (!flux_up || !flux_dn || !flux_dir))
        return arg_error("sw_solver: NULL output");
    if (has_dif_bc && !inc_dif) return arg_error("sw_solver: has_dif_bc without inc_flux_dif");
    hipStream_t s = as_stream(stream);
    const int64_t ncv = ncol * (nlay + 1), ng = ncol * nlay * ngpt;
    Real *up = flux_up, *dn = flux_dn, *dr = flux_dir;
    Real *R = nullptr, *T = nullptr, *SU = nullptr, *SD = nullptr;
    if (do_broadband)
    {
        RRTMGP_TRY(scratch_alloc(&up, size_t(ncv * ngpt), s));
        RRTMGP_TRY(scratch_alloc(&dn, size_t(ncv * ngpt), s));
        RRTMGP_TRY(scratch_alloc(&dr, size_t(ncv * ngpt), s));
    }
    RRTMGP_TRY(scratch_alloc(&R, size_t(ng), s));
    RRTMGP_TRY(scratch_alloc(&T, size_t(ng), s));
    RRTMGP_TRY(scratch_alloc(&SU, size_t(ng), s));
    RRTMGP_TRY(scratch_alloc(&SD, size_t(ng), s));
    const dim3 grid(grid1(ncol, SV_BX), grid1(ngpt, SV_BY));
    {
        RRTMGP_TIMED("sw_solver_kernel", s);
        hipLaunchKernelGGL(sw_solver_kernel<Real>, grid, dim3(SV_BX, SV_BY), 0, s, ncol, nlay, ngpt, top_at_1,
                           tau, ssa, g, mu0, alb_dir, alb_dif, inc_dir, has_dif_bc ? inc_dif : nullptr,
                           up, dn, dr, R, T, SU, SD);
    }
    int rc = check_hip(hipGetLastError(), "sw_solver_kernel");
    if (rc == RRTMGP_OK && do_broadband)
    {
        const Real* src[3] = {up, dn, dr};
        Real* dst[3] = {bb_up, bb_dn, bb_dir};
        for (int i = 0; i < 3; ++i)
        {
            RRTMGP_TIMED("sum_gpt_kernel", s);
            hipLaunchKernelGGL(sum_gpt_kernel<Real>, dim3(grid1(ncv, 256)), dim3(256), 0, s, ncv, ngpt, src[i], dst[i]);
        }
        rc = check_hip(hipGetLastError(), "sum_gpt_kernel");
    }
    scratch_free(R, s); scratch_free(T, s); scratch_free(SU, s); scratch_free(SD, s);
    if (do_broadband) { scratch_free(up, s); scratch_free(dn, s); scratch_free(dr, s); }
    return rc;
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
                        lev_source_inc, lev_source_dec, sfc_emis, sfc_src, inc_flux, flux_up,        \
                        flux_dn, do_broadband, broadband_up, broadband_dn, do_jacobians,             \
                        sfc_src_jac, flux_up_jac, stream);                                           \
}                                                                                                    \
extern "C" int rrtmgp_hip_sw_solver_2stream_##S(int64_t ncol, int64_t nlay, int64_t ngpt,         \
    int top_at_1, const T* tau, const T* ssa, const T* g, const T* mu0, const T* sfc_alb_dir,       \
    const T* sfc_alb_dif, const T* inc_flux_dir, T* flux_up, T* flux_dn, T* flux_dir,               \
    int has_dif_bc, const T* inc_flux_dif, int do_broadband, T* broadband_up, T* broadband_dn,      \
    T* broadband_dir, void* stream)                                                                  \
{                                                                                                    \
    return sw_solver<T>(ncol, nlay, ngpt, top_at_1, tau, ssa, g, mu0, sfc_alb_dir, sfc_alb_dif,     \
                        inc_flux_dir, flux_up, flux_dn, flux_dir, has_dif_bc, inc_flux_dif,          \
                        do_broadband, broadband_up, broadband_dn, broadband_dir, stream);            \
}

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
