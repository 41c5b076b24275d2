// Adjoint (reverse-mode) kernels of the RTE solvers.  No counterpart in the reference, whose LW
// path has only the surface-temperature Jacobian (src/Rte_lw.cpp:140, do_jacobians).
//
// Longwave.  The forward solver (lw_solver_noscat, F90:58-376) runs, per column, g-point and
// angle, with layers j = 0..n-1 counted from the top and levels j = 0..n:
//     down  I^d_0 = inc / fac,   I^d_{j+1} = T_j I^d_j + Sd_j
//     sfc   I^u_n = (1 - e) I^d_n + e B_s
//     up    I^u_j = T_j I^u_{j+1} + Su_j
// with T = e^{-t}, t = D tau, f = (1 - T)/t - T, Sd = (1 - T) Bd + 2 f (Bl - Bd), Su likewise
// with Bu; broadband fluxes are sums of fac I over g-points and angles.  For cotangents a^u, a^d
// of the broadband fluxes at every level, the adjoint intensities obey
//     lu_0 = fac a^u_0,          lu_{j+1} = fac a^u_{j+1} + T_j lu_j        (top-down)
//     ld_n = fac a^d_n + (1 - e) lu_n,   ld_j = fac a^d_j + T_j ld_{j+1}   (bottom-up)
// and layer j receives dL/dT_j = lu_j I^u_{j+1} + ld_{j+1} I^d_j, dL/dSu_j = lu_j,
// dL/dSd_j = ld_{j+1}.  lu needs only T, so it runs top-down next to the forward down sweep;
// a bottom-up sweep then carries I^u and ld together and replays each K-layer segment's I^d and
// lu from checkpoints (the two-sided scheme of the forward SW solver).  Per (column, g-point)
// the inputs are read twice; no per-layer state is stored.
//
// Shortwave.  The forward is the two-sided adding solver of the SW solver kernel, per column and
// g-point, with layers j counted from the top: the direct beam dir_{j+1} = dir_j Tn_j with layer
// sources SU_j = Rdir_j dir_j, SD_j = Tdir_j dir_j; top-down D_{j+1} = SD + T (D + Ru SU) e1,
// Ru_{j+1} = R + T^2 Ru e1, e1 = 1/(1 - R Ru); bottom-up A_j = R + T^2 A' e2,
// S_j = SU + T (S' + A' SD) e2, e2 = 1/(1 - R A') (A' = A_{j+1}); fluxes dn = (D + Ru S)/(1 - Ru A),
// up = A dn + S, total down dn + dir.  The adjoint stores these states (in the build's precision, per column,
// g-point and level, in a scratch buffer sized by the caller's g-point range), then runs the
// reverse of the bottom-up recursion top-down and the reverse of the top-down recursion and the
// direct beam bottom-up.  The layer Jacobian d(R, T, Rdir, Tdir, Tn)/d(tau, w0, g) comes from the
// FP64 layer formulas evaluated on dual numbers.
#include <algorithm>

#include "common.h"
#include "two_stream.h"

namespace rrtmgp_hip
{
constexpr int ADJ_COLS = 64;
constexpr int ADJ_GPTS = 16;
template<typename Real> struct AdjSeg;
template<> struct AdjSeg<float> { static constexpr int K = 4; };
template<> struct AdjSeg<double> { static constexpr int K = 4; };

template<typename Real>
__device__ __forceinline__ const Real& ld32(const Real* base, uint32_t i)
{
    return *reinterpret_cast<const Real*>(reinterpret_cast<const char*>(base) + i * uint32_t(sizeof(Real)));
}

template<typename Real>
__device__ __forceinline__ Real& st32(Real* base, uint32_t i)
{
    return *reinterpret_cast<Real*>(reinterpret_cast<char*>(base) + i * uint32_t(sizeof(Real)));
}

// T, f and df/dt of the linear-in-tau source (lw_source_noscat, F90:616-653).  Below the switch
// point both f and f' come from the series f = t/2 - t^2/3 + t^3/8 - t^4/30 + t^5/144 - t^6/840
// + t^7/5760; above it from the closed forms, f' = (t T - 1 + T)/t^2 + T.  The switch (1/4 in
// FP32, 1/100 in FP64) keeps both the series truncation and the cancellation of the closed
// form below the precision's rounding.
template<typename Real>
__device__ __forceinline__ void lw_layer_d(Real t, Real& T, Real& f, Real& df)
{
    constexpr Real sw = sizeof(Real) == 4 ? Real(0.25) : Real(0.01);
    T = exp(-t);
    Real q = Real(1. / 840.) - t * Real(1. / 5760.);
    q = Real(1. / 144.) - t * q;
    q = Real(1. / 30.) - t * q;
    q = Real(1. / 8.) - t * q;
    q = Real(1. / 3.) - t * q;
    q = Real(.5) - t * q;
    Real dq = Real(6. / 840.) - t * Real(7. / 5760.);
    dq = Real(5. / 144.) - t * dq;
    dq = Real(4. / 30.) - t * dq;
    dq = Real(3. / 8.) - t * dq;
    dq = Real(2. / 3.) - t * dq;
    dq = Real(.5) - t * dq;
    if (t < sw)
    {
        f = t * q;
        df = dq;
    }
    else
    {
        const Real r = Real(1) / t;
        f = (Real(1) - T) * r - T;
        df = (t * T - Real(1) + T) * r * r + T;
    }
}

// One (column, g-point) per thread, 64 columns x 16 g-points per work-group.
template<typename Real, int NMU>
__global__ __launch_bounds__(ADJ_COLS * ADJ_GPTS)
void lw_adjoint_kernel(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1, int nmus,
                       const Real* __restrict__ secants, const Real* __restrict__ weights,
                       const Real* __restrict__ tau, const Real* __restrict__ lay_src,
                       const Real* __restrict__ lev_inc, const Real* __restrict__ lev_dec,
                       const Real* __restrict__ emis, const Real* __restrict__ sfc_src,
                       const Real* __restrict__ inc_flux, const Real* __restrict__ a_up,
                       const Real* __restrict__ a_dn, Real* __restrict__ ckpt,
                       Real* __restrict__ d_tau, Real* __restrict__ d_lay, Real* __restrict__ d_lev_inc,
                       Real* __restrict__ d_lev_dec, Real* __restrict__ d_sfc_src, Real* __restrict__ d_emis)
{
    constexpr int K = AdjSeg<Real>::K;
    const int lane = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.y));
    const int64_t icol = int64_t(blockIdx.x) * ADJ_COLS + lane;
    const int64_t igpt = int64_t(blockIdx.y) * ADJ_GPTS + w;
    if (icol >= ncol || igpt >= ngpt) return;     // no barriers in this kernel
    const int64_t n = nlay;
    const int64_t nseg = (n + K - 1) / K;
    const int64_t c2 = icol + ncol * igpt;
    const int64_t lslab = ncol * nlay * igpt;
    const Real* tau_g = tau + lslab;
    const Real* lay_g = lay_src + lslab;
    const Real* bd_g = (top_at_1 ? lev_inc : lev_dec) + lslab;     // source at the layer's lower level
    const Real* bu_g = (top_at_1 ? lev_dec : lev_inc) + lslab;     // source at the layer's upper level
    Real* dtau_g = d_tau + lslab;
    Real* dlay_g = d_lay + lslab;
    Real* dbd_g = (top_at_1 ? d_lev_inc : d_lev_dec) + lslab;
    Real* dbu_g = (top_at_1 ? d_lev_dec : d_lev_inc) + lslab;
    const uint32_t ucol = uint32_t(icol), uncol = uint32_t(ncol);
    auto lay_of = [&](int64_t j) -> uint32_t { return uint32_t(top_at_1 ? j : n - 1 - j); };   // layer j from the top
    auto lev_of = [&](int64_t j) -> uint32_t { return uint32_t(top_at_1 ? j : n - j); };       // level j from the top
    // Checkpoint q of segment s: q < NMU: I^d, NMU <= q: lu, at the segment's top level.
    auto ck = [&](int64_t s, int q) -> Real& { return ckpt[((s * 2 * NMU + q) * ngpt + igpt) * ncol + icol]; };
    const Real pi = Real(3.14159265358979323846);

    Real D[NMU], fac[NMU], Id[NMU], Lu[NMU];
    const Real inc = inc_flux ? inc_flux[c2] : Real(0);
#pragma unroll
    for (int mu = 0; mu < NMU; ++mu)
    {
        const bool on = mu < nmus;
        D[mu] = on ? secants[c2 + ncol * ngpt * mu] : Real(0);
        fac[mu] = on ? Real(2) * pi * weights[mu] : Real(0);
        Id[mu] = on ? inc / fac[mu] : Real(0);
        Lu[mu] = fac[mu] * ld32(a_up, ucol + uncol * lev_of(0));
    }

    // Pass 1, top-down: forward I^d and the adjoint lu.
    for (int64_t s = 0; s < nseg; ++s)
    {
        const int64_t j0 = s * K;
        const int nk = int(std::min<int64_t>(K, n - j0));
#pragma unroll
        for (int mu = 0; mu < NMU; ++mu) { ck(s, mu) = Id[mu]; ck(s, NMU + mu) = Lu[mu]; }
        Real tv[K], lv[K], bv[K], au[K];
#pragma unroll
        for (int kk = 0; kk < K; ++kk)
            if (kk < nk)
            {
                const uint32_t l = ucol + uncol * lay_of(j0 + kk);
                tv[kk] = ld32(tau_g, l); lv[kk] = ld32(lay_g, l); bv[kk] = ld32(bd_g, l);
                au[kk] = ld32(a_up, ucol + uncol * lev_of(j0 + kk + 1));
            }
#pragma unroll
        for (int kk = 0; kk < K; ++kk)
            if (kk < nk)
#pragma unroll
                for (int mu = 0; mu < NMU; ++mu)
                {
                    Real T, f, df;
                    lw_layer_d(tv[kk] * D[mu], T, f, df);
                    Id[mu] = T * Id[mu] + ((Real(1) - T) * bv[kk] + Real(2) * f * (lv[kk] - bv[kk]));
                    Lu[mu] = fac[mu] * au[kk] + T * Lu[mu];
                }
    }

    // Surface.
    const Real e = emis[c2], bs = sfc_src[c2];
    const Real ad_s = ld32(a_dn, ucol + uncol * lev_of(n));
    Real Iu[NMU], Ld[NMU];
    Real g_e = Real(0), g_bs = Real(0);
#pragma unroll
    for (int mu = 0; mu < NMU; ++mu)
    {
        Iu[mu] = (Real(1) - e) * Id[mu] + e * bs;
        g_e += Lu[mu] * (bs - Id[mu]);
        g_bs += e * Lu[mu];
        Ld[mu] = fac[mu] * ad_s + (Real(1) - e) * Lu[mu];
    }
    d_emis[c2] = g_e;
    d_sfc_src[c2] = g_bs;

    // Pass 2, bottom-up: forward I^u and the adjoint ld, replaying I^d and lu per segment.
    for (int64_t s = nseg - 1; s >= 0; --s)
    {
        const int64_t j0 = s * K;
        const int nk = int(std::min<int64_t>(K, n - j0));
        Real tv[K], lv[K], bdv[K], buv[K], au[K], adv[K];
#pragma unroll
        for (int kk = 0; kk < K; ++kk)
            if (kk < nk)
            {
                const uint32_t l = ucol + uncol * lay_of(j0 + kk);
                tv[kk] = ld32(tau_g, l); lv[kk] = ld32(lay_g, l);
                bdv[kk] = ld32(bd_g, l); buv[kk] = ld32(bu_g, l);
                au[kk] = ld32(a_up, ucol + uncol * lev_of(j0 + kk + 1));
                adv[kk] = ld32(a_dn, ucol + uncol * lev_of(j0 + kk));
            }
        Real sId[K][NMU], sLu[K][NMU], sT[K][NMU], sF[K][NMU], sDF[K][NMU];
        {
            Real id[NMU], lu[NMU];
#pragma unroll
            for (int mu = 0; mu < NMU; ++mu) { id[mu] = ck(s, mu); lu[mu] = ck(s, NMU + mu); }
#pragma unroll
            for (int kk = 0; kk < K; ++kk)
                if (kk < nk)
#pragma unroll
                    for (int mu = 0; mu < NMU; ++mu)
                    {
                        Real T, f, df;
                        lw_layer_d(tv[kk] * D[mu], T, f, df);
                        sId[kk][mu] = id[mu]; sLu[kk][mu] = lu[mu];
                        sT[kk][mu] = T; sF[kk][mu] = f; sDF[kk][mu] = df;
                        id[mu] = T * id[mu] + ((Real(1) - T) * bdv[kk] + Real(2) * f * (lv[kk] - bdv[kk]));
                        lu[mu] = fac[mu] * au[kk] + T * lu[mu];
                    }
        }
#pragma unroll
        for (int kk = K - 1; kk >= 0; --kk)
            if (kk < nk)
            {
                Real g_tau = Real(0), g_bl = Real(0), g_bd = Real(0), g_bu = Real(0);
#pragma unroll
                for (int mu = 0; mu < NMU; ++mu)
                {
                    const Real T = sT[kk][mu], f = sF[kk][mu];
                    const Real gSu = sLu[kk][mu], gSd = Ld[mu];
                    const Real gT = sLu[kk][mu] * Iu[mu] + Ld[mu] * sId[kk][mu]
                                  - bdv[kk] * gSd - buv[kk] * gSu;
                    const Real gf = Real(2) * ((lv[kk] - bdv[kk]) * gSd + (lv[kk] - buv[kk]) * gSu);
                    g_tau += D[mu] * (-T * gT + sDF[kk][mu] * gf);
                    g_bd += gSd * ((Real(1) - T) - Real(2) * f);
                    g_bu += gSu * ((Real(1) - T) - Real(2) * f);
                    g_bl += Real(2) * f * (gSd + gSu);
                    Iu[mu] = T * Iu[mu] + ((Real(1) - T) * buv[kk] + Real(2) * f * (lv[kk] - buv[kk]));
                    Ld[mu] = fac[mu] * adv[kk] + T * Ld[mu];
                }
                const uint32_t l = ucol + uncol * lay_of(j0 + kk);
                st32(dtau_g, l) = g_tau;
                st32(dlay_g, l) = g_bl;
                st32(dbd_g, l) = g_bd;
                st32(dbu_g, l) = g_bu;
            }
    }
}

template<typename Real>
int lw_adjoint(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1, int64_t nmus, const Real* secants,
               const Real* weights, const Real* tau, const Real* lay, const Real* linc, const Real* ldec,
               const Real* emis, const Real* sfc_src, const Real* inc_flux, const Real* a_up, const Real* a_dn,
               Real* d_tau, Real* d_lay, Real* d_linc, Real* d_ldec, Real* d_sfc_src, Real* d_emis, void* stream)
{
    if (ncol <= 0 || nlay <= 0 || ngpt <= 0) return arg_error("lw_adjoint: empty extents");
    if (nmus < 1 || nmus > 4) return arg_error("lw_adjoint: n_gauss_angles must be 1..4");
    if (!secants || !weights || !tau || !lay || !linc || !ldec || !emis || !sfc_src || !a_up || !a_dn ||
        !d_tau || !d_lay || !d_linc || !d_ldec || !d_sfc_src || !d_emis)
        return arg_error("lw_adjoint: NULL array");
    if (ncol * (nlay + 1) * int64_t(sizeof(Real)) >= (int64_t(1) << 32))
        return arg_error("lw_adjoint: ncol * (nlay + 1) values of one g-point must stay below 4 GiB");
    hipStream_t s = as_stream(stream);
    constexpr int K = AdjSeg<Real>::K;
    const int nmu_t = nmus == 1 ? 1 : 4;
    const int64_t nseg = (nlay + K - 1) / K;
    Real* ck = nullptr;
    RRTMGP_TRY(scratch_alloc(&ck, size_t(nseg * 2 * nmu_t * ngpt * ncol), s));
    const dim3 grid(grid1(ncol, ADJ_COLS), grid1(ngpt, ADJ_GPTS));
    const dim3 block(ADJ_COLS, ADJ_GPTS);
    int rc;
    {
        RRTMGP_TIMED("lw_adjoint_kernel", s);
        auto launch = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, grid, block, 0, s, ncol, nlay, ngpt, top_at_1, int(nmus), secants, weights,
                               tau, lay, linc, ldec, emis, sfc_src, inc_flux, a_up, a_dn, ck, d_tau, d_lay, d_linc,
                               d_ldec, d_sfc_src, d_emis);
        };
        if (nmus == 1) launch(lw_adjoint_kernel<Real, 1>);
        else launch(lw_adjoint_kernel<Real, 4>);
        rc = check_hip(hipGetLastError(), "lw_adjoint_kernel");
    }
    scratch_free(ck, s);
    return rc;
}
// ---------------------------------------------------------------------------------------------
// Shortwave adjoint: one (column, g-point) per thread, 64 columns x 4 g-points per work-group.
// State store per (column, g-point): the layer properties R, T, Rdir, Tdir, Tn and the level
// states D, Ru, dir, A, S and the reverse sweep's adjoints lA, lS, all in the build's precision.
// The recursions carry them in FP64; the stored values are the FP64 ones rounded, and every
// sweep computes in FP64 from what it reads.  The four sweeps are four kernels: they communicate
// through the store only, and the three light ones then run at the occupancy their own registers
// allow instead of that of the dual-number sweep.
constexpr int SWA_GPTS = 4;
constexpr int SWA_NL = 5;    // stored layer properties
constexpr int SWA_NV = 7;    // stored level states

template<typename Real>
struct SwAdjArgs
{
    int64_t ncol, nlay, ngpt;
    int top_at_1;
    const Real *tau, *ssa, *gasy, *mu0, *alb_dir, *alb_dif, *inc_dir, *inc_dif, *a_up, *a_dn, *a_dir;
    Real *sl, *sv, *d_tau, *d_ssa, *d_g, *d_alb_dir, *d_alb_dif;
};

// Indexing of one thread's (column, g-point).  Layer properties: 0 R, 1 T, 2 Rdir, 3 Tdir, 4 Tn.
// Level states: 0 D, 1 Ru, 2 dir, 3 A, 4 S, 5 lA, 6 lS (adjoints of A and S at the level, from
// the top-down reverse sweep; slot n of 5 and 6 holds the sweep's final values).
template<typename Real>
struct SwAdjAt
{
    const SwAdjArgs<Real>& a;
    int64_t icol, igpt, n, nlev, c2, lb;
    __device__ SwAdjAt(const SwAdjArgs<Real>& args, int64_t ic, int64_t ig)
        : a(args), icol(ic), igpt(ig), n(args.nlay), nlev(args.nlay + 1), c2(ic + args.ncol * ig),
          lb(ic + args.ncol * args.nlay * ig) {}
    __device__ int64_t lay(int64_t j) const { return lb + a.ncol * (a.top_at_1 ? j : n - 1 - j); }
    __device__ int64_t lev(int64_t j) const { return icol + a.ncol * (a.top_at_1 ? j : n - j); }
    __device__ Real& L(int q, int64_t jj) const { return a.sl[((int64_t(q) * nlev + jj) * a.ngpt + igpt) * a.ncol + icol]; }
    __device__ Real& V(int q, int64_t jj) const { return a.sv[((int64_t(q) * nlev + jj) * a.ngpt + igpt) * a.ncol + icol]; }

    // Flux cotangents at level j: gradients with respect to D, Ru, A, S.
    __device__ void level_partials(int64_t jj, double D, double Ru, double A, double S, double& gD, double& gRu,
                                   double& gA, double& gS) const
    {
        const double au = double(a.a_up[lev(jj)]), adn = double(a.a_dn[lev(jj)]);
        const double q = D + Ru * S, den = 1. / (1. - Ru * A), dn = q * den;
        const double gdn = adn + au * A;
        gD = gdn * den;
        gRu = gdn * (S * den + q * A * den * den);
        gA = gdn * q * Ru * den * den + au * dn;
        gS = gdn * Ru * den + au;
    }
};

// Contributions of the bottom-up recursion's adjoint at layer j to the layer's R, T, SU, SD
// cotangents, from the adjoints lA, lS of A_j, S_j.
__device__ __forceinline__ void bottom_up_partials(double lA, double lS, double R, double T, double SD, double A1,
                                                   double S1, double& gR, double& gT, double& gSU, double& gSD)
{
    const double e2 = 1. / (1. - R * A1);
    const double w = S1 + A1 * SD;
    gR = lA * (1. + T * T * A1 * A1 * e2 * e2) + lS * T * w * A1 * e2 * e2;
    gT = lA * 2. * T * A1 * e2 + lS * w * e2;
    gSU = lS;
    gSD = lS * T * A1 * e2;
}

template<typename Real>
__device__ __forceinline__ bool sw_adj_thread(const SwAdjArgs<Real>& a, int64_t& icol, int64_t& igpt)
{
    icol = int64_t(blockIdx.x) * ADJ_COLS + threadIdx.x;
    igpt = int64_t(blockIdx.y) * SWA_GPTS + threadIdx.y;
    return icol < a.ncol && igpt < a.ngpt;
}

// Sweep 1, forward top-down: layer properties, direct beam, column above each level.
template<typename Real>
__global__ __launch_bounds__(ADJ_COLS * SWA_GPTS)
void sw_adjoint_down_kernel(SwAdjArgs<Real> a)
{
    int64_t icol, igpt;
    if (!sw_adj_thread(a, icol, igpt)) return;
    const SwAdjAt<Real> t(a, icol, igpt);
    const double m0 = double(a.mu0[icol]);
    double dir = double(a.inc_dir[t.c2]) * m0, D = a.inc_dif ? double(a.inc_dif[t.c2]) : 0., Ru = 0.;
    for (int64_t jj = 0; jj < t.n; ++jj)
    {
        t.V(0, jj) = Real(D); t.V(1, jj) = Real(Ru); t.V(2, jj) = Real(dir);
        const int64_t l = t.lay(jj);
        double R, T, Rd, Td, Tn;
        sw_two_stream_t<double>(double(a.tau[l]), double(a.ssa[l]), double(a.gasy[l]), m0, R, T, Rd, Td, Tn);
        // The stored (build-precision) values are the ones every later sweep uses.
        t.L(0, jj) = Real(R); t.L(1, jj) = Real(T); t.L(2, jj) = Real(Rd); t.L(3, jj) = Real(Td); t.L(4, jj) = Real(Tn);
        R = double(Real(R)); T = double(Real(T)); Rd = double(Real(Rd)); Td = double(Real(Td));
        Tn = double(Real(Tn));
        const double SU = Rd * dir, SD = Td * dir;
        const double e1 = 1. / (1. - R * Ru);
        D = SD + T * (D + Ru * SU) * e1;
        Ru = R + T * T * Ru * e1;
        dir *= Tn;
    }
    t.V(0, t.n) = Real(D); t.V(1, t.n) = Real(Ru); t.V(2, t.n) = Real(dir);
}

// Sweep 2, forward bottom-up: column below each level.
template<typename Real>
__global__ __launch_bounds__(ADJ_COLS * SWA_GPTS)
void sw_adjoint_up_kernel(SwAdjArgs<Real> a)
{
    int64_t icol, igpt;
    if (!sw_adj_thread(a, icol, igpt)) return;
    const SwAdjAt<Real> t(a, icol, igpt);
    double A = double(a.alb_dif[t.c2]), S = double(t.V(2, t.n)) * double(a.alb_dir[t.c2]);
    t.V(3, t.n) = Real(A); t.V(4, t.n) = Real(S);
    for (int64_t jj = t.n - 1; jj >= 0; --jj)
    {
        const double R = double(t.L(0, jj)), T = double(t.L(1, jj)), dir = double(t.V(2, jj));
        const double SU = double(t.L(2, jj)) * dir, SD = double(t.L(3, jj)) * dir;
        const double e2 = 1. / (1. - R * A);
        S = SU + T * (S + A * SD) * e2;
        A = R + T * T * A * e2;
        t.V(3, jj) = Real(A); t.V(4, jj) = Real(S);
    }
}

// Sweep 3: reverse of the bottom-up recursion, top-down; albedo gradients.
template<typename Real>
__global__ __launch_bounds__(ADJ_COLS * SWA_GPTS)
void sw_adjoint_rev_down_kernel(SwAdjArgs<Real> a)
{
    int64_t icol, igpt;
    if (!sw_adj_thread(a, icol, igpt)) return;
    const SwAdjAt<Real> t(a, icol, igpt);
    const int64_t n = t.n;
    double lA = 0., lS = 0.;
    double A = double(t.V(3, 0)), S = double(t.V(4, 0));
    for (int64_t jj = 0; jj <= n; ++jj)
    {
        double gD, gRu, gA, gS;
        t.level_partials(jj, double(t.V(0, jj)), double(t.V(1, jj)), A, S, gD, gRu, gA, gS);
        lA += gA; lS += gS;
        if (jj == n) break;
        t.V(5, jj) = Real(lA); t.V(6, jj) = Real(lS);
        const double R = double(t.L(0, jj)), T = double(t.L(1, jj)), dir = double(t.V(2, jj));
        const double SD = double(t.L(3, jj)) * dir;
        const double A1 = double(t.V(3, jj + 1)), S1 = double(t.V(4, jj + 1));
        const double e2 = 1. / (1. - R * A1);
        const double w = S1 + A1 * SD;
        const double nA = lA * T * T * e2 * e2 + lS * T * e2 * (SD + w * R * e2);
        lS = lS * T * e2;
        lA = nA;
        A = A1; S = S1;
    }
    t.V(5, n) = Real(lA); t.V(6, n) = Real(lS);
    a.d_alb_dif[t.c2] = Real(lA);
    a.d_alb_dir[t.c2] = Real(lS * double(t.V(2, n)));
}

// Sweep 4: reverse of the top-down recursion and the direct beam, bottom-up; layer gradients.
template<typename Real>
__global__ __launch_bounds__(ADJ_COLS * SWA_GPTS)
void sw_adjoint_rev_up_kernel(SwAdjArgs<Real> a)
{
    int64_t icol, igpt;
    if (!sw_adj_thread(a, icol, igpt)) return;
    const SwAdjAt<Real> t(a, icol, igpt);
    const int64_t n = t.n;
    const double m0 = double(a.mu0[icol]);
    const double ad = double(a.alb_dir[t.c2]);
    double lD, lRu, ldir;
    double A1 = double(t.V(3, n)), S1 = double(t.V(4, n));
    {
        double gD, gRu, gA, gS;
        t.level_partials(n, double(t.V(0, n)), double(t.V(1, n)), A1, S1, gD, gRu, gA, gS);
        lD = gD; lRu = gRu;
        ldir = double(a.a_dn[t.lev(n)]) + double(a.a_dir[t.lev(n)]) + double(t.V(6, n)) * ad;
    }
    for (int64_t jj = n - 1; jj >= 0; --jj)
    {
        // Layer properties and their Jacobian by forward mode on (tau, w0, g); the values are
        // rounded to the stored precision, as the earlier sweeps used them.
        const int64_t l = t.lay(jj);
        Dual3 oR, oT, oRd, oTd, oTn;
        sw_two_stream_t<Dual3>(mk(double(a.tau[l]), 1., 0., 0.), mk(double(a.ssa[l]), 0., 1., 0.),
                               mk(double(a.gasy[l]), 0., 0., 1.), m0, oR, oT, oRd, oTd, oTn);
        const double R = double(Real(oR.v)), T = double(Real(oT.v)), Rd = double(Real(oRd.v));
        const double Td = double(Real(oTd.v)), Tn = double(Real(oTn.v));
        const double D = double(t.V(0, jj)), Ru = double(t.V(1, jj)), dir = double(t.V(2, jj));
        const double SU = Rd * dir, SD = Td * dir;
        // A_j, S_j rebuilt from A_{j+1}, S_{j+1} exactly as the bottom-up forward sweep formed them.
        const double e2f = 1. / (1. - R * A1);
        const double S = SU + T * (S1 + A1 * SD) * e2f;
        const double A = R + T * T * A1 * e2f;
        double bR, bT, bSU, bSD;
        bottom_up_partials(double(t.V(5, jj)), double(t.V(6, jj)), R, T, SD, A1, S1, bR, bT, bSU, bSD);
        const double e1 = 1. / (1. - R * Ru);
        const double v = D + Ru * SU;
        const double gSD = bSD + lD;
        const double gSU = bSU + lD * T * Ru * e1;
        const double gT = bT + lD * v * e1 + lRu * 2. * T * Ru * e1;
        const double gR = bR + lD * T * v * Ru * e1 * e1 + lRu * (1. + T * T * Ru * Ru * e1 * e1);
        const double gRd = gSU * dir, gTd = gSD * dir, gTn = ldir * dir;
        const double nD = lD * T * e1;
        const double nRu = lD * (T * SU * e1 + T * v * R * e1 * e1) + lRu * T * T * e1 * e1;
        const double ndir = ldir * Tn + gSU * Rd + gSD * Td;
        double g3[3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
            g3[c] = gR * oR.d[c] + gT * oT.d[c] + gRd * oRd.d[c] + gTd * oTd.d[c] + gTn * oTn.d[c];
        a.d_tau[l] = Real(g3[0]); a.d_ssa[l] = Real(g3[1]); a.d_g[l] = Real(g3[2]);
        double gD, gRu, gA, gS;
        t.level_partials(jj, D, Ru, A, S, gD, gRu, gA, gS);
        lD = nD + gD;
        lRu = nRu + gRu;
        ldir = ndir + double(a.a_dn[t.lev(jj)]) + double(a.a_dir[t.lev(jj)]);
        A1 = A; S1 = S;
    }
}

template<typename Real>
int sw_adjoint(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1, const Real* tau, const Real* ssa,
               const Real* g, const Real* mu0, const Real* alb_dir, const Real* alb_dif, const Real* inc_dir,
               const Real* inc_dif, const Real* a_up, const Real* a_dn, const Real* a_dir, Real* d_tau,
               Real* d_ssa, Real* d_g, Real* d_alb_dir, Real* d_alb_dif, void* stream)
{
    if (ncol <= 0 || nlay <= 0 || ngpt <= 0) return arg_error("sw_adjoint: empty extents");
    if (!tau || !ssa || !g || !mu0 || !alb_dir || !alb_dif || !inc_dir || !a_up || !a_dn || !a_dir || !d_tau ||
        !d_ssa || !d_g || !d_alb_dir || !d_alb_dif)
        return arg_error("sw_adjoint: NULL array");
    hipStream_t s = as_stream(stream);
    const size_t per = size_t((nlay + 1) * ngpt * ncol);
    Real* sl = nullptr;
    Real* sv = nullptr;
    RRTMGP_TRY(scratch_alloc(&sl, SWA_NL * per, s));
    if (int rc0 = scratch_alloc(&sv, SWA_NV * per, s); rc0 != RRTMGP_OK)
    {
        scratch_free(sl, s);
        return rc0;
    }
    const SwAdjArgs<Real> args{ncol, nlay, ngpt, top_at_1, tau, ssa, g, mu0, alb_dir, alb_dif, inc_dir, inc_dif,
                               a_up, a_dn, a_dir, sl, sv, d_tau, d_ssa, d_g, d_alb_dir, d_alb_dif};
    const dim3 grid(grid1(ncol, ADJ_COLS), grid1(ngpt, SWA_GPTS));
    const dim3 block(ADJ_COLS, SWA_GPTS);
    int rc = RRTMGP_OK;
    {
        RRTMGP_TIMED("sw_adjoint_down_kernel", s);
        hipLaunchKernelGGL(sw_adjoint_down_kernel<Real>, grid, block, 0, s, args);
        rc = check_hip(hipGetLastError(), "sw_adjoint_down_kernel");
    }
    if (rc == RRTMGP_OK)
    {
        RRTMGP_TIMED("sw_adjoint_up_kernel", s);
        hipLaunchKernelGGL(sw_adjoint_up_kernel<Real>, grid, block, 0, s, args);
        rc = check_hip(hipGetLastError(), "sw_adjoint_up_kernel");
    }
    if (rc == RRTMGP_OK)
    {
        RRTMGP_TIMED("sw_adjoint_rev_down_kernel", s);
        hipLaunchKernelGGL(sw_adjoint_rev_down_kernel<Real>, grid, block, 0, s, args);
        rc = check_hip(hipGetLastError(), "sw_adjoint_rev_down_kernel");
    }
    if (rc == RRTMGP_OK)
    {
        RRTMGP_TIMED("sw_adjoint_rev_up_kernel", s);
        hipLaunchKernelGGL(sw_adjoint_rev_up_kernel<Real>, grid, block, 0, s, args);
        rc = check_hip(hipGetLastError(), "sw_adjoint_rev_up_kernel");
    }
    scratch_free(sv, s);
    scratch_free(sl, s);
    return rc;
}


// ---------------------------------------------------------------------------------------------
// Assembly of the gas-optics gradients.

// col_gas(gas) = vmr(gas) col_dry for gas >= 1 and col_gas(0) = col_dry, so
//   d vmr(gas) = d col_gas(gas) col_dry,   d col_dry = d col_gas(0) + sum_gas d col_gas(gas) vmr(gas).
// d_col (ncol, nlay, ngas+1) -> d_vmr (ncol, nlay, ngas) (overwritten), d_col_dry (ncol, nlay).
template<typename Real>
__global__ void col_chain_kernel(int64_t ncol, int64_t nlay, int ngas, const rrtmgp_gas_vmr gases,
                                 const Real* __restrict__ col_dry, const Real* __restrict__ d_col,
                                 Real* __restrict__ d_vmr, Real* __restrict__ d_col_dry)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t ncl = ncol * nlay;
    if (i >= ncl) return;
    const Real cd = col_dry[i];
    for (int gas = 0; gas < ngas; ++gas)
        d_vmr[i + ncl * gas] = d_col[i + ncl * (gas + 1)] * cd;
    // The derivative at fixed mixing ratios, formed by the gas-optics adjoint (slot ngas + 1).
    d_col_dry[i] = d_col[i + ncl * (ngas + 1)];
}

// d_tlev (ncol, nlev): level i receives d_lo of layer i and d_hi of layer i-1.
template<typename Real>
__global__ void levels_kernel(int64_t ncol, int64_t nlay, const Real* __restrict__ lo, const Real* __restrict__ hi,
                              Real* __restrict__ d_tlev)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncol * (nlay + 1)) return;
    const int64_t ilev = i / ncol;
    Real v = Real(0);
    if (ilev < nlay) v += lo[i];
    if (ilev > 0) v += hi[i - ncol];
    d_tlev[i] = v;
}

// out (nbnd, ncol) = sums over each band's g-points of in (ncol, ngpt); band_lims (2, nbnd) 1-based.
template<typename Real>
__global__ void gpt_to_band_kernel(int64_t ncol, int64_t nbnd, const int* __restrict__ band_lims,
                                   const Real* __restrict__ in, Real* __restrict__ out)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncol * nbnd) return;
    const int64_t b = i % nbnd, icol = i / nbnd;
    Real s = Real(0);
    for (int g = band_lims[2 * b] - 1; g < band_lims[2 * b + 1]; ++g) s += in[icol + ncol * g];
    out[b + nbnd * icol] = s;
}
}  // namespace rrtmgp_hip

using namespace rrtmgp_hip;

#define RRTMGP_DEF_ADJOINT(S, T)                                                                   \
extern "C" int rrtmgp_hip_lw_solver_noscat_adjoint_##S(int64_t ncol, int64_t nlay, int64_t ngpt,   \
    int top_at_1, int64_t nmus, const T* secants, const T* weights, const T* tau,                   \
    const T* lay_source, const T* lev_source_inc, const T* lev_source_dec, const T* sfc_emis,       \
    const T* sfc_src, const T* inc_flux, const T* bb_up_cot, const T* bb_dn_cot, T* d_tau,          \
    T* d_lay_source, T* d_lev_source_inc, T* d_lev_source_dec, T* d_sfc_src, T* d_sfc_emis,         \
    void* stream)                                                                                   \
{                                                                                                   \
    return lw_adjoint<T>(ncol, nlay, ngpt, top_at_1, nmus, secants, weights, tau, lay_source,       \
                         lev_source_inc, lev_source_dec, sfc_emis, sfc_src, inc_flux, bb_up_cot,    \
                         bb_dn_cot, d_tau, d_lay_source, d_lev_source_inc, d_lev_source_dec,        \
                         d_sfc_src, d_sfc_emis, stream);                                            \
}                                                                                                   \
extern "C" int rrtmgp_hip_sw_solver_2stream_adjoint_##S(int64_t ncol, int64_t nlay, int64_t ngpt,  \
    int top_at_1, const T* tau, const T* ssa, const T* g, const T* mu0, const T* sfc_alb_dir,       \
    const T* sfc_alb_dif, const T* inc_flux_dir, const T* inc_flux_dif, const T* bb_up_cot,         \
    const T* bb_dn_cot, const T* bb_dir_cot, T* d_tau, T* d_ssa, T* d_g, T* d_sfc_alb_dir,           \
    T* d_sfc_alb_dif, void* stream)                                                                 \
{                                                                                                   \
    return sw_adjoint<T>(ncol, nlay, ngpt, top_at_1, tau, ssa, g, mu0, sfc_alb_dir, sfc_alb_dif,    \
                         inc_flux_dir, inc_flux_dif, bb_up_cot, bb_dn_cot, bb_dir_cot, d_tau, d_ssa,\
                         d_g, d_sfc_alb_dir, d_sfc_alb_dif, stream);                                \
}

RRTMGP_DEF_ADJOINT(f32, float)
RRTMGP_DEF_ADJOINT(f64, double)

#define RRTMGP_DEF_ADJ_HELPERS(S, T)                                                               \
extern "C" int rrtmgp_hip_adjoint_col_chain_##S(int64_t ncol, int64_t nlay,                        \
    const rrtmgp_gas_vmr* gases, const T* col_dry, const T* d_col_gas, T* d_vmr, T* d_col_dry,       \
    void* stream)                                                                                   \
{                                                                                                   \
    if (!gases || gases->ngas < 1 || gases->ngas > RRTMGP_MAX_GAS || !col_dry || !d_col_gas ||     \
        !d_vmr || !d_col_dry || ncol <= 0 || nlay <= 0)                                             \
        return arg_error("adjoint_col_chain: bad arguments");                                      \
    const int64_t n = ncol * nlay;                                                                  \
    RRTMGP_TIMED("col_chain_kernel", as_stream(stream));                                            \
    hipLaunchKernelGGL(col_chain_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),  \
                       ncol, nlay, int(gases->ngas), *gases, col_dry, d_col_gas, d_vmr, d_col_dry); \
    RRTMGP_LAUNCHED("col_chain_kernel");                                                           \
    return RRTMGP_OK;                                                                               \
}                                                                                                   \
extern "C" int rrtmgp_hip_adjoint_levels_##S(int64_t ncol, int64_t nlay, const T* d_lo,            \
    const T* d_hi, T* d_tlev, void* stream)                                                         \
{                                                                                                   \
    if (!d_lo || !d_hi || !d_tlev || ncol <= 0 || nlay <= 0) return arg_error("adjoint_levels: bad arguments"); \
    const int64_t n = ncol * (nlay + 1);                                                            \
    RRTMGP_TIMED("levels_kernel", as_stream(stream));                                               \
    hipLaunchKernelGGL(levels_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),     \
                       ncol, nlay, d_lo, d_hi, d_tlev);                                             \
    RRTMGP_LAUNCHED("levels_kernel");                                                              \
    return RRTMGP_OK;                                                                               \
}                                                                                                   \
extern "C" int rrtmgp_hip_gpt_to_band_##S(int64_t ncol, int64_t nbnd, const int* band_lims_gpt,    \
    const T* in, T* out, void* stream)                                                              \
{                                                                                                   \
    if (!band_lims_gpt || !in || !out || ncol <= 0 || nbnd <= 0) return arg_error("gpt_to_band: bad arguments"); \
    const int64_t n = ncol * nbnd;                                                                  \
    RRTMGP_TIMED("gpt_to_band_kernel", as_stream(stream));                                          \
    hipLaunchKernelGGL(gpt_to_band_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),\
                       ncol, nbnd, band_lims_gpt, in, out);                                         \
    RRTMGP_LAUNCHED("gpt_to_band_kernel");                                                         \
    return RRTMGP_OK;                                                                               \
}

RRTMGP_DEF_ADJ_HELPERS(f32, float)
RRTMGP_DEF_ADJ_HELPERS(f64, double)
