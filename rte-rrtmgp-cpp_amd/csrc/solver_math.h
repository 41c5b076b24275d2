// Arithmetic building blocks of the RTE solvers: FP32 forms on the hardware approximations
// (v_exp_f32, v_rcp_f32, each within about 1 ulp) with series where FP32 would lose accuracy to
// cancellation, the FP64 reciprocal of the adding denominators, and 32-bit-offset addressing.
#pragma once

#include "common.h"

namespace rrtmgp_hip
{
__device__ __forceinline__ float rcp_f32(float x) { return __builtin_amdgcn_rcpf(x); }

// e^{-y} for y >= 0 through 2^x.
__device__ __forceinline__ float exp_neg_f32(float y) { return __builtin_amdgcn_exp2f(-y * 1.44269504088896341f); }

// 1 - e^{-y} for 0 <= y < 0.25 from its series (truncation < y^8/8! relative).
__device__ __forceinline__ float one_minus_exp_neg_series_f32(float y)
{
    float t = 1.f / 720.f - y * (1.f / 5040.f);
    t = 1.f / 120.f - y * t;
    t = 1.f / 24.f - y * t;
    t = 1.f / 6.f - y * t;
    t = .5f - y * t;
    t = 1.f - y * t;
    return y * t;
}

// 1 - e^{-y} for y >= 0 given e = e^{-y}: the series below 0.25, else 1 - e.
__device__ __forceinline__ float one_minus_exp_neg_f32(float y, float e)
{
    return (y < .25f) ? one_minus_exp_neg_series_f32(y) : 1.f - e;
}

// Array element i at a uniform base with a 32-bit lane offset: addressed as scalar base +
// 32-bit VGPR offset (the launchers keep the offsets below 2^32 bytes).
template<typename Real>
__device__ __forceinline__ const Real& at32(const Real* base, uint32_t i)
{
    return *reinterpret_cast<const Real*>(reinterpret_cast<const char*>(base) + i * uint32_t(sizeof(Real)));
}

template<typename Real>
__device__ __forceinline__ Real& at32(Real* base, uint32_t i)
{
    return *reinterpret_cast<Real*>(reinterpret_cast<char*>(base) + i * uint32_t(sizeof(Real)));
}

// The same element read as a stream (non-temporal): inputs read once per sweep that no later
// access finds in the cache.
template<typename Real>
__device__ __forceinline__ Real ld_nt(const Real* base, uint32_t i)
{
    return __builtin_nontemporal_load(&at32(base, i));
}

// Layer transmittance T = e^{-t} and the linear-in-tau source factor of lw_source_noscat
// (F90:616-653): fact = (1 - T)/t - T, with t (1/2 - t/3) below sqrt(epsilon).  The FP32 build
// evaluates fact for t < 1/4 from its series t/2 - t^2/3 + t^3/8 - t^4/30 + t^5/144 - t^6/840
// + t^7/5760 (the reference's small-t form is its first two terms; truncation < 3e-9 relative),
// which avoids the cancellation of (1 - T)/t - T, and uses the hardware exp2 and reciprocal.
__device__ __forceinline__ void lw_layer(double t, double& T, double& fact)
{
    T = exp(-t);
    fact = (t > sqrt(Limits<double>::eps)) ? (1. - T) / t - T : t * (.5 - (1. / 3.) * t);
}

__device__ __forceinline__ void lw_layer(float t, float& T, float& fact)
{
    T = exp_neg_f32(t);
    float q = 1.f / 840.f - t * (1.f / 5760.f);
    q = 1.f / 144.f - t * q;
    q = 1.f / 30.f - t * q;
    q = 1.f / 8.f - t * q;
    q = 1.f / 3.f - t * q;
    q = .5f - t * q;
    fact = (t < .25f) ? t * q : (1.f - T) * rcp_f32(t) - T;
}

// 1 - e^{-t} given T = e^{-t}, without the cancellation of 1 - T for small t.
__device__ __forceinline__ double one_minus_T(double t, double) { return -expm1(-t); }
__device__ __forceinline__ float one_minus_T(float t, float T) { return one_minus_exp_neg_f32(t, T); }

// 1/x in FP64 for the adding denominators, x in (0, 1]: v_rcp_f64 (relative error up to 4.6e-8
// on gfx950) and one Newton step, which leaves at most 2.3e-15, far below the FP32 rounding of
// the fluxes and of the carried states' inputs.
__device__ __forceinline__ double rcp_f64(double x)
{
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.), r);
}
}  // namespace rrtmgp_hip
