// Shortwave two-stream layer properties (sw_dif_and_source, F90:1043-1104, PIFM coefficients;
// the Rdir/Tdir clamp of src_kernels_cuda/rte_solver_kernels.cu:581-583), written once for any
// scalar type with the arithmetic below: double for the FP64 solver, Dual3 (value and the
// derivatives with respect to tau, w0 and g) for the adjoint's layer Jacobian.
#pragma once

#include "common.h"

namespace rrtmgp_hip
{
// Forward-mode dual number with three tangents.
struct Dual3
{
    double v, d[3];
};

__device__ __forceinline__ Dual3 mk(double v, double d0, double d1, double d2) { return Dual3{v, {d0, d1, d2}}; }
__device__ __forceinline__ Dual3 operator+(Dual3 a, Dual3 b) { return mk(a.v + b.v, a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]); }
__device__ __forceinline__ Dual3 operator-(Dual3 a, Dual3 b) { return mk(a.v - b.v, a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]); }
__device__ __forceinline__ Dual3 operator-(Dual3 a) { return mk(-a.v, -a.d[0], -a.d[1], -a.d[2]); }
__device__ __forceinline__ Dual3 operator*(Dual3 a, Dual3 b)
{
    return mk(a.v * b.v, a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]);
}
__device__ __forceinline__ Dual3 operator/(Dual3 a, Dual3 b)
{
    const double r = 1. / b.v, q = a.v * r;
    return mk(q, (a.d[0] - q * b.d[0]) * r, (a.d[1] - q * b.d[1]) * r, (a.d[2] - q * b.d[2]) * r);
}
__device__ __forceinline__ Dual3 operator+(Dual3 a, double b) { return mk(a.v + b, a.d[0], a.d[1], a.d[2]); }
__device__ __forceinline__ Dual3 operator+(double a, Dual3 b) { return b + a; }
__device__ __forceinline__ Dual3 operator-(Dual3 a, double b) { return mk(a.v - b, a.d[0], a.d[1], a.d[2]); }
__device__ __forceinline__ Dual3 operator-(double a, Dual3 b) { return mk(a - b.v, -b.d[0], -b.d[1], -b.d[2]); }
__device__ __forceinline__ Dual3 operator*(Dual3 a, double b) { return mk(a.v * b, a.d[0] * b, a.d[1] * b, a.d[2] * b); }
__device__ __forceinline__ Dual3 operator*(double a, Dual3 b) { return b * a; }
__device__ __forceinline__ Dual3 operator/(Dual3 a, double b) { return a * (1. / b); }
__device__ __forceinline__ Dual3 operator/(double a, Dual3 b) { return mk(a, 0., 0., 0.) / b; }

__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double val(Dual3 x) { return x.v; }
__device__ __forceinline__ double ts_exp(double x) { return exp(x); }
__device__ __forceinline__ Dual3 ts_exp(Dual3 x)
{
    const double e = exp(x.v);
    return mk(e, e * x.d[0], e * x.d[1], e * x.d[2]);
}
__device__ __forceinline__ double ts_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ Dual3 ts_sqrt(Dual3 x)
{
    const double s = sqrt(x.v), h = .5 / s;
    return mk(s, h * x.d[0], h * x.d[1], h * x.d[2]);
}
// max/min pick one operand; the tangent follows the picked operand (as fmax/fmin select it).
__device__ __forceinline__ double ts_max(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ Dual3 ts_max(Dual3 a, double b) { return a.v >= b ? a : mk(b, 0., 0., 0.); }
__device__ __forceinline__ Dual3 ts_max(double a, Dual3 b) { return ts_max(b, a); }
__device__ __forceinline__ double ts_min(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ Dual3 ts_min(Dual3 a, Dual3 b) { return a.v <= b.v ? a : b; }

// The layer properties.  Returns Rdif, Tdif, Rdir, Tdir (per unit incident direct flux) and the
// direct-beam transmittance Tnoscat.
template<typename S>
__device__ __forceinline__ void sw_two_stream_t(S tau, S w0, S g, double mu0, S& Rdif, S& Tdif, S& Rdir,
                                                S& Tdir, S& Tnoscat)
{
    const double eps = Limits<double>::eps;
    const S gamma1 = (8. - w0 * (5. + 3. * g)) * .25;
    const S gamma2 = 3. * (w0 * (1. - g)) * .25;
    const S gamma3 = (2. - 3. * mu0 * g) * .25;
    const S gamma4 = 1. - gamma3;
    const S alpha1 = gamma1 * gamma4 + gamma2 * gamma3;
    const S alpha2 = gamma1 * gamma3 + gamma2 * gamma4;
    const S k = ts_sqrt(ts_max((gamma1 - gamma2) * (gamma1 + gamma2), Limits<double>::k_min));
    const S k_mu = k * mu0, k_gamma3 = k * gamma3, k_gamma4 = k * gamma4;
    const S e1 = ts_exp(-tau * k);
    const S e2 = e1 * e1;
    const S RT = 1. / (k * (1. + e2) + gamma1 * (1. - e2));
    Rdif = RT * gamma2 * (1. - e2);
    Tdif = RT * 2. * k * e1;
    const S d = 1. - k_mu * k_mu;
    const S denom = (fabs(val(d)) >= eps) ? d : d * 0. + eps;
    const S RT2 = w0 * RT / denom;
    Tnoscat = ts_exp(-tau / mu0);
    S rd = RT2 * ((1. - k_mu) * (alpha2 + k_gamma3) - (1. + k_mu) * (alpha2 - k_gamma3) * e2 -
                  2. * (k_gamma3 - alpha2 * k_mu) * e1 * Tnoscat);
    S td = -RT2 * ((1. + k_mu) * (alpha1 + k_gamma4) * Tnoscat -
                   (1. - k_mu) * (alpha1 - k_gamma4) * e2 * Tnoscat -
                   2. * (k_gamma4 + alpha1 * k_mu) * e1);
    rd = ts_max(ts_min(rd, 1. - Tnoscat), eps);
    td = ts_max(ts_min(td, 1. - Tnoscat - rd), eps);
    Rdir = rd;
    Tdir = td;
}
}  // namespace rrtmgp_hip
