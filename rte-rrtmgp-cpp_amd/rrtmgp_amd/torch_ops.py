"""torch.autograd bindings of the column-radiation path.

No reference counterpart: the reference's Python binding (python/radiation.pyx) is forward only.
Forward passes run the library's HIP kernels (fused gas optics, cloud optics, solvers); backward
passes run its adjoint drivers (rrtmgp_api_adjoint_lw / _sw) and the heating-rate adjoint kernel.
torch only holds the tensors: they are copied device to device into the library's Fortran-order
arrays and back.

Shapes (tensors on the HIP device, any strides): profiles (ncol, nlay) or (ncol, nlev); t_sfc,
mu0 and tsi_scaling (ncol); surface emissivity and albedo (nbnd, ncol).  Gases other than h2o and
o3 come as a dict of constant vmr (scalars or (ncol, nlay) arrays) and receive no gradient;
clouds (lwp, iwp, rel, rei) are held fixed.
"""
from __future__ import annotations

import ctypes as C
from types import SimpleNamespace

import numpy as np
import torch

from . import _lib, api
from .device import DeviceArray, synchronize


def _np_dtype(precision):
    return np.float64 if precision == 8 else np.float32


def _torch_dtype(precision):
    return torch.float64 if precision == 8 else torch.float32


def to_device_array(t: torch.Tensor, precision: int) -> DeviceArray:
    """Copy a device tensor into a new DeviceArray of the same shape (Fortran order)."""
    x = t.detach().to(_torch_dtype(precision))
    x = x.permute(*reversed(range(x.dim()))).contiguous()      # C order of the reversed dims
    d = DeviceArray(tuple(t.shape), _np_dtype(precision), zero=False)
    torch.cuda.synchronize(x.device)
    if d.nbytes:
        _lib.check_hip(_lib.hip().rrtmgp_hip_memcpy_d2d(d.ptr, C.c_void_p(x.data_ptr()), C.c_size_t(d.nbytes), None))
    synchronize()
    return d


def to_tensor(d: DeviceArray, device) -> torch.Tensor:
    """A tensor of d's shape on `device` holding a copy of d."""
    dt = torch.float64 if d.dtype == np.float64 else torch.float32
    x = torch.empty(tuple(reversed(d.shape)), dtype=dt, device=device)
    synchronize()
    if d.nbytes:
        _lib.check_hip(_lib.hip().rrtmgp_hip_memcpy_d2d(C.c_void_p(x.data_ptr()), d.ptr, C.c_size_t(d.nbytes), None))
    synchronize()
    return x.permute(*reversed(range(x.dim())))


def _atmosphere(precision, ncol, nlay, arrays: dict, vmr: dict, clouds):
    """A DeviceAtmosphere-like namespace over device copies of the inputs."""
    a = SimpleNamespace(ncol=ncol, nlay=nlay, precision=precision)
    for k, t in arrays.items():
        setattr(a, k, to_device_array(t, precision))
    dev_vmr = {k: (to_device_array(v, precision) if isinstance(v, torch.Tensor)
                   else np.atleast_2d(np.asarray(v, _np_dtype(precision)))) for k, v in vmr.items()}
    a.gas_concs = api.GasConcs(dev_vmr, precision)
    a._vmr = dev_vmr                                          # keeps the device copies alive
    a.cloudy = clouds is not None
    for k in ("lwp", "iwp", "rel", "rei"):
        setattr(a, k, to_device_array(clouds[k], precision) if clouds is not None else None)
    return a


def _refuse_surrogate(solver):
    """The backward runs the adjoint of the plain gas optics; a forward through an MLP surrogate
    would pair fluxes of one function with gradients of another."""
    if getattr(solver.kd, "_surrogate", None) is not None:
        raise ValueError("autograd through a k-distribution with an MLP surrogate set is not supported: the "
                         "adjoint differentiates the plain gas optics; clear the surrogate first")


class _Longwave(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, other_vmr, clouds, t_lay, t_lev, t_sfc, p_lay, p_lev, h2o, o3, emis_sfc):
        _refuse_surrogate(solver)
        ncol, nlay = t_lay.shape
        vmr = dict(other_vmr)
        vmr["h2o"], vmr["o3"] = h2o, o3
        da = _atmosphere(solver.precision, ncol, nlay,
                         {"t_lay": t_lay, "t_lev": t_lev, "t_sfc": t_sfc, "p_lay": p_lay, "p_lev": p_lev,
                          "emis_sfc": emis_sfc}, vmr, clouds)
        o = solver.solve(da)
        ctx.solver, ctx.da, ctx.dev = solver, da, t_lay.device
        ctx.save_for_backward(p_lay)
        return to_tensor(o["flux_up"], t_lay.device), to_tensor(o["flux_dn"], t_lay.device)

    @staticmethod
    def backward(ctx, g_up, g_dn):
        (p_lay,) = ctx.saved_tensors
        P = ctx.solver.precision
        g = ctx.solver.adjoint(ctx.da, to_device_array(g_up, P), to_device_array(g_dn, P))
        t = {k: to_tensor(v, ctx.dev) for k, v in g.items()}
        names = ctx.solver.kd.gas_names
        return (None, None, None, t["t_lay"], t["t_lev"], t["t_sfc"], t["lnp_lay"] / p_lay.to(t["lnp_lay"].dtype),
                t["p_lev"], t["vmr"][:, :, names.index("h2o")], t["vmr"][:, :, names.index("o3")], t["emis_sfc"])


class _Shortwave(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, other_vmr, clouds, delta_cloud, mu0, tsi_scaling, t_lay, p_lay, p_lev, h2o, o3,
                sfc_alb):
        _refuse_surrogate(solver)
        ncol, nlay = t_lay.shape
        vmr = dict(other_vmr)
        vmr["h2o"], vmr["o3"] = h2o, o3
        P = solver.precision
        # t_lev is not read by the SW path; a copy of p_lev's shape keeps the interface uniform.
        da = _atmosphere(P, ncol, nlay, {"t_lay": t_lay, "t_lev": p_lev, "p_lay": p_lay, "p_lev": p_lev,
                                         "sfc_alb": sfc_alb, "mu0": mu0}, vmr, clouds)
        tsi = to_device_array(tsi_scaling, P)
        o = solver.solve(da, tsi, delta_cloud=delta_cloud)
        ctx.solver, ctx.da, ctx.tsi, ctx.delta, ctx.dev = solver, da, tsi, delta_cloud, t_lay.device
        ctx.save_for_backward(p_lay)
        return tuple(to_tensor(o[k], t_lay.device) for k in ("flux_up", "flux_dn", "flux_dn_dir"))

    @staticmethod
    def backward(ctx, g_up, g_dn, g_dir):
        (p_lay,) = ctx.saved_tensors
        P = ctx.solver.precision
        g = ctx.solver.adjoint(ctx.da, ctx.tsi, to_device_array(g_up, P), to_device_array(g_dn, P),
                               to_device_array(g_dir, P), delta_cloud=ctx.delta)
        t = {k: to_tensor(v, ctx.dev) for k, v in g.items()}
        names = ctx.solver.kd.gas_names
        return (None, None, None, None, None, None, t["t_lay"], t["lnp_lay"] / p_lay.to(t["lnp_lay"].dtype),
                t["p_lev"], t["vmr"][:, :, names.index("h2o")], t["vmr"][:, :, names.index("o3")],
                t["sfc_alb_dir"] + t["sfc_alb_dif"])


class _HeatingRate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, flux_up, flux_dn, p_lev, precision):
        ncol, nlev = flux_up.shape
        d = [to_device_array(x, precision) for x in (flux_up, flux_dn, p_lev)]
        hr = DeviceArray((ncol, nlev - 1), _np_dtype(precision))
        fn = getattr(_lib.hip(), "rrtmgp_hip_heating_rate_" + ("f64" if precision == 8 else "f32"))
        _lib.check_hip(fn(C.c_int64(ncol), C.c_int64(nlev - 1), d[0].ptr, d[1].ptr, d[2].ptr, hr.ptr, None))
        ctx.d, ctx.precision, ctx.dev = d, precision, flux_up.device
        return to_tensor(hr, flux_up.device)

    @staticmethod
    def backward(ctx, g_hr):
        P = ctx.precision
        up, dn, pl = ctx.d
        ncol, nlev = up.shape
        gu, gd, gp = (DeviceArray((ncol, nlev), _np_dtype(P)) for _ in range(3))
        gh = to_device_array(g_hr, P)
        fn = getattr(_lib.hip(), "rrtmgp_hip_heating_rate_adjoint_" + ("f64" if P == 8 else "f32"))
        _lib.check_hip(fn(C.c_int64(ncol), C.c_int64(nlev - 1), up.ptr, dn.ptr, pl.ptr, gh.ptr, gu.ptr, gd.ptr,
                          gp.ptr, None))
        return to_tensor(gu, ctx.dev), to_tensor(gd, ctx.dev), to_tensor(gp, ctx.dev), None


def lw_fluxes(solver: api.SolverLW, t_lay, t_lev, t_sfc, p_lay, p_lev, h2o, o3, emis_sfc, other_vmr=None,
              clouds=None):
    """Broadband LW (flux_up, flux_dn) (ncol, nlev), differentiable in every tensor argument."""
    return _Longwave.apply(solver, other_vmr or {}, clouds, t_lay, t_lev, t_sfc, p_lay, p_lev, h2o, o3, emis_sfc)


def sw_fluxes(solver: api.SolverSW, t_lay, p_lay, p_lev, h2o, o3, sfc_alb, mu0, tsi_scaling, other_vmr=None,
              clouds=None, delta_cloud=True):
    """Broadband SW (flux_up, flux_dn, flux_dn_dir) (ncol, nlev); differentiable in t_lay, p_lay,
    p_lev, h2o, o3 and sfc_alb (used for the direct and the diffuse albedo)."""
    return _Shortwave.apply(solver, other_vmr or {}, clouds, delta_cloud, mu0, tsi_scaling, t_lay, p_lay, p_lev,
                            h2o, o3, sfc_alb)


def heating_rate(flux_up, flux_dn, p_lev, precision=8):
    """Heating rate (ncol, nlay) [K/day] of rrtmgp_hip_heating_rate_*, differentiable."""
    return _HeatingRate.apply(flux_up, flux_dn, p_lev, precision)
