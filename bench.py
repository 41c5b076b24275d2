#!/usr/bin/env python3
"""Throughput of the all-sky LW+SW column-radiation forward on one MI355X.

Workload: 65536 columns (256 x 256) x 128 layers, all-sky (liquid and ice clouds), FP32, plain
RRTMGP gas optics, seeded synthetic k-distributions (LW 16 bands / 256 g-points, SW 14 bands /
224 g-points), cloud LUTs and profiles.  One step = Radiation_solver_longwave_gpu::solve_gpu +
Radiation_solver_shortwave_gpu::solve_gpu over all columns (gas optics, cloud optics, solvers,
broadband fluxes), inputs already resident in HBM.

metric: col-gpt-layers/s = ncol * nlay * (256 + 224) / (time of one step).

Also reported, on the same JSON line:
  roofline      the kernel with the largest share of GPU time: its algorithmic bytes (the
                class-API arrays it must read or write, per launch) over its average launch
                duration, measured with HIP events on its launch stream during the timed steps,
                against the 8 TB/s HBM peak; traffic from rocprofv3 FETCH_SIZE/WRITE_SIZE when a
                summary of them for this workload is committed under profiles/.
  cpu_baseline  the FP64 CPU oracle (a restatement of the reference CPU path) on the first
                columns of the same workload, timed on this host: serial (one core, as the
                reference runs) and with all cores (the build's OpenMP over column blocks), each
                the median of three runs.
  parity        max |flux - FP64 oracle| over those columns (W m-2).
  nn            the NN gas-optics path: LW and SW MLP surrogates trained on the GPU on the plain
                gas optics, then the same step timed with them, its speed-up over the plain path
                and its flux RMSE / max error against the plain path (--no-nn skips it).
  large         the same plain (and NN) step at 262144 columns x 128 layers, the size the
                north-star roofline target is stated at (--no-large skips it).
  adjoint       adjoint (reverse-mode) time over forward time at 16384 columns, and the peak
                device memory of the adjoint alone (--no-adjoint).
  kernels       per-kernel launches and time, from a separate pass of the same number of steps
                with per-launch HIP events (the timed steps run without them); class-API GB/s
                (the bytes the reference's class API would move, which a fused kernel may
                exceed the peak with) and, where profiles/pmc_traffic.json has the kernel, the
                measured-HBM GB/s.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "rte-rrtmgp-cpp_amd"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

METRIC = "LW+SW col-gpt-layers/s, plain vs NN gas optics at ≤0.1 W/m² flux err; adj/fwd"
HBM_PEAK_GBS = 8000.0
AVAILABLE = ["h2o", "co2", "o3", "n2o", "ch4"]
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(ncol, nlay, E, kd_lw, kd_sw, n2d_gases, cloudy):
    """Algorithmic bytes of each kernel over one LW+SW step, as {kernel: bytes}.

    Compulsory traffic of the arrays the reference's class API materializes: each
    (ncol, nlay|nlev, ngpt) boundary array is written once by its producer and read once by each
    consumer, E bytes per element (r = nlev/nlay).  A fused kernel is credited with the arrays of
    every class-API step it replaces:
      LW gas optics + cloud add_to:  tau, lay_src, lev_src_inc, lev_src_dec written (4), tau
                                     read and written again by add_to (2)         -> 6 E per U
      SW gas optics + cloud add_to:  tau, ssa, g written (3), add_to reads and writes all
                                     three (6)                                    -> 9 E per U
      LW solver + reduce:            4 inputs read, flux_up/dn per g-point written (2r) and
                                     read by the reduction (2r)                   -> (4 + 4r) E
      SW solver + reduce:            3 inputs read, flux_up/dn/dir per g-point written (3r)
                                     and read (3r)                                -> (3 + 6r) E
    so one all-sky step is (10 + 4r) E per LW U and (12 + 6r) E per SW U.  Per-(col, lay)
    inputs and (ncol, ngpt) boundary values are added; k-tables and LUTs (read through L2) are not.
    """
    nlev = nlay + 1
    cl, cv = ncol * nlay, ncol * nlev
    glw, gsw = kd_lw.ngpt, kd_sw.ngpt
    blw, bsw = kd_lw.nbnd, kd_sw.nbnd
    per_lay_in = (3 + n2d_gases) * E * cl            # play, tlay, col_dry, 2-D vmr
    b = {}
    b["gas_optics_lw_kernel"] = ((6 if cloudy else 4) * E * cl * glw + 2 * E * ncol * glw + per_lay_in
                                 + E * cv + E * ncol + (E * cl * blw if cloudy else 0))
    b["gas_optics_sw_kernel"] = ((9 if cloudy else 3) * E * cl * gsw + per_lay_in
                                 + (3 * E * cl * bsw if cloudy else 0))
    b["lw_solver_kernel"] = (4 * E * cl * glw + 4 * E * cv * glw + E * ncol * glw + E * ncol * blw
                             + 2 * E * cv)
    b["sw_solver_kernel"] = (3 * E * cl * gsw + 6 * E * cv * gsw + E * ncol * (gsw + 2 * bsw + 1)
                             + 3 * E * cv)
    if cloudy:
        b["cloud_optics_kernel"] = 4 * E * cl * 2 + E * cl * blw + 3 * E * cl * bsw
    b["col_dry_kernel"] = 2 * (E * cv + 2 * E * cl)
    b["net_kernel"] = 2 * 3 * E * cv
    b["scale_by_col_kernel"] = 2 * E * ncol * gsw + E * ncol
    return b


def adjoint_measure(kl, ks, ll, ls, prec, ncol=16384, nlay=128, reps=3):
    """Forward and adjoint (LW + SW, all-sky) of the 16384-column workload: mean times, their ratio
    and the adjoint's own peak device memory.  For the latter the library's cache is released after
    the forward runs, so what the cache holds after one adjoint (it only grows between releases)
    less what stayed live (the forward's workspaces) is the adjoint's scratch peak; the inputs and
    cotangents (allocated outside the cache) are added."""
    import ctypes as C
    from rrtmgp_amd import _lib, api, device
    from rrtmgp_amd import synthetic as syn
    atm = syn.make_atmosphere(syn.SEED_ATMOS[1], ncol, nlay, cloudy=True)
    kd_lw, kd_sw = api.KDist(kl, AVAILABLE, prec), api.KDist(ks, AVAILABLE, prec)
    slw, ssw = api.SolverLW(kd_lw, api.Cloud(ll, prec)), api.SolverSW(kd_sw, api.Cloud(ls, prec))
    dt = api.real_dtype(prec)
    da = api.DeviceAtmosphere(atm, prec)
    tsi = device.DeviceArray.from_host(np.full(ncol, atm.tsi / kd_sw.scalar("tsi")), dt)
    rng = np.random.default_rng(0)
    cot = [device.DeviceArray.from_host(rng.normal(size=(ncol, nlay + 1)), dt) for _ in range(5)]

    def fwd():
        slw.solve(da)
        ssw.solve(da, tsi)

    def adj():
        slw.adjoint(da, cot[0], cot[1])
        ssw.adjoint(da, tsi, cot[2], cot[3], cot[4])

    def stats():
        live, cached = C.c_size_t(0), C.c_size_t(0)
        _lib.check_hip(_lib.hip().rrtmgp_hip_cache_stats(C.byref(live), C.byref(cached)))
        return int(live.value), int(cached.value)

    out = {}
    fwd()
    device.synchronize()
    _lib.check_hip(_lib.hip().rrtmgp_hip_cache_release())
    live0, _ = stats()
    adj()
    device.synchronize()
    live1, cached1 = stats()
    peak_scratch = live1 + cached1 - live0
    inputs = sum(a.nbytes for a in vars(da).values() if isinstance(a, device.DeviceArray))
    inputs += tsi.nbytes + sum(c.nbytes for c in cot)
    for name, fn in (("forward", fwd), ("adjoint", adj)):
        fn()
        device.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        device.synchronize()
        out[name] = (time.perf_counter() - t0) * 1e3 / reps
    return {"workload": f"all-sky LW+SW, {ncol} columns x {nlay} layers, FP{8 * prec}",
            "forward_ms": round(out["forward"], 3), "adjoint_ms": round(out["adjoint"], 3),
            "adjoint_over_forward": round(out["adjoint"] / out["forward"], 3),
            "adjoint_peak_scratch_bytes": int(peak_scratch),
            "adjoint_peak_bytes_with_inputs": int(peak_scratch + inputs),
            "gradients": "T_lay, T_lev, T_sfc, ln p_lay, p_lev, vmr of every gas, emissivity, albedos"}


def train_surrogates(kd_lw, kd_sw, train_steps=60000, n_train_atm=32):
    """One surrogate per spectral range trained on the GPU on the plain gas optics of
    n_train_atm x 1024 x 128 perturbed clear-sky pairs (about 4M samples)."""
    from rrtmgp_amd import device, surrogate as S, synthetic as syn
    t0 = time.perf_counter()
    atms = [syn.make_atmosphere(500 + s, 1024, 128) for s in range(n_train_atm)]
    nets, losses, rows = {}, {}, 0
    for lw, kd in ((True, kd_lw), (False, kd_sw)):
        ts = S.build_training_set(kd, atms, seed=3 if lw else 4)
        net = S.MlpGasOptics(kd, seed=9)
        net.fit_normalisation(ts)
        hold = np.random.default_rng(10).integers(0, ts.rows, 16384)
        net.train(ts, steps=train_steps, batch=4096, lr=1e-3, lr_final=1e-5)
        losses["lw" if lw else "sw"] = round(net.loss(ts, hold), 5)
        nets[lw] = net
        rows = ts.rows
        del ts
    device.synchronize()
    train_s = time.perf_counter() - t0
    log(f"NN: trained LW and SW surrogates in {train_s:.1f} s; normalised losses {losses}")
    return nets, {"samples_per_range": rows, "steps": train_steps, "batch": 4096, "optimizer": "Adam lr 1e-3",
                  "seconds_lw_sw": round(train_s, 1), "holdout_loss": losses}


def nn_measure(kd_lw, kd_sw, nets, step, out_lw, out_sw, glw, gsw, steps, units, plain_kernels):
    """The same all-sky step timed with the surrogates set (no per-launch events in the timed
    steps; a second pass records the gas-optics kernel times), and its fluxes compared with the
    plain path's.  Training and the comparison are outside the timed steps."""
    from rrtmgp_amd import device, profiling, surrogate as S
    S.set_surrogate(kd_lw, nets[True])
    S.set_surrogate(kd_sw, nets[False])
    try:
        step()
        device.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        device.synchronize()
        wall = time.perf_counter() - t0
        profiling.enable(True)
        for _ in range(steps):
            step()
        device.synchronize()
        profiling.enable(False)
        kt = {k.name: k.avg_ms for k in profiling.collect()}
        fb = {"lw": S.last_fallback_units(kd_lw), "sw": S.last_fallback_units(kd_sw)}
        nlw = {k: v.to_host() for k, v in out_lw.items()}
        nsw = {k: v.to_host() for k, v in out_sw.items()}
    finally:
        S.set_surrogate(kd_lw, None)
        S.set_surrogate(kd_sw, None)
    rmse, emax = {}, {}
    for tag, got, ref in (("lw", nlw, glw), ("sw", nsw, gsw)):
        for k in ("flux_up", "flux_dn", "flux_net"):
            d = got[k].astype(np.float64) - ref[k]
            rmse[f"{tag}_{k}"] = round(float(np.sqrt(np.mean(d * d))), 4)
            emax[f"{tag}_{k}"] = round(float(np.abs(d).max()), 4)
    ms = wall * 1e3 / steps
    return {"ms_per_step": round(ms, 3), "value": round(units * steps / wall, 1),
            "gas_optics_ms": {"plain_lw": round(plain_kernels.get("gas_optics_lw_kernel", 0.0), 3),
                              "nn_lw": round(kt.get("mlp_gas_optics_lw_kernel", 0.0), 3),
                              "plain_sw": round(plain_kernels.get("gas_optics_sw_kernel", 0.0), 3),
                              "nn_sw": round(kt.get("mlp_gas_optics_sw_kernel", 0.0), 3)},
            "flux_rmse_wm2": rmse, "flux_err_max_wm2": emax, "fallback_units_last_call": fb,
            "network": "8 inputs (T, ln p, 5 ln vmr, lower-atmosphere flag) -> 128 -> 128 -> 2 ngpt, softplus, "
                       "BF16 weights/activations, FP32 accumulation"}


def cpu_baseline(atm, kl, ks, ll, ls, glw, gsw, ncols, threads, runs=3):
    """FP64 oracle on the first ncols columns with `threads` OpenMP threads, `runs` times; returns
    the median rate (col-gpt-layers/s), the median seconds and the max flux error of the GPU
    fluxes over those columns."""
    import oracle as O
    from rrtmgp_amd import synthetic as syn
    sub = syn.subset_columns(atm, slice(0, ncols))
    okl, oks = O.Kdist(kl, AVAILABLE), O.Kdist(ks, AVAILABLE)
    ocl, ocs = O.Cloud(ll), O.Cloud(ls)
    tsi = np.full(sub.ncol, sub.tsi / oks.scalar("tsi"))
    prev = O.set_threads(threads)
    secs = []
    try:
        for _ in range(runs):
            t0 = time.perf_counter()
            rlw = O.solve_lw(okl, sub, ocl)
            rsw = O.solve_sw(oks, sub, ocs, tsi_scaling=tsi)
            secs.append(time.perf_counter() - t0)
    finally:
        O.set_threads(prev)
    dt = float(np.median(secs))
    units = sub.ncol * sub.nlay * (okl.ngpt + oks.ngpt)
    err = {}
    for name, ref, got in (("lw", rlw, glw), ("sw", rsw, gsw)):
        for k, v in ref.items():
            if k in got:
                err[f"{name}_{k}"] = float(np.abs(got[k][:ncols] - v).max())
    return units / dt, dt, err


def precision_study(ncol, nlay, block, ocols, threads):
    """The FP32 bulk against an all-FP64 GPU run on the same inputs, with 10% of the cloudy layers
    forced to ssa = 1 - 1e-7, g = 0.85 before delta-scaling (near-conservative scattering): the
    shortwave through gas optics, cloud optics, delta-scale, add_to and the broadband solver, all in
    co-albedo form (1 - ssa carried in place of ssa, as the shortwave driver runs them); the
    longwave through the driver.  Max and RMS flux error (W m-2) and heating-rate error (K/day,
    from each run's net fluxes differenced in FP64 on the host) over all columns; both GPU runs
    against the FP64 oracle on the first ocols columns.  Columns go in blocks of `block`."""
    import oracle as O
    from rrtmgp_amd import api, device, synthetic as syn
    kl, ks = syn.make_kdist(True, syn.SEED_KDIST_LW), syn.make_kdist(False, syn.SEED_KDIST_SW)
    ll, ls = syn.make_cloud_lut(True, syn.SEED_CLOUD_LW), syn.make_cloud_lut(False, syn.SEED_CLOUD_SW)
    t0 = time.perf_counter()
    atm = syn.make_atmosphere(syn.SEED_ATMOS[2], ncol, nlay, cloudy=True, conservative_frac=0.1)
    cons = atm.extra["conservative"]
    log(f"precision study: {ncol} x {nlay}, {int(cons.sum())} near-conservative cloud layers "
        f"({time.perf_counter() - t0:.1f} s to generate)")
    g_cp = 9.80665 / 1004.64 * 86400.0

    def heating(up, dn, plev):
        net = up - dn
        return g_cp * (net[:, 1:] - net[:, :-1]) / (plev[:, 1:] - plev[:, :-1])

    res = {}
    for prec in (4, 8):
        t0 = time.perf_counter()
        kd = api.KDist(ks, AVAILABLE, prec)
        cloud = api.Cloud(ls, prec)
        scale = atm.tsi / kd.scalar("tsi")
        sw = {k: np.empty((ncol, nlay + 1)) for k in ("up", "dn", "dir")}
        sw["hr"] = np.empty((ncol, nlay))
        for c0 in range(0, ncol, block):
            sub = syn.subset_columns(atm, slice(c0, min(ncol, c0 + block)))
            n = sub.ncol
            vmr = {g: np.broadcast_to(np.asarray(sub.vmr[g], np.float64), (n, nlay)).copy(order="F")
                   for g in sub.vmr}
            gc = api.GasConcs(vmr, prec)
            cd = api.get_col_dry(vmr["h2o"], sub.p_lev, prec)
            # The whole shortwave chain in co-albedo form (1 - ssa), as the driver runs it.
            go = api.gas_optics(kd, sub, gc, cd, coalbedo=True)
            ct, cc, cg = (a.to_host().astype(np.float64)
                          for a in cloud.optics(sub.lwp, sub.iwp, sub.rel, sub.rei, True, coalbedo=True))
            m = sub.extra["conservative"][:, :, None]
            cc, cg = np.where(m, 1e-7, cc), np.where(m, 0.85, cg)
            f = cg * cg                                  # delta-scaling in co-albedo form, FP64
            d = (1.0 - f) + f * cc
            ct, cc, cg = ct * d, cc / np.maximum(d, 1e-300), (cg - f) / np.maximum(1.0 - f, 1e-12)
            dt = api.real_dtype(prec)
            ct, cc, cg = (device.DeviceArray.from_host(a, dt) for a in (ct, cc, cg))
            api.add_to(kd, go["tau"], ct, go["coalbedo"], go["g"], cc, cg, coalbedo=True)
            toa = device.DeviceArray.from_host(go["toa_src"].to_host() * scale, dt)
            o = api.rte_sw_bnd(kd, sub.top_at_1, go["tau"], go["coalbedo"], go["g"], sub.mu0, toa, sub.sfc_alb,
                               0.5 * sub.sfc_alb, p_lev=sub.p_lev, coalbedo=True)
            for k, v in (("up", "flux_up"), ("dn", "flux_dn"), ("dir", "flux_dn_dir"), ("hr", "heating_rate")):
                sw[k][c0:c0 + n] = o[v].to_host()
            del go, ct, cc, cg, o, toa
        kdl = api.KDist(kl, AVAILABLE, prec)
        slw = api.SolverLW(kdl, api.Cloud(ll, prec))
        slw.set_block(block)
        lw = slw.solve(api.DeviceAtmosphere(atm, prec), heating_rate=True)
        lw = {"up": lw["flux_up"].to_host().astype(np.float64), "dn": lw["flux_dn"].to_host().astype(np.float64),
              "hr": lw["heating_rate"].to_host().astype(np.float64)}
        device.synchronize()
        res[prec] = {"sw": sw, "lw": lw}
        log(f"FP{8 * prec} run in {time.perf_counter() - t0:.1f} s")
        del kd, cloud, kdl, slw

    def errs(a, b):
        d = a - b
        return {"max": float(np.abs(d).max()), "rms": float(np.sqrt(np.mean(d * d)))}

    out = {}
    plev = np.asarray(atm.p_lev, np.float64)
    for rng, names in (("sw", ("up", "dn", "dir")), ("lw", ("up", "dn"))):
        f4, f8 = res[4][rng], res[8][rng]
        for k in names:
            out[f"{rng}_flux_{k}"] = errs(f4[k], f8[k])
        out[f"{rng}_flux_net"] = errs(f4["up"] - f4["dn"], f8["up"] - f8["dn"])
        # Heating rates from the solvers' FP64 island (each layer's net-flux convergence formed per
        # g-point and summed in FP64), on all layers.
        out[f"{rng}_heating_rate_kday"] = errs(f4["hr"], f8["hr"])
        # For comparison: heating rates differenced on the host from each run's rounded level fluxes.
        # Above 1 hPa the layers are ~1 Pa thick, and the FP32 rounding of the fluxes alone
        # (~1e-4 W m-2) moves those by K/day.
        h4, h8 = heating(f4["up"], f4["dn"], plev), heating(f8["up"], f8["dn"], plev)
        deep = (0.5 * (plev[:, 1:] + plev[:, :-1])) > 100.0
        out[f"{rng}_heating_rate_kday_from_level_fluxes"] = errs(h4, h8)
        out[f"{rng}_heating_rate_kday_from_level_fluxes_p_gt_1hPa"] = errs(h4[deep], h8[deep])
    # Oracle on the first ocols columns (SW with the same forced layers).
    sub = syn.subset_columns(atm, slice(0, ocols))
    okd = O.Kdist(ks, AVAILABLE)
    prev = O.set_threads(threads)
    try:
        ogo = O.gas_optics(okd, sub)
        ct, cs, cg = O.Cloud(ls).optics(sub.lwp, sub.iwp, sub.rel, sub.rei, True)
        m = sub.extra["conservative"][:, :, None]
        ct, cs, cg = O.delta_scale(ct, np.where(m, 1.0 - 1e-7, cs), np.where(m, 0.85, cg))
        tau, ssa, g = O.inc_bybnd(ogo["tau"], ct, okd.band_lims, ogo["ssa"], ogo["g"], cs, cg)
        rup, rdn, rdr = O.rte_sw(tau, ssa, g, sub.mu0, ogo["toa_src"] * (atm.tsi / okd.scalar("tsi")), sub.sfc_alb,
                                 0.5 * sub.sfc_alb, okd.band_lims, sub.top_at_1)
        rlw = O.solve_lw(O.Kdist(kl, AVAILABLE), sub, O.Cloud(ll))
    finally:
        O.set_threads(prev)
    oref = {"up": rup.sum(2), "dn": rdn.sum(2), "dir": rdr.sum(2)}
    oracle = {}
    for prec in (4, 8):
        f = res[prec]
        oracle[f"fp{8 * prec}"] = {
            "sw_flux_max": max(float(np.abs(f["sw"][k][:ocols] - oref[k]).max()) for k in oref),
            "lw_flux_max": max(float(np.abs(f["lw"][k][:ocols] - rlw["flux_" + k]).max()) for k in ("up", "dn"))}
    return {"metric": "precision study: FP32 bulk vs all-FP64 GPU (W m-2, K/day)",
            "config": {"workload": f"all-sky LW+SW, {ncol} columns x {nlay} layers, 10% of cloud layers at "
                                   "ssa = 1 - 1e-7, g = 0.85", "ncol": ncol, "nlay": nlay,
                       "near_conservative_layers": int(cons.sum()), "column_block": block},
            "fp32_vs_fp64_gpu": out, "vs_fp64_oracle": {"columns": ocols, **oracle}}


def copy_calibration():
    """One device copy of 1 GiB (copy_cols_kernel): read and written bytes are known exactly, so
    the FETCH_SIZE / WRITE_SIZE counters of this dispatch calibrate those of the others."""
    import ctypes as C
    from rrtmgp_amd import _lib, device
    n, m = 65536, 4096
    a = device.DeviceArray((n, m), np.float32, zero=True)
    b = device.DeviceArray((n, m), np.float32, zero=False)
    _lib.check_hip(_lib.hip().rrtmgp_hip_copy_cols_f32(a.ptr, C.c_int64(n), C.c_int64(0), b.ptr, C.c_int64(n),
                                                       C.c_int64(0), C.c_int64(n), C.c_int64(m), None))
    device.synchronize()
    a.free()
    b.free()


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ncol", type=int, default=None, help="columns (65536; 262144 for --precision-study)")
    ap.add_argument("--nlay", type=int, default=128)
    ap.add_argument("--precision", type=int, choices=(32, 64), default=32)
    ap.add_argument("--n-col-block", type=int, default=0, help="columns per solver block (0: all)")
    ap.add_argument("--cpu-cols", type=int, default=4096, help="columns of the all-cores CPU baseline sample")
    ap.add_argument("--cpu-cols-serial", type=int, default=1024, help="columns of the serial CPU baseline sample")
    ap.add_argument("--large-ncol", type=int, default=262144, help="columns of the second (large) step")
    ap.add_argument("--large-steps", type=int, default=5, help="timed steps at the large size")
    ap.add_argument("--no-large", action="store_true", help="skip the large-size step")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0: OMP_NUM_THREADS, max 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward measurement")
    ap.add_argument("--no-nn", action="store_true", help="skip the NN gas-optics (surrogate) measurement")
    ap.add_argument("--precision-study", action="store_true",
                    help="instead of the throughput line: FP32 vs FP64 flux and heating-rate errors with "
                         "near-conservative clouds (262144 columns unless --ncol is given)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="launch one untimed 1 GiB device copy first (reference dispatch for PMC counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): columns are independent, "
              "so a multi-GPU run would shard columns with no exchange; only the one-GPU "
              "throughput is measured here.", file=sys.stderr, flush=True)
        sys.exit(2)

    from rrtmgp_amd import api, device, profiling, synthetic as syn

    if args.precision_study:
        threads = args.cpu_threads or min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        blk = args.n_col_block if args.n_col_block > 0 else 65536
        res = precision_study(args.ncol or 262144, args.nlay, blk, min(args.cpu_cols, args.ncol or 262144), threads)
        print(json.dumps(res), flush=True)
        return

    prec = 4 if args.precision == 32 else 8
    ncol, nlay = args.ncol or 65536, args.nlay
    kl, ks = syn.make_kdist(True, syn.SEED_KDIST_LW), syn.make_kdist(False, syn.SEED_KDIST_SW)
    ll, ls = syn.make_cloud_lut(True, syn.SEED_CLOUD_LW), syn.make_cloud_lut(False, syn.SEED_CLOUD_SW)
    kd_lw, kd_sw = api.KDist(kl, AVAILABLE, prec), api.KDist(ks, AVAILABLE, prec)

    if args.copy_calibration:
        copy_calibration()
        log("calibration copy done (1 GiB read, 1 GiB written)")

    m = measure_step(ncol, nlay, prec, args, kd_lw, kd_sw, ll, ls)
    atm, units, glw, gsw = m["atm"], m["units"], m["glw"], m["gsw"]

    cpu = None
    parity = {"finite": m["finite"]}
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        n1 = min(args.cpu_cols_serial, ncol)
        log(f"CPU baseline: FP64 oracle, serial on {n1} columns, median of 3")
        r1, s1, err1 = cpu_baseline(atm, kl, ks, ll, ls, glw, gsw, n1, 1)
        nn_ = min(args.cpu_cols, ncol)
        log(f"CPU baseline: FP64 oracle, {threads} threads on {nn_} columns, median of 3")
        rn, sn, err = cpu_baseline(atm, kl, ks, ll, ls, glw, gsw, nn_, threads)
        cpu = {"value": round(r1, 1), "unit": "col-gpt-layers/s", "cores": 1, "kind": "port",
               "sample": f"first {n1} of the {ncol} columns x {nlay} layers, all-sky LW+SW, FP64 oracle "
                         f"(restated reference CPU path, serial as the reference, 12-column blocks), median of "
                         f"3 runs of {s1:.2f} s",
               "all_cores": {"value": round(rn, 1), "cores": threads,
                             "sample": f"first {nn_} columns, OpenMP over 12-column blocks (this build's "
                                       f"parallelisation, not the reference's), median of 3 runs of {sn:.2f} s"}}
        parity.update({"columns": nn_, "flux_err_max_wm2": round(max(err.values()), 6),
                       "per_output": {k: round(v, 6) for k, v in err.items()}})
        log(f"CPU baseline {r1:.3e} U/s serial, {rn:.3e} U/s on {threads} threads; "
            f"max flux error {max(err.values()):.3e} W m-2")

    nn, nets = None, None
    if not args.no_nn and prec == 4:
        log("NN gas optics: training surrogates and timing the same step with them")
        nets, training = train_surrogates(kd_lw, kd_sw)
        nn = nn_measure(kd_lw, kd_sw, nets, m["step"], m["out_lw"], m["out_sw"], glw, gsw, args.steps, units,
                        m["kernel_avg_ms"])
        nn["training"] = training
        nn["speedup_vs_plain"] = round(m["ms_per_step"] / nn["ms_per_step"], 3)
        log(f"NN path {nn['ms_per_step']:.2f} ms/step ({nn['speedup_vs_plain']:.2f}x); flux RMSE {nn['flux_rmse_wm2']}")

    line = {
        "metric": METRIC, "value": round(m["value"], 1), "unit": "col-gpt-layers/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(m["ms_per_step"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec == 4 else "f64",
        "data": "synthetic: seeded RRTMGP-layout k-distributions, cloud LUTs and profiles",
        "config": {"workload": f"all-sky LW+SW forward, plain gas optics, {ncol} columns x {nlay} layers",
                   "ncol": ncol, "nlay": nlay, "ngpt_lw": kd_lw.ngpt, "ngpt_sw": kd_sw.ngpt,
                   "n_col_block": m["block"], "gpu_ms_per_step": m["gpu_ms_per_step"]},
        "roofline": m["roofline"], "cpu_baseline": cpu, "parity": parity, "nn": nn,
    }
    kernels = m["kernels"]
    del m, atm, glw, gsw

    large = None
    if not args.no_large and args.large_ncol > ncol:
        _release_device_cache()
        ml = measure_step(args.large_ncol, nlay, prec, args, kd_lw, kd_sw, ll, ls, steps=args.large_steps,
                          warmup=min(args.warmup, 2))
        large = {"workload": f"all-sky LW+SW forward, plain gas optics, {args.large_ncol} columns x {nlay} layers",
                 "value": round(ml["value"], 1), "ms_per_step": round(ml["ms_per_step"], 3),
                 "steps": args.large_steps, "roofline": ml["roofline"], "finite": ml["finite"],
                 "kernels": ml["kernels"][:6]}
        if nets is not None:
            nl = nn_measure(kd_lw, kd_sw, nets, ml["step"], ml["out_lw"], ml["out_sw"], ml["glw"], ml["gsw"],
                            args.large_steps, ml["units"], ml["kernel_avg_ms"])
            nl["speedup_vs_plain"] = round(ml["ms_per_step"] / nl["ms_per_step"], 3)
            large["nn"] = nl
            log(f"NN at {args.large_ncol}: {nl['ms_per_step']:.2f} ms/step ({nl['speedup_vs_plain']:.2f}x)")
        del ml
        _release_device_cache()
    line["large"] = large

    adjoint = None
    if not args.no_adjoint:
        log("adjoint: LW + SW forward and adjoint at 16384 columns")
        adjoint = adjoint_measure(kl, ks, ll, ls, prec)
        log(f"adjoint {adjoint['adjoint_ms']:.1f} ms vs forward {adjoint['forward_ms']:.1f} ms")
    line["adjoint"] = adjoint
    line["kernels"] = kernels
    print(json.dumps(line), flush=True)


def _release_device_cache():
    from rrtmgp_amd import _lib, device
    device.synchronize()
    _lib.check_hip(_lib.hip().rrtmgp_hip_cache_release())


def _traffic():
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            return json.load(f)
    return None


def measure_step(ncol, nlay, prec, args, kd_lw, kd_sw, ll, ls, steps=None, warmup=None):
    """Set up the all-sky step at ncol columns (inputs resident), warm up, time `steps` steps with
    no per-launch instrumentation, then run the same number of steps again with per-launch HIP
    events for the per-kernel table and the roofline of the dominant kernel."""
    from rrtmgp_amd import api, device, profiling, synthetic as syn
    steps = steps or args.steps
    warmup = args.warmup if warmup is None else warmup
    E = prec
    log(f"workload: all-sky LW+SW, {ncol} columns x {nlay} layers, FP{8 * prec}")
    t0 = time.perf_counter()
    atm = syn.make_atmosphere(syn.SEED_ATMOS[1], ncol, nlay, cloudy=True)
    log(f"profiles generated in {time.perf_counter() - t0:.1f} s")
    solver_lw = api.SolverLW(kd_lw, api.Cloud(ll, prec))
    solver_sw = api.SolverSW(kd_sw, api.Cloud(ls, prec))
    blk = args.n_col_block if args.n_col_block > 0 else ncol
    solver_lw.set_block(blk)
    solver_sw.set_block(blk)
    da = api.DeviceAtmosphere(atm, prec)
    dt = api.real_dtype(prec)
    tsi = device.DeviceArray.from_host(np.full(ncol, atm.tsi / kd_sw.scalar("tsi")), dt)
    nlev = nlay + 1
    out_lw = {k: device.DeviceArray((ncol, nlev), dt) for k in ("flux_up", "flux_dn", "flux_net")}
    out_sw = {k: device.DeviceArray((ncol, nlev), dt) for k in ("flux_up", "flux_dn", "flux_dn_dir", "flux_net")}
    device.synchronize()
    log("inputs resident on the device")

    def step():
        solver_lw.solve(da, outputs=out_lw)
        solver_sw.solve(da, tsi, outputs=out_sw)

    for i in range(max(warmup, 1)):
        step()
        device.synchronize()
        log(f"warmup step {i + 1} done")

    start, stop = device.Event(), device.Event()
    device.synchronize()
    t0 = time.perf_counter()
    start.record()
    for _ in range(steps):
        step()
    stop.record()
    device.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = start.elapsed_ms(stop)
    log(f"{steps} timed steps: {wall * 1e3:.1f} ms wall, {gpu_ms:.1f} ms between events")

    profiling.enable(True)
    for _ in range(steps):
        step()
    device.synchronize()
    profiling.enable(False)
    kernels = profiling.collect()
    log(f"{steps} profiled steps done ({len(kernels)} kernels)")

    units = ncol * nlay * (kd_lw.ngpt + kd_sw.ngpt)
    ms_per_step = wall * 1e3 / steps
    n2d = sum(1 for v in atm.vmr.values() if v.shape[0] > 1)
    abytes = algorithmic_bytes(ncol, nlay, E, kd_lw, kd_sw, n2d, True)
    tr = _traffic()
    tr_ok = tr is not None and tr.get("workload") == {"ncol": ncol, "nlay": nlay, "precision": 8 * prec}
    krows = []
    for k in kernels:
        ab = abytes.get(k.name)
        row = {"kernel": k.name, "launches": k.launches, "total_ms": round(k.total_ms, 4),
               "avg_ms": round(k.avg_ms, 5),
               "class_api_GBps": (round(ab * steps / (k.total_ms * 1e-3) / 1e9, 1)
                                  if ab is not None and k.total_ms > 0 else None)}
        rec = tr.get("kernels", {}).get(k.name) if tr_ok else None
        if rec and k.avg_ms > 0:
            row["hbm_GBps_measured"] = round(rec["hbm_bytes_per_launch"] / (k.avg_ms * 1e-3) / 1e9, 1)
        if row["class_api_GBps"] is not None and row["class_api_GBps"] > HBM_PEAK_GBS:
            row["note"] = ("class-API credit above the HBM peak: the fused kernel does not move all the "
                           "arrays the reference's class API would (see hbm_GBps_measured)")
        krows.append(row)
    krows.sort(key=lambda r: -r["total_ms"])
    dom = next((r for r in krows if r["class_api_GBps"] is not None), None)
    roofline = None
    if dom is None:
        log("no kernel timings were recorded; roofline left empty")
    else:
        dom_bytes_per_launch = abytes[dom["kernel"]] * steps / dom["launches"]
        rec = tr.get("kernels", {}).get(dom["kernel"]) if tr_ok else None
        step_bytes = sum(abytes.values())
        step_gbs = step_bytes / (ms_per_step * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": dom["kernel"], "achieved": dom["class_api_GBps"],
                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(dom["class_api_GBps"] / HBM_PEAK_GBS, 4),
                    "traffic": rec.get("hbm_bytes_per_launch") if rec else None,
                    "algorithmic_bytes_per_launch": int(dom_bytes_per_launch),
                    "avg_launch_ms": dom["avg_ms"],
                    "step": {"algorithmic_bytes": int(step_bytes), "achieved": round(step_gbs, 1),
                             "frac": round(step_gbs / HBM_PEAK_GBS, 4)}}
    glw = {k: v.to_host() for k, v in out_lw.items()}
    gsw = {k: v.to_host() for k, v in out_sw.items()}
    finite = all(np.isfinite(a).all() for a in list(glw.values()) + list(gsw.values()))
    return {"atm": atm, "units": units, "ms_per_step": ms_per_step, "value": units * steps / wall,
            "gpu_ms_per_step": round(gpu_ms / steps, 3), "block": blk, "roofline": roofline, "kernels": krows,
            "kernel_avg_ms": {k.name: k.avg_ms for k in kernels}, "glw": glw, "gsw": gsw, "finite": bool(finite),
            "step": step, "out_lw": out_lw, "out_sw": out_sw}


if __name__ == "__main__":
    main()
