"""MLP gas-optics surrogate: training data from the plain gas optics, BF16 matrix-core training and
fused inference behind a k-distribution.

No reference counterpart.  Per (column, layer) pair the network maps
    x = [T, ln p, ln h2o, ln o3, ln co2, ln ch4, ln n2o, f],
f = 1 in the lower atmosphere (p > p_trop, where the k-distribution switches to its lower tables,
a jump in ln p that a smooth network otherwise fits poorly) and 0 above, normalised with the
training set's mean and standard deviation, through two softplus layers of
128 units to 2 ngpt outputs, interleaved per g-point:
    longwave   (ln(tau / col_dry), Planck fraction)
    shortwave  (ln(tau_abs / col_dry), ln(tau_ray / col_dry))
Targets come from the library's own fused gas optics (FP32) on atmospheres whose gas amounts are
perturbed by factors in [0.5, 2] per pair.  Training runs entirely on the device through the
library's kernels (rrtmgp_hip_mlp_*): BF16 operands on matrix cores, FP32 accumulation, FP32
master weights and Adam moments; longwave outputs are weighted in the loss by their band's Planck
emission, the Planck fractions four times more (MlpGasOptics.loss_weights).  Inference is one fused kernel per call
(rrtmgp_hip_mlp_gas_optics_lw/sw) behind KDist via rrtmgp_api_kdist_set_surrogate, so the
solvers use it unchanged; pairs outside the training envelope fall back to the plain kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib, api
from .device import DeviceArray, synchronize

GASES = ("h2o", "o3", "co2", "ch4", "n2o")
HIDDEN = 128
N_IN = 8            # T, ln p, 5 ln vmr, lower-atmosphere flag
N_INP = 16          # inputs padded to one matrix-core K step
MAX_GAS = 32        # RRTMGP_MAX_GAS


class GasVmrC(C.Structure):
    _fields_ = [("ngas", C.c_int64), ("vmr", C.c_void_p * MAX_GAS), ("dim1", C.c_int64 * MAX_GAS),
                ("dim2", C.c_int64 * MAX_GAS)]


class MlpNetC(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("W1", "W2", "W3", "b1", "b2", "b3", "x_mean", "x_std", "x_lo",
                                          "x_hi", "y_mean", "y_std")] + [("n_out", C.c_int64), ("p_trop", C.c_double)]


def _L():
    return _lib.hip()


def _chk(rc):
    _lib.check_hip(rc)


def _i(x):
    return C.c_int64(int(x))


def _f(x):
    return C.c_float(float(x))


def gas_desc(vmr: dict, ncol: int, nlay: int):
    """rrtmgp_gas_vmr of the five input gases (FP32 DeviceArrays of shape (ncol|1, nlay|1))."""
    d = GasVmrC()
    d.ngas = len(GASES)
    for i, g in enumerate(GASES):
        a = vmr[g]
        if not isinstance(a, DeviceArray) or a.dtype != np.float32 or len(a.shape) != 2:
            raise TypeError(f"{g}: expected a 2-D FP32 DeviceArray")
        d.vmr[i] = a.ptr.value
        d.dim1[i], d.dim2[i] = a.shape
    return d


def perturbed_vmr(atm, rng, lo=0.5, hi=2.0):
    """Host vmr (ncol, nlay) per gas: the atmosphere's amounts times log-uniform factors in [lo, hi]."""
    ncol, nlay = atm.p_lay.shape
    out = {}
    for g in GASES:
        base = np.broadcast_to(np.asarray(atm.vmr[g], dtype=np.float64), (ncol, nlay))
        f = np.exp(rng.uniform(np.log(lo), np.log(hi), (ncol, nlay)))
        out[g] = np.asfortranarray(base * f)
    return out


class TrainingSet:
    """Raw inputs X (rows, 16) and targets Y (rows, 2 ngpt), FP32 on the device (row-major)."""

    def __init__(self, X: DeviceArray, Y: DeviceArray, rows: int, n_out: int):
        self.X, self.Y, self.rows, self.n_out = X, Y, rows, n_out


def build_training_set(kd: api.KDist, atmospheres, seed=0, lo=0.5, hi=2.0) -> TrainingSet:
    """Runs the plain FP32 gas optics (clear sky) on each atmosphere, gases perturbed per pair."""
    if kd.precision != 4:
        raise ValueError("the surrogate trains on the FP32 gas optics")
    rng = np.random.Generator(np.random.PCG64(seed))
    set_surrogate(kd, None)
    rows = sum(a.p_lay.size for a in atmospheres)
    no = 2 * kd.ngpt
    X = DeviceArray((N_INP, rows), np.float32)          # Fortran (16, rows) = row-major (rows, 16)
    Y = DeviceArray((no, rows), np.float32)
    g2b = gpt2band(kd)
    planck = planck_table(kd) if kd.is_lw else None
    r0 = 0
    for atm in atmospheres:
        ncol, nlay = atm.p_lay.shape
        host_vmr = perturbed_vmr(atm, rng, lo, hi)
        vmr = {g: DeviceArray.from_host(v, np.float32) for g, v in host_vmr.items()}
        gc = api.GasConcs(vmr, 4)
        cd = api.get_col_dry(vmr["h2o"], atm.p_lev, 4)
        o = api.gas_optics(kd, atm, gc, cd)
        play = DeviceArray.from_host(atm.p_lay, np.float32)
        tlay = DeviceArray.from_host(atm.t_lay, np.float32)
        desc = gas_desc(vmr, ncol, nlay)
        xp = C.c_void_p(X.ptr.value + 4 * N_INP * r0)
        yp = C.c_void_p(Y.ptr.value + 4 * no * r0)
        _chk(_L().rrtmgp_hip_mlp_inputs(_i(ncol), _i(nlay), play.ptr, tlay.ptr, C.byref(desc),
                                        C.c_double(kd.src.press_ref_trop), xp, None))
        b = o["lay_src"] if kd.is_lw else o["ssa"]
        _chk(_L().rrtmgp_hip_mlp_targets(
            C.c_int(int(kd.is_lw)), _i(ncol), _i(nlay), _i(kd.ngpt), g2b.ptr, tlay.ptr, cd.ptr, o["tau"].ptr, b.ptr,
            planck[0].ptr if planck else None, _i(planck[1] if planck else 2),
            C.c_double(planck[2] if planck else 0.0), C.c_double(planck[3] if planck else 1.0), yp, None))
        synchronize()
        r0 += ncol * nlay
    return TrainingSet(X, Y, rows, no)


def gpt2band(kd: api.KDist) -> DeviceArray:
    """0-based band of each g-point (int32 on the device)."""
    lims = kd.src.band2gpt
    g2b = np.zeros(kd.ngpt, dtype=np.int32)
    for b in range(kd.nbnd):
        g2b[lims[0, b] - 1:lims[1, b]] = b
    return DeviceArray.from_host(g2b, np.int32)


def planck_table(kd: api.KDist):
    """(totplnk FP32 (nplanck, nbnd), nplanck, temp_ref_min, delta) of a longwave k-distribution."""
    tp = np.asarray(kd.src.totplnk)
    tmin, tmax = kd.scalar("temp_ref_min"), kd.scalar("temp_ref_max")
    return (DeviceArray.from_host(tp, np.float32), tp.shape[0], tmin, (tmax - tmin) / (tp.shape[0] - 1))


def set_surrogate(kd: api.KDist, net: "MlpGasOptics | None"):
    """Routes kd's gas optics (and its solvers') through net; None restores the plain kernels."""
    kd._check(kd.L.rrtmgp_api_kdist_set_surrogate(kd.h, C.byref(net.c_net()) if net is not None else None))
    kd._surrogate = net                   # keeps the network's device arrays alive while in use


def last_fallback_units(kd: api.KDist) -> int:
    n = C.c_int64(0)
    kd._check(kd.L.rrtmgp_api_kdist_last_fallback(kd.h, C.byref(n)))
    return n.value


class _Param:
    """FP32 master weights with Adam moments and an optional BF16 copy on the device."""

    def __init__(self, host: np.ndarray, bf16: bool):
        self.shape = host.shape
        self.n = host.size
        self.p = DeviceArray.from_host(np.ascontiguousarray(host, np.float32).ravel(), np.float32)
        self.g = DeviceArray((self.n,), np.float32)
        self.m = DeviceArray((self.n,), np.float32)
        self.v = DeviceArray((self.n,), np.float32)
        self.b16 = DeviceArray((self.n,), np.uint16, zero=False) if bf16 else None
        if bf16:
            _chk(_L().rrtmgp_hip_mlp_to_bf16(_i(self.n), self.p.ptr, self.b16.ptr, None))

    def host(self):
        return self.p.to_host().reshape(self.shape)


class MlpGasOptics:
    """The surrogate of one k-distribution: parameters, normalisation, training and inference.

    Weights are row-major (out, in): W1 (128, 16) (inputs 7..15 are padding and stay zero), W2 (128,
    128), W3 (2 ngpt, 128)."""

    def __init__(self, kd: api.KDist, seed=0):
        if kd.precision != 4:
            raise ValueError("the surrogate runs with the FP32 library")
        if (2 * kd.ngpt) % 32:
            raise ValueError("2 ngpt must be a multiple of 32")
        self.kd = kd
        self.lw = kd.is_lw
        self.n_out = 2 * kd.ngpt
        self.p_trop = float(kd.src.press_ref_trop)
        rng = np.random.Generator(np.random.PCG64(seed))

        def glorot(o, i, fan_in):
            return rng.normal(0.0, np.sqrt(2.0 / (fan_in + o)), (o, i))

        w1 = np.zeros((HIDDEN, N_INP))
        w1[:, :N_IN] = glorot(HIDDEN, N_IN, N_IN)
        self.W1 = _Param(w1, True)
        self.W2 = _Param(glorot(HIDDEN, HIDDEN, HIDDEN), True)
        self.W3 = _Param(glorot(self.n_out, HIDDEN, HIDDEN), True)
        self.b1 = _Param(np.zeros(HIDDEN), False)
        self.b2 = _Param(np.zeros(HIDDEN), False)
        self.b3 = _Param(np.zeros(self.n_out), False)
        self.params = [self.W1, self.b1, self.W2, self.b2, self.W3, self.b3]
        self.step = 0
        self.stats = None
        self._ws = None

    # -- normalisation -------------------------------------------------------------------------
    def loss_weights(self, t_ref=260.0, pf_weight=4.0) -> np.ndarray | None:
        """Per-output weights of the training loss.  Longwave: each band's Planck emission at t_ref
        (clipped to [0.1, 10] times the mean) on both outputs of its g-points, so that the error
        budget goes where the flux is, and pf_weight more on the Planck fractions, whose errors
        dominate the upward flux's.  On the synthetic 16-band set the clear-sky upward-flux RMSE
        was 1.35 W m-2 unweighted, 0.92 / 1.57 / 1.79 over three seeds with the band weights
        alone, and 0.86 / 0.94 / 0.89 with both.  Shortwave: none (weighting by band solar energy
        raised the upward-flux error there)."""
        if not self.lw:
            return None
        tp = np.asarray(self.kd.src.totplnk, np.float64)
        tmin, tmax = self.kd.scalar("temp_ref_min"), self.kd.scalar("temp_ref_max")
        x = (t_ref - tmin) / (tmax - tmin) * (tp.shape[0] - 1)
        i0 = min(int(x), tp.shape[0] - 2)
        e = tp[i0] + (x - i0) * (tp[i0 + 1] - tp[i0])
        lims = np.asarray(self.kd.src.band2gpt)
        w = np.empty(self.kd.ngpt)
        for b in range(self.kd.nbnd):
            w[lims[0, b] - 1:lims[1, b]] = np.clip(e[b] / e.mean(), 0.1, 10.0)
        w = np.repeat(w, 2) * np.tile([1.0, pf_weight], self.kd.ngpt)
        return w / w.mean()

    def fit_normalisation(self, ts: TrainingSet, sample=65536, seed=1, weights="auto"):
        """Mean/std of inputs and targets over a random sample of rows; input envelope over all rows.
        Output o is normalised by std_o / sqrt(w_o), which makes the training's mean squared error
        a w-weighted one; weights="auto" takes loss_weights(), None none."""
        X = ts.X.to_host().T[:, :N_IN].astype(np.float64)                 # (rows, N_IN)
        rng = np.random.Generator(np.random.PCG64(seed))
        idx = np.sort(rng.choice(ts.rows, size=min(sample, ts.rows), replace=False))
        Ys = self._rows(ts.Y, ts.n_out, idx).astype(np.float64)
        x_std = X.std(axis=0)
        y_std = np.where(Ys.std(axis=0) > 1e-6, Ys.std(axis=0), 1.0)
        w = self.loss_weights() if isinstance(weights, str) and weights == "auto" else weights
        if w is not None:
            y_std = y_std / np.sqrt(np.asarray(w, np.float64))
        self.set_normalisation(X.mean(axis=0), np.where(x_std > 1e-6, x_std, 1.0), X.min(axis=0), X.max(axis=0),
                               Ys.mean(axis=0), y_std)

    def set_normalisation(self, x_mean, x_std, x_lo, x_hi, y_mean, y_std):
        f = lambda a: DeviceArray.from_host(np.asarray(a, np.float32).ravel(), np.float32)  # noqa: E731
        self.host_stats = dict(x_mean=np.asarray(x_mean), x_std=np.asarray(x_std), x_lo=np.asarray(x_lo),
                               x_hi=np.asarray(x_hi), y_mean=np.asarray(y_mean), y_std=np.asarray(y_std))
        self.stats = {k: f(v) for k, v in self.host_stats.items()}
        self._cnet = None

    @staticmethod
    def _rows(d: DeviceArray, n: int, idx: np.ndarray) -> np.ndarray:
        rows = len(idx)
        di = DeviceArray.from_host(idx.astype(np.int64), np.int64)
        out = DeviceArray((n, rows), np.float32)
        _chk(_L().rrtmgp_hip_mlp_gather_rows(_i(rows), C.c_int(n), di.ptr, d.ptr, out.ptr, None))
        return out.to_host().T

    # -- training ------------------------------------------------------------------------------
    def _workspace(self, B):
        if self._ws is not None and self._ws["B"] == B:
            return self._ws
        H, no = HIDDEN, self.n_out
        f32 = lambda *s: DeviceArray(s, np.float32)                         # noqa: E731
        b16 = lambda *s: DeviceArray(s, np.uint16)                          # noqa: E731
        ws = dict(B=B, idx=DeviceArray((B,), np.int64), xr=f32(B * N_INP), yr=f32(B * no), X16=b16(B * N_INP),
                  Yn=f32(B * no), Z1=f32(B * H), H1=b16(B * H), Z2=f32(B * H), H2=b16(B * H), Yp=f32(B * no),
                  dY16=b16(B * no), dYf=f32(B * no), loss=f32((B * no + 255) // 256), dH=f32(B * H),
                  dZ16=b16(B * H), dZf=f32(B * H))
        self._ws = ws
        return ws

    def _gemm(self, M, N, K, A, sAm, sAk, B, sBk, sBn, Cm, sCm, sCn, acc=0):
        _chk(_L().rrtmgp_hip_mlp_gemm_bf16(_i(M), _i(N), _i(K), A.ptr, _i(sAm), _i(sAk), B.ptr, _i(sBk), _i(sBn),
                                           Cm.ptr, _i(sCm), _i(sCn), C.c_int(acc), None))

    def forward_backward(self, ts: TrainingSet, idx: np.ndarray) -> DeviceArray:
        """One forward/backward pass over rows idx; leaves gradients in the parameters' g buffers and
        returns the per-block loss partial sums (sum of squared normalised errors)."""
        L = _L()
        B, H, no = len(idx), HIDDEN, self.n_out
        ws = self._workspace(B)
        st = self.stats
        _chk(L.rrtmgp_hip_memcpy_h2d(ws["idx"].ptr, np.ascontiguousarray(idx, np.int64).ctypes.data_as(C.c_void_p),
                                     C.c_size_t(8 * B), None))
        _chk(L.rrtmgp_hip_mlp_gather_rows(_i(B), C.c_int(N_INP), ws["idx"].ptr, ts.X.ptr, ws["xr"].ptr, None))
        _chk(L.rrtmgp_hip_mlp_normalize_rows(_i(B), C.c_int(N_INP), C.c_int(N_IN), ws["xr"].ptr, st["x_mean"].ptr,
                                             st["x_std"].ptr, ws["X16"].ptr, None, None))
        _chk(L.rrtmgp_hip_mlp_gather_rows(_i(B), C.c_int(no), ws["idx"].ptr, ts.Y.ptr, ws["yr"].ptr, None))
        _chk(L.rrtmgp_hip_mlp_normalize_rows(_i(B), C.c_int(no), C.c_int(no), ws["yr"].ptr, st["y_mean"].ptr,
                                             st["y_std"].ptr, None, ws["Yn"].ptr, None))
        # Forward: Z1 = X W1^T, H1 = softplus(Z1 + b1), ..., Y = H2 W3^T + b3.
        self._gemm(B, H, N_INP, ws["X16"], N_INP, 1, self.W1.b16, 1, N_INP, ws["Z1"], H, 1)
        _chk(L.rrtmgp_hip_mlp_bias_softplus(_i(B), C.c_int(H), ws["Z1"].ptr, self.b1.p.ptr, ws["H1"].ptr, None))
        self._gemm(B, H, H, ws["H1"], H, 1, self.W2.b16, 1, H, ws["Z2"], H, 1)
        _chk(L.rrtmgp_hip_mlp_bias_softplus(_i(B), C.c_int(H), ws["Z2"].ptr, self.b2.p.ptr, ws["H2"].ptr, None))
        self._gemm(B, no, H, ws["H2"], H, 1, self.W3.b16, 1, H, ws["Yp"], no, 1)
        _chk(L.rrtmgp_hip_mlp_mse_grad(_i(B), C.c_int(no), ws["Yp"].ptr, self.b3.p.ptr, ws["Yn"].ptr,
                                       ws["dY16"].ptr, ws["dYf"].ptr, ws["loss"].ptr, None))
        # Backward.
        self._gemm(no, H, B, ws["dY16"], 1, no, ws["H2"], H, 1, self.W3.g, H, 1)          # dW3 = dY^T H2
        _chk(L.rrtmgp_hip_mlp_bias_grad(_i(B), C.c_int(no), ws["dYf"].ptr, self.b3.g.ptr, None))
        self._gemm(B, H, no, ws["dY16"], no, 1, self.W3.b16, H, 1, ws["dH"], H, 1)        # dH2 = dY W3
        _chk(L.rrtmgp_hip_mlp_softplus_back(_i(B * H), ws["dH"].ptr, ws["Z2"].ptr, ws["dZ16"].ptr, ws["dZf"].ptr,
                                            None))
        self._gemm(H, H, B, ws["dZ16"], 1, H, ws["H1"], H, 1, self.W2.g, H, 1)            # dW2 = dZ2^T H1
        _chk(L.rrtmgp_hip_mlp_bias_grad(_i(B), C.c_int(H), ws["dZf"].ptr, self.b2.g.ptr, None))
        self._gemm(B, H, H, ws["dZ16"], H, 1, self.W2.b16, H, 1, ws["dH"], H, 1)          # dH1 = dZ2 W2
        _chk(L.rrtmgp_hip_mlp_softplus_back(_i(B * H), ws["dH"].ptr, ws["Z1"].ptr, ws["dZ16"].ptr, ws["dZf"].ptr,
                                            None))
        self._gemm(H, N_INP, B, ws["dZ16"], 1, H, ws["X16"], N_INP, 1, self.W1.g, N_INP, 1)  # dW1 = dZ1^T X
        _chk(L.rrtmgp_hip_mlp_bias_grad(_i(B), C.c_int(H), ws["dZf"].ptr, self.b1.g.ptr, None))
        return ws["loss"]

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        for p in self.params:
            _chk(_L().rrtmgp_hip_mlp_adam(_i(p.n), p.p.ptr, p.g.ptr, p.m.ptr, p.v.ptr,
                                          p.b16.ptr if p.b16 is not None else None, _f(lr), _f(beta1), _f(beta2),
                                          _f(eps), _i(self.step), None))
        self._cnet = None

    def train(self, ts: TrainingSet, steps=60000, batch=4096, lr=1e-3, lr_final=1e-5, decay_from=0.5, seed=2,
              log_every=0):
        """Adam over random mini-batches: learning rate lr for the first decay_from of the steps, then
        a geometric decay to lr_final.  Returns the loss history (mean squared normalised error) at
        the logged steps."""
        if self.stats is None:
            self.fit_normalisation(ts)
        rng = np.random.Generator(np.random.PCG64(seed))
        hist = []
        s0 = int(decay_from * steps)
        for s in range(steps):
            idx = rng.integers(0, ts.rows, batch)
            loss = self.forward_backward(ts, idx)
            rate = lr if s < s0 else lr * (lr_final / lr) ** ((s - s0) / max(1, steps - 1 - s0))
            self.adam_step(rate)
            if log_every and (s % log_every == 0 or s == steps - 1):
                hist.append((s, float(loss.to_host().sum()) / (batch * self.n_out)))
        synchronize()
        return hist

    def loss(self, ts: TrainingSet, idx: np.ndarray) -> float:
        """Mean squared normalised error on rows idx (gradients are computed and discarded)."""
        l = self.forward_backward(ts, idx)
        return float(l.to_host().sum()) / (len(idx) * self.n_out)

    # -- inference -----------------------------------------------------------------------------
    def c_net(self) -> MlpNetC:
        if self.stats is None:
            raise RuntimeError("the surrogate has no normalisation (train it or load weights first)")
        if getattr(self, "_cnet", None) is None:
            st = self.stats
            self._cnet = MlpNetC(self.W1.b16.ptr.value, self.W2.b16.ptr.value, self.W3.b16.ptr.value,
                                 self.b1.p.ptr.value, self.b2.p.ptr.value, self.b3.p.ptr.value,
                                 st["x_mean"].ptr.value, st["x_std"].ptr.value, st["x_lo"].ptr.value,
                                 st["x_hi"].ptr.value, st["y_mean"].ptr.value, st["y_std"].ptr.value, self.n_out,
                                 self.p_trop)
        return self._cnet

    def host_weights(self) -> dict:
        w = {k: getattr(self, k).host() for k in ("W1", "b1", "W2", "b2", "W3", "b3")}
        w.update(self.host_stats)
        w["p_trop"] = np.float64(self.p_trop)
        return w

    def save(self, path):
        np.savez(path, **self.host_weights())

    @classmethod
    def load(cls, kd: api.KDist, path):
        d = np.load(path)
        net = cls(kd)
        for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
            p = getattr(net, k)
            a = np.asarray(d[k], np.float32)
            if a.shape != p.shape:
                raise ValueError(f"{k}: shape {a.shape}, expected {p.shape}")
            setattr(net, k, _Param(a, p.b16 is not None))
        net.params = [net.W1, net.b1, net.W2, net.b2, net.W3, net.b3]
        net.set_normalisation(*(d[k] for k in ("x_mean", "x_std", "x_lo", "x_hi", "y_mean", "y_std")))
        net.p_trop = float(d["p_trop"])
        return net
