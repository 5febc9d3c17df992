"""Learned stand-in for a micro population inside micro-macro coupling.

An MLP maps a 50-bin (1 ms bins) history of the drive a population receives
from its ROI (rate_scale * E, Hz) plus its previous outputs to the next bin's
population rate (Hz) and mean soma voltage (mV). Training windows come from
plain micro-macro runs; the network is trained on the GPU by libmind's MFMA
forward / backward / Adam kernels. At run time libmind integrates the ROI graph
with the network in place of every micro population (mind_run_macro_surrogate)
and counts windows whose inputs leave the training range; a run with any such
window is redone with the plain micro simulation, which stays the reference.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from . import macro as M

HISTORY = 50
N_IN = HISTORY + 2
N_OUT = 2
HIDDEN = 256
LAYERS = 3
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "surrogate_weights.npz")


def _lib():
    L = _abi.lib()
    if not getattr(L, "_mlp_ready", False):
        P = C.POINTER
        L.mind_mlp_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int64,
                                      P(C.c_void_p)]
        L.mind_mlp_destroy.argtypes = [C.c_void_p]
        L.mind_mlp_destroy.restype = None
        L.mind_mlp_param_count.argtypes = [C.c_void_p, P(C.c_int64)]
        L.mind_mlp_set_params.argtypes = [C.c_void_p, P(C.c_float), C.c_int64]
        L.mind_mlp_get_params.argtypes = [C.c_void_p, P(C.c_float), C.c_int64]
        L.mind_mlp_forward.argtypes = [C.c_void_p, P(C.c_float), C.c_int64, P(C.c_float)]
        L.mind_mlp_forward_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        L.mind_mlp_grad.argtypes = [C.c_void_p, P(C.c_float), P(C.c_float), C.c_int64,
                                    P(C.c_float), P(C.c_float)]
        L.mind_mlp_train.argtypes = [C.c_void_p, P(C.c_float), P(C.c_float), C.c_int64, C.c_int64,
                                     C.c_int32, C.c_float, C.c_uint64, P(C.c_float)]
        L.mind_build_surrogate.argtypes = [P(M.MacroDesc), P(SurrogateDescC), C.c_int32, P(C.c_void_p)]
        L.mind_initialize_surrogate.argtypes = [C.c_void_p]
        L.mind_run_surrogate.argtypes = [C.c_void_p, C.c_double, P(M.MacroResultC), P(C.c_int64),
                                         P(C.c_double)]
        L.mind_surrogate_destroy.argtypes = [C.c_void_p]
        L.mind_surrogate_destroy.restype = None
        L._mlp_ready = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


class SurrogateDescC(C.Structure):
    _fields_ = [
        ("n_in", C.c_int32), ("hidden", C.c_int32), ("layers", C.c_int32), ("n_out", C.c_int32),
        ("history", C.c_int32), ("params", C.POINTER(C.c_float)), ("n_params", C.c_int64),
        ("x_mean", C.POINTER(C.c_float)), ("x_std", C.POINTER(C.c_float)),
        ("y_mean", C.POINTER(C.c_float)), ("y_std", C.POINTER(C.c_float)),
        ("x_lo", C.POINTER(C.c_float)), ("x_hi", C.POINTER(C.c_float)),
        ("skip_w", C.POINTER(C.c_float)), ("skip_b", C.POINTER(C.c_float)), ("v_rest", C.c_double),
    ]


class MLP:
    """Handle on libmind's MFMA MLP."""

    def __init__(self, n_in=N_IN, hidden=HIDDEN, layers=LAYERS, n_out=N_OUT, max_batch=1024):
        self.dims = (n_in, hidden, layers, n_out)
        self.max_batch = int(max_batch)
        self.h = C.c_void_p()
        _abi.check(_lib().mind_mlp_create(n_in, hidden, layers, n_out, max_batch, C.byref(self.h)))
        n = C.c_int64()
        _lib().mind_mlp_param_count(self.h, C.byref(n))
        self.nparam = n.value

    def __del__(self):
        if getattr(self, "h", None):
            _lib().mind_mlp_destroy(self.h)
            self.h = None

    def set_params(self, p):
        p = _f32(p)
        _abi.check(_lib().mind_mlp_set_params(self.h, _abi.as_ptr(p, C.c_float), p.size))

    def params(self):
        out = np.zeros(self.nparam, np.float32)
        _abi.check(_lib().mind_mlp_get_params(self.h, _abi.as_ptr(out, C.c_float), out.size))
        return out

    def forward(self, x):
        x = _f32(x)
        y = np.zeros((x.shape[0], self.dims[3]), np.float32)
        _abi.check(_lib().mind_mlp_forward(self.h, _abi.as_ptr(x, C.c_float), x.shape[0],
                                           _abi.as_ptr(y, C.c_float)))
        return y

    def grad(self, x, y):
        x, y = _f32(x), _f32(y)
        loss = C.c_float()
        g = np.zeros(self.nparam, np.float32)
        _abi.check(_lib().mind_mlp_grad(self.h, _abi.as_ptr(x, C.c_float), _abi.as_ptr(y, C.c_float),
                                        x.shape[0], C.byref(loss), _abi.as_ptr(g, C.c_float)))
        return loss.value, g

    def train(self, X, Y, batch=512, epochs=10, lr=1e-3, seed=0):
        X, Y = _f32(X), _f32(Y)
        losses = np.zeros(epochs, np.float32)
        _abi.check(_lib().mind_mlp_train(self.h, _abi.as_ptr(X, C.c_float), _abi.as_ptr(Y, C.c_float),
                                         X.shape[0], batch, epochs, float(lr), int(seed),
                                         _abi.as_ptr(losses, C.c_float)))
        return losses


# --- training windows --------------------------------------------------------
def bins_from_run(res, rate_scale, bin_steps=10):
    """Per-ROI 1 ms bins of a micro-macro result: drive (mean over the bin of
    rate_scale * E at each macro step's start), population rate and mean soma
    voltage at the bin's end."""
    E = res.E[:, :-1]  # boundaries 0..S-1 start macro steps 0..S-1
    S = E.shape[1] // bin_steps * bin_steps
    drive = (rate_scale[:, None] * E[:, :S]).reshape(E.shape[0], -1, bin_steps).mean(axis=2)
    idx = np.arange(bin_steps, S + 1, bin_steps)
    return drive, res.rate[:, idx], res.mean_v[:, idx]


def windows(drive, rate, volt, v_rest=-65.0, prev=None):
    """(X, Y): X = [drive bins t-49..t, rate t-1, V t-1], Y = [rate t, V t] for every bin t.
    Bins before the first are padded with the first drive bin, rate 0 and rest voltage,
    which is the state a population starts from. prev = (rates, voltages) replaces the
    observed previous outputs (e.g. by the network's own closed-loop outputs)."""
    R, T = drive.shape
    if prev is None:
        prev_r = np.concatenate([np.zeros((R, 1)), rate[:, :-1]], axis=1)
        prev_v = np.concatenate([np.full((R, 1), v_rest), volt[:, :-1]], axis=1)
    else:
        prev_r, prev_v = prev
    X = np.zeros((R, T, N_IN))
    X[:, :, :HISTORY] = history_windows(drive)
    X[:, :, HISTORY] = prev_r
    X[:, :, HISTORY + 1] = prev_v
    Y = np.stack([rate, volt], axis=2)
    return X.reshape(-1, N_IN), Y.reshape(-1, N_OUT)


def history_windows(drive):
    """[R, T, HISTORY]: the drive bins t-49..t of every bin t, first bin repeated before the start."""
    R, T = drive.shape
    dpad = np.concatenate([np.repeat(drive[:, :1], HISTORY - 1, axis=1), drive], axis=1)
    idx = np.arange(T)[:, None] + np.arange(HISTORY)[None, :]
    return dpad[:, idx]


def rollout(predict, drive, v_rest=-65.0):
    """Closed loop over given drive sequences: each bin's previous outputs are the
    predictor's own outputs of the bin before (rate clamped at 0), as in the
    coupled run. Returns (previous rates, previous voltages, rates, voltages), [R, T] each."""
    R, T = drive.shape
    hist = history_windows(drive)
    rate, volt = np.zeros(R), np.full(R, v_rest)
    out = np.zeros((4, R, T))
    for t in range(T):
        out[0, :, t], out[1, :, t] = rate, volt
        y = predict(np.concatenate([hist[:, t], rate[:, None], volt[:, None]], axis=1))
        rate, volt = np.maximum(y[:, 0], 0.0), y[:, 1]
        out[2, :, t], out[3, :, t] = rate, volt
    return out


@dataclass
class Normalizer:
    x_mean: np.ndarray
    x_std: np.ndarray
    y_mean: np.ndarray
    y_std: np.ndarray
    x_lo: np.ndarray
    x_hi: np.ndarray

    @classmethod
    def fit(cls, X, Y, margin=0.05):
        xm, xs = X.mean(axis=0), X.std(axis=0) + 1e-6
        # the drive history shares one scale so the network sees one unit
        xm[:HISTORY], xs[:HISTORY] = X[:, :HISTORY].mean(), X[:, :HISTORY].std() + 1e-6
        lo, hi = X.min(axis=0), X.max(axis=0)
        span = hi - lo
        return cls(xm, xs, Y.mean(axis=0), Y.std(axis=0) + 1e-6, lo - margin * span, hi + margin * span)

    def x(self, X):
        return (X - self.x_mean) / self.x_std

    def y(self, Y):
        return (Y - self.y_mean) / self.y_std

    def y_inv(self, Yn):
        return Yn * self.y_std + self.y_mean

    def ood(self, X):
        return np.any((X < self.x_lo) | (X > self.x_hi), axis=1)


@dataclass
class Network:
    """A trained surrogate: flat MLP parameters (mind_mlp order), input / output
    normalisation and training range, and the least-squares linear map fitted on
    the same normalised windows ([n_in + 1, n_out], last row the intercept). With
    residual=True the network's output is added to that linear map (a linear skip
    path), so the MLP learns what the linear map misses; the linear map alone is the
    method without the network."""
    params: np.ndarray
    norm: Normalizer
    dims: tuple = (N_IN, HIDDEN, LAYERS, N_OUT)
    v_rest: float = -65.0
    linear: np.ndarray = None
    meta: dict = field(default_factory=dict)
    residual: bool = False

    @property
    def history(self):
        return self.dims[0] - 2

    def save(self, path):
        n = self.norm
        np.savez(path, params=_f32(self.params), dims=np.asarray(self.dims, np.int64), x_mean=n.x_mean,
                 x_std=n.x_std, y_mean=n.y_mean, y_std=n.y_std, x_lo=n.x_lo, x_hi=n.x_hi,
                 v_rest=np.float64(self.v_rest), residual=np.bool_(self.residual),
                 linear=self.linear if self.linear is not None else np.zeros(0),
                 meta=np.array(repr(sorted(self.meta.items()))))

    @classmethod
    def load(cls, path=WEIGHTS):
        z = np.load(path, allow_pickle=False)
        norm = Normalizer(z["x_mean"], z["x_std"], z["y_mean"], z["y_std"], z["x_lo"], z["x_hi"])
        lin = z["linear"] if z["linear"].size else None
        residual = bool(z["residual"]) if "residual" in z.files else False
        return cls(z["params"], norm, tuple(int(x) for x in z["dims"]), float(z["v_rest"]), lin,
                   {"text": str(z["meta"])}, residual)

    def predict(self, X, mlp=None):
        """One-step predictions of raw windows on the device (mind_mlp_forward)."""
        if mlp is None:
            mlp = MLP(*self.dims, max_batch=1024)
            mlp.set_params(self.params)
        Xn = self.norm.x(np.asarray(X, np.float64))
        B = mlp.max_batch
        out = np.concatenate([mlp.forward(Xn[i:i + B]) for i in range(0, len(Xn), B)]).astype(np.float64)
        if self.residual:
            out = out + predict_linear(self.linear, Xn)
        return self.norm.y_inv(out)

    def predict_linear(self, X):
        return self.norm.y_inv(predict_linear(self.linear, self.norm.x(np.asarray(X, np.float64))))


@dataclass
class SurrogateRun:
    result: "M.MacroResult"
    ood_windows: int
    first_ood_time: float


class CoupledSurrogate:
    """The ROI graph of a macro model driven by a trained network instead of its
    micro populations (mind_build_surrogate / mind_run_surrogate)."""

    def __init__(self, model: "M.MacroModel", net: Network, dt_macro=0.1, device=0):
        self.model, self.net = model, net
        self.n_roi = len(model.connectivity.labels)
        mdesc, self._mkeep = M.macro_descriptor(model, dt_macro)
        n = net.norm
        self._keep = [_f32(net.params), _f32(n.x_mean), _f32(n.x_std), _f32(n.y_mean), _f32(n.y_std),
                      _f32(n.x_lo), _f32(n.x_hi)]
        if net.residual:
            self._keep += [_f32(net.linear[:-1]), _f32(net.linear[-1])]
        d = SurrogateDescC()
        d.n_in, d.hidden, d.layers, d.n_out = (int(x) for x in net.dims)
        d.history = net.history
        d.params = _abi.as_ptr(self._keep[0], C.c_float)
        d.n_params = self._keep[0].size
        for name, arr in zip(("x_mean", "x_std", "y_mean", "y_std", "x_lo", "x_hi", "skip_w", "skip_b"),
                             self._keep[1:]):
            setattr(d, name, _abi.as_ptr(arr, C.c_float))
        d.v_rest = float(net.v_rest)
        self.h = C.c_void_p()
        _abi.check(_lib().mind_build_surrogate(C.byref(mdesc), C.byref(d), int(device), C.byref(self.h)))

    def initialize(self):
        _abi.check(_lib().mind_initialize_surrogate(self.h))
        return self

    def run(self, stop) -> SurrogateRun:
        res = M.MacroResultC()
        ood = C.c_int64()
        first = C.c_double()
        _abi.check(_lib().mind_run_surrogate(self.h, float(stop), C.byref(res), C.byref(ood), C.byref(first)))
        try:
            out = M.result_from_c(res, self.n_roi)
        finally:
            _lib().mind_macro_result_free(C.byref(res))
        return SurrogateRun(out, int(ood.value), float(first.value))

    def close(self):
        if getattr(self, "h", None):
            _lib().mind_surrogate_destroy(self.h)
            self.h = None

    __del__ = close


@dataclass
class CoupledResult:
    result: "M.MacroResult"
    source: str             # "surrogate" or "micro" (the plain fallback ran)
    ood_windows: int
    first_ood_time: float
    surrogate: "M.MacroResult"


def run_coupled(model: "M.MacroModel", stop, net: Network = None, fallback=True, dt_micro=0.025,
                dt_macro=0.1, seed=1, method="batched") -> CoupledResult:
    """Micro-macro coupling through the surrogate; when any bin's inputs leave the
    training range the plain micro-macro simulation is run instead and returned."""
    net = net if net is not None else Network.load()
    cs = CoupledSurrogate(model, net, dt_macro)
    try:
        out = cs.initialize().run(stop)
    finally:
        cs.close()
    if out.ood_windows and fallback:
        M.build(model, dt_micro=dt_micro, dt_macro=dt_macro, seed=seed, method=method)
        M.initialize(model)
        res = M.run(model, stop)
        return CoupledResult(res, "micro", out.ood_windows, out.first_ood_time, out.result)
    return CoupledResult(out.result, "surrogate", out.ood_windows, out.first_ood_time, out.result)


def fit_linear(X, Y):
    """Least-squares linear map on the same normalised inputs: the method without
    the learned nonlinearity, used to show what the MLP gains."""
    A = np.concatenate([X, np.ones((X.shape[0], 1))], axis=1)
    coef, *_ = np.linalg.lstsq(A, Y, rcond=None)
    return coef


def predict_linear(coef, X):
    return np.concatenate([X, np.ones((X.shape[0], 1))], axis=1) @ coef


def linear_only(net: Network) -> Network:
    """The same trained surrogate with its MLP switched off (output layer zero), so
    that only the linear skip map drives the ROIs: the method without the network,
    for closed-loop comparisons."""
    if not net.residual or net.linear is None:
        raise ValueError("the network has no linear skip map")
    p = np.array(net.params, np.float32, copy=True)
    n_in, hidden, layers, n_out = net.dims
    last = p.size - (hidden * n_out + n_out)  # W then b of the output layer, last in the flat order
    p[last:] = 0.0
    return Network(p, net.norm, net.dims, net.v_rest, net.linear, dict(net.meta), net.residual)
