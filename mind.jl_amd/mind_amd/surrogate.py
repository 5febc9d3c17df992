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
from dataclasses import dataclass

import numpy as np

from . import _abi

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
        L.mind_mlp_grad.argtypes = [C.c_void_p, P(C.c_float), P(C.c_float), C.c_int64,
                                    P(C.c_float), P(C.c_float)]
        L.mind_mlp_train.argtypes = [C.c_void_p, P(C.c_float), P(C.c_float), C.c_int64, C.c_int64,
                                     C.c_int32, C.c_float, C.c_uint64, P(C.c_float)]
        L._mlp_ready = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


class MLP:
    """Handle on libmind's MFMA MLP."""

    def __init__(self, n_in=N_IN, hidden=HIDDEN, layers=LAYERS, n_out=N_OUT, max_batch=1024):
        self.dims = (n_in, hidden, layers, n_out)
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


def windows(drive, rate, volt, v_rest=-65.0):
    """(X, Y): X = [drive bins t-49..t, rate t-1, V t-1], Y = [rate t, V t] for every bin t.
    Bins before the first are padded with the first drive bin, rate 0 and rest voltage,
    which is the state a population starts from."""
    R, T = drive.shape
    pad = np.repeat(drive[:, :1], HISTORY - 1, axis=1)
    dpad = np.concatenate([pad, drive], axis=1)
    prev_r = np.concatenate([np.zeros((R, 1)), rate[:, :-1]], axis=1)
    prev_v = np.concatenate([np.full((R, 1), v_rest), volt[:, :-1]], axis=1)
    X = np.zeros((R, T, N_IN))
    for t in range(T):
        X[:, t, :HISTORY] = dpad[:, t:t + HISTORY]
    X[:, :, HISTORY] = prev_r
    X[:, :, HISTORY + 1] = prev_v
    Y = np.stack([rate, volt], axis=2)
    return X.reshape(-1, N_IN), Y.reshape(-1, N_OUT)


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


def save(path, params, norm: Normalizer, meta: dict):
    np.savez(path, params=params, x_mean=norm.x_mean, x_std=norm.x_std, y_mean=norm.y_mean,
             y_std=norm.y_std, x_lo=norm.x_lo, x_hi=norm.x_hi,
             meta=np.array(repr(sorted(meta.items()))))


def load(path=WEIGHTS):
    z = np.load(path, allow_pickle=False)
    norm = Normalizer(z["x_mean"], z["x_std"], z["y_mean"], z["y_std"], z["x_lo"], z["x_hi"])
    return z["params"], norm, str(z["meta"])


def fit_linear(X, Y):
    """Least-squares linear map on the same normalised inputs: the method without
    the learned nonlinearity, used to show what the MLP gains."""
    A = np.concatenate([X, np.ones((X.shape[0], 1))], axis=1)
    coef, *_ = np.linalg.lstsq(A, Y, rcond=None)
    return coef


def predict_linear(coef, X):
    return np.concatenate([X, np.ones((X.shape[0], 1))], axis=1) @ coef
