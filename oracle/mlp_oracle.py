"""CPU oracle of the surrogate MLP -- TEST INFRASTRUCTURE ONLY.

Forward and backward of the n_in -> hidden^L (GELU) -> n_out MLP with mean
squared error, in float64 numpy. With bf16=True it rounds exactly the tensors
the HIP kernels round to bfloat16 (inputs, weights, hidden activations and the
output-side gradients fed to the matrix cores), so the GPU's FP32-accumulated
results must agree to accumulation-order tolerance. With bf16=False it is the
plain float64 network whose gradient is pinned by central finite differences.
"""
from __future__ import annotations

import math

import numpy as np

_erf = np.vectorize(math.erf)


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    out = (r.astype(np.uint32) << 16).view(np.float32)
    nan = np.isnan(a)
    out = np.where(nan, a, out)
    return out.astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    cdf = 0.5 * (1.0 + _erf(x / math.sqrt(2.0)))
    return cdf + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def shapes(n_in, hidden, layers, n_out):
    dims = [n_in] + [hidden] * layers + [n_out]
    return [(dims[i], dims[i + 1]) for i in range(layers + 1)]


def unflatten(params, n_in, hidden, layers, n_out):
    out, off = [], 0
    for k, n in shapes(n_in, hidden, layers, n_out):
        W = np.asarray(params[off:off + k * n], np.float64).reshape(k, n)
        off += k * n
        b = np.asarray(params[off:off + n], np.float64)
        off += n
        out.append((W, b))
    return out


def init_params(n_in, hidden, layers, n_out, seed=0):
    """He-uniform weights, zero biases (flat float32)."""
    g = np.random.default_rng(seed)
    flat = []
    for k, n in shapes(n_in, hidden, layers, n_out):
        bound = math.sqrt(6.0 / k)
        flat.append(g.uniform(-bound, bound, k * n))
        flat.append(np.zeros(n))
    return np.concatenate(flat).astype(np.float32)


def forward(params, x, dims, bf16_ops=True):
    layers = unflatten(params, *dims)
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    a = r(x)
    acts, zs = [a], []
    for li, (W, b) in enumerate(layers):
        z = a @ r(W) + b
        zs.append(z)
        if li < len(layers) - 1:
            a = r(gelu(z))
            acts.append(a)
    return zs[-1], acts, zs


def loss_and_grad(params, x, y, dims, bf16_ops=True):
    layers = unflatten(params, *dims)
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    out, acts, zs = forward(params, x, dims, bf16_ops)
    B, n_out = y.shape
    err = out - y
    loss = float(np.mean(err * err))
    dz = 2.0 * err / (B * n_out)
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        dzr = r(dz)
        dW = acts[li].T @ dzr
        db = dz.sum(axis=0)
        grads[li] = (dW, db)
        if li == 0:
            break
        dh = dzr @ r(W).T
        dz = dh * gelu_grad(zs[li - 1])
    flat = np.concatenate([np.concatenate([g[0].reshape(-1), g[1]]) for g in grads])
    return loss, flat


def adam_step(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update in float64 (Kingma & Ba 2015, Algorithm 1): biased moments
    m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2, bias corrections 1 - b^t, then
    w - lr * m_hat / (sqrt(v_hat) + eps). Returns (w, m, v)."""
    w, g, m, v = (np.asarray(a, np.float64) for a in (w, g, m, v))
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    m_hat = m / (1.0 - b1 ** t)
    v_hat = v / (1.0 - b2 ** t)
    return w - lr * m_hat / (np.sqrt(v_hat) + eps), m, v
