"""The MFMA MLP kernels against the CPU oracle (forward, loss, gradient, training)."""
import numpy as np
import pytest

import mlp_oracle as MO
from mind_amd import surrogate as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,batch", [((52, 256, 3, 2), 16), ((7, 40, 2, 3), 37), ((52, 256, 3, 2), 300)])
def test_forward_and_gradient_match_oracle(gpu, dims, batch):
    print(f"case: dims={dims} batch={batch}", flush=True)
    g = np.random.default_rng(batch)
    p = MO.init_params(*dims, seed=5)
    p = (p + 0.02 * g.standard_normal(p.size)).astype(np.float32)
    x = g.standard_normal((batch, dims[0])).astype(np.float32)
    y = g.standard_normal((batch, dims[3])).astype(np.float32)
    net = S.MLP(*dims, max_batch=512)
    net.set_params(p)
    assert np.array_equal(net.params(), p)
    out = net.forward(x)
    ref, _, _ = MO.forward(p, x, dims, bf16_ops=True)
    scale = np.max(np.abs(ref))
    # same BF16 operands, FP32 accumulation in another order
    assert np.max(np.abs(out - ref)) <= 2e-3 * scale
    loss, grad = net.grad(x, y)
    rloss, rgrad = MO.loss_and_grad(p, x, y, dims, bf16_ops=True)
    assert abs(loss - rloss) <= 1e-3 * rloss
    # per tensor, relative to its largest entry; a transposed or misplaced
    # MFMA operand breaks this by O(1)
    off = 0
    for k, n in MO.shapes(*dims):
        for size in (k * n, n):
            a, b = grad[off:off + size], rgrad[off:off + size]
            assert np.max(np.abs(a - b)) <= 2e-2 * np.max(np.abs(b)) + 1e-7, (k, n, size)
            off += size
    # and the exact float64 network is within BF16 distance
    _, xgrad = MO.loss_and_grad(p, x, y, dims, bf16_ops=False)
    assert np.corrcoef(grad, xgrad)[0, 1] > 0.99


def test_training_fits_a_smooth_map(gpu):
    g = np.random.default_rng(7)
    dims = (8, 64, 2, 2)
    x = g.uniform(-1, 1, (8192, 8)).astype(np.float32)
    y = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2], np.tanh(x[:, 3] - x[:, 4])], axis=1).astype(np.float32)
    net = S.MLP(*dims, max_batch=256)
    net.set_params(MO.init_params(*dims, seed=1))
    losses = net.train(x, y, batch=256, epochs=40, lr=3e-3, seed=3)
    print("epoch losses", losses[[0, 9, 19, 39]], flush=True)
    assert losses[-1] < 0.05 * losses[0]
    # the trained weights reproduce the reported loss on the oracle
    p = net.params()
    ref, _, _ = MO.forward(p, x, dims, bf16_ops=True)
    assert np.mean((ref - y) ** 2) < 2 * losses[-1] + 1e-3
    # and beat the best linear fit of the same data
    coef = S.fit_linear(x, y)
    assert np.mean((ref - y) ** 2) < 0.5 * np.mean((S.predict_linear(coef, x) - y) ** 2)


def test_adam_steps_match_oracle(gpu):
    """Two training steps through mind_mlp_train (k_adam on FP32 master weights)
    against the float64 Adam oracle applied to the GPU's own gradients. The batch
    is one row repeated, so the shuffled minibatch has the gradient that
    mind_mlp_grad returns for it. A missing bias correction or swapped betas moves
    weights by O(lr); FP32 rounding stays below 1e-7."""
    dims, B, lr = (52, 256, 3, 2), 16, 1e-3
    print(f"case: adam dims={dims} batch={B}", flush=True)
    g = np.random.default_rng(4)
    p0 = MO.init_params(*dims, seed=9)
    p0 = (p0 + 0.01 * g.standard_normal(p0.size)).astype(np.float32)
    net = S.MLP(*dims, max_batch=64)
    net.set_params(p0)
    w, m, v = p0.astype(np.float64), 0.0, 0.0
    for t in (1, 2):
        x = np.repeat(g.standard_normal((1, dims[0])).astype(np.float32), B, axis=0)
        y = np.repeat(g.standard_normal((1, dims[3])).astype(np.float32), B, axis=0)
        _, grad = net.grad(x, y)
        assert np.count_nonzero(grad) > 0.5 * grad.size
        net.train(x, y, batch=B, epochs=1, lr=lr, seed=t)
        got = net.params().astype(np.float64)
        w, m, v = MO.adam_step(w, grad, m, v, t, lr)
        np.testing.assert_allclose(got, w, rtol=0, atol=1e-7)
        w = got  # continue from the device's FP32 weights
    assert np.max(np.abs(got - p0)) > 1.5 * lr
