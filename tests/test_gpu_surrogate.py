"""The persistent surrogate coupling kernel against the CPU oracle.

Random (untrained) networks: the recorded trajectory is replayed on the host --
delayed coupling and Wilson-Cowan RK4 must agree bit for bit with the oracle's
macro arithmetic, and every bin's MLP output with the BF16 oracle forward of the
inputs the kernel saw.
"""
import os

import numpy as np
import pytest

import mlp_oracle as MO
import oracle
from mind_amd import configs, destroy
from mind_amd import macro as M
from mind_amd import surrogate as S

pytestmark = pytest.mark.gpu

DIMS = (52, 256, 3, 2)


def random_net(seed=0, x_hi=1e9, skip=False):
    p = MO.init_params(*DIMS, seed=seed)
    xm = np.full(DIMS[0], 40.0)
    xs = np.full(DIMS[0], 25.0)
    xm[-2:], xs[-2:] = (20.0, -60.0), (10.0, 5.0)
    hi = np.full(DIMS[0], x_hi)
    norm = S.Normalizer(xm.astype(np.float32), xs.astype(np.float32), np.array([20.0, -60.0], np.float32),
                        np.array([10.0, 5.0], np.float32), np.full(DIMS[0], -1e9, np.float32),
                        hi.astype(np.float32))
    if not skip:
        return S.Network(p, norm, DIMS)
    lin = np.random.default_rng(seed + 100).normal(0.0, 0.3, (DIMS[0] + 1, 2))
    return S.Network(p, norm, DIMS, linear=lin, residual=True)


def replay(mm, net, res, bin_steps=10):
    n = len(mm.connectivity.labels)
    W, D = mm.connectivity.weights, mm.connectivity.delays
    Wp = np.zeros((n, n))
    for (t, s) in mm.projections:
        Wp[t, s] = W[t, s]
    ds = np.rint(D / 0.1).astype(np.int64)
    E, I, rate, cpl, V = res.E, res.I, res.rate, res.coupling, res.mean_v
    steps = E.shape[1] - 1
    wc = [[mm.regions[r][k] for k in M.WC_PARAMS] for r in range(n)]
    H = net.history
    xm, xs = net.norm.x_mean.astype(np.float32), net.norm.x_std.astype(np.float32)
    ym, ys = net.norm.y_mean.astype(np.float32), net.norm.y_std.astype(np.float32)
    drive = [[] for _ in range(n)]
    acc = np.zeros(n)
    feats, outs = [], []
    for s in range(steps):
        for i in range(n):
            c = 0.0
            for j in range(n):
                if Wp[i, j] == 0.0:
                    continue
                c = c + Wp[i, j] * E[j, max(s - ds[i, j], 0)]
            assert c == cpl[i, s + 1], ("coupling", i, s)
            acc[i] = acc[i] + wc[i][M.WC_PARAMS.index("rate_scale")] * E[i, s]
        if (s + 1) % bin_steps == 0:
            for i in range(n):
                drive[i].append(acc[i] / bin_steps)
                acc[i] = 0.0
                h = drive[i][-H:]
                h = [h[0]] * (H - len(h)) + h
                feats.append(np.array(h + [rate[i, s], V[i, s]], np.float64).astype(np.float32))
                outs.append((rate[i, s + 1], V[i, s + 1]))
        for i in range(n):
            u = cpl[i, s + 1] + wc[i][M.WC_PARAMS.index("k_micro")] * (rate[i, s + 1] / wc[i][M.WC_PARAMS.index("rate_scale")])
            e, ii = oracle.wc_rk4(wc[i], E[i, s], I[i, s], u, 0.1)
            assert e == E[i, s + 1] and ii == I[i, s + 1], ("rk4", i, s)
    F = np.array(feats)
    Xn = (F - xm) / xs  # float32, as the kernel normalises
    ref, _, _ = MO.forward(net.params, Xn, net.dims, bf16_ops=True)
    if net.residual:  # linear skip on the same normalised inputs
        ref = ref + Xn.astype(np.float64) @ net.linear[:-1] + net.linear[-1]
    ref32 = ref.astype(np.float32)
    y = ref32 * ys + ym
    return F, np.array(outs), y, ref


@pytest.mark.parametrize("n_roi,stop,skip", [(16, 120.0, False), (21, 80.0, True)])
def test_surrogate_kernel_matches_oracle(gpu, n_roi, stop, skip):
    print(f"case: surrogate n_roi={n_roi} stop={stop} skip={skip}", flush=True)
    mm, _ = configs.micro_macro(n_roi=n_roi, cells=1, nseg=4, seed=5)
    net = random_net(seed=n_roi, skip=skip)
    cs = S.CoupledSurrogate(mm, net).initialize()
    out = cs.run(stop)
    assert out.ood_windows == 0 and out.first_ood_time == -1.0
    res = out.result
    assert res.E.shape == (n_roi, int(round(stop / 0.1)) + 1)
    F, got, y, ref = replay(mm, net, res)
    # rate is clamped at zero, V is not
    scale = np.max(np.abs(ref), axis=0)
    gv = (got[:, 1] - net.norm.y_mean[1]) / net.norm.y_std[1]
    assert np.max(np.abs(gv - ref[:, 1])) <= 5e-3 * scale[1] + 1e-4
    pos = got[:, 0] > 0
    assert pos.sum() > 0.2 * len(pos)
    gr = (got[pos, 0] - net.norm.y_mean[0]) / net.norm.y_std[0]
    assert np.max(np.abs(gr - ref[pos, 0])) <= 5e-3 * scale[0] + 1e-4
    assert np.all(y[~pos, 0] <= 1e-2 * net.norm.y_std[0])
    # continuing a run is the same as one longer run
    more = cs.run(stop + 20.0).result
    cs2 = S.CoupledSurrogate(mm, net).initialize()
    whole = cs2.run(stop + 20.0).result
    assert np.array_equal(more.E, whole.E[:, -more.E.shape[1]:])
    assert np.array_equal(more.rate, whole.rate[:, -more.E.shape[1]:])
    cs.close()
    cs2.close()


def test_surrogate_counts_out_of_range_bins(gpu):
    mm, _ = configs.micro_macro(n_roi=5, cells=1, nseg=4, seed=6)
    net = random_net(seed=1, x_hi=0.0)  # every positive drive is out of range
    cs = S.CoupledSurrogate(mm, net).initialize()
    out = cs.run(30.0)
    assert out.ood_windows == 5 * 30
    assert out.first_ood_time == 1.0
    cs.close()
    ok = S.CoupledSurrogate(mm, random_net(seed=1)).initialize().run(30.0)
    assert ok.ood_windows == 0


TRAINED = pytest.mark.skipif(not os.path.exists(S.WEIGHTS),
                             reason="no trained surrogate weights in the package")


@TRAINED
def test_trained_surrogate_against_detailed_micro_macro(gpu):
    """Evaluation seed 1 (training seeds are 1000+): per-node time-averaged
    population rate within 5% of the detailed run, at least 10x its simulated
    s per wall s, and no window outside the training range."""
    net = S.Network.load()
    mm, _ = configs.micro_macro(seed=1)
    stop = 1000.0
    print(f"case: detailed micro-macro seed 1, {stop} ms", flush=True)
    M.build(mm, seed=1, method="batched")
    M.initialize(mm)
    plain = M.run(mm, stop, record_spikes=False)
    print(f"case: surrogate seed 1, {stop} ms", flush=True)
    cs = S.CoupledSurrogate(mm, net).initialize()
    sur = cs.run(stop)
    cs.close()
    assert sur.ood_windows == 0
    rp = plain.rate[:, 10::10].mean(axis=1)
    rs = sur.result.rate[:, 10::10].mean(axis=1)
    err = np.abs(rs - rp) / rp
    vp = plain.mean_v[:, 10::10].mean(axis=1)
    vs = sur.result.mean_v[:, 10::10].mean(axis=1)
    speed = plain.run_seconds / sur.result.run_seconds
    print(f"rate error max {err.max():.4f} mean {err.mean():.4f}; V error max {np.abs(vs - vp).max():.3f} mV; "
          f"speedup {speed:.1f}x", flush=True)
    assert err.max() <= 0.05
    assert np.abs(vs - vp).max() <= 0.5
    assert speed >= 10.0
    # the ROI state the surrogate drives stays close to the detailed run's
    assert np.max(np.abs(sur.result.E.mean(axis=1) - plain.E.mean(axis=1))) <= 0.02
    # what the network adds to the linear map alone (the method without it), one step
    # ahead on this run's own windows
    rs_ = np.array([mm.regions[r]["rate_scale"] for r in range(16)])
    X, Y = S.windows(*S.bins_from_run(plain, rs_))
    e_net = np.sqrt(np.mean((net.predict(X) - Y) ** 2, axis=0))
    e_lin = np.sqrt(np.mean((net.predict_linear(X) - Y) ** 2, axis=0))
    print(f"one-step RMSE rate {e_net[0]:.4f} vs linear {e_lin[0]:.4f}; V {e_net[1]:.4f} vs {e_lin[1]:.4f}",
          flush=True)
    assert e_net[0] < e_lin[0] and e_net[1] <= e_lin[1] * 1.001
    destroy(mm)
    mm.micro.state = None


@TRAINED
def test_out_of_range_input_falls_back_to_detailed_simulation(gpu):
    """A drive far above anything seen in training (external input P raised) is
    flagged; run_coupled then returns the detailed micro-macro run itself."""
    net = S.Network.load()
    mm, _ = configs.micro_macro(n_roi=4, cells=8, nseg=40, seed=2, wc={"P": 6.0, "rate_scale": 400.0})
    print("case: fallback 4 ROIs x 8 cells, 60 ms", flush=True)
    out = S.run_coupled(mm, 60.0, net)
    assert out.source == "micro" and out.ood_windows > 0 and out.first_ood_time > 0
    # the returned result is the detailed simulation, bit for bit
    M.build(mm, seed=1, method="batched")
    M.initialize(mm)
    ref = M.run(mm, 60.0)
    assert np.array_equal(out.result.E, ref.E) and np.array_equal(out.result.rate, ref.rate)
    assert np.array_equal(out.result.spike_times, ref.spike_times)
    # in range, the surrogate result is returned
    mm2, _ = configs.micro_macro(n_roi=4, cells=8, nseg=40, seed=2)
    ok = S.run_coupled(mm2, 60.0, net)
    assert ok.source == "surrogate" and ok.ood_windows == 0
