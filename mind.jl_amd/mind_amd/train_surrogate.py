"""Generate surrogate training windows from plain micro-macro runs and train the
network on the GPU with libmind's MFMA forward / backward / Adam kernels.

    python -m mind_amd.train_surrogate --seeds 1000-1006 --val-seed 1007 --out weights.npz

Every seed builds the 16-ROI micro-macro model of configs.micro_macro (its
connectome, morphologies and Poisson streams all follow the seed), runs it for
--stop ms with the batched micro path and cuts every ROI's run into 1 ms bins;
each bin is one window (50 drive bins + previous rate and voltage -> rate and
voltage). The validation seed is held out of training and normalisation.
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import configs
from . import macro as M
from . import surrogate as S
from ._abi import lib


def seed_bins(seed, stop, method="batched", log=print):
    mm, _ = configs.micro_macro(seed=seed)
    M.build(mm, seed=seed, method=method)
    M.initialize(mm)
    t0 = time.time()
    res = M.run(mm, stop, record_spikes=False)
    lib().mind_state_destroy(mm.state)
    mm.state = None
    rs = np.array([mm.regions[r]["rate_scale"] for r in range(len(mm.regions))])
    drive, rate, volt = S.bins_from_run(res, rs)
    log(f"seed {seed}: {stop:.0f} ms in {time.time() - t0:.1f} s; drive {drive.min():.1f}..{drive.max():.1f} Hz, "
        f"rate mean {rate.mean():.2f} Hz, V {volt.min():.1f}..{volt.max():.1f} mV", flush=True)
    return drive, rate, volt


def parse_seeds(text):
    out = []
    for part in text.split(","):
        if "-" in part:
            a, b = part.split("-")
            out.extend(range(int(a), int(b) + 1))
        else:
            out.append(int(part))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--seeds", default="1000-1006")
    ap.add_argument("--val-seed", type=int, default=1007)
    ap.add_argument("--stop", type=float, default=2000.0)
    ap.add_argument("--epochs", default="20,10")
    ap.add_argument("--lrs", default="1e-3,3e-4")
    ap.add_argument("--closed-loop-epochs", default="4,4")
    ap.add_argument("--closed-loop-lrs", default="3e-4,1e-4")
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--out", default=S.WEIGHTS)
    ap.add_argument("--bins-out", default="")
    ap.add_argument("--bins-in", default="", help="reuse bins written by --bins-out")
    a = ap.parse_args(argv)
    seeds = parse_seeds(a.seeds)
    if a.val_seed in seeds or any(1 <= s <= 16 for s in seeds + [a.val_seed]):
        sys.exit("training and validation seeds must be disjoint from each other and from the evaluation seeds 1-16")
    if a.bins_in:
        z = np.load(a.bins_in, allow_pickle=False)
        data = {s: (z[f"drive_{s}"], z[f"rate_{s}"], z[f"volt_{s}"]) for s in seeds + [a.val_seed]}
    else:
        data = {s: seed_bins(s, a.stop) for s in seeds + [a.val_seed]}
    if a.bins_out:
        np.savez_compressed(a.bins_out, **{f"{k}_{s}": v for s, (d, r, u) in data.items()
                                           for k, v in (("drive", d), ("rate", r), ("volt", u))})
    ints = lambda t: [int(x) for x in t.split(",")]
    floats = lambda t: [float(x) for x in t.split(",")]
    net = fit([data[s] for s in seeds], data[a.val_seed], ints(a.epochs), floats(a.lrs),
              ints(a.closed_loop_epochs), floats(a.closed_loop_lrs), a.batch,
              meta={"seeds": a.seeds, "val_seed": a.val_seed, "stop_ms": a.stop})
    net.save(a.out)
    print(f"saved {a.out}", flush=True)


def closed_loop_error(predict, seq):
    """Per-ROI relative error of the time-averaged rate and absolute error of the
    time-averaged voltage when the predictor runs closed loop on a sequence's drive."""
    d, r, v = seq
    out = S.rollout(predict, d)
    rate_err = np.abs(out[2].mean(axis=1) - r.mean(axis=1)) / r.mean(axis=1)
    v_err = np.abs(out[3].mean(axis=1) - v.mean(axis=1))
    return rate_err, v_err


def fit(train, val, epochs, lrs, cl_epochs=(), cl_lrs=(), batch=512, seed=0, meta=None, log=print):
    """Normalise, fit the least-squares linear map, then train the MLP on what the
    linear map misses (its output layer starts at zero, so training starts from the
    linear solution): first on teacher-forced windows in stages of decreasing
    learning rate, then optionally on the union of those and closed-loop windows
    whose previous outputs are the network's own (the inputs it sees inside the
    coupled run). The closed-loop stage is kept only if it lowers the validation
    seed's closed-loop rate error. Reports one-step and closed-loop validation errors."""
    X, Y = map(np.concatenate, zip(*[S.windows(*seq) for seq in train]))
    Xv, Yv = S.windows(*val)
    log(f"{len(X)} training windows, {len(Xv)} validation windows", flush=True)
    norm = S.Normalizer.fit(X, Y)
    coef = S.fit_linear(norm.x(X), norm.y(Y))
    mlp = S.MLP(S.N_IN, S.HIDDEN, S.LAYERS, S.N_OUT, max_batch=batch)
    p0 = he_init(seed)
    p0[-(S.HIDDEN * S.N_OUT + S.N_OUT):] = 0.0  # last layer: start from the linear map
    mlp.set_params(p0)
    net = S.Network(p0, norm, linear=coef, meta=dict(meta or {}), residual=True)
    predict = lambda Z: net.predict(Z, mlp)
    t0 = time.time()
    losses = []

    def stages(Xt, Yt, eps, rates, tag):
        Xn = norm.x(Xt)
        Rn = norm.y(Yt) - S.predict_linear(coef, Xn)
        for k, (ep, lr) in enumerate(zip(eps, rates)):
            mlp.set_params(mlp.params())  # keep the weights, restart Adam's moments
            losses.extend(mlp.train(Xn, Rn, batch=batch, epochs=ep, lr=lr, seed=seed + len(losses)).tolist())
            log(f"{tag} stage {k}: lr {lr:g}, {ep} epochs, residual loss {losses[-1]:.4f} "
                f"({time.time() - t0:.1f} s)", flush=True)

    stages(X, Y, epochs, lrs, "teacher-forced")
    best, best_loss = mlp.params(), losses[-1]
    best_err, _ = closed_loop_error(predict, val)
    log(f"closed loop after teacher forcing: rate error max {best_err.max():.4f} mean {best_err.mean():.4f}",
        flush=True)
    net.meta["closed_loop_stage_kept"] = False
    if len(cl_epochs):
        D = np.concatenate([seq[0] for seq in train])
        out = S.rollout(predict, D)
        lo = np.cumsum([0] + [len(seq[0]) for seq in train])
        Xc, Yc = map(np.concatenate, zip(*[
            S.windows(*seq, prev=(out[0][lo[i]:lo[i + 1]], out[1][lo[i]:lo[i + 1]]))
            for i, seq in enumerate(train)]))
        stages(np.concatenate([X, Xc]), np.concatenate([Y, Yc]), cl_epochs, cl_lrs, "closed-loop")
        err, _ = closed_loop_error(predict, val)
        log(f"closed loop after closed-loop stage: rate error max {err.max():.4f} mean {err.mean():.4f}",
            flush=True)
        if err.mean() < best_err.mean():
            best, best_err, best_loss = mlp.params(), err, losses[-1]
            net.meta["closed_loop_stage_kept"] = True
    mlp.set_params(best)
    net.params = best
    pv, pl = net.predict(Xv, mlp), net.predict_linear(Xv)
    for j, name in enumerate(("rate", "V")):
        e_mlp = np.sqrt(np.mean((pv[:, j] - Yv[:, j]) ** 2))
        e_lin = np.sqrt(np.mean((pl[:, j] - Yv[:, j]) ** 2))
        log(f"validation one-step RMSE {name}: network {e_mlp:.4f}, linear map alone {e_lin:.4f}", flush=True)
        net.meta[f"val_rmse_{name}_mlp"] = float(e_mlp)
        net.meta[f"val_rmse_{name}_linear"] = float(e_lin)
    for tag, f in (("mlp", predict), ("linear", net.predict_linear)):
        re, ve = closed_loop_error(f, val)
        log(f"validation closed loop {tag}: rate error max {re.max():.4f} mean {re.mean():.4f}, "
            f"V error max {ve.max():.3f} mV", flush=True)
        net.meta[f"val_closed_loop_rate_err_max_{tag}"] = float(re.max())
        net.meta[f"val_closed_loop_rate_err_mean_{tag}"] = float(re.mean())
    net.meta["train_windows"] = int(len(X))
    net.meta["final_residual_loss"] = float(best_loss)
    net.meta["train_seconds"] = float(time.time() - t0)
    return net


def he_init(seed=0, dims=(S.N_IN, S.HIDDEN, S.LAYERS, S.N_OUT)):
    """He-uniform weights (bound sqrt(6 / fan_in)), zero biases, flat mind_mlp order."""
    g = np.random.default_rng(seed)
    n_in, hidden, layers, n_out = dims
    sizes = [n_in] + [hidden] * layers + [n_out]
    flat = []
    for k, n in zip(sizes[:-1], sizes[1:]):
        b = np.sqrt(6.0 / k)
        flat += [g.uniform(-b, b, k * n), np.zeros(n)]
    return np.concatenate(flat).astype(np.float32)


if __name__ == "__main__":
    main()
