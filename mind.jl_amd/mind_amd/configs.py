"""Seeded synthetic workloads of the simulation step, built through the model API.

Randomness comes from numpy's Philox bit generator keyed by (seed, purpose),
so every machine with the same numpy builds the same model. Sizes follow the
benchmark definitions:
  single_cell   soma (L 20, diam 20 um, hh) + one 100-segment pas dendrite
                (L 1000, diam 2 um, g 1e-4 S/cm2, e -65 mV), Ra 100, cm 1,
                IClamp 0.5 nA at soma(0.5) from 10 to 110 ms, threshold -20 mV.
  ball_stick    N ball-and-stick cells (hh soma, 50-segment pas dendrite) with an
                Exp2Syn (tau1 0.5, tau2 5, e 0) at dendrite segment 25; fixed
                in-degree recurrent NetCons (presynaptic cells uniform with
                replacement, weights lognormal(ln 2e-4, 0.5) uS, delays U(1, 5) ms)
                plus one Poisson NetStim per cell (10 Hz, weight 5e-4 uS, 1 ms).
  trees         N seed-generated branched cells: compartment count
                lognormal(ln 400, 0.8) clipped to [50, 4000]; a hh soma and
                single-segment pas sections (L U(10, 200) um, diam U(0.5, 4) um)
                attached by random binary branching; 20 Exp2Syn at random
                dendritic sections; in-degree 50 recurrent NetCons (delays
                U(1, 5) ms) and a 5 Hz Poisson NetStim per cell (1 ms delay).
Choices the benchmark text leaves open (dendrite length of the ball-and-stick
cells, NetStim delay, NetStim and recurrent targets within a cell) are stated
above and fixed here.
"""
from __future__ import annotations

import numpy as np

from .model import Morphology, MicroModel, Section

HH = "hh"
PAS = "pas"
PURPOSE = {"morphology": 1, "connectivity": 2, "synapses": 3}


def rng(seed: int, purpose: str) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=[int(seed), PURPOSE[purpose]]))


def ball_and_stick_sections(nseg: int, length=1000.0, diameter=2.0):
    return [
        Section("soma", label="soma", nseg=1, length=20.0, diameter=20.0),
        Section("dend", label="dend", parent="soma", parent_x=1.0, nseg=nseg, length=length,
                diameter=diameter),
    ]


def _biophysics(model, cell):
    model.insert(cell, HH, sections="soma")
    model.insert(cell, PAS, sections="dend", params={"g": 1e-4, "e": -65.0})


def single_cell(threshold=-20.0):
    """One soma + 100-segment dendrite with an IClamp (the CPU-path case)."""
    model = MicroModel()
    (cell,) = model.add_population(ball_and_stick_sections(100), 1)
    _biophysics(model, cell)
    clamp = model.add_pointprocess(cell, "IClamp", section="soma", x=0.5,
                                   params={"amp": 0.5, "del": 10.0, "dur": 100.0})
    model.register_spike_source(cell, section="soma", x=0.5, threshold=threshold)
    return model, dict(cells=[cell], clamp=clamp, tstop=200.0)


def ball_stick_network(ncells=1000, nseg=50, indegree=100, seed=2, rate_hz=10.0,
                       threshold=-20.0, netstim_weight=5e-4):
    model = MicroModel()
    cells = model.add_population(ball_and_stick_sections(nseg), ncells)
    syn_x = (25 - 0.5) / 50.0
    syns, sources = [], []
    for cell in cells:
        _biophysics(model, cell)
        syns.append(model.add_pointprocess(cell, "Exp2Syn", section="dend", x=syn_x,
                                           params={"tau1": 0.5, "tau2": 5.0, "e": 0.0}))
        sources.append(model.register_spike_source(cell, section="soma", x=0.5,
                                                   threshold=threshold))
    g = rng(seed, "connectivity")
    pre = g.integers(0, ncells, size=(ncells, indegree))
    w = np.exp(g.normal(np.log(2e-4), 0.5, size=(ncells, indegree)))
    dl = g.uniform(1.0, 5.0, size=(ncells, indegree))
    for target in range(ncells):  # target-major creation fixes the edge order
        for k in range(indegree):
            model.connect(sources[pre[target, k]], syns[target], weight=float(w[target, k]),
                          delay=float(dl[target, k]))
    stims = []
    for target in range(ncells):
        ns = model.add_pointprocess(model, "NetStim", params={
            "interval": 1000.0 / rate_hz, "number": 1e9, "start": 0.0, "noise": 1.0})
        model.connect(ns, syns[target], weight=netstim_weight, delay=1.0)
        stims.append(ns)
    return model, dict(cells=cells, syns=syns, stims=stims, tstop=1000.0)


def tree_morphology(g: np.random.Generator, compartments: int, fanout: int = 2) -> Morphology:
    """Soma plus (compartments - 1) one-segment sections by random branching, each
    section accepting up to `fanout` children at x = 1 (binary by default)."""
    nd = compartments - 1
    parent = np.empty(nd + 1, np.int32)
    parent[0] = -1
    # open attachment slots: each section accepts `fanout` children at x = 1
    slots = [0] * fanout
    picks = g.random(nd)
    for i in range(1, nd + 1):
        j = int(picks[i - 1] * len(slots))
        p = slots[j]
        slots[j] = slots[-1]
        slots.pop()
        parent[i] = p
        slots.extend([i] * fanout)
    length = np.concatenate([[20.0], g.uniform(10.0, 200.0, nd)])
    diam = np.concatenate([[20.0], g.uniform(0.5, 4.0, nd)])
    names = ["soma"] + [f"dend[{i}]" for i in range(nd)]
    labels = ["soma"] + ["dend"] * nd
    n = nd + 1
    return Morphology(names, labels, parent, np.ones(n), np.zeros(n), np.ones(n, np.int32),
                      length, diam, np.full(n, 100.0), np.ones(n))


def tree_network(ncells=5000, median=400, sigma=0.8, lo=50, hi=4000, synapses=20, indegree=50,
                 seed=1, rate_hz=5.0, threshold=-20.0, netstim_weight=5e-4, fanout=2):
    model = MicroModel()
    gm = rng(seed, "morphology")
    sizes = np.clip(np.round(np.exp(gm.normal(np.log(median), sigma, ncells))), lo, hi).astype(int)
    gs = rng(seed, "synapses")
    cells, syns, sources = [], [], []
    for c in range(ncells):
        morph = tree_morphology(gm, int(sizes[c]), fanout)
        (cell,) = model.add_population(morph, 1)
        cells.append(cell)
        model.insert(cell, HH, sections="soma")
        model.insert(cell, PAS, sections="dend", params={"g": 1e-4, "e": -65.0})
        sites = gs.integers(1, len(morph.names), size=synapses)
        syns.append([model.add_pointprocess(cell, "Exp2Syn", section=morph.names[s], x=0.5,
                                            params={"tau1": 0.5, "tau2": 5.0, "e": 0.0})
                     for s in sites])
        sources.append(model.register_spike_source(cell, section="soma", x=0.5,
                                                   threshold=threshold))
    g = rng(seed, "connectivity")
    pre = g.integers(0, ncells, size=(ncells, indegree))
    which = g.integers(0, synapses, size=(ncells, indegree))
    w = np.exp(g.normal(np.log(2e-4), 0.5, size=(ncells, indegree)))
    dl = g.uniform(1.0, 5.0, size=(ncells, indegree))
    for t in range(ncells):
        for k in range(indegree):
            model.connect(sources[pre[t, k]], syns[t][which[t, k]], weight=float(w[t, k]),
                          delay=float(dl[t, k]))
    drive = g.integers(0, synapses, size=ncells)
    for t in range(ncells):
        ns = model.add_pointprocess(model, "NetStim", params={
            "interval": 1000.0 / rate_hz, "number": 1e9, "start": 0.0, "noise": 1.0})
        model.connect(ns, syns[t][drive[t]], weight=netstim_weight, delay=1.0)
    return model, dict(cells=cells, syns=syns, sizes=sizes, tstop=500.0)


def micro_macro(n_roi=16, cells=256, nseg=200, seed=3, input_weight=5e-3, rate_scale=200.0,
                k_micro=0.5, threshold=-20.0, wc=None):
    """Wilson-Cowan ROIs coupled all-to-all (W U(0, 0.2), D U(2, 10) ms), each bound to a
    population of ball-and-stick variants (hh soma, nseg-segment pas dendrite of
    length U(800, 1200) um and diameter U(1.5, 2.5) um, one Exp2Syn at dend(0.1))
    driven by an inhomogeneous Poisson input at rate_scale * E and read back as a
    1 ms binned population rate."""
    from . import macro as M

    g = rng(seed, "connectivity")
    W = g.uniform(0.0, 0.2, (n_roi, n_roi))
    np.fill_diagonal(W, 0.0)
    D = g.uniform(2.0, 10.0, (n_roi, n_roi))
    mm = M.MacroModel(M.Connectivity(W, D))
    micro = MicroModel()
    gm = rng(seed, "morphology")
    pops, syns, sources = [], [], []
    for r in range(n_roi):
        region = mm.roi(r)
        params = dict(wc or {})
        params.setdefault("k_micro", k_micro)
        params.setdefault("rate_scale", rate_scale)
        M.use_macro(region, "WilsonCowan", params=params,
                    initial_state={"E": 0.05 + 0.2 * g.random(), "I": 0.05 + 0.1 * g.random()})
        members = []
        for c in range(cells):
            L = gm.uniform(800.0, 1200.0)
            d = gm.uniform(1.5, 2.5)
            (cell,) = micro.add_population(ball_and_stick_sections(nseg, L, d), 1)
            _biophysics(micro, cell)
            syn = micro.add_pointprocess(cell, "Exp2Syn", section="dend", x=0.1,
                                         params={"tau1": 0.5, "tau2": 5.0, "e": 0.0})
            src = micro.register_spike_source(cell, section="soma", x=0.5, threshold=threshold)
            members.append(cell)
            syns.append(syn)
            sources.append(src)
        pops.append(members)
        M.use_micro(region, micro, exposures=("rate",))
    for r in range(n_roi):
        for q in range(n_roi):
            if r != q:
                M.connect(mm.roi(q), mm.roi(r), "LinearDelay")
    for r in range(n_roi):
        for c in range(cells):
            k = r * cells + c
            M.macro2micro(mm.roi(r), "PoissonDrive", target=syns[k], weight=input_weight, delay=0.0)
            M.micro2macro(mm.roi(r), "PopRate", source=sources[k], params={"bin": 1.0})
    return mm, dict(populations=pops, syns=syns, sources=sources, tstop=2000.0)
