"""Pins of the CPU oracle against closed forms, published vectors and invariants.

Each test fixes something the reference's own code determines (cited), using a
source independent of the oracle: mpmath, numpy/scipy solvers, cable theory,
Random123's published known-answer vectors, NEURON's published hh resting
values, or orders computed by hand.
"""
import math

import mpmath as mp
import numpy as np
import pytest

import oracle
from mind_amd import MicroModel, Section, configs, macro as M, mechanism_record, voltage_record
from mind_amd.model import pack_nodes

mp.mp.dps = 60


def ulp_error(got, exact):
    exact_d = float(exact)
    if exact_d == 0.0:
        return abs(got)
    return abs(got - exact_d) / math.ulp(exact_d)


# --- elementary functions shared by the kernels and the oracle ---------------
def test_exp_within_one_ulp():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-745, 709.7, 4000), rng.uniform(-1, 1, 2000),
                         rng.uniform(-1e-6, 1e-6, 500), [0.0, 1.0, -1.0, 0.5 * math.log(2),
                                                         700.0, -700.0, 1e-300]])
    worst = max(ulp_error(oracle.exp(float(x)), mp.exp(mp.mpf(float(x)))) for x in xs)
    assert worst <= 1.0, worst
    assert oracle.exp(710.0) == math.inf and oracle.exp(-746.0) == 0.0
    assert math.isnan(oracle.exp(math.nan))


def test_log_within_one_ulp():
    rng = np.random.default_rng(1)
    xs = np.concatenate([10.0 ** rng.uniform(-300, 300, 4000), rng.uniform(0.5, 2.0, 2000),
                         1.0 + rng.uniform(-1e-7, 1e-7, 200), [1.0, 2.0, 0.5, 5e-324, 1e308]])
    worst = max(ulp_error(oracle.log(float(x)), mp.log(mp.mpf(float(x)))) for x in xs)
    assert worst <= 1.0, worst
    assert oracle.log(1.0) == 0.0 and oracle.log(0.0) == -math.inf
    assert math.isnan(oracle.log(-1.0))


# --- Random123 Philox4x32-10 (src/runtime/cuda/micro.jl:3-51) ------------------
KAT = [  # Random123 kat_vectors, philox4x32 10 rounds
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(ctr, key, expect):
    assert tuple(oracle.philox(ctr, key)) == expect


def test_uniform_stream_layout():
    """Counter block (ctr>>2, key3, id lo, id hi), key (seed lo, seed hi), (w+1)/(2^32+1)."""
    seed, lid, digest = 0x1234_5678_9ABC, (7 << 32) | 99, 0xA186833B569FF3B2
    u = oracle.uniform_stream(seed, lid, digest, 12)
    key = (seed & 0xFFFFFFFF, seed >> 32)
    for block in range(3):
        words = oracle.philox((block, digest & 0xFFFFFFFF, lid & 0xFFFFFFFF, lid >> 32), key)
        expect = [(w + 1.0) * (1.0 / 4294967297.0) for w in words]
        assert list(u[4 * block:4 * block + 4]) == expect


# --- Hines elimination (src/runtime/cuda/micro.jl:688-749) --------------------
@pytest.mark.parametrize("seed", range(6))
def test_hines_matches_dense_solve(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 200))
    parent = np.array([-1] + [int(rng.integers(0, i)) for i in range(1, n)], np.int32)
    g = rng.uniform(0.01, 5.0, n)
    g[0] = 0.0
    A = np.zeros((n, n))
    for i in range(1, n):
        A[i, parent[i]] = A[parent[i], i] = -g[i]
    d = -A.sum(axis=1) + rng.uniform(0.1, 3.0, n)
    A[np.arange(n), np.arange(n)] = d
    b = rng.normal(size=n)
    x = oracle.hines(parent, g, d, b)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


# --- geometry (src/model/micro/morphology.jl:406-484) --------------------------
def test_cylinder_geometry_closed_form():
    L, dia, nseg, Ra, cm = 300.0, 3.0, 7, 150.0, 0.9
    m = MicroModel()
    m.add_population([Section("a", nseg=nseg, length=L, diameter=dia, Ra=Ra, cm=cm)], 1)
    geo = oracle.OracleSim(m).geometry()
    h = L / nseg
    area = math.pi * dia * h
    r_half = 0.04 * Ra * (h / 2) / (math.pi * dia * dia)
    # root, nseg segment nodes, endpoint
    assert geo["parent"].tolist() == [-1] + list(range(nseg + 1))
    np.testing.assert_allclose(geo["area"], [100.0] + [area] * nseg + [100.0], rtol=1e-13)
    np.testing.assert_allclose(geo["cap"], [0.0] + [cm * area * 1e-5] * nseg + [0.0], rtol=1e-13)
    expect_g = [0.0, 1 / r_half] + [1 / (2 * r_half)] * (nseg - 1) + [1 / r_half]
    np.testing.assert_allclose(geo["axial"], expect_g, rtol=1e-13)
    adiag = np.zeros(nseg + 2)
    for i in range(1, nseg + 2):
        adiag[i] += expect_g[i]
        adiag[i - 1] += expect_g[i]
    np.testing.assert_allclose(geo["adiag"], adiag, rtol=1e-13)


def test_tapered_pt3d_geometry():
    """A frustum: lateral area pi (r1 + r2) s and resistance 4 Ra L/(pi d1 d2)."""
    d1, d2, L = 4.0, 2.0, 50.0
    m = MicroModel()
    m.add_population([Section("t", nseg=1, points=[(0, 0, 0, d1), (L, 0, 0, d2)])], 1)
    geo = oracle.OracleSim(m).geometry()
    dm = (d1 + d2) / 2
    s1 = math.hypot(L / 2, (d1 - dm) / 2)
    s2 = math.hypot(L / 2, (dm - d2) / 2)
    area = math.pi * ((d1 + dm) / 2 * s1 + (dm + d2) / 2 * s2)
    np.testing.assert_allclose(geo["area"][1], area, rtol=1e-13)
    r_left = 0.04 * 100.0 / math.pi * (L / 2) / (d1 * dm)
    np.testing.assert_allclose(geo["axial"][1], 1 / r_left, rtol=1e-13)


@pytest.mark.parametrize("which", ["single", "ball", "trees"])
def test_product_packing_matches_oracle(which):
    if which == "single":
        m, _ = configs.single_cell()
    elif which == "ball":
        m, _ = configs.ball_stick_network(ncells=20, indegree=5)
    else:
        m, _ = configs.tree_network(ncells=25, median=150)
    got = pack_nodes(m)
    ref = oracle.OracleSim(m).geometry()
    for k in ("parent", "area", "cap", "axial", "adiag"):
        assert np.array_equal(got[k], ref[k]), k


# --- membrane dynamics ---------------------------------------------------------
def passive_cell(v0=-60.0, nseg=1, length=20.0, diameter=20.0):
    m = MicroModel()
    (cell,) = m.add_population([Section("soma", label="soma", nseg=nseg, length=length,
                                        diameter=diameter)], 1, voltage=v0)
    m.insert(cell, "pas", params={"g": 1e-4, "e": -65.0})
    return m, cell


def test_isopotential_decay_is_backward_euler():
    """tau = cm/g = 10 ms; each step contracts v - e by 1/(1 + dt/tau)."""
    m, cell = passive_cell()
    rec = [voltage_record(m, cell)]
    r = oracle.OracleSim(m).initialize().run(50.0, rec)
    n = np.arange(len(r.time))
    expect = -65.0 + 5.0 * (1.0 / (1.0 + 0.025 / 10.0)) ** n
    np.testing.assert_allclose(r.traces["v_0"], expect, atol=1e-9)


def test_isopotential_decay_secondorder_is_crank_nicolson():
    """secondorder 1 and 2 (assembly at dt/2, v += 2 dv): each step contracts v - e
    by (1 - a)/(1 + a) with a = dt/(2 tau), the trapezoidal rule's factor."""
    for so in (1, 2):
        m, cell = passive_cell()
        rec = [voltage_record(m, cell)]
        r = oracle.OracleSim(m, secondorder=so).initialize().run(50.0, rec)
        n = np.arange(len(r.time))
        a = 0.025 / (2.0 * 10.0)
        expect = -65.0 + 5.0 * ((1.0 - a) / (1.0 + a)) ** n
        np.testing.assert_allclose(r.traces["v_0"], expect, atol=1e-9)
    with pytest.raises(ValueError):
        oracle.OracleSim(passive_cell()[0], secondorder=3)


def test_rest_is_a_fixed_point():
    m, cell = passive_cell(v0=-65.0, nseg=5, length=200.0, diameter=2.0)
    sim = oracle.OracleSim(m).initialize()
    sim.run(10.0)
    assert np.all(sim.voltage() == -65.0)


def test_cable_input_resistance():
    """Passive soma + sealed finite cable: R_in = 1/(G_soma + tanh(L/lambda)/(r_a lambda))."""
    m = MicroModel()
    (cell,) = m.add_population(configs.ball_and_stick_sections(200), 1, voltage=-65.0)
    m.insert(cell, "pas", params={"g": 1e-4, "e": -65.0})
    amp = 0.05
    m.add_pointprocess(cell, "IClamp", params={"amp": amp, "del": 0.0, "dur": 1e9})
    rec = [voltage_record(m, cell)]
    r = oracle.OracleSim(m).initialize().run(250.0, rec)
    dv = r.traces["v_0"][-1] + 65.0
    Rm, Ra, d, L = 1e4, 100.0, 2e-4, 0.1  # ohm cm2, ohm cm, cm, cm
    lam = math.sqrt(d * Rm / (4 * Ra))
    ra = 4 * Ra / (math.pi * d * d)  # ohm / cm
    g_cable = math.tanh(L / lam) / (ra * lam)
    g_soma = math.pi * 20e-4 * 20e-4 / Rm
    r_in_mohm = 1.0 / (g_cable + g_soma) / 1e6
    np.testing.assert_allclose(dv / amp, r_in_mohm, rtol=2e-3)


def test_hh_resting_gates_match_neuron():
    """NEURON's hh at -65 mV, 6.3 degC: m 0.0529325, h 0.596121, n 0.317677."""
    m, info = configs.single_cell()
    sim = oracle.OracleSim(m).initialize()
    f = {k: sim.field(0, i, 1)[0] for k, i in (("m", 4), ("h", 5), ("n", 6))}
    assert abs(f["m"] - 0.0529325) < 1e-6
    assert abs(f["h"] - 0.596121) < 1e-6
    assert abs(f["n"] - 0.317677) < 1e-6


@pytest.mark.parametrize("v", [-90.0, -65.0, -55.0, -40.0, 0.0, 30.0])
def test_hh_steady_state_against_mpmath(v):
    """After many steps clamped by a huge passive load the gates sit at x_inf(v)."""
    vv = mp.mpf(v)

    def vtrap(x, y):
        return x / (mp.e ** (x / y) - 1) if abs(x / y) >= 1e-6 else y * (1 - x / y / 2)

    am, bm = mp.mpf("0.1") * vtrap(-(vv + 40), 10), 4 * mp.e ** (-(vv + 65) / 18)
    ah, bh = mp.mpf("0.07") * mp.e ** (-(vv + 65) / 20), 1 / (mp.e ** (-(vv + 35) / 10) + 1)
    an, bn = mp.mpf("0.01") * vtrap(-(vv + 55), 10), mp.mpf("0.125") * mp.e ** (-(vv + 65) / 80)
    m = MicroModel()
    (cell,) = m.add_population([Section("soma", label="soma", length=20.0, diameter=20.0)], 1,
                               voltage=v)
    m.insert(cell, "hh", params={"gnabar": 0.0, "gkbar": 0.0, "gl": 0.0})
    sim = oracle.OracleSim(m).initialize()
    for i, (a, b) in zip((4, 5, 6), ((am, bm), (ah, bh), (an, bn))):
        assert abs(sim.field(0, i, 1)[0] - float(a / (a + b))) < 1e-14


def test_exp2syn_peak_is_weight_and_decay_is_exact():
    m, cell = passive_cell(v0=-65.0)
    syn = m.add_pointprocess(cell, "Exp2Syn", params={"tau1": 0.5, "tau2": 5.0, "e": 0.0})
    ns = m.add_pointprocess(m, "NetStim", params={"interval": 10.0, "number": 1, "start": 2.0,
                                                  "noise": 0.0})
    w = 3e-3
    m.connect(ns, syn, weight=w, delay=1.0)
    recs = [mechanism_record(m, syn, "g"), mechanism_record(m, syn, "B", name="B")]
    r = oracle.OracleSim(m).initialize().run(40.0, recs)
    g = r.traces["g"]
    assert abs(g.max() - w) / w < 0.01  # factor normalises the peak to the weight
    B = r.traces["B"]
    k = int(np.argmax(B))
    ratio = B[k + 10:k + 200] / B[k + 9:k + 199]
    np.testing.assert_allclose(ratio, math.exp(-0.025 / 5.0), rtol=1e-13)


def _exp2syn_closed_form(t, t0, w, tau1=0.5, tau2=5.0):
    """Exp2Syn state difference B - A = w * factor * (exp(-(t-t0)/tau2) - exp(-(t-t0)/tau1))
    after an event of weight w applied at t0 (Exp2Syn.mod:56-85), zero before. (The
    recorded `g` is this value from the last current evaluation, a step behind.)"""
    tp = tau1 * tau2 / (tau2 - tau1) * math.log(tau2 / tau1)
    factor = 1.0 / (-math.exp(-tp / tau1) + math.exp(-tp / tau2))
    dt_ = np.maximum(t - t0, 0.0)
    return np.where(t > t0, w * factor * (np.exp(-dt_ / tau2) - np.exp(-dt_ / tau1)), 0.0)


def test_injected_event_is_applied_at_its_time():
    """inject_event!: the connection's event arrives at the given time without its delay,
    applied at the first delivery whose cutoff reaches it (here the end of the step
    ending at 5.0 ms), after which the conductance follows the closed form."""
    m, cell = passive_cell(v0=-65.0)
    syn = m.add_pointprocess(cell, "Exp2Syn", params={"tau1": 0.5, "tau2": 5.0, "e": 0.0})
    src = m.register_spike_source(m, sid=7)
    w = 2e-3
    conn = m.connect(src, syn, weight=w, delay=3.0)
    recs = [mechanism_record(m, syn, "A"), mechanism_record(m, syn, "B")]
    sim = oracle.OracleSim(m).initialize()
    with pytest.raises(ValueError):
        sim.inject_event(conn + 1, 1.0)  # unknown connection
    sim.inject_event(conn, 5.0)
    r = sim.run(30.0, recs)
    expect = _exp2syn_closed_form(r.time, 5.0, w)
    np.testing.assert_allclose(r.traces["B"] - r.traces["A"], expect, rtol=1e-9, atol=1e-15)
    assert len(r.spike_times) == 0
    with pytest.raises(ValueError):
        sim.inject_event(conn, 29.0)  # before the current time


def test_injected_spike_routes_through_every_connection():
    """inject_spike!: an external input's spike at 3 ms is recorded with its SID and
    reaches each target after that connection's delay."""
    m, cell = passive_cell(v0=-65.0)
    s1 = m.add_pointprocess(cell, "Exp2Syn", params={"tau1": 0.5, "tau2": 5.0, "e": 0.0})
    s2 = m.add_pointprocess(cell, "Exp2Syn", params={"tau1": 0.5, "tau2": 5.0, "e": 0.0})
    src = m.register_spike_source(m, sid=42)
    m.connect(src, s1, weight=1e-3, delay=1.0)
    m.connect(src, s2, weight=4e-3, delay=2.5)
    recs = [mechanism_record(m, p, f, name=f"{f}{k}") for k, p in ((1, s1), (2, s2)) for f in ("A", "B")]
    sim = oracle.OracleSim(m).initialize()
    sim.inject_spike(src, 3.0)
    r = sim.run(30.0, recs)
    assert r.spike_times.tolist() == [3.0] and r.spike_sources.tolist() == [42]
    g1, g2 = (r.traces[f"B{k}"] - r.traces[f"A{k}"] for k in (1, 2))
    np.testing.assert_allclose(g1, _exp2syn_closed_form(r.time, 4.0, 1e-3), rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(g2, _exp2syn_closed_form(r.time, 5.5, 4e-3), rtol=1e-9, atol=1e-15)
    det = m.register_spike_source(cell)
    with pytest.raises(ValueError):
        oracle.OracleSim(m).initialize().inject_spike(det, 1.0)  # not an external input


def test_pattern_stim_raster_and_delivery():
    """PatternStim (builtin ARTIFICIAL_CELL with spike_times): every given spike is
    recorded with the SID of the point's source, sorted by (time, SID), including one at
    t = 0 and a tie between two points; each reaches its synapse after the connection's
    delay, so the conductance is the sum of the Exp2Syn closed forms. A point without a
    source emits nothing."""
    m, cell = passive_cell(v0=-65.0)
    syn = m.add_pointprocess(cell, "Exp2Syn", params={"tau1": 0.5, "tau2": 5.0, "e": 0.0})
    p1 = m.add_pointprocess(m, "PatternStim", spike_times=[0.0, 5.0, 12.3])
    p2 = m.add_pointprocess(m, "PatternStim", spike_times=[5.0, 7.5])
    m.add_pointprocess(m, "PatternStim", spike_times=[1.0])  # no source
    s2 = m.register_spike_source(p2, sid=11)
    s1 = m.register_spike_source(p1, sid=20)
    m.connect(s1, syn, weight=1e-3, delay=1.0)
    m.connect(s2, syn, weight=3e-3, delay=2.0)
    recs = [mechanism_record(m, syn, "A"), mechanism_record(m, syn, "B")]
    r = oracle.OracleSim(m).initialize().run(30.0, recs)
    assert r.spike_times.tolist() == [0.0, 5.0, 5.0, 7.5, 12.3]
    assert r.spike_sources.tolist() == [20, 11, 20, 11, 20]
    arrivals = [(1.0, 1e-3), (6.0, 1e-3), (13.3, 1e-3), (7.0, 3e-3), (9.5, 3e-3)]
    # an arrival at t is applied at the end of the step whose cutoff reaches it
    expect = sum(_exp2syn_closed_form(r.time, 0.025 * math.ceil(t / 0.025 - 1e-9), w) for t, w in arrivals)
    np.testing.assert_allclose(r.traces["B"] - r.traces["A"], expect, rtol=1e-9, atol=1e-15)


def test_pattern_stim_validation():
    m, cell = passive_cell(v0=-65.0)
    with pytest.raises(ValueError, match="sorted"):
        m.add_pointprocess(m, "PatternStim", spike_times=[2.0, 1.0])
    with pytest.raises(ValueError, match="invalid"):
        m.add_pointprocess(m, "PatternStim", spike_times=[-1.0])
    with pytest.raises(ValueError, match="invalid"):
        m.add_pointprocess(m, "PatternStim", spike_times=[float("nan")])
    with pytest.raises(ValueError, match="PatternStim only"):
        m.add_pointprocess(cell, "ExpSyn", spike_times=[1.0])
    placed = m.add_pointprocess(cell, "PatternStim", spike_times=[1.0])  # location unused
    assert m.points[placed.index][1] == -1
    with pytest.raises(ValueError, match="NET_RECEIVE"):
        m.connect(m.register_spike_source(cell), m.add_pointprocess(m, "PatternStim"))


def test_seclamp_on_a_passive_compartment_is_the_implicit_recurrence():
    """SEClamp (SEClamp.mod) on an isopotential passive compartment (level changes off the
    step grid, so that the evaluation time matters): with the clamp level
    vc(t) and on(t) taken at each step's current evaluation time t + dt/2 (amp1 before
    dur1, amp2 before dur1+dur2, amp3 before dur1+dur2+dur3, then off), a backward-Euler
    step solves (C/dt + G + on/rs) v' = C/dt v + G e + on vc/rs exactly, since every current
    is linear in v; the recorded i = on (vc - v')/rs is the after_cvode update at the new v,
    and at t = 0 it is the initialization's (amp1 - v0)/rs."""
    m, cell = passive_cell(v0=-65.0)
    p = {"rs": 0.05, "dur1": 2.01, "amp1": -40.0, "dur2": 3.0, "amp2": -20.0, "dur3": 1.0, "amp3": -50.0}
    clamp = m.add_pointprocess(cell, "SEClamp", params=p)
    recs = [voltage_record(m, cell), mechanism_record(m, clamp, "i"), mechanism_record(m, clamp, "vc")]
    r = oracle.OracleSim(m).initialize().run(10.0, recs)
    dt, area = 0.025, math.pi * 20.0 * 20.0
    C, G, e = 1.0 * area * 1e-5, 1e-4 * area * 0.01, -65.0
    v, vs, iis, vcs = -65.0, [-65.0], [(p["amp1"] + 65.0) / p["rs"]], [p["amp1"]]
    for k in range(len(r.time) - 1):
        te = k * dt + 0.5 * dt
        levels = [(p["dur1"], p["amp1"]), (p["dur1"] + p["dur2"], p["amp2"]),
                  (p["dur1"] + p["dur2"] + p["dur3"], p["amp3"])]
        on, vc = 0.0, 0.0
        for end, amp in levels:
            if te < end:
                on, vc = 1.0, amp
                break
        gc = on / p["rs"]
        v = (C / dt * v + G * e + gc * vc) / (C / dt + G + gc)
        vs.append(v)
        iis.append(gc * (vc - v))
        vcs.append(vc)
    np.testing.assert_allclose(r.traces["v_0"], vs, rtol=0, atol=1e-8)
    np.testing.assert_allclose(r.traces["i"], iis, rtol=0, atol=1e-6)
    assert r.traces["vc"].tolist() == vcs
    assert abs(r.traces["v_0"][int(round(4.0 / dt))] - p["amp2"]) < 0.1  # clamped
    assert r.traces["i"][-1] == 0.0  # off after dur1 + dur2 + dur3


def test_vclamp_holds_each_level_and_balances_the_leak():
    """VClamp (VClamp.mod, gain 1e5, rstim 1 MOhm, tau1 1 us) on an isopotential passive
    compartment: at the end of each of the three levels the voltage sits at the level
    and the electrode current (positive depolarises) equals the leak G (v - e) the
    compartment loses there; after dur1 + dur2 + dur3 the clamp is off (i = 0) and the
    compartment relaxes toward e. Array parameters are given as sequences."""
    m, cell = passive_cell(v0=-65.0)
    dur, amp = [2.01, 3.0, 1.0], [-40.0, -20.0, -50.0]
    clamp = m.add_pointprocess(cell, "VClamp", params={"dur": dur, "amp": amp})
    recs = [voltage_record(m, cell), mechanism_record(m, clamp, "i"), mechanism_record(m, clamp, "stim"),
            mechanism_record(m, clamp, "amp", index=3, name="amp3")]
    r = oracle.OracleSim(m).initialize().run(12.0, recs)
    area = math.pi * 20.0 * 20.0
    G, e = 1e-4 * area * 0.01, -65.0
    v, i, t = r.traces["v_0"], r.traces["i"], r.time
    for end, level in zip(np.cumsum(dur), amp):
        k = int(np.searchsorted(t, end)) - 2  # the last sample inside the level
        assert abs(v[k] - level) < 1e-3, (end, v[k])
        assert abs(i[k] - G * (v[k] - e)) < 1e-3 * abs(G * (level - e)), (end, i[k])
    assert r.traces["amp3"][0] == -50.0
    off = t > sum(dur) + 0.05
    assert np.all(i[off] == 0.0)
    assert v[-1] < -50.0 and np.all(np.diff(v[off]) < 0)  # relaxing toward e = -65
    with pytest.raises(ValueError):
        m.add_pointprocess(cell, "VClamp", params={"dur": [1.0, 2.0]})  # dur has three elements
    with pytest.raises(ValueError):
        mechanism_record(m, clamp, "amp", index=4)


def test_alpha_synapse_conductance_is_the_alpha_function():
    """AlphaSynapse (AlphaSynapse.mod): g = gmax x e^(1-x), x = (t - onset)/tau, at the
    current evaluation t + dt/2, zero outside 0 <= x <= 10; peak gmax at x = 1. The
    recorded g is the last evaluation's, half a step before the sample."""
    m, cell = passive_cell(v0=-65.0)
    syn = m.add_pointprocess(cell, "AlphaSynapse", params={"onset": 5.0, "tau": 2.0, "gmax": 1e-3, "e": 0.0})
    r = oracle.OracleSim(m).initialize().run(40.0, [mechanism_record(m, syn, "g"), voltage_record(m, cell)])
    tc = r.time[1:] - 0.5 * 0.025
    x = (tc - 5.0) / 2.0
    expect = np.where((x < 0) | (x > 10), 0.0, 1e-3 * x * np.exp(1.0 - x))
    np.testing.assert_allclose(r.traces["g"][1:], expect, rtol=1e-13, atol=1e-18)
    assert abs(r.traces["g"].max() - 1e-3) < 1e-3 * 1e-4
    assert r.traces["v_0"].max() > -65.0 + 1.0  # the synapse depolarises toward e = 0


def test_netstim_regular_train():
    m = MicroModel()
    ns = m.add_pointprocess(m, "NetStim", params={"interval": 7.0, "number": 5, "start": 3.0,
                                                  "noise": 0.0})
    m.register_spike_source(ns)
    r = oracle.OracleSim(m).initialize().run(100.0)
    assert r.spike_times.tolist() == [3.0, 10.0, 17.0, 24.0, 31.0]


def test_netstim_poisson_statistics():
    from scipy import stats

    m = MicroModel()
    ns = m.add_pointprocess(m, "NetStim", params={"interval": 2.0, "number": 1e9, "start": 0.0,
                                                  "noise": 1.0})
    m.register_spike_source(ns)
    r = oracle.OracleSim(m).initialize().run(4000.0)
    isi = np.diff(np.concatenate([[0.0], r.spike_times]))
    assert abs(isi.mean() - 2.0) < 0.06
    assert stats.kstest(isi, "expon", args=(0, 2.0)).pvalue > 1e-3


def test_spike_detection_rule_matches_trace():
    """A spike is the first step whose starting v exceeds threshold (time t + 1e-10)."""
    m, info = configs.single_cell()
    rec = [voltage_record(m, info["cells"][0])]
    r = oracle.OracleSim(m).initialize().run(200.0, rec)
    v = r.traces["v_0"]
    above = v > -20.0
    k = np.nonzero(above[:-1] & ~np.concatenate([[False], above[:-2]]))[0]
    expect = k * 0.025 + 1e-10
    assert np.array_equal(r.spike_times, expect)


def test_equal_time_events_follow_source_then_edge_order():
    """Two cells crossing threshold at the same step, equal delays into one synapse:
    the heap tickets put source 0's edges first, each source's in creation order."""
    m = MicroModel()
    cells = m.add_population(configs.ball_and_stick_sections(10), 3)
    srcs = []
    for c in cells:
        m.insert(c, "hh", sections="soma")
        m.insert(c, "pas", sections="dend", params={"g": 1e-4, "e": -65.0})
        m.add_pointprocess(c, "IClamp", params={"amp": 0.5, "del": 1.0, "dur": 5.0})
        srcs.append(m.register_spike_source(c, threshold=-20.0))
    syn = m.add_pointprocess(cells[2], "ExpSyn", section="dend", params={"tau": 2.0})
    # edges: cell1 -> syn (w 1), cell0 -> syn (w 2), cell0 -> syn (w 3), cell1 -> syn (w 4)
    for s, w in ((1, 1e-5), (0, 2e-5), (0, 3e-5), (1, 4e-5)):
        m.connect(srcs[s], syn, weight=w, delay=2.0)
    sim = oracle.OracleSim(m)
    sim.log_deliveries(True)
    sim.initialize()
    r = sim.run(12.0)
    t0 = r.spike_times[r.spike_sources == 0][0]
    t1 = r.spike_times[r.spike_sources == 1][0]
    assert t0 == r.spike_times[0]
    assert t0 == t1  # identical cells spike at the same step
    t, ticket, edge, target = sim.delivery_log()
    first = t == t[0]
    assert first.sum() == 4 and np.all(target == 3)
    # CSR edges by source: source 0 holds creation #2, #3 (edges 1, 2), source 1
    # holds #1, #4 (edges 3, 4); the heap delivers source 0's edges first.
    assert edge[first].tolist() == [1, 2, 3, 4]
    assert np.all(np.diff(ticket[first]) > 0)


def test_run_rejects_fractional_stop():
    m, _ = configs.single_cell()
    sim = oracle.OracleSim(m).initialize()
    with pytest.raises(RuntimeError, match="integer number of steps"):
        sim.run(0.0301)


# --- micro-macro coupling --------------------------------------------------------
def test_wilson_cowan_rk4_step_against_scipy():
    from scipy.integrate import solve_ivp

    from mind_amd.macro import WC_DEFAULTS, WC_PARAMS

    p = dict(WC_DEFAULTS)
    vec = [p[k] for k in WC_PARAMS]
    u = 0.3

    def S(x, a, th):
        return 1 / (1 + np.exp(-a * (x - th))) - 1 / (1 + np.exp(a * th))

    def f(t, y):
        e, i = y
        return [(-e + S(p["c_ee"] * e - p["c_ei"] * i + p["P"] + u, p["a_e"], p["theta_e"])) / p["tau_e"],
                (-i + S(p["c_ie"] * e - p["c_ii"] * i + p["Q"], p["a_i"], p["theta_i"])) / p["tau_i"]]

    def local_error(h, e, i):
        exact = solve_ivp(f, (0, h), [e, i], rtol=1e-14, atol=1e-16, method="DOP853").y[:, -1]
        return np.abs(np.array(oracle.wc_rk4(vec, e, i, u, h)) - exact), exact

    e, i = 0.2, 0.1
    for _ in range(20):
        err, exact = local_error(0.1, e, i)
        assert err.max() < 1e-9
        e, i = exact
    # fourth-order method: the local error scales as h^5
    big, _ = local_error(0.8, 0.2, 0.1)
    small, _ = local_error(0.4, 0.2, 0.1)
    assert 20 < big.max() / small.max() < 45


def test_macro_coupling_reads_the_delayed_history():
    """coupling_0(t) = W_01 E_1(t - D_01): a delay of 3.3 ms is 33 macro steps."""
    from mind_amd import MicroModel

    W = np.array([[0.0, 0.15], [0.0, 0.0]])
    D = np.array([[0.0, 3.3], [0.0, 0.0]])
    mm = M.MacroModel(M.Connectivity(W, D))
    for r in range(2):
        M.use_macro(mm.roi(r), "WilsonCowan", initial_state={"E": 0.1 * (r + 1), "I": 0.05})
    mm.micro = MicroModel()
    M.connect(mm.roi(1), mm.roi(0), "LinearDelay")
    r = oracle.OracleMacro(mm).initialize().run(20.0)
    E1 = r.E[1]
    c0 = r.coupling[0]
    # coupling recorded after step s uses E_1 at boundary s - 33 (the initial value before 0)
    for s in range(1, len(c0)):
        src = E1[max(s - 1 - 33, 0)]
        assert c0[s] == 0.15 * src


def test_micro_macro_runs_and_is_reproducible():
    mm, _ = configs.micro_macro(n_roi=3, cells=4, nseg=20, input_weight=5e-3)
    a = oracle.OracleMacro(mm).initialize().run(30.0)
    b = oracle.OracleMacro(mm).initialize().run(30.0)
    assert np.array_equal(a.E, b.E) and np.array_equal(a.spike_times, b.spike_times)
    assert a.E.shape == (3, 301)
    assert np.all(np.isfinite(a.E)) and np.all(a.rate >= 0)


# --- ExpSyn (mod/ExpSyn.mod:1-42) -------------------------------------------------
def test_expsyn_jump_is_weight_and_decay_is_exact():
    """ExpSyn: NET_RECEIVE g = g + w, then DERIVATIVE g' = -g/tau by cnexp. One event
    of weight w due at 3.0 ms (NetStim at 2.0 + delay 1.0) is applied by the
    delivery that follows the 3.0 ms sample, so g(3.0) = 0 in the trace, g = w just
    after, and g(3.0 + k dt) = w exp(-k dt / tau); i = g (v - e)."""
    m, cell = passive_cell(v0=-65.0)
    tau, e, w, dt = 2.0, 0.0, 3e-3, 0.025
    syn = m.add_pointprocess(cell, "ExpSyn", params={"tau": tau, "e": e})
    ns = m.add_pointprocess(m, "NetStim", params={"interval": 10.0, "number": 1, "start": 2.0, "noise": 0.0})
    m.connect(ns, syn, weight=w, delay=1.0)
    recs = [mechanism_record(m, syn, "g"), mechanism_record(m, syn, "i"), voltage_record(m, cell)]
    sim = oracle.OracleSim(m).initialize()
    r = sim.run(3.0, recs)
    assert np.all(r.traces["g"] == 0.0)
    from mind_amd.mechanisms import builtin

    mech = builtin("ExpSyn")
    assert sim.field(mech.id, mech.field_index("g"))[0] == w  # the jump, exactly w
    r2 = sim.run(8.0, recs)
    k = np.arange(1, len(r2.time))
    np.testing.assert_allclose(r2.traces["g"][1:], w * np.exp(-k * dt / tau), rtol=1e-12)
    # the current is evaluated at t + dt/2 from g before that step's decay
    g_prev = np.concatenate([[w], r2.traces["g"][1:-1]])
    v_prev = r2.traces["v_0"][:-1]
    np.testing.assert_allclose(r2.traces["i"][1:], g_prev * (v_prev - e), rtol=1e-6)


# --- MACRO2MICRO PoissonDrive and MICRO2MACRO PopRate (project MIND roles) --------
def _held_macro(n_roi, cells, rate_hz, k_micro=0.0, clamp_amps=None, bin_ms=1.0):
    """ROIs whose Wilson-Cowan state is held (tau 1e15 ms: E + h dE rounds to E), so
    the drive rate is rate_scale * E exactly; each ROI owns `cells` one-compartment
    cells with an ExpSyn of weight 0 fed by PoissonDrive, a detector at -20 mV and
    PopRate. With clamp_amps the cells are hh with an IClamp of that amplitude."""
    W = np.zeros((n_roi, n_roi))
    mm = M.MacroModel(M.Connectivity(W, np.full((n_roi, n_roi), 2.0)))
    micro = MicroModel()
    for r in range(n_roi):
        M.use_macro(mm.roi(r), "WilsonCowan", params={"tau_e": 1e15, "tau_i": 1e15, "k_micro": k_micro,
                                                      "rate_scale": rate_hz / 0.5},
                    initial_state={"E": 0.5, "I": 0.1})
    sources = []
    for r in range(n_roi):
        for c in range(cells):
            (cell,) = micro.add_population([Section("soma", label="soma", length=20.0, diameter=20.0)], 1)
            if clamp_amps is None:
                micro.insert(cell, "pas", params={"g": 1e-4, "e": -65.0})
            else:
                micro.insert(cell, "hh")
                micro.add_pointprocess(cell, "IClamp", params={"del": 1.0 + c, "dur": 1e9,
                                                               "amp": clamp_amps[r][c]})
            syn = micro.add_pointprocess(cell, "ExpSyn", params={"tau": 2.0, "e": 0.0})
            src = micro.register_spike_source(cell, section="soma", x=0.5, threshold=-20.0)
            sources.append((r, src))
            if r == 0 and c == 0:
                M.use_micro(mm.roi(0), micro)
            elif c == 0:
                M.use_micro(mm.roi(r), micro)
            M.macro2micro(mm.roi(r), "PoissonDrive", target=syn, weight=0.0, delay=0.0)
            M.micro2macro(mm.roi(r), "PopRate", source=src, params={"bin": bin_ms})
    return mm, sources


def test_poisson_drive_counts_and_intervals():
    """PoissonDrive at a held rate of 200 Hz per binding (NET_RECEIVE: tn = start +
    negexp * 1000 / rate while tn < stop, rekeyed every macro step): events per
    0.1 ms step have mean and variance rate * dt / 1000 = 0.02, consecutive steps
    are uncorrelated (a stream that was not rekeyed would repeat itself every
    step), and the intervals of each binding's train are exponential with mean
    5 ms (Kolmogorov-Smirnov). A Hz/kHz slip changes the mean 1000-fold."""
    from scipy import stats

    n_roi, cells, rate, stop = 2, 6, 200.0, 500.0
    mm, _ = _held_macro(n_roi, cells, rate)
    res = oracle.OracleMacro(mm).initialize().run(stop)
    assert np.all(res.E == 0.5)
    virtual = res.spike_sources < 0
    times, sids = res.spike_times[virtual], -res.spike_sources[virtual] - 1
    nb = n_roi * cells
    assert set(np.unique(sids)) == set(range(nb))
    steps = int(round(stop / 0.1))
    counts = np.zeros((nb, steps))
    np.add.at(counts, (sids, np.minimum((times / 0.1).astype(int), steps - 1)), 1)
    lam = rate * 0.1 / 1000.0
    n = counts.size
    assert abs(counts.mean() - lam) < 5 * math.sqrt(lam / n)
    assert abs(counts.var() / counts.mean() - 1.0) < 0.08
    a, b = counts[:, :-1].ravel(), counts[:, 1:].ravel()
    assert abs(np.corrcoef(a, b)[0, 1]) < 5 / math.sqrt(a.size)
    isi = np.concatenate([np.diff(np.sort(times[sids == k])) for k in range(nb)])
    assert stats.kstest(isi, "expon", args=(0, 1000.0 / rate)).pvalue > 1e-3
    # reproducible for a seed, different for another
    again = oracle.OracleMacro(mm).initialize().run(stop)
    assert np.array_equal(again.spike_times, res.spike_times)
    other = oracle.OracleMacro(mm, seed=2).initialize().run(stop)
    assert not np.array_equal(other.spike_times[other.spike_sources < 0], times)


def test_pop_rate_bins_match_hand_count():
    """PopRate: at the end of every 1 ms bin an ROI publishes (spikes of its sources
    in the bin) / (sources x 1 ms) in Hz, held until the next bin end. Counted here
    from the raster by hand, for regularly firing hh cells of different rates."""
    amps = [[0.3, 0.6, 1.2], [0.8, 0.0, 2.0]]
    mm, sources = _held_macro(2, 3, 0.0, clamp_amps=amps)
    stop = 60.0
    res = oracle.OracleMacro(mm).initialize().run(stop)
    cells = res.spike_sources >= 0
    t, sid = res.spike_times[cells], res.spike_sources[cells]
    assert len(t) > 20
    sid_roi = {src.sid: r for r, src in sources}
    for r in range(2):
        mine = np.array([sid_roi[s] == r for s in sid])
        for b in range(int(stop)):
            n = np.count_nonzero(mine & (t >= b) & (t < b + 1))
            expect = n / (3 * 1.0) * 1000.0
            sample = 10 * (b + 1)  # the macro step that closes the bin
            assert res.rate[r, sample] == expect, (r, b)
            if b + 1 < int(stop):
                assert np.all(res.rate[r, sample:sample + 10] == expect)
    assert res.rate.max() > 0


# --- short and zero NetCon delays, NetStim inputs (events.jl:57-165) ---------------
def test_delivery_call_of_short_delays():
    """A spike detected at the start of step k (time k dt + 1e-10) reaches ExpSyns over
    delays 0, 0.6 dt, 1.4 dt and 1.6 dt. The reference delivers through t + dt/2
    before the currents and through t + dt after recording, so g first shows the
    event (decayed once by nrn_state) at sample k+1 for delay 0 (midpoint of step k),
    k+2 for 0.6 dt (end of step k) and 1.4 dt (midpoint of step k+1), and k+3 for
    1.6 dt (end of step k+1)."""
    dt, w, tau = 0.025, 1e-3, 2.0
    model, info = configs.single_cell()
    src = model.register_spike_source(info["cells"][0], section="soma", x=0.5, threshold=-20.0)
    syns = []
    for factor in (0.0, 0.6, 1.4, 1.6):
        (cell,) = model.add_population([Section("soma", label="soma", length=20.0, diameter=20.0)], 1)
        model.insert(cell, "pas", params={"g": 1e-4, "e": -65.0})
        syn = model.add_pointprocess(cell, "ExpSyn", params={"tau": tau, "e": 0.0})
        model.connect(src, syn, weight=w, delay=factor * dt)
        syns.append(syn)
    recs = [mechanism_record(model, s, "g", name=f"g{j}") for j, s in enumerate(syns)]
    r = oracle.OracleSim(model).initialize().run(30.0, recs)
    first = [t for t, sid in zip(r.spike_times, r.spike_sources) if sid == src.sid][0]
    k = int(round((first - 1e-10) / dt))
    assert first == k * dt + 1e-10
    for j, lag in enumerate((1, 2, 2, 3)):
        g = r.traces[f"g{j}"]
        assert np.all(g[:k + lag] == 0.0), j
        assert g[k + lag] == pytest.approx(w * math.exp(-dt / tau), rel=1e-12), j


def test_netstim_input_starts_and_stops_a_train():
    """NetStim NET_RECEIVE with flag 0 (NetStim.mod:77-105): an event of positive weight
    at 5 ms switches on a NetStim that INITIAL left off (start < 0); with noise 0 its
    first self-event is due at once (event - interval = 0), so it fires at 5, 7, ...
    every interval; an event of negative weight at 8 ms switches it off and the 9 ms
    self-event then does nothing. Zero delays keep every event inside its step."""
    m = MicroModel()
    ns = m.add_pointprocess(m, "NetStim", params={"interval": 2.0, "number": 10, "start": -1.0, "noise": 0.0})
    out = m.register_spike_source(ns, sid=3)
    on, off = m.register_spike_source(m, sid=1), m.register_spike_source(m, sid=2)
    m.connect(on, ns, weight=1.0, delay=0.0)
    m.connect(off, ns, weight=-1.0, delay=0.0)
    sim = oracle.OracleSim(m).initialize()
    sim.inject_spike(on, 5.0)
    sim.inject_spike(off, 8.0)
    r = sim.run(20.0)
    ns_times = [t for t, s in zip(r.spike_times, r.spike_sources) if s == out.sid]
    assert ns_times == [5.0, 7.0]
    assert sorted(r.spike_sources[r.spike_sources != out.sid]) == [1, 2]
