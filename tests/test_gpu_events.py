"""The event-exact path on the GPU against the oracle: NetCon delays shorter than
1.5 dt (including zero) and connections into NetStims (feedback receivers), in
both kernels; injected input and a checkpoint on that path. Voltages, synapse
states, NetStim records and rasters bit for bit."""
import numpy as np
import pytest

import oracle
from mind_amd import (MicroModel, build, checkpoint, configs, destroy, initialize, inject_spike, mechanism_record,
                      read_field, read_voltage, restore, run, voltage_record)
from mind_amd.mechanisms import builtin

pytestmark = pytest.mark.gpu
DT = 0.025


def _network(ncells=30, indegree=8, seed=11, short=True, feedback=False):
    """Ball-and-stick cells with Exp2Syn, recurrent connections and Poisson NetStim
    drive. short: delays U(0, 2 dt), every fourth exactly 0. feedback: six NetStims
    switched on and off by cell spikes (weights +1 / -1, delay 0), each driving
    three synapses."""
    m = MicroModel()
    cells = m.add_population(configs.ball_and_stick_sections(20), ncells)
    syns, srcs = [], []
    for c in cells:
        m.insert(c, builtin("hh"), sections="soma")
        m.insert(c, builtin("pas"), sections="dend", params={"g": 1e-4, "e": -65.0})
        syns.append(m.add_pointprocess(c, "Exp2Syn", section="dend", x=0.3, params={"tau1": 0.5, "tau2": 5.0,
                                                                                   "e": 0.0}))
        srcs.append(m.register_spike_source(c, section="soma", x=0.5, threshold=-20.0))
    g = np.random.default_rng(seed)
    n = 0
    for t in range(ncells):
        for s in g.integers(0, ncells, indegree):
            delay = float(g.uniform(0.0, 2.0 * DT)) if short else float(g.uniform(1.0, 3.0))
            if short and n % 4 == 0:
                delay = 0.0
            m.connect(srcs[s], syns[t], weight=float(g.uniform(2e-3, 6e-3)), delay=delay)
            n += 1
        ns = m.add_pointprocess(m, "NetStim", params={"interval": 20.0, "number": 1e9, "start": 0.0, "noise": 1.0})
        m.connect(ns, syns[t], weight=2e-2, delay=0.0 if short else 1.0)
    gated = []
    if feedback:
        for k in range(6):
            ns = m.add_pointprocess(m, "NetStim", params={"interval": 3.0, "number": 1e9, "start": -1.0,
                                                          "noise": 0.3 * (k % 2)})
            m.connect(srcs[k], ns, weight=1.0, delay=0.0 if k % 2 else 2.0)  # on
            m.connect(srcs[k + 6], ns, weight=-1.0, delay=1.0)            # off
            for j in range(3):
                m.connect(ns, syns[(5 * k + j) % ncells], weight=1.5e-2, delay=0.5 if short else 1.5)
            gated.append(ns)
    return m, cells, syns, gated


def _compare(m, cells, syns, gated, method, stop=80.0):
    recs = [voltage_record(m, cells[k], "soma", 0.5, name=f"v{k}") for k in range(3)]
    recs.append(mechanism_record(m, syns[0], "A", name="A0"))
    recs += [mechanism_record(m, ns, "on", name=f"on{k}") for k, ns in enumerate(gated)]
    print(f"case: event-exact {method} cells={len(cells)} gated={len(gated)} stop={stop}", flush=True)
    build(m, dt_micro=DT, seed=1, method=method)
    initialize(m)
    got = run(m, stop, recs)
    v_got, a_got = read_voltage(m), read_field(m, "Exp2Syn", "A")
    destroy(m)
    sim = oracle.OracleSim(m).initialize()
    ref = sim.run(stop, recs)
    exp2 = builtin("Exp2Syn")
    assert np.array_equal(got.spike_times, ref.spike_times)
    assert np.array_equal(got.spike_sources, ref.spike_sources)
    assert np.array_equal(v_got, sim.voltage())
    assert np.array_equal(a_got, sim.field(exp2.id, exp2.field_index("A")))
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    return got


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_short_and_zero_delays(gpu, method):
    m, cells, syns, gated = _network()
    got = _compare(m, cells, syns, gated, method)
    assert len(got.spike_times) > 30


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_connections_into_netstims(gpu, method):
    m, cells, syns, gated = _network(short=False, feedback=True)
    got = _compare(m, cells, syns, gated, method, stop=100.0)
    on = np.array([got.traces[f"on{k}"] for k in range(6)])
    assert on.max() == 1.0  # some gated NetStim was switched on


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_short_delays_with_feedback(gpu, method):
    m, cells, syns, gated = _network(feedback=True)
    _compare(m, cells, syns, gated, method, stop=60.0)


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_injection_and_checkpoint_on_event_exact_path(gpu, method):
    """A NetStim switched on at 5 ms and off at 8 ms by injected spikes fires at 5 and 7
    ms; a checkpoint taken at 6 ms and restored into a second state continues bit for
    bit."""
    m = MicroModel()
    (cell,) = m.add_population(configs.ball_and_stick_sections(10), 1)
    m.insert(cell, builtin("pas"), sections="all", params={"g": 1e-4, "e": -65.0})
    syn = m.add_pointprocess(cell, "ExpSyn", section="soma", params={"tau": 2.0, "e": 0.0})
    ns = m.add_pointprocess(m, "NetStim", params={"interval": 2.0, "number": 10, "start": -1.0, "noise": 0.0})
    out = m.register_spike_source(ns, sid=3)
    on, off = m.register_spike_source(m, sid=1), m.register_spike_source(m, sid=2)
    m.connect(on, ns, weight=1.0, delay=0.0)
    m.connect(off, ns, weight=-1.0, delay=0.0)
    m.connect(out, syn, weight=1e-3, delay=0.0)
    recs = [mechanism_record(m, syn, "g", name="g"), mechanism_record(m, ns, "ispike", name="ispike")]
    print(f"case: event-exact injection/checkpoint {method}", flush=True)
    build(m, dt_micro=DT, seed=1, method=method)
    initialize(m)
    inject_spike(m, on, 5.0)
    inject_spike(m, off, 8.0)
    first = run(m, 6.0, recs)
    blob = checkpoint(m)
    rest = run(m, 20.0, recs)
    sim = oracle.OracleSim(m).initialize()
    sim.inject_spike(on, 5.0)
    sim.inject_spike(off, 8.0)
    ref1 = sim.run(6.0, recs)
    ref2 = sim.run(20.0, recs)
    for got, ref in ((first, ref1), (rest, ref2)):
        assert np.array_equal(got.spike_times, ref.spike_times)
        assert np.array_equal(got.spike_sources, ref.spike_sources)
        for r in recs:
            assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    ns_times = [t for t, s in zip(np.concatenate([first.spike_times, rest.spike_times]),
                                  np.concatenate([first.spike_sources, rest.spike_sources])) if s == 3]
    assert ns_times == [5.0, 7.0]
    # restore into a fresh state of the same model and continue from 6 ms
    m2 = m
    build(m2, dt_micro=DT, seed=1, method=method)
    initialize(m2)
    restore(m2, blob)
    again = run(m2, 20.0, recs)
    assert np.array_equal(again.spike_times, rest.spike_times)
    for r in recs:
        assert np.array_equal(again.traces[r.name], rest.traces[r.name]), r.name
    destroy(m2)


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_more_netstims_due_in_a_window_than_lds_holds(gpu, method):
    """Window path (delays of 1 ms): 3000 NetStims fall due together at 1 ms and every
    2 ms after, half of them with noise, more than the 2048 events the exchange kernel
    orders in LDS, so it orders them in its HBM heap. Rasters, voltages, synapse
    states and a NetStim record bit for bit."""
    m = MicroModel()
    cells = m.add_population(configs.ball_and_stick_sections(10), 20)
    syns, stims = [], []
    for c in cells:
        m.insert(c, builtin("hh"), sections="soma")
        m.insert(c, builtin("pas"), sections="dend", params={"g": 1e-4, "e": -65.0})
        syns.append(m.add_pointprocess(c, "Exp2Syn", section="dend", x=0.5, params={"tau1": 0.5, "tau2": 5.0,
                                                                                  "e": 0.0}))
    for k in range(3000):
        ns = m.add_pointprocess(m, "NetStim", params={"interval": 2.0, "number": 1e9, "start": 1.0,
                                                      "noise": 0.5 * (k % 2)})
        m.connect(ns, syns[k % len(syns)], weight=2e-4, delay=1.0)
        stims.append(ns)
    recs = [voltage_record(m, cells[k], "soma", 0.5, name=f"v{k}") for k in range(2)]
    recs.append(mechanism_record(m, stims[1], "ispike", name="ispike"))
    print(f"case: 3000 NetStims due in one window, {method}", flush=True)
    build(m, dt_micro=DT, seed=3, method=method)
    initialize(m)
    got = run(m, 12.0, recs)
    v_got, a_got = read_voltage(m), read_field(m, "Exp2Syn", "A")
    destroy(m)
    sim = oracle.OracleSim(m, seed=3).initialize()
    ref = sim.run(12.0, recs)
    exp2 = builtin("Exp2Syn")
    assert np.array_equal(got.spike_times, ref.spike_times)
    assert np.array_equal(got.spike_sources, ref.spike_sources)
    assert np.array_equal(v_got, sim.voltage())
    assert np.array_equal(a_got, sim.field(exp2.id, exp2.field_index("A")))
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    first = m.register_spike_source(stims[0]).sid  # the source its connection made
    assert np.count_nonzero(got.spike_sources == first) == 6  # noiseless: 1, 3, ..., 11 ms
    assert len(got.spike_times) > 3000 * 5
