"""Runtime-compiled mechanisms on the GPU: MOD files lowered by mind_compile_mod,
their cell kernels compiled with hipRTC at build, against the oracle running a
g++ build of the same lowering, bit for bit, in both kernels; and against the
builtin mechanisms they restate (hh's n gate, ExpSyn), bit for bit."""
import os

import numpy as np
import pytest

import oracle
from mind_amd import (MicroModel, build, configs, destroy, initialize, load_mod, mechanism_record, read_field,
                      read_voltage, run, voltage_record)
from mind_amd.mechanisms import builtin

pytestmark = pytest.mark.gpu
DT = 0.025
MODS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mod")


def _mod(name):
    return load_mod(os.path.join(MODS, f"{name}.mod"))


def _k_network(user, ncells=24, delay=1.0):
    """Ball-and-stick cells with a delayed-rectifier K channel on the soma (the
    compiled kdr, or hh with gnabar = gl = 0), pas everywhere, a current clamp and
    an exponential synapse (compiled ExpDecaySyn, or ExpSyn) driven by a Poisson
    NetStim and by the somata of two other cells."""
    kdr, syn2 = _mod("kdr"), _mod("expsyn2")
    m = MicroModel()
    cells = m.add_population(configs.ball_and_stick_sections(12), ncells)
    syns, srcs = [], []
    for k, c in enumerate(cells):
        if user:
            m.insert(c, kdr, sections="soma", params={"gkbar": 0.03 + 0.001 * (k % 5)})
        else:
            m.insert(c, builtin("hh"), sections="soma",
                     params={"gnabar": 0.0, "gl": 0.0, "gkbar": 0.03 + 0.001 * (k % 5)})
        m.insert(c, builtin("pas"), sections="all", params={"g": 1e-4, "e": -65.0})
        m.add_pointprocess(c, "IClamp", section="soma", params={"del": 1.0 + k % 3, "dur": 40.0, "amp": 0.25})
        syns.append(m.add_pointprocess(c, syn2 if user else "ExpSyn", section="dend", x=0.5,
                                       params={"tau": 1.5 + 0.1 * (k % 4), "e": 0.0}))
        srcs.append(m.register_spike_source(c, section="soma", x=0.5, threshold=-50.0))
    for k in range(ncells):
        ns = m.add_pointprocess(m, "NetStim", params={"interval": 4.0, "number": 1e9, "start": 0.5, "noise": 0.7})
        m.connect(ns, syns[k], weight=0.012, delay=delay)
        for j in (1, 5):
            m.connect(srcs[(k + j) % ncells], syns[k], weight=0.004, delay=delay + 0.5)
    return m, cells, syns


def _run_gpu(m, recs, method, stop):
    build(m, dt_micro=DT, seed=1, method=method)
    initialize(m)
    got = run(m, stop, recs)
    v = read_voltage(m)
    return got, v


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_compiled_k_channel_and_synapse_match_oracle_and_builtins(gpu, method):
    m, cells, syns = _k_network(True)
    recs = [voltage_record(m, cells[k], "soma", 0.5, name=f"v{k}") for k in range(4)]
    recs.append(mechanism_record(m, syns[0], "g", name="g0"))
    print(f"case: compiled kdr + ExpDecaySyn, {method}", flush=True)
    got, v_got = _run_gpu(m, recs, method, 60.0)
    syn2 = _mod("expsyn2")
    g_got = read_field(m, syn2, "g")
    destroy(m)
    sim = oracle.OracleSim(m).initialize()
    ref = sim.run(60.0, recs)
    assert np.array_equal(v_got, sim.voltage())
    assert np.array_equal(g_got, sim.field(syn2.id, syn2.field_index("g")))
    assert np.array_equal(got.spike_times, ref.spike_times)
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    # the builtin restatement of the same model on the GPU
    b, bcells, bsyns = _k_network(False)
    brecs = [voltage_record(b, bcells[k], "soma", 0.5, name=f"v{k}") for k in range(4)]
    print(f"case: builtin hh(K only) + ExpSyn, {method}", flush=True)
    bgot, v_b = _run_gpu(b, brecs, method, 60.0)
    destroy(b)
    assert np.array_equal(v_got, v_b)
    assert np.array_equal(got.spike_times, bgot.spike_times)
    assert len(got.spike_times) > 0


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_compiled_euler_channel_with_array_field(gpu, method):
    """nap: forward Euler, an array parameter set in INITIAL, a FUNCTION; with hh
    spiking cells and Exp2Syn connections."""
    nap = _mod("nap")
    m = MicroModel()
    cells = m.add_population(configs.ball_and_stick_sections(16), 20)
    syns, srcs = [], []
    for k, c in enumerate(cells):
        m.insert(c, builtin("hh"), sections="soma")
        m.insert(c, nap, sections="all", params={"gbar": 2e-4 + 1e-5 * k})
        m.insert(c, builtin("pas"), sections="dend", params={"g": 1e-4, "e": -65.0})
        syns.append(m.add_pointprocess(c, "Exp2Syn", section="dend", x=0.3,
                                       params={"tau1": 0.5, "tau2": 4.0, "e": 0.0}))
        srcs.append(m.register_spike_source(c, section="soma", x=0.5, threshold=-20.0))
    g = np.random.default_rng(3)
    for k in range(20):
        for s in g.integers(0, 20, 4):
            m.connect(srcs[s], syns[k], weight=float(g.uniform(2e-3, 5e-3)), delay=float(g.uniform(1.0, 3.0)))
        ns = m.add_pointprocess(m, "NetStim", params={"interval": 8.0, "number": 1e9, "start": 0.0, "noise": 1.0})
        m.connect(ns, syns[k], weight=0.02, delay=1.0)
    recs = [voltage_record(m, cells[k], "soma", 0.5, name=f"v{k}") for k in range(3)]
    recs.append(mechanism_record(m, cells[2], nap, "p", section="dend", x=0.5, name="p"))
    print(f"case: compiled nap (euler, array) + hh, {method}", flush=True)
    got, v_got = _run_gpu(m, recs, method, 50.0)
    p_got = read_field(m, nap, "p")
    destroy(m)
    sim = oracle.OracleSim(m).initialize()
    ref = sim.run(50.0, recs)
    assert np.array_equal(v_got, sim.voltage())
    assert np.array_equal(p_got, sim.field(nap.id, nap.field_index("p")))
    assert np.array_equal(got.spike_times, ref.spike_times)
    assert np.array_equal(got.spike_sources, ref.spike_sources)
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    assert len(got.spike_times) > 10


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_compiled_synapse_on_event_exact_path(gpu, method):
    """Zero and sub-step NetCon delays into the compiled synapse: the event-exact
    path's queue kernel runs the compiled NET_RECEIVE."""
    m, cells, syns = _k_network(True, ncells=12, delay=0.0)
    recs = [voltage_record(m, cells[k], "soma", 0.5, name=f"v{k}") for k in range(3)]
    print(f"case: compiled synapse, zero delays, {method}", flush=True)
    got, v_got = _run_gpu(m, recs, method, 30.0)
    destroy(m)
    sim = oracle.OracleSim(m).initialize()
    ref = sim.run(30.0, recs)
    assert np.array_equal(v_got, sim.voltage())
    assert np.array_equal(got.spike_times, ref.spike_times)
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_compiled_globals_and_global_record(gpu, method):
    """A compiled leak whose conductance and reversal are GLOBALs set for every
    instance: global_record reads the shared value at every sample, and voltages,
    spikes and records equal the oracle's and the same model with pas, bit for bit."""
    from mind_amd import global_record

    def network(user, ncells=16, g=3e-4, e=-61.0):
        m = MicroModel()
        cells = m.add_population(configs.ball_and_stick_sections(10), ncells)
        leak = _mod("leakg") if user else builtin("pas")
        if user:
            m.set_global(leak, "gl", g)
            m.set_global(leak, "el", e)
        for k, c in enumerate(cells):
            m.insert(c, builtin("hh"), sections="soma")
            m.insert(c, leak, sections="all", params={} if user else {"g": g, "e": e})
            m.add_pointprocess(c, "IClamp", section="soma", params={"del": 1.0 + k % 4, "dur": 20.0, "amp": 0.3})
            m.register_spike_source(c, section="soma", x=0.5, threshold=-20.0)
        return m, cells, leak

    m, cells, leak = network(True)
    recs = [global_record(m, leak, "gl", name="gl"), global_record(m, leak, "el", name="el"),
            mechanism_record(m, cells[5], leak, "i", section="dend", name="i5")]
    recs += [voltage_record(m, cells[k], "soma", 0.5, name=f"v{k}") for k in range(4)]
    print(f"case: compiled GLOBAL leak, {method}", flush=True)
    got, v_got = _run_gpu(m, recs, method, 30.0)
    destroy(m)
    sim = oracle.OracleSim(m).initialize()
    ref = sim.run(30.0, recs)
    assert np.all(got.traces["gl"] == 3e-4) and np.all(got.traces["el"] == -61.0)
    assert np.array_equal(v_got, sim.voltage())
    assert np.array_equal(got.spike_times, ref.spike_times)
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    p, pcells, _ = network(False)
    prec = [voltage_record(p, pcells[k], "soma", 0.5, name=f"v{k}") for k in range(4)]
    pgot, pv = _run_gpu(p, prec, method, 30.0)
    destroy(p)
    assert np.array_equal(v_got, pv)
    assert np.array_equal(got.spike_times, pgot.spike_times)
    assert len(got.spike_times) > 10
