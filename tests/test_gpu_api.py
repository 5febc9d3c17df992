"""The state API beyond run!/step! through libmind on cuda:0, against the oracle:
read_value, connection_weights, section/ion selectors, and injected events and
spikes (inject_event!, inject_spike!) across split runs, in both kernels."""
import numpy as np
import pytest

import oracle
from mind_amd import (build, checkpoint, configs, connection_weights, destroy, initialize, inject_event,
                      inject_spike, ion_record, mechanism_record, read_value, read_voltage, restore, run,
                      section, voltage_record)
from mind_amd.mechanisms import builtin

pytestmark = pytest.mark.gpu


def _network():
    """A small ball-and-stick network with NetStim drive, recurrent connections and an
    external spike input (SID 900) connected to four synapses with delays 1-4 ms."""
    model, info = configs.ball_stick_network(ncells=20, indegree=5, rate_hz=40.0, netstim_weight=2e-2)
    ext = model.register_spike_source(model, sid=900)
    conns = [model.connect(ext, info["syns"][k], weight=3e-2, delay=1.0 + k) for k in range(4)]
    return model, info, ext, conns


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_injected_events_and_spikes_match_oracle(gpu, method):
    model, info, ext, conns = _network()
    cells = info["cells"]
    recs = [voltage_record(model, c, "soma", 0.5) for c in cells[:4]]
    recs.append(mechanism_record(model, info["syns"][1], "B"))
    print(f"case: injection {method}", flush=True)
    build(model, dt_micro=0.025, seed=1, method=method)
    initialize(model)
    sim = oracle.OracleSim(model).initialize()
    got, ref = [], []
    # two spikes in one window, one on a step boundary, one beyond the delivery
    # ring's horizon; events near and far, one at the current time after a run
    plan = [(10.0, [("s", 12.3), ("s", 12.6), ("s", 15.0), ("s", 150.0), ("e", conns[0], 11.0),
                    ("e", conns[2], 100.0), ("e", conns[3], 10.0)]),
            (60.0, [("s", 60.0), ("e", conns[1], 60.0)]),
            (200.0, [])]
    for stop, injections in plan:
        got.append(run(model, stop, recs))
        ref.append(sim.run(stop, recs))
        for inj in injections:
            if inj[0] == "s":
                inject_spike(model, ext, inj[1])
                sim.inject_spike(ext, inj[1])
            else:
                inject_event(model, inj[1], inj[2])
                sim.inject_event(inj[1], inj[2])
    for g, r in zip(got, ref):
        for rec in recs:
            assert np.array_equal(g.traces[rec.name], r.traces[rec.name]), rec.name
        assert np.array_equal(g.spike_times, r.spike_times)
        assert np.array_equal(g.spike_sources, r.spike_sources)
    raster = np.concatenate([g.spike_times[g.spike_sources == 900] for g in got])
    assert raster.tolist() == [12.3, 12.6, 15.0, 60.0, 150.0]
    with pytest.raises(Exception):
        inject_spike(model, ext, 100.0)  # before the current time
    # initialize clears pending input
    inject_spike(model, ext, 250.0)
    initialize(model)
    again = run(model, 10.0, recs)
    assert np.array_equal(again.traces[recs[0].name], got[0].traces[recs[0].name])
    destroy(model)


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_thousands_of_injected_spikes_in_one_window(gpu, method):
    """3000 injected spikes due within 0.3 ms (more than one window's exchange takes):
    the run shortens its windows so that each hands at most 1024 to the exchange, with
    results identical to the oracle's; more than 1024 due in one step are refused."""
    model, info, ext, conns = _network()
    recs = [voltage_record(model, c, "soma", 0.5) for c in info["cells"][:4]]
    print(f"case: injection burst {method}", flush=True)
    build(model, dt_micro=0.025, seed=1, method=method)
    initialize(model)
    sim = oracle.OracleSim(model).initialize()
    times = 20.0 + 1e-4 * np.arange(3000)
    got, ref = [run(model, 15.0, recs)], [sim.run(15.0, recs)]
    for t in times:
        inject_spike(model, ext, float(t))
        sim.inject_spike(ext, float(t))
    got.append(run(model, 40.0, recs))
    ref.append(sim.run(40.0, recs))
    for g, r in zip(got, ref):
        for rec in recs:
            assert np.array_equal(g.traces[rec.name], r.traces[rec.name]), rec.name
        assert np.array_equal(g.spike_times, r.spike_times)
        assert np.array_equal(g.spike_sources, r.spike_sources)
    assert np.array_equal(np.sort(got[1].spike_times[got[1].spike_sources == 900]), times)
    for k in range(1100):  # all due within one step
        inject_spike(model, ext, 50.0 + 1e-6 * k)
    with pytest.raises(Exception):
        run(model, 60.0, recs)
    destroy(model)


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_read_value_and_selectors_match_oracle(gpu, method):
    model, info, ext, conns = _network()
    cell = info["cells"][3]
    recs = [voltage_record(model, section(cell, "dend"), x=0.3),
            mechanism_record(model, section(cell, "soma"), "hh", "n"),
            ion_record(model, cell, "ina"), ion_record(model, section(cell, "soma"), "ek", name="ek"),
            mechanism_record(model, info["syns"][3], "A")]
    assert recs[0].name == f"{cell.gid}.dend.v" and recs[1].name == f"{cell.gid}.soma.n"
    print(f"case: selectors {method}", flush=True)
    build(model, dt_micro=0.025, seed=1, method=method)
    initialize(model)
    got = run(model, 30.0, recs)
    sim = oracle.OracleSim(model).initialize()
    ref = sim.run(30.0, recs)
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    assert np.all(got.traces["ek"] == -77.0)
    # read_value: the current state, voltage and fields
    v = sim.voltage()
    node = sim.location_node(cell.index, 1, 0.3)
    assert read_value(model, recs[0]) == v[node]
    syn = info["syns"][3]
    a = sim.field(builtin("Exp2Syn").id, builtin("Exp2Syn").field_index("A"), len(info["syns"]))
    assert read_value(model, recs[4]) == a[sim.point_row(syn)]
    ns = info["stims"][5]
    ev = mechanism_record(model, ns, "event")
    nsf = sim.field(builtin("NetStim").id, builtin("NetStim").field_index("event"), len(info["stims"]))
    assert read_value(model, ev) == nsf[sim.point_row(ns)]
    # connection weights: the built weight of each connection
    assert connection_weights(model, conns[2]).tolist() == [3e-2]
    assert connection_weights(model, 1).tolist() == [model.connections[0][2]]
    destroy(model)


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_checkpoint_resumes_bit_for_bit(gpu, method):
    """A checkpoint taken between runs (with events in flight and injected input still
    pending) restored into a second state of the same model continues exactly as the
    original; other models and truncated blobs are rejected."""
    def records(model, info):
        return [voltage_record(model, c, "soma", 0.5) for c in info["cells"][:3]] + \
               [mechanism_record(model, info["syns"][2], "A")]

    model, info, ext, conns = _network()
    recs = records(model, info)
    print(f"case: checkpoint {method}", flush=True)
    build(model, dt_micro=0.025, seed=1, method=method)
    initialize(model)
    run(model, 20.0, recs)
    inject_spike(model, ext, 150.0)
    inject_event(model, conns[2], 90.0)
    blob = checkpoint(model)
    a = run(model, 200.0, recs)
    va = read_voltage(model)

    model2, info2, ext2, _ = _network()
    recs2 = records(model2, info2)
    build(model2, dt_micro=0.025, seed=1, method=method)
    restore(model2, blob)
    b = run(model2, 200.0, recs2)
    for r, r2 in zip(recs, recs2):
        assert np.array_equal(a.traces[r.name], b.traces[r2.name]), r.name
    assert np.array_equal(a.spike_times, b.spike_times) and np.array_equal(a.spike_sources, b.spike_sources)
    assert 150.0 in b.spike_times.tolist()
    assert np.array_equal(va, read_voltage(model2))
    with pytest.raises(Exception):
        restore(model2, blob[:-8])
    other, _ = configs.single_cell()
    build(other, dt_micro=0.025, seed=1, method=method)
    with pytest.raises(Exception):
        restore(other, blob)
    for m in (model, model2, other):
        destroy(m)


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_pattern_stim_matches_oracle(gpu, method):
    """PatternStim drives the network alongside NetStim: spikes at t = 0, on step
    boundaries, tied across points, dense (about 300 in one window) and beyond the
    delivery ring's horizon; traces and raster bit-identical to the oracle, and a second
    initialize replays the pattern."""
    model, info = configs.ball_stick_network(ncells=20, indegree=5, rate_hz=40.0, netstim_weight=2e-2)
    rng = np.random.default_rng(3)
    dense = np.sort(np.round(rng.uniform(30.0, 31.0, 300), 3))
    patterns = [[0.0, 5.0, 12.3, 150.0], [5.0, 7.5, 40.0], dense.tolist()]
    for k, times in enumerate(patterns):
        p = model.add_pointprocess(model, "PatternStim", spike_times=times)
        src = model.register_spike_source(p, sid=800 + k)
        for j in range(3):
            model.connect(src, info["syns"][3 * k + j], weight=1e-2, delay=0.5 + j)
    recs = [voltage_record(model, c, "soma", 0.5) for c in info["cells"][:9]]
    recs.append(mechanism_record(model, info["syns"][0], "B"))
    print(f"case: PatternStim {method}", flush=True)
    build(model, dt_micro=0.025, seed=1, method=method)
    initialize(model)
    got = [run(model, 20.0, recs), run(model, 200.0, recs)]
    sim = oracle.OracleSim(model).initialize()
    ref = [sim.run(20.0, recs), sim.run(200.0, recs)]
    for g, r in zip(got, ref):
        for rec in recs:
            assert np.array_equal(g.traces[rec.name], r.traces[rec.name]), rec.name
        assert np.array_equal(g.spike_times, r.spike_times)
        assert np.array_equal(g.spike_sources, r.spike_sources)
    for k, times in enumerate(patterns):
        emitted = np.concatenate([g.spike_times[g.spike_sources == 800 + k] for g in got])
        assert emitted.tolist() == list(times)
    initialize(model)
    again = run(model, 20.0, recs)
    assert np.array_equal(again.spike_times, got[0].spike_times)
    destroy(model)


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_network_of_artificial_cells_only(gpu, method):
    """A model without compartments: a PatternStim and two NetStims (one regular, one
    Poisson). Rasters and NetStim field records bit-identical to the oracle in both
    kernels, on the window path and, once a connection goes into a NetStim, on the
    event-exact path."""
    from mind_amd import MicroModel
    m = MicroModel()
    ps = m.add_pointprocess(m, "PatternStim", spike_times=[1.0, 2.5, 30.0])
    regular = m.add_pointprocess(m, "NetStim", params={"interval": 5.0, "number": 3, "start": 0.0, "noise": 0.0})
    poisson = m.add_pointprocess(m, "NetStim", params={"interval": 4.0, "number": 4, "start": 50.0, "noise": 0.5})
    for k, p in enumerate((ps, regular, poisson)):
        m.register_spike_source(p, sid=10 + k)
    recs = []
    print(f"case: artificial cells only {method}", flush=True)
    build(m, dt_micro=0.025, seed=3, method=method)
    initialize(m)
    got = run(m, 100.0, recs)
    ref = oracle.OracleSim(m, seed=3).initialize().run(100.0, recs)
    assert np.array_equal(got.spike_times, ref.spike_times)
    assert np.array_equal(got.spike_sources, ref.spike_sources)
    assert got.spike_sources.tolist().count(10) == 3 and got.spike_sources.tolist().count(11) == 3
    assert got.spike_sources.tolist().count(12) == 4
    recs = [mechanism_record(m, poisson, "event", name="event"), mechanism_record(m, regular, "on", name="on"),
            mechanism_record(m, poisson, "ispike", name="ispike")]
    initialize(m)
    got = run(m, 110.0, recs)
    ref = oracle.OracleSim(m, seed=3).initialize().run(110.0, recs)
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    assert got.traces["ispike"][-1] == 4.0
    destroy(m)
    # an input into a NetStim (a feedback receiver) moves the model to the event-exact path
    m.connect(m.register_spike_source(ps), regular, weight=1.0, delay=0.5)
    build(m, dt_micro=0.025, seed=3, method=method)
    initialize(m)
    got = run(m, 110.0, recs)
    ref = oracle.OracleSim(m, seed=3).initialize().run(110.0, recs)
    assert np.array_equal(got.spike_times, ref.spike_times)
    assert np.array_equal(got.spike_sources, ref.spike_sources)
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    with pytest.raises(ValueError, match="NET_RECEIVE"):  # a PatternStim receives no events
        m.connect(m.register_spike_source(regular), ps, weight=1.0, delay=1.0)
    destroy(m)
