"""Parity with the oracle at the benchmark workloads' own sizes.

* The 1000-cell ball-and-stick network (hh soma, 50-segment pas dendrite, Exp2Syn,
  in-degree 100, lognormal weights, delays U(1, 5) ms, 10 Hz Poisson NetStim per
  cell) over its whole 1000 ms, in both kernels.
* The 5000-tree network (lognormal 50-4000 compartments, 20 Exp2Syn per cell,
  in-degree 50, 5 Hz NetStim) over 10 ms (400 steps, so that events from every
  delay in 1-5 ms are delivered), in the batched kernel with its autotuned plan
  and in the plain kernel.
* ExpSyn synapses driven by NetStims and recurrent connections, both kernels.

Compared: every node's voltage at the end (bit for bit), soma traces, the raster
with its source IDs, and the synapse state columns.
"""
import numpy as np
import pytest

import oracle
from mind_amd import (MicroModel, batch_plan, build, configs, destroy, initialize, mechanism_record,
                      read_field, read_voltage, run, voltage_record)
from mind_amd.mechanisms import builtin

pytestmark = pytest.mark.gpu


def _check(model, stop, records, method, fields=()):
    print(f"case: {method} stop={stop} records={len(records)}", flush=True)
    build(model, dt_micro=0.025, seed=1, method=method)
    if method == "batched":
        print(f"plan: {batch_plan(model)}", flush=True)
    initialize(model)
    got = run(model, stop, records)
    v_got = read_voltage(model)
    f_got = {f: read_field(model, *f) for f in fields}
    destroy(model)
    return got, v_got, f_got


def _reference(model, stop, records, fields=()):
    sim = oracle.OracleSim(model).initialize()
    ref = sim.run(stop, records)
    out = {}
    for mech, fname in fields:
        m = builtin(mech)
        out[(mech, fname)] = sim.field(m.id, m.field_index(fname))
    return ref, sim.voltage(), out


def _same(got, ref, records, v_got, v_ref, f_got, f_ref):
    assert np.array_equal(got.spike_times, ref.spike_times)
    assert np.array_equal(got.spike_sources, ref.spike_sources)
    assert np.max(np.abs(v_got - v_ref)) <= 1e-6
    assert np.array_equal(v_got, v_ref)
    for r in records:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    for k in f_ref:
        assert np.array_equal(f_got[k], f_ref[k]), k


@pytest.fixture(scope="module")
def ball_stick():
    model, info = configs.ball_stick_network()
    recs = [voltage_record(model, c, "soma", 0.5, name=f"v{i}") for i, c in enumerate(info["cells"][:4])]
    recs.append(mechanism_record(model, info["syns"][0], "g", name="g0"))
    fields = (("Exp2Syn", "A"), ("Exp2Syn", "B"), ("hh", "m"))
    ref, v_ref, f_ref = _reference(model, info["tstop"], recs, fields)
    print(f"oracle: {len(ref.spike_times)} spikes in {info['tstop']} ms ({ref.run_seconds:.1f} s)", flush=True)
    return model, info, recs, fields, ref, v_ref, f_ref


@pytest.mark.parametrize("method", ["batched", "plain"])
def test_ball_stick_1000_cells_full_run(gpu, ball_stick, method):
    """1000 ball-and-stick cells, in-degree 100, for the whole 1000 ms."""
    model, info, recs, fields, ref, v_ref, f_ref = ball_stick
    assert len(info["cells"]) == 1000 and info["tstop"] == 1000.0
    got, v_got, f_got = _check(model, info["tstop"], recs, method, fields)
    assert len(ref.spike_times) > 1000
    _same(got, ref, recs, v_got, v_ref, f_got, f_ref)


@pytest.fixture(scope="module")
def trees():
    model, info = configs.tree_network()
    recs = [voltage_record(model, c, "soma", 0.5, name=f"v{i}") for i, c in enumerate(info["cells"][:4])]
    fields = (("Exp2Syn", "A"), ("Exp2Syn", "B"), ("hh", "h"))
    stop = 10.0
    ref, v_ref, f_ref = _reference(model, stop, recs, fields)
    print(f"oracle: {len(ref.spike_times)} spikes in {stop} ms ({ref.run_seconds:.1f} s)", flush=True)
    return model, info, recs, fields, stop, ref, v_ref, f_ref


@pytest.mark.parametrize("method", ["batched", "plain"])
def test_trees_5000_cells(gpu, trees, method):
    """5000 uneven trees (2.7 M compartments) over 400 steps, so that NetCon and
    NetStim events of every delay are delivered."""
    model, info, recs, fields, stop, ref, v_ref, f_ref = trees
    assert len(info["cells"]) == 5000
    got, v_got, f_got = _check(model, stop, recs, method, fields)
    assert len(ref.spike_times) > 100
    _same(got, ref, recs, v_got, v_ref, f_got, f_ref)


def _expsyn_network(ncells=40, indegree=10, seed=5):
    model = MicroModel()
    cells = model.add_population(configs.ball_and_stick_sections(50), ncells)
    syns, srcs = [], []
    for c in cells:
        model.insert(c, builtin("hh"), sections="soma")
        model.insert(c, builtin("pas"), sections="dend", params={"g": 1e-4, "e": -65.0})
        syns.append(model.add_pointprocess(c, "ExpSyn", section="dend", x=0.1,
                                           params={"tau": 2.0, "e": 0.0}))
        srcs.append(model.register_spike_source(c, section="soma", x=0.5, threshold=-20.0))
    g = np.random.default_rng(seed)
    for t in range(ncells):
        for s in g.integers(0, ncells, indegree):
            model.connect(srcs[s], syns[t], weight=float(g.uniform(1e-3, 4e-3)), delay=float(g.uniform(1.0, 3.0)))
        ns = model.add_pointprocess(model, "NetStim", params={"interval": 25.0, "number": 1e9, "start": 0.0,
                                                             "noise": 1.0})
        model.connect(ns, syns[t], weight=2e-2, delay=1.0)
    return model, cells, syns


@pytest.mark.parametrize("method", ["batched", "plain"])
def test_expsyn_driven_by_netstim(gpu, method):
    """ExpSyn (g += w on an event, g decays with tau) under NetStim drive and
    recurrent input: state column g, current i and the raster."""
    model, cells, syns = _expsyn_network()
    recs = [voltage_record(model, cells[0], "soma", 0.5, name="v0"),
            mechanism_record(model, syns[0], "g", name="g0"), mechanism_record(model, syns[1], "i", name="i1")]
    fields = (("ExpSyn", "g"), ("ExpSyn", "i"))
    ref, v_ref, f_ref = _reference(model, 100.0, recs, fields)
    got, v_got, f_got = _check(model, 100.0, recs, method, fields)
    assert len(ref.spike_times) > 40 and ref.traces["g0"].max() > 1e-2
    _same(got, ref, recs, v_got, v_ref, f_got, f_ref)
