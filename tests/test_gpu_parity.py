"""Parity of the HIP path with the CPU oracle (bit-exact voltages and rasters)."""
import numpy as np
import pytest

import oracle
from mind_amd import (build, configs, destroy, initialize, mechanism_record, read_field,
                      read_voltage, run, voltage_record)

pytestmark = pytest.mark.gpu


def _compare(model, stop, records, method="plain", **kw):
    print(f"case: {method} stop={stop} records={len(records)} {kw}", flush=True)
    build(model, dt_micro=0.025, seed=1, method=method, **kw)
    initialize(model)
    got = run(model, stop, records)
    ref_sim = oracle.OracleSim(model).initialize()
    ref = ref_sim.run(stop, records)
    for r in records:
        a, b = got.traces[r.name], ref.traces[r.name]
        assert np.max(np.abs(a - b)) <= 1e-6, r.name
    assert np.array_equal(got.spike_times, ref.spike_times)
    assert np.array_equal(got.spike_sources, ref.spike_sources)
    v_got, v_ref = read_voltage(model), ref_sim.voltage()
    assert np.max(np.abs(v_got - v_ref)) <= 1e-6
    return got, ref, v_got, v_ref, ref_sim


def test_single_cell_plain(gpu):
    model, info = configs.single_cell()
    cell = info["cells"][0]
    recs = [voltage_record(model, cell, "soma", 0.5),
            voltage_record(model, cell, "dend", 0.5, name="vd"),
            mechanism_record(model, cell, "hh", "m", section="soma", x=0.5),
            mechanism_record(model, info["clamp"], "i")]
    got, ref, vg, vr, _ = _compare(model, 200.0, recs)
    assert len(got.spike_times) == 9
    # bit-exact: identical IEEE operation sequence on both sides
    assert np.array_equal(vg, vr)
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    destroy(model)


def test_ball_stick_small_plain(gpu):
    model, info = configs.ball_stick_network(ncells=40, indegree=10, rate_hz=40.0,
                                             netstim_weight=2e-2)
    recs = [voltage_record(model, c, "soma", 0.5) for c in info["cells"][:4]]
    recs.append(mechanism_record(model, info["syns"][0], "A"))
    got, ref, vg, vr, ref_sim = _compare(model, 60.0, recs)
    assert np.array_equal(vg, vr)
    assert len(got.spike_times) > 0
    a = read_field(model, "Exp2Syn", "A")
    b = ref_sim.field(3, 3, len(a))
    assert np.array_equal(a, b)
    destroy(model)


def test_single_cell_batched(gpu):
    model, info = configs.single_cell()
    cell = info["cells"][0]
    recs = [voltage_record(model, cell, "soma", 0.5),
            mechanism_record(model, cell, "hh", "n", section="soma", x=0.5)]
    got, ref, vg, vr, _ = _compare(model, 200.0, recs, method="batched")
    assert np.array_equal(vg, vr)
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    destroy(model)


def test_ball_stick_small_batched(gpu):
    model, info = configs.ball_stick_network(ncells=40, indegree=10, rate_hz=40.0,
                                             netstim_weight=2e-2)
    recs = [voltage_record(model, c, "soma", 0.5) for c in info["cells"][:4]]
    got, ref, vg, vr, ref_sim = _compare(model, 60.0, recs, method="batched")
    assert np.array_equal(vg, vr)
    a = read_field(model, "Exp2Syn", "B")
    assert np.array_equal(a, ref_sim.field(3, 4, len(a)))
    destroy(model)


def test_trees_batched_and_plain(gpu):
    """Branched cells of uneven size: batched and plain agree with the oracle bit for bit."""
    model, info = configs.tree_network(ncells=30, median=120, lo=20, hi=600, indegree=10,
                                       rate_hz=50.0, netstim_weight=3e-2)
    recs = [voltage_record(model, c, "soma", 0.5) for c in info["cells"][:3]]
    got_b, ref, vg_b, vr, _ = _compare(model, 40.0, recs, method="batched")
    assert np.array_equal(vg_b, vr)
    got_p, _, vg_p, _, _ = _compare(model, 40.0, recs, method="plain")
    assert np.array_equal(vg_p, vg_b)
    assert np.array_equal(got_p.spike_times, got_b.spike_times)
    destroy(model)


def test_trees_oversized_cells_batched(gpu):
    """Cells above the LDS capacity run with v (and rhs/diag) in HBM; results unchanged."""
    model, info = configs.tree_network(ncells=4, median=4500, sigma=0.3, lo=3500, hi=6000,
                                       indegree=3, synapses=4, rate_hz=50.0, netstim_weight=3e-2)
    assert info["sizes"].max() > 5200 and info["sizes"].min() < 5100  # both HBM modes
    recs = [voltage_record(model, info["cells"][0], "soma", 0.5)]
    _, _, vg, vr, _ = _compare(model, 5.0, recs, method="batched", cells_per_group=1)
    assert np.array_equal(vg, vr)
    destroy(model)
