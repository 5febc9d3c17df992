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
