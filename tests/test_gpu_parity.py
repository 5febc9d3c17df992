"""Parity of the HIP path with the CPU oracle (bit-exact voltages and rasters)."""
import numpy as np
import pytest

import oracle
from mind_amd import (batch_plan, build, configs, destroy, initialize, mechanism_record,
                      read_field, read_voltage, run, set_batch_plan, voltage_record)

pytestmark = pytest.mark.gpu


def _compare(model, stop, records, method="plain", plan=None, **kw):
    print(f"case: {method} stop={stop} records={len(records)} plan={plan} {kw}", flush=True)
    build(model, dt_micro=0.025, seed=1, method=method, **kw)
    if plan is not None:
        set_batch_plan(model, *plan)
        assert batch_plan(model)["layout"] == plan[2]
    initialize(model)
    got = run(model, stop, records)
    ref_sim = oracle.OracleSim(model, secondorder=kw.get("secondorder", 0)).initialize()
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


@pytest.mark.parametrize("plan", [(300, 128, 0), (600, 64, 4), (2400, 512, 4), (500, 64, 5), (3000, 256, 5),
                                  (1200, 1024, 0), (4000, 1024, 5)])
def test_trees_fixed_plans_batched(gpu, plan):
    """Both layouts of the window kernel (all in LDS; lean, with the solve's node
    records streamed from HBM) and block sizes below and above a level's width:
    bit-identical voltages and raster."""
    model, info = configs.tree_network(ncells=30, median=120, lo=20, hi=600, indegree=10,
                                       rate_hz=50.0, netstim_weight=3e-2)
    recs = [voltage_record(model, c, "soma", 0.5) for c in info["cells"][:3]]
    _, _, vg, vr, _ = _compare(model, 40.0, recs, method="batched", plan=plan)
    assert np.array_equal(vg, vr)
    destroy(model)


@pytest.mark.parametrize("plan", [None, (300, 128, 0), (600, 64, 4), (500, 64, 5), (3000, 1024, 5)])
def test_trees_with_many_children_batched(gpu, plan):
    """Sections with up to four children at one point (nodes with three and four
    children take the CSR path of assembly and elimination) in every layout:
    bit-identical to the plain kernel and within tolerance of the oracle."""
    model, info = configs.tree_network(ncells=30, median=120, lo=20, hi=600, indegree=10,
                                       rate_hz=50.0, netstim_weight=3e-2, fanout=4)
    recs = [voltage_record(model, c, "soma", 0.5) for c in info["cells"][:3]]
    got_b, _, vg_b, vr, _ = _compare(model, 40.0, recs, method="batched", plan=plan)
    assert np.array_equal(vg_b, vr)
    destroy(model)
    if plan is None:
        got_p, _, vg_p, _, _ = _compare(model, 40.0, recs, method="plain")
        assert np.array_equal(vg_p, vg_b)
        assert np.array_equal(got_p.spike_times, got_b.spike_times)
        destroy(model)


@pytest.mark.parametrize("plan", [None, (4000, 256, 4), (2000, 512, 5)])
def test_trees_oversized_cells_batched(gpu, plan):
    """Cells above the LDS capacity of the all-in-LDS layout run lean with v in
    LDS, larger ones lean with v in HBM, the largest with everything in HBM;
    results unchanged."""
    model, info = configs.tree_network(ncells=4, median=6000, sigma=0.8, lo=2000, hi=11500,
                                       indegree=3, synapses=4, rate_hz=50.0, netstim_weight=3e-2)
    sizes = np.sort(info["sizes"])
    # about two nodes per compartment: one cell lean with v in LDS (< 6300 nodes),
    # one lean (< 9700), two all in HBM
    assert 1800 < sizes[0] < 3100 and 3200 < sizes[1] < 4800 and sizes[2] > 4900
    recs = [voltage_record(model, c, "soma", 0.5) for c in info["cells"]]
    _, _, vg, vr, _ = _compare(model, 5.0, recs, method="batched", cells_per_group=1, plan=plan)
    assert np.array_equal(vg, vr)
    destroy(model)


@pytest.mark.parametrize("secondorder", [1, 2])
@pytest.mark.parametrize("method,plan", [("plain", None), ("batched", (300, 128, 0)), ("batched", (600, 64, 4)),
                                         ("batched", (3000, 256, 5))])
def test_secondorder_matches_oracle(gpu, secondorder, method, plan):
    """Crank-Nicolson steps (build! secondorder 1 and 2: assembly at dt/2, v += 2 dv)
    in the plain kernel and every batched layout: bit-identical to the oracle."""
    model, info = configs.tree_network(ncells=30, median=120, lo=20, hi=600, indegree=10,
                                       rate_hz=50.0, netstim_weight=3e-2)
    recs = [voltage_record(model, c, "soma", 0.5) for c in info["cells"][:3]]
    got, ref, vg, vr, _ = _compare(model, 40.0, recs, method=method, plan=plan, secondorder=secondorder)
    assert np.array_equal(vg, vr)
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    destroy(model)


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_alpha_synapse_matches_oracle(gpu, method):
    """AlphaSynapse on the single cell, in both kernels: bit-identical to the oracle."""
    model, info = configs.single_cell()
    cell = info["cells"][0]
    syn = model.add_pointprocess(cell, "AlphaSynapse", section="dend", x=0.2,
                                 params={"onset": 20.0, "tau": 1.5, "gmax": 5e-3, "e": 0.0})
    recs = [voltage_record(model, cell, "soma", 0.5), mechanism_record(model, syn, "g"),
            mechanism_record(model, syn, "i")]
    got, ref, vg, vr, _ = _compare(model, 60.0, recs, method=method)
    assert np.array_equal(vg, vr)
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    assert got.traces["g"].max() > 4.9e-3
    destroy(model)


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_seclamp_matches_oracle(gpu, method):
    """SEClamp at the soma of the single cell (levels switching off the step grid, the
    clamp releasing into hh spiking), in both kernels: bit-identical to the oracle."""
    model, info = configs.single_cell()
    cell = info["cells"][0]
    clamp = model.add_pointprocess(cell, "SEClamp", section="soma", x=0.5,
                                   params={"rs": 0.01, "dur1": 5.01, "amp1": -70.0, "dur2": 10.0,
                                           "amp2": -10.0, "dur3": 4.99, "amp3": -80.0})
    recs = [voltage_record(model, cell, "soma", 0.5), voltage_record(model, cell, "dend", 0.5, name="vd"),
            mechanism_record(model, clamp, "i"), mechanism_record(model, clamp, "vc"),
            mechanism_record(model, clamp, "on")]
    got, ref, vg, vr, _ = _compare(model, 60.0, recs, method=method)
    assert np.array_equal(vg, vr)
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    assert got.traces["on"][-1] == 0.0 and got.traces["i"].min() < 0.0
    destroy(model)


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_vclamp_matches_oracle(gpu, method):
    """VClamp at the soma of the single cell (levels switching off the step grid, the
    clamp releasing into hh spiking), in both kernels: bit-identical to the oracle."""
    model, info = configs.single_cell()
    cell = info["cells"][0]
    clamp = model.add_pointprocess(cell, "VClamp", section="soma", x=0.5,
                                   params={"dur": [5.01, 10.0, 4.99], "amp": [-70.0, -10.0, -80.0],
                                           "tau2": 0.002, "fac": 0.5})
    recs = [voltage_record(model, cell, "soma", 0.5), voltage_record(model, cell, "dend", 0.5, name="vd"),
            mechanism_record(model, clamp, "i"), mechanism_record(model, clamp, "e"),
            mechanism_record(model, clamp, "vo"), mechanism_record(model, clamp, "stim")]
    got, ref, vg, vr, _ = _compare(model, 60.0, recs, method=method)
    assert np.array_equal(vg, vr)
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    assert got.traces["i"][-1] == 0.0
    destroy(model)
