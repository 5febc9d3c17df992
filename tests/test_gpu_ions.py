"""Ion dynamics on the GPU against the oracle, bit for bit: calcium entering
through a GHK current (cal), a calcium shell writing cai (cadecay, Euler), a
calcium-activated potassium current reading it (kca), a point process writing
ica in nA and reading cai in NET_RECEIVE (CaPump), hh with a Nernst sodium
species, and a section reversal for potassium; both kernels, secondorder 0 and
2 (ion currents corrected by di/dv dv), the window path and the event-exact path
(zero-delay connections). Voltages, rasters, ion rows (cai, eca, ica, dica/dv,
ena, ek) and fields."""
import os

import numpy as np
import pytest

import oracle
from mind_amd import (IonSpecies, MicroModel, build, configs, destroy, initialize, ion_record, load_mod,
                      mechanism_record, read_field, read_voltage, run, voltage_record)
from mind_amd.mechanisms import builtin

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODS = os.path.join(ROOT, "tests", "golden", "mod")


def _model(ncells=12, delay=1.0, seed=5):
    cal, cad, kca, pump = (load_mod(os.path.join(MODS, f"{n}.mod")) for n in ("cal", "cadecay", "kca", "capump"))
    m = MicroModel(ions=[IonSpecies("na", 1.0, 10.0, 140.0, 50.0, nernst=True)])
    cells = m.add_population(configs.ball_and_stick_sections(12), ncells, ions={"dend": {"k": {"reversal": -85.0}}})
    g = np.random.default_rng(seed)
    syns, srcs, pumps = [], [], []
    for c in cells:
        m.insert(c, builtin("hh"), sections="soma")
        m.insert(c, builtin("pas"), sections="dend", params={"g": 1e-4, "e": -65.0})
        m.insert(c, cal, sections="all", params={"pbar": float(g.uniform(1e-4, 4e-4))})
        m.insert(c, cad, sections="all", params={"taur": float(g.uniform(40, 120))})
        m.insert(c, kca, sections="soma")
        syns.append(m.add_pointprocess(c, "Exp2Syn", section="dend", x=0.2,
                                       params={"tau1": 0.5, "tau2": 4.0, "e": 0.0}))
        pumps.append(m.add_pointprocess(c, pump, section="dend", x=0.6, params={"amp": -0.005}))
        srcs.append(m.register_spike_source(c, section="soma", x=0.5, threshold=-20.0))
        m.add_pointprocess(c, "IClamp", section="soma", params={"del": float(g.uniform(1, 5)), "dur": 40.0,
                                                                "amp": float(g.uniform(0.15, 0.35))})
    for t in range(ncells):
        for s in g.integers(0, ncells, 4):
            m.connect(srcs[s], syns[t], weight=float(g.uniform(1e-3, 4e-3)), delay=delay)
        ns = m.add_pointprocess(m, "NetStim", params={"interval": 6.0, "number": 1e9, "start": 1.0,
                                                      "noise": 0.5})
        m.connect(ns, pumps[t], weight=1.0, delay=delay)
    return m, cells, pumps


def _records(m, cells, pumps):
    recs = [voltage_record(m, cells[k], "soma", 0.5, name=f"v{k}") for k in range(2)]
    for f in ("cai", "eca", "ica", "dica_dv_", "ena", "ek", "ik", "dik_dv_"):
        recs.append(ion_record(m, cells[0], f))
    recs += [ion_record(m, cells[1], f, section="dend", x=0.6, name=f"d_{f}") for f in ("cai", "ica", "ek")]
    recs += [mechanism_record(m, pumps[1], "seen", name="seen"), mechanism_record(m, pumps[1], "on", name="on")]
    return recs


@pytest.mark.parametrize("method,secondorder", [("plain", 0), ("batched", 0), ("plain", 2), ("batched", 2)])
def test_calcium_dynamics_match_oracle(gpu, method, secondorder):
    m, cells, pumps = _model()
    recs = _records(m, cells, pumps)
    print(f"case: ions {method} secondorder={secondorder}", flush=True)
    build(m, dt_micro=0.025, seed=1, method=method, secondorder=secondorder)
    initialize(m)
    got = run(m, 60.0, recs)
    cad = load_mod(os.path.join(MODS, "cadecay.mod"))
    v_got, c_got = read_voltage(m), read_field(m, cad, "cai")
    destroy(m)
    sim = oracle.OracleSim(m, secondorder=secondorder).initialize()
    ref = sim.run(60.0, recs)
    assert len(ref.spike_times) > 10
    assert np.array_equal(got.spike_times, ref.spike_times)
    assert np.array_equal(got.spike_sources, ref.spike_sources)
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
    assert np.array_equal(v_got, sim.voltage())
    assert np.array_equal(c_got, sim.field(cad.id, cad.field_index("cai")))
    assert ref.traces["cai"].max() > 2 * ref.traces["cai"][0]  # calcium entered
    assert ref.traces["seen"].max() > 0


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_calcium_dynamics_on_the_event_exact_path(gpu, method):
    m, cells, pumps = _model(ncells=6, delay=0.0)
    recs = _records(m, cells, pumps)
    print(f"case: ions event-exact {method}", flush=True)
    build(m, dt_micro=0.025, seed=1, method=method)
    initialize(m)
    got = run(m, 30.0, recs)
    destroy(m)
    ref = oracle.OracleSim(m).initialize().run(30.0, recs)
    assert np.array_equal(got.spike_times, ref.spike_times)
    for r in recs:
        assert np.array_equal(got.traces[r.name], ref.traces[r.name]), r.name
