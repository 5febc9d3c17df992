"""Runtime MOD compilation at the C ABI (mind_compile_mod and the mechanism
queries), on the host.

* The library's C++ NMODL frontend refuses malformed inputs naming the file
  and line.
* compile_mod / load_mod: slots, field columns, defaults (PARAMETER values and
  ion values), refusals with file and line, the same text compiling to the
  same slot.
* The cell kernels compile with hipRTC for gfx950 with the compiled
  mechanisms added (no device needed).
* The oracle, running a g++ build of the same lowering, gives bit for bit the
  voltages of the builtin mechanisms it restates by hand, for MOD files that
  restate those builtins (a K channel with hh's n-gate kinetics against hh
  with gnabar = gl = 0; an exponential synapse against ExpSyn).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
from mind_amd import MicroModel, compile_mod, configs, load_mod, voltage_record
from mind_amd import _abi
from mind_amd.mechanisms import builtin, lowering

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mind.jl_amd", "csrc")
MODS = os.path.join(ROOT, "tests", "golden", "mod")
REF_MOD = os.environ.get("MIND_MOD_DIR", "/root/reference/mod")
BUILTINS = ("hh", "pas", "ExpSyn", "Exp2Syn", "IClamp", "NetStim", "AlphaSynapse", "SEClamp", "VClamp")


@pytest.fixture(scope="module")
def nmodlc(tmp_path_factory):
    exe = tmp_path_factory.mktemp("nmodlc") / "nmodlc"
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(exe), os.path.join(CSRC, "nmodl.cpp"),
                    os.path.join(CSRC, "nmodlc.cpp")], check=True, capture_output=True, text=True)
    return str(exe)


def _lower_cpp(nmodlc, tmp_path, files):
    out = tmp_path / "out.h"
    r = subprocess.run([nmodlc, *files, "-o", str(out)], capture_output=True, text=True)
    return ("ok", out.read_text()) if r.returncode == 0 else ("err", r.stderr.strip())


BAD = [
    "NEURON { SUFFIX a }\nPARAMETER { g = 1 }\nBREAKPOINT { i = q*g }\nASSIGNED { i }",
    "NEURON { SUFFIX a POINTER p }",
    "NEURON { SUFFIX a }\nVERBATIM\nENDVERBATIM",
    "NEURON { SUFFIX a }\nSTATE { s }\nBREAKPOINT { SOLVE d METHOD cnexp }\nDERIVATIVE d { s' = s*s }",
    "PARAMETER { g = 1 }",
    "NEURON { SUFFIX a }\nPARAMETER { g[0] }",
    "NEURON { SUFFIX a NONSPECIFIC_CURRENT i }\nASSIGNED { i }\nBREAKPOINT { i = f(1) }\nFUNCTION f(x) { f = f(x) }",
    "NEURON { SUFFIX a }\nBREAKPOINT { SOLVE s METHOD runge }",
    "NEURON { SUFFIX a }\nPARAMETER { g = 1 $ }",
    "NEURON { SUFFIX a }\nCOMMENT\nnever closed",
]


@pytest.mark.parametrize("text", BAD)
def test_cpp_frontend_reports_errors_with_file_and_line(nmodlc, tmp_path, text):
    f = tmp_path / "bad.mod"
    f.write_text(text)
    got = _lower_cpp(nmodlc, tmp_path, [str(f)])
    assert got[0] == "err" and got[1].startswith(f"{f}:")


def test_cpp_frontend_lowers_the_test_mods(nmodlc, tmp_path):
    for name in ("kdr", "expsyn2", "nap"):
        got = _lower_cpp(nmodlc, tmp_path, [os.path.join(MODS, f"{name}.mod")])
        assert got[0] == "ok" and f"_current_eval(" in got[1], name


def test_compile_mod_slots_fields_and_defaults():
    kdr = load_mod(os.path.join(MODS, "kdr.mod"))
    assert kdr.compiled and kdr.id >= 10 and kdr.kind == "density" and kdr.has_state
    assert kdr.fields == ("gkbar", "n", "ek", "gk", "ik", "ninf", "ntau")
    assert kdr.defaults[kdr.field_index("gkbar")] == 0.036
    assert kdr.defaults[kdr.field_index("ek")] == -77.0  # default_ions potassium reversal
    assert load_mod(os.path.join(MODS, "kdr.mod")) is kdr  # same text and file: same slot
    nap = load_mod(os.path.join(MODS, "nap.mod"))
    assert nap.fields[:3] == ("gbar", "shape[1]", "shape[2]")
    assert nap.field_index("shape", 2) == 2
    assert nap.defaults[nap.field_index("ena")] == 50.0
    syn = load_mod(os.path.join(MODS, "expsyn2.mod"))
    assert syn.kind == "point" and syn.has_net_receive and syn.name == "ExpDecaySyn"
    assert len({kdr.id, nap.id, syn.id}) == 3
    assert "u%d_current" % syn.id in lowering(syn)
    # the C ABI's own view, and lookups by name (the compiled one shadows nothing here)
    L = _abi.lib()
    mech = C.c_int32(-1)
    _abi.check(L.mind_mech_lookup(b"ExpDecaySyn", C.byref(mech)))
    assert mech.value == syn.id
    _abi.check(L.mind_mech_lookup(b"hh", C.byref(mech)))
    assert mech.value == 0
    desc = _abi.MechDesc()
    _abi.check(L.mind_mech_info(0, C.byref(desc)))
    assert desc.name == b"hh" and desc.n_fields == 20 and desc.has_state == 1 and desc.compiled == 0
    names = (C.c_char_p * 20)()
    vals = (C.c_double * 20)()
    count = C.c_int32()
    _abi.check(L.mind_mech_fields(0, names, vals, 20, C.byref(count)))
    assert [names[k].decode() for k in range(count.value)] == list(builtin("hh").fields)
    assert vals[builtin("hh").field_index("gnabar")] == 0.12


@pytest.mark.parametrize("text,msg", [
    ("NEURON { ARTIFICIAL_CELL a }", "ARTIFICIAL_CELL"),
    ("NEURON { POINT_PROCESS a RANDOM r }", "RANDOM"),
    ("NEURON { POINT_PROCESS a }\nNET_RECEIVE(w) { net_event(t) }", "net_send and net_event"),
    ("NEURON { SUFFIX a USEION ca READ cao WRITE ena }\nASSIGNED { cao ena }", "not an ion variable of ca"),
    ("NEURON { SUFFIX a USEION ca READ cai }", "cai: not a declared scalar variable"),
    ("NEURON { POINT_PROCESS a }\nASSIGNED { x }\nNET_RECEIVE(w) { x = celsius }", "celsius in NET_RECEIVE"),
    ("NEURON { SUFFIX a }\nBREAKPOINT { x = 1 }", "cannot assign x"),
    ("NEURON { SUFFIX a GLOBAL g }\nASSIGNED { g }\nINITIAL { g = 2 }", "GLOBAL g is assigned"),
    ("NEURON { SUFFIX a GLOBAL g }\nASSIGNED { g }\nPROCEDURE p() { g = 1 }\nINITIAL { p() }",
     "GLOBAL g is assigned"),
    ("NEURON { POINT_PROCESS a GLOBAL g }\nASSIGNED { g }\nNET_RECEIVE(w) { g = w }", "GLOBAL g is assigned"),
])
def test_compile_mod_refusals_name_file_and_line(text, msg):
    with pytest.raises(ValueError, match=msg) as err:
        compile_mod(text, "user.mod")
    assert str(err.value).startswith("user.mod:")


def test_cell_kernels_compile_with_hiprtc_for_gfx950():
    mechs = [load_mod(os.path.join(MODS, f"{n}.mod")) for n in ("kdr", "expsyn2", "nap")]
    ids = (C.c_int32 * 3)(*[m.id for m in mechs])
    _abi.check(_abi.lib().mind_rtc_check(ids, 3, b"gfx950"))


def _k_network(user, ncells=6, stop=40.0):
    kdr = load_mod(os.path.join(MODS, "kdr.mod"))
    syn2 = load_mod(os.path.join(MODS, "expsyn2.mod"))
    m = MicroModel()
    cells = m.add_population(configs.ball_and_stick_sections(10), ncells)
    syns = []
    for c in cells:
        if user:
            m.insert(c, kdr, sections="soma", params={"gkbar": 0.04})
        else:
            m.insert(c, builtin("hh"), sections="soma", params={"gnabar": 0.0, "gl": 0.0, "gkbar": 0.04})
        m.insert(c, builtin("pas"), sections="all", params={"g": 1e-4, "e": -65.0})
        m.add_pointprocess(c, "IClamp", section="soma", params={"del": 2.0, "dur": 30.0, "amp": 0.3})
        syns.append(m.add_pointprocess(c, syn2 if user else "ExpSyn", section="dend", x=0.5,
                                       params={"tau": 2.0, "e": 0.0}))
    for k in range(ncells):
        ns = m.add_pointprocess(m, "NetStim", params={"interval": 5.0, "number": 1e9, "start": 1.0, "noise": 0.5})
        m.connect(ns, syns[k], weight=0.01, delay=1.0)
    return m, cells


def test_oracle_user_lowering_equals_its_builtin_restatement():
    out = []
    for user in (True, False):
        m, cells = _k_network(user)
        recs = [voltage_record(m, cells[k], "soma", 0.5, name=f"v{k}") for k in range(3)]
        sim = oracle.OracleSim(m).initialize()
        res = sim.run(40.0, recs)
        out.append((res, sim.voltage()))
    (a, va), (b, vb) = out
    assert np.array_equal(va, vb)
    for name in a.traces:
        assert np.array_equal(a.traces[name], b.traces[name]), name
    assert va.max() - va.min() > 15.0  # the clamp and the synapses move the membrane


def _leak_network(user, ncells=5, g=3e-4, e=-61.0):
    """Ball-and-stick cells with hh on the soma and a leak everywhere: the compiled
    leakg with its GLOBALs set for every instance, or pas with the same values."""
    m = MicroModel()
    cells = m.add_population(configs.ball_and_stick_sections(10), ncells)
    leak = load_mod(os.path.join(MODS, "leakg.mod")) if user else builtin("pas")
    if user:
        m.set_global(leak, "gl", g)
        m.set_global(leak, "el", e)
    for k, c in enumerate(cells):
        m.insert(c, builtin("hh"), sections="soma")
        m.insert(c, leak, sections="all", params={} if user else {"g": g, "e": e})
        m.add_pointprocess(c, "IClamp", section="soma", params={"del": 1.0 + k, "dur": 20.0, "amp": 0.3})
    return m, cells, leak


def test_globals_are_shared_values_read_by_every_instance():
    from mind_amd import global_record, mechanism_record
    from mind_amd.mechanisms import globals_of

    m, cells, leak = _leak_network(True)
    assert globals_of(leak) == ("gl", "el") and globals_of(builtin("pas")) == ()
    with pytest.raises(ValueError, match="GLOBAL of leakg"):
        m.insert(cells[0], leak, params={"gl": 1e-3})
    with pytest.raises(ValueError, match="not a shared GLOBAL"):
        global_record(m, leak, "i")
    with pytest.raises(ValueError, match="not a shared GLOBAL"):
        m.set_global(leak, "i", 1.0)
    recs = [global_record(m, leak, "gl", name="gl"), global_record(m, leak, "el", name="el"),
            mechanism_record(m, cells[3], leak, "i", section="dend", name="i3")]
    recs += [voltage_record(m, c, "soma", 0.5, name=f"v{k}") for k, c in enumerate(cells)]
    sim = oracle.OracleSim(m).initialize()
    a = sim.run(25.0, recs)
    # every sample of a GLOBAL is the value set for every instance
    assert np.all(a.traces["gl"] == 3e-4) and np.all(a.traces["el"] == -61.0)
    assert np.any(a.traces["i3"] != 0.0)
    # and the compiled leak with those GLOBALs is pas with the same values, bit for bit
    p, pcells, _ = _leak_network(False)
    psim = oracle.OracleSim(p).initialize()
    b = psim.run(25.0, [voltage_record(p, c, "soma", 0.5, name=f"v{k}") for k, c in enumerate(pcells)])
    assert np.array_equal(sim.voltage(), psim.voltage())
    for k in range(len(cells)):
        assert np.array_equal(a.traces[f"v{k}"], b.traces[f"v{k}"])
    assert max(a.traces[f"v{k}"].max() for k in range(len(cells))) > 0.0  # the clamps fire the somata
