"""NMODL -> HIP C lowering (the C++ frontend csrc/nmodl.cpp through its driver
nmodlc), checked on the host with g++.

* Small MOD inputs written here: parse errors, field layout, and the numbers the
  generated C++ computes against closed forms.
* The six builtin MOD files, read where the reference tree is present: the
  generated phases against the hand lowering the kernels were pinned with, bit
  for bit over randomised states (the oracle tests pin the hand lowering).
"""
import os
import re
import subprocess
import textwrap

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mind.jl_amd", "csrc")
MOD_DIR = "/root/reference/mod"
BUILTINS = ("hh", "pas", "ExpSyn", "Exp2Syn", "IClamp", "NetStim", "AlphaSynapse", "SEClamp", "VClamp")

LEAKGATE = textwrap.dedent("""
    TITLE toy gated leak
    COMMENT
    a one-gate channel written for these tests
    ENDCOMMENT
    NEURON {
        SUFFIX leakgate
        NONSPECIFIC_CURRENT i
        RANGE gbar, erev, tau0
    }
    PARAMETER {
        gbar = .002 (S/cm2) <0,1e9>
        erev = -70 (mV)
        tau0 = 4 (ms)
    }
    STATE { s }
    ASSIGNED { v (mV) i (mA/cm2) sinf }
    BREAKPOINT {
        SOLVE gate METHOD cnexp
        i = gbar*s^2*(v - erev)   : quadratic gate
    }
    INITIAL { settle(v)  s = sinf }
    DERIVATIVE gate {
        settle(v)
        s' = (sinf - s)/tau0
    }
    PROCEDURE settle(vm (mV)) {
        sinf = 1/(1 + exp(-(vm + 40)/5))
    }
""")


class ModError(Exception):
    """A lowering failure reported by nmodlc ("<file>:<line>: <message>")."""


@pytest.fixture(scope="module")
def nmodlc(tmp_path_factory):
    exe = tmp_path_factory.mktemp("nmodlc") / "nmodlc"
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(exe), os.path.join(CSRC, "nmodl.cpp"),
                    os.path.join(CSRC, "nmodlc.cpp")], check=True, capture_output=True, text=True)
    return str(exe)


def lower(nmodlc, tmp_path, sources):
    """mech_gen.h text of (filename, text) sources, or ModError."""
    files = []
    for name, text in sources:
        f = tmp_path / os.path.basename(name)
        f.write_text(text)
        files.append(str(f))
    return lower_files(nmodlc, tmp_path, files)


def lower_files(nmodlc, tmp_path, files):
    out = tmp_path / "lowered.h"
    r = subprocess.run([nmodlc, *files, "-o", str(out)], capture_output=True, text=True)
    if r.returncode:
        raise ModError(r.stderr.strip())
    return out.read_text()


def field_columns(header, name):
    """(first column, field name) of mechanism `name`'s fields in the header's enum."""
    enum = re.search(rf"enum {{ ({name}_\w+ = \d+(?:, {name}_\w+ = \d+)*) }};", header).group(1)
    cols = [(int(v), k[len(name) + 1:]) for k, v in (e.split(" = ") for e in enum.split(", "))]
    return sorted(c for c in cols if c[1] != "NFIELDS")


def columns(header, name):
    return [n for _, n in field_columns(header, name)]


def gxx_run(tmp_path, header_text, main_src, name="prog"):
    hdr = tmp_path / "mech_gen.h"
    hdr.write_text(header_text)
    src = tmp_path / f"{name}.cpp"
    src.write_text(main_src)
    exe = tmp_path / name
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", f"-I{CSRC}",
                    f"-I{tmp_path}", "-o", str(exe), str(src)], check=True, capture_output=True, text=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout


def test_parse_layout_and_defaults(nmodlc, tmp_path):
    hdr = lower(nmodlc, tmp_path, [("leakgate.mod", LEAKGATE)])
    # PARAMETER, then STATE, then ASSIGNED names without v
    assert columns(hdr, "leakgate") == ["gbar", "erev", "tau0", "s", "i", "sinf"]
    assert "f[leakgate_erev] = -70.0;" in hdr


def test_generated_toy_mechanism_matches_closed_forms(nmodlc, tmp_path):
    hdr = lower(nmodlc, tmp_path, [("leakgate.mod", LEAKGATE)])
    out = gxx_run(tmp_path, hdr, textwrap.dedent("""
        #include "mech_gen.h"
        #include <cstdio>
        using namespace mind;
        int main() {
          double f[gen::leakgate_NFIELDS] = {0};
          gen::leakgate_defaults(f);
          Cols c{f, 1};
          gen::leakgate_initial(c, 0, -50.0, 6.3, nullptr, 0);
          double s0 = f[gen::leakgate_s];
          gen::leakgate_advance(c, 0, -30.0, 0.025, 0.025, 6.3);
          double s1 = f[gen::leakgate_s];
          double val, sh;
          gen::leakgate_current(c, 0, -30.0, 0.0, 0.025, val, sh, true, 6.3, nullptr);
          std::printf("%.17g %.17g %.17g %.17g %.17g\\n", s0, s1, val, sh, f[gen::leakgate_i]);
        }
    """))
    s0, s1, val, sh, i = map(float, out.split())
    sinf = lambda vm: 1 / (1 + np.exp(-(vm + 40) / 5))
    assert s0 == pytest.approx(sinf(-50.0), rel=1e-15)
    tau = 4.0
    s1_ref = s0 + (1 - np.exp(-0.025 / tau)) * (sinf(-30.0) - s0)  # cnexp relaxation
    assert s1 == pytest.approx(s1_ref, rel=1e-14)
    assert val == pytest.approx(0.002 * s1 ** 2 * (-30.0 + 70.0), rel=1e-14)
    assert sh == pytest.approx(0.002 * s1 ** 2 * (-30.0 + 0.001 + 70.0), rel=1e-14)
    assert i == val


@pytest.mark.parametrize("text,msg", [
    ("NEURON { SUFFIX a }\nPARAMETER { g = 1 }\nBREAKPOINT { i = q*g }\nASSIGNED { i }", "unknown name q"),
    ("NEURON { SUFFIX a POINTER p }", "POINTER"),
    ("NEURON { SUFFIX a }\nVERBATIM\nENDVERBATIM", "VERBATIM"),
    ("NEURON { SUFFIX a }\nSTATE { s }\nBREAKPOINT { SOLVE d METHOD cnexp }\nDERIVATIVE d { s' = s*s }",
     "cnexp form"),
    ("PARAMETER { g = 1 }", "no NEURON block"),
])
def test_errors_name_the_problem(nmodlc, tmp_path, text, msg):
    with pytest.raises(ModError, match=msg):
        lower(nmodlc, tmp_path, [("bad.mod", text)])


HARNESS = textwrap.dedent("""
    #include "mech_gen.h"
    #include <cstdio>
    #include <cstring>
    #include <random>
    using namespace mind;
    static int bad = 0;
    static void same(const void* a, const void* b, size_t n, const char* what, int mech, int trial) {
      if (std::memcmp(a, b, n)) {
        if (bad < 10) std::printf("mismatch %s mech %d trial %d\\n", what, mech, trial);
        ++bad;
      }
    }
    static void same_out(const NetStimOut& a, const NetStimOut& b, const char* what, int trial) {
      // member-wise: the struct's padding bytes are unspecified
      same(&a.spiked, &b.spiked, sizeof a.spiked, what, 5, trial);
      same(&a.spike_time, &b.spike_time, 8, what, 5, trial);
      same(&a.send, &b.send, sizeof a.send, what, 5, trial);
      same(&a.send_delay, &b.send_delay, 8, what, 5, trial);
      same(&a.send_flag, &b.send_flag, 8, what, 5, trial);
    }
    int main() {
      std::mt19937_64 g(12345);
      auto U = [&](double a, double b) { return std::uniform_real_distribution<double>(a, b)(g); };
      const int nf[MIND_MECH_COUNT] = {HH_NFIELDS, PAS_NFIELDS, EXPSYN_NFIELDS, EXP2SYN_NFIELDS,
                                       ICLAMP_NFIELDS, NETSTIM_NFIELDS, ALPHASYNAPSE_NFIELDS,
                                       PATTERNSTIM_NFIELDS, SECLAMP_NFIELDS, VCLAMP_NFIELDS};
      for (int trial = 0; trial < 3000; ++trial) {
        for (int mech = 0; mech < MIND_MECH_COUNT; ++mech) {
          if (mech == MIND_MECH_NETSTIM) continue;  // through its stream, below
          if (mech == MIND_MECH_PATTERNSTIM) continue;  // no code: its spike times enter the queue
          double a[32], b[32];
          for (int k = 0; k < nf[mech]; ++k) a[k] = U(0.0, 1.0);
          if (mech == MIND_MECH_HH) { a[HH_GNABAR] = U(0, 0.3); a[HH_EL] = U(-80, -40);
                                      a[HH_ENA] = U(30, 60); a[HH_EK] = U(-100, -60); }
          if (mech == MIND_MECH_PAS) { a[PAS_G] = U(0, 1e-3); a[PAS_E] = U(-90, -50); }
          if (mech == MIND_MECH_EXPSYN) { a[EXPSYN_TAU] = U(0.1, 10); a[EXPSYN_E] = U(-80, 0); }
          if (mech == MIND_MECH_EXP2SYN) { a[EXP2SYN_TAU1] = U(0.05, 5); a[EXP2SYN_TAU2] = U(0.05, 20);
                                           a[EXP2SYN_E] = U(-80, 0); }
          if (mech == MIND_MECH_ICLAMP) { a[ICLAMP_DEL] = U(0, 100); a[ICLAMP_DUR] = U(0, 100);
                                          a[ICLAMP_AMP] = U(-1, 1); }
          if (mech == MIND_MECH_ALPHASYNAPSE) { a[ALPHASYNAPSE_ONSET] = U(0, 250); a[ALPHASYNAPSE_TAU] = U(0.05, 20);
                                                a[ALPHASYNAPSE_GMAX] = U(0, 1e-2); a[ALPHASYNAPSE_E] = U(-80, 0); }
          if (mech == MIND_MECH_SECLAMP) { a[SECLAMP_RS] = U(1e-3, 10); a[SECLAMP_DUR1] = U(0, 100);
                                           a[SECLAMP_AMP1] = U(-100, 50); a[SECLAMP_DUR2] = trial % 3 ? U(0, 100) : 0.0;
                                           a[SECLAMP_AMP2] = U(-100, 50); a[SECLAMP_DUR3] = U(0, 100);
                                           a[SECLAMP_AMP3] = U(-100, 50); }
          if (mech == MIND_MECH_VCLAMP) { for (int k = 0; k < 3; ++k) { a[VCLAMP_DUR + k] = U(0, 100);
                                                                        a[VCLAMP_AMP + k] = U(-100, 50); }
                                          a[VCLAMP_GAIN] = U(1, 1e5); a[VCLAMP_RSTIM] = U(0.01, 10);
                                          a[VCLAMP_TAU1] = U(1e-4, 0.1); a[VCLAMP_TAU2] = trial % 2 ? U(0, 0.1) : 0.0;
                                          a[VCLAMP_FAC] = trial % 3 ? U(0, 10) : 0.0;
                                          a[VCLAMP_E0] = U(-10, 10); a[VCLAMP_VO0] = U(-100, 50);
                                          a[VCLAMP_VI0] = U(-100, 50); }
          std::memcpy(b, a, sizeof a);
          Cols ca{a, 1}, cb{b, 1};
          const double v = U(-100, 60), t = U(0, 250), dt = trial % 2 ? 0.025 : U(0.001, 0.2),
                       cel = U(6, 37), w = U(-1e-3, 1e-3);
          double va, sa, vb, sb;
          mech_initial(mech, ca, 0, v, cel);
          gen::initial(mech, cb, 0, v, cel);
          same(a, b, sizeof a, "initial", mech, trial);
          double ia[16] = {0}, ib[16] = {0};  // written ion currents at v + 0.001, then at v
          mech_current(mech, ca, 0, v, t, dt, va, sa, true, ia);
          gen::current(mech, cb, 0, v, t, dt, vb, sb, true, 6.3, ib);
          same(ia, ib, sizeof ia, "ion current probe", mech, trial);
          same(&va, &vb, 8, "current", mech, trial);
          same(&sa, &sb, 8, "current probe", mech, trial);
          same(a, b, sizeof a, "current fields", mech, trial);
          mech_advance(mech, ca, 0, v, t + dt, dt, cel);
          gen::advance(mech, cb, 0, v, t + dt, dt, cel);
          same(a, b, sizeof a, "advance", mech, trial);
          synapse_receive(mech, ca, 0, w);
          gen::receive(mech, cb, 0, w);
          same(a, b, sizeof a, "receive", mech, trial);
        }
        // NetStim through its stream: INITIAL, then a random sequence of events
        double a[NETSTIM_NFIELDS], b[NETSTIM_NFIELDS];
        a[NETSTIM_INTERVAL] = U(-1, 50); a[NETSTIM_NUMBER] = (double)(g() % 6); a[NETSTIM_START] = U(-5, 20);
        a[NETSTIM_NOISE] = U(-0.2, 1.2); a[NETSTIM_EVENT] = 0; a[NETSTIM_ON] = 0; a[NETSTIM_ISPIKE] = 0;
        std::memcpy(b, a, sizeof a);
        Cols ca{a, 1}, cb{b, 1};
        Stream sa{}, sb{};
        sa.key[0] = g(); sa.key[1] = g(); sa.key[2] = g() & 0xffffffffu;
        sb = sa;
        const uint64_t seed = g();
        NetStimOut oa = netstim_initial(ca, 0, sa, seed);
        NetStimOut ob = gen::NetStim_initial(cb, 0, 0.0, 0.0, &sb, seed);
        same_out(oa, ob, "netstim initial out", trial);
        same(a, b, sizeof a, "netstim initial", 5, trial);
        same(&sa, &sb, sizeof sa, "netstim initial stream", 5, trial);
        double t = 0;
        for (int e = 0; e < 12; ++e) {
          const double flags[3] = {0.0, 1.0, 3.0};
          const double flag = flags[g() % 3], w = U(-1, 1);
          t += U(0, 20);
          oa = netstim_receive(ca, 0, sa, seed, t, w, flag);
          ob = gen::NetStim_receive(cb, 0, t, w, flag, &sb, seed);
          same_out(oa, ob, "netstim receive out", trial);
          same(a, b, sizeof a, "netstim receive", 5, trial);
          same(&sa, &sb, sizeof sa, "netstim stream", 5, trial);
        }
      }
      std::printf("mismatches %d\\n", bad);
      return 0;
    }
""")


@pytest.mark.skipif(not os.path.isdir(MOD_DIR), reason="the builtin MOD files are not present")
def test_builtin_mod_files_lower_to_the_hand_lowering_bit_for_bit(nmodlc, tmp_path):
    paths = [os.path.join(MOD_DIR, f"{n}.mod") for n in BUILTINS]
    hdr = lower_files(nmodlc, tmp_path, paths)
    # NetStim's stream key comes from its source digest
    assert "constexpr uint64_t NetStim_digest = 0xa186833b569ff3b2ull;" in hdr
    out = gxx_run(tmp_path, hdr, HARNESS, "equiv")
    assert out.strip() == "mismatches 0", out


@pytest.mark.skipif(not os.path.isdir(MOD_DIR), reason="the builtin MOD files are not present")
def test_parameter_defaults_match_the_packing_defaults(nmodlc, tmp_path):
    paths = [os.path.join(MOD_DIR, f"{n}.mod") for n in BUILTINS]
    hdr = lower_files(nmodlc, tmp_path, paths)
    (tmp_path / "mech_gen.h").write_text(hdr)
    src = tmp_path / "defaults.cpp"
    src.write_text(textwrap.dedent("""
        #include "mech_gen.h"
        #include "layout.h"
        #include <cstdio>
        #include <string>
        namespace mind { int fail(int code, const std::string&) { return code; } }
        using namespace mind;
        int main() {
          double f[32], h[32];
          #define SHOW(M, ID) for (int k = 0; k < 32; ++k) f[k] = 0; gen::M##_defaults(f); \\
            field_defaults(ID, h); \\
            for (int k = 0; k < gen::M##_NFIELDS; ++k) std::printf("%s %d %.17g %.17g\\n", #M, k, f[k], h[k]);
          SHOW(hh, MIND_MECH_HH) SHOW(pas, MIND_MECH_PAS) SHOW(ExpSyn, MIND_MECH_EXPSYN)
          SHOW(Exp2Syn, MIND_MECH_EXP2SYN) SHOW(IClamp, MIND_MECH_ICLAMP) SHOW(NetStim, MIND_MECH_NETSTIM)
          SHOW(AlphaSynapse, MIND_MECH_ALPHASYNAPSE) SHOW(SEClamp, MIND_MECH_SECLAMP)
          SHOW(VClamp, MIND_MECH_VCLAMP)
        }
    """))
    exe = tmp_path / "defaults"
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{CSRC}", f"-I{tmp_path}", "-o", str(exe), str(src),
                    *(os.path.join(CSRC, f) for f in ("layout.cpp", "mechreg.cpp", "nmodl.cpp"))],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    for line in out.strip().split("\n"):
        name, k, gen_default, packed = line.split()
        # an array's elements follow its first column
        fname = [n for c, n in field_columns(hdr, name) if c <= int(k)][-1]
        if fname in ("ena", "ek"):  # ion reversal potentials come from the ion defaults
            continue
        assert float(gen_default) == float(packed), (name, fname)


@pytest.mark.skipif(not os.path.isdir(MOD_DIR), reason="the builtin MOD files are not present")
def test_equivalence_check_sees_a_reassociated_product(nmodlc, tmp_path):
    """The comparison is sensitive to operation order: regrouping hh's m*m*m*h
    product (same value in exact arithmetic) is reported."""
    paths = [os.path.join(MOD_DIR, f"{n}.mod") for n in BUILTINS]
    hdr = lower_files(nmodlc, tmp_path, paths)
    exact = "((((in.gnabar * in.m) * in.m) * in.m) * in.h)"  # hh_current_eval, from the loaded columns
    assert exact in hdr
    hdr = hdr.replace(exact, "(in.gnabar * ((in.m * in.m) * (in.m * in.h)))")
    out = gxx_run(tmp_path, hdr, HARNESS, "planted")
    count = int(out.strip().split()[-1])
    assert count > 0 and "current" in out
