"""The macro step's MIND-role mechanisms as MOD files (mind.jl_amd/mod/), lowered
by the library's NMODL frontend (nmodlc --macro), on the host.

* MIND blocks (ROLE, EXPOSURE, READ_/WRITE_ bindings, frontend.jl:514-548) parse,
  and malformed ones are refused with file and line.
* The lowered WilsonCowan REGION (RK4 in a PROCEDURE) advances (E, I) bit for bit
  as the oracle's own Runge-Kutta restatement, which is pinned against scipy's
  DOP853 (test_oracle.py); LinearDelay, PoissonDrive and PopRate give their
  closed forms bit for bit.
* Runtime compilation refuses role mechanisms: their lowering belongs in the
  macro kernels built with the library.
"""
import os
import subprocess
import textwrap

import numpy as np
import pytest

import oracle
from mind_amd import compile_mod

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mind.jl_amd", "csrc")
MOD = os.path.join(ROOT, "mind.jl_amd", "mod")
ROLES = ("WilsonCowan", "LinearDelay", "PoissonDrive", "PopRate")
WC_DEFAULTS = (10.0, 10.0, 16.0, 12.0, 15.0, 3.0, 1.3, 4.0, 2.0, 3.7, 1.25, 0.0, 0.0, 100.0, 0.1, 0.05)


@pytest.fixture(scope="module")
def nmodlc(tmp_path_factory):
    exe = tmp_path_factory.mktemp("nmodlc") / "nmodlc"
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(exe), os.path.join(CSRC, "nmodl.cpp"),
                    os.path.join(CSRC, "nmodlc.cpp")], check=True, capture_output=True, text=True)
    return str(exe)


def _lower(nmodlc, tmp_path, files):
    out = tmp_path / "macro_gen.h"
    r = subprocess.run([nmodlc, "--macro", *files, "-o", str(out)], capture_output=True, text=True)
    return (r.returncode, out.read_text() if r.returncode == 0 else r.stderr.strip())


HARNESS = textwrap.dedent("""
    #include "macro_gen.h"
    #include <cstdio>
    #include <cstdlib>
    using namespace mind;
    using namespace mind::gen_macro;
    int main(int argc, char** argv) {
      // stdin: lines "p0..p15 E I coupling rate dt"; stdout: E' I' as hex bit patterns
      double p[16], e, i, c, rate, dt;
      while (true) {
        for (int k = 0; k < 16; ++k) if (std::scanf("%lf", &p[k]) != 1) return 0;
        if (std::scanf("%lf %lf %lf %lf %lf", &e, &i, &c, &rate, &dt) != 5) return 1;
        double f[WilsonCowan_NFIELDS] = {0};
        for (int k = 0; k < 16; ++k) f[k] = p[k];
        f[WilsonCowan_E] = e; f[WilsonCowan_I] = i; f[WilsonCowan_coupling] = c; f[WilsonCowan_rate] = rate;
        WilsonCowan_advance(Cols{f, 1}, 0, 0.0, 0.0, dt, 0.0);
        double g[LinearDelay_NFIELDS] = {0};
        g[LinearDelay_weight] = c; g[LinearDelay_x] = e;
        LinearDelay_evaluate(Cols{g, 1}, 0, 0.0, 0.0);
        double h[PoissonDrive_NFIELDS] = {0};
        h[PoissonDrive_e] = e; h[PoissonDrive_scale] = p[13];
        PoissonDrive_evaluate(Cols{h, 1}, 0, 0.0, 0.0);
        double q[PopRate_NFIELDS] = {0};
        q[PopRate_count] = rate; q[PopRate_sources] = p[2]; q[PopRate_width] = dt;
        PopRate_evaluate(Cols{q, 1}, 0, 0.0, 0.0);
        std::printf("%a %a %a %a %a %a\\n", f[WilsonCowan_E], f[WilsonCowan_I], g[LinearDelay_y],
                    h[PoissonDrive_rate], q[PopRate_rate], q[PopRate_count]);
      }
    }
""")


def test_role_mods_lower_and_match_the_oracle(nmodlc, tmp_path):
    code, hdr = _lower(nmodlc, tmp_path, [os.path.join(MOD, f"{n}.mod") for n in ROLES])
    assert code == 0, hdr
    for n in ROLES:
        assert f"{n}_evaluate" in hdr and f"{n}_NFIELDS" in hdr
    assert "WilsonCowan_rk4step" in hdr  # RK4 as a PROCEDURE, the state phase's SOLVE
    src = tmp_path / "h.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "h"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", f"-I{CSRC}",
                    f"-I{tmp_path}", "-o", str(exe), str(src)], check=True, capture_output=True, text=True)
    g = np.random.default_rng(7)
    cases = []
    for k in range(600):
        p = np.array(WC_DEFAULTS) * g.uniform(0.5, 1.5, 16)
        p[11] = g.uniform(-1, 1)             # Q
        p[12] = g.uniform(0, 2) * (k % 2)    # k_micro
        e, i = g.uniform(0, 1, 2)
        c = g.uniform(-2, 2)
        rate = float(g.integers(0, 40)) * (k % 3 != 0)
        dt = [0.1, 0.05, 0.025][k % 3]
        if k % 7 == 0:
            p[2] = 0.0                       # PopRate with no sources
        cases.append((p, e, i, c, rate, dt))
    text = "".join(" ".join(repr(float(v)) for v in [*p, e, i, c, rate, dt]) + "\n" for p, e, i, c, rate, dt in cases)
    out = subprocess.run([str(exe)], input=text, check=True, capture_output=True, text=True).stdout.split("\n")
    for (p, e, i, c, rate, dt), line in zip(cases, out):
        got = [float.fromhex(x) for x in line.split()]
        u = c + p[12] * (rate / p[13])  # the REGION's input, as the oracle's callers form it
        e_ref, i_ref = oracle.wc_rk4(list(p), e, i, u, dt)
        assert got[0] == e_ref and got[1] == i_ref
        assert got[2] == c * e                                 # LinearDelay: weight x E
        assert got[3] == p[13] * e                             # PoissonDrive: rate_scale x E
        assert got[4] == (rate / (p[2] * dt) * 1000.0 if p[2] > 0 else 0.0)  # PopRate
        assert got[5] == 0.0                                   # the count restarts


@pytest.mark.parametrize("text,msg", [
    ("NEURON { SUFFIX a }\nMIND { ROLE SYNAPSE }", "unknown MIND role synapse"),
    ("NEURON { SUFFIX a }\nMIND { ROLE REGION FLAVOR x }", "unknown MIND declaration FLAVOR"),
    ("NEURON { SUFFIX a }\nASSIGNED { y }\nBREAKPOINT { y = 1 }", "no MIND ROLE"),
])
def test_role_errors_name_file_and_line(nmodlc, tmp_path, text, msg):
    f = tmp_path / "bad.mod"
    f.write_text(text)
    code, err = _lower(nmodlc, tmp_path, [str(f)])
    assert code != 0 and msg in err and err.startswith(f"{f}:")


def test_runtime_compilation_refuses_role_mechanisms():
    with pytest.raises(ValueError, match="MIND ROLE region"):
        compile_mod(open(os.path.join(MOD, "WilsonCowan.mod")).read(), "WilsonCowan.mod")
