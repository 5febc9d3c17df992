"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to
check and time the HIP path; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "mind.jl_amd"))
from mind_amd import _abi  # noqa: E402

LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"oracle not built ({LIB})")
        L = C.CDLL(LIB)
        P = C.POINTER
        L.oracle_build.restype = C.c_void_p
        L.oracle_build.argtypes = [P(_abi.ModelDesc), C.c_double, C.c_uint64]
        L.oracle_nernst.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double]
        L.oracle_nernst.restype = C.c_double
        L.oracle_free.argtypes = [C.c_void_p]
        L.oracle_set_secondorder.argtypes = [C.c_void_p, C.c_int32]
        L.oracle_set_threads.argtypes = [C.c_void_p, C.c_int32]
        L.oracle_row_count.argtypes = [C.c_void_p, C.c_int, P(C.c_int64)]
        L.oracle_inject_event.argtypes = [C.c_void_p, C.c_int64, C.c_double]
        L.oracle_inject_spike.argtypes = [C.c_void_p, C.c_int64, C.c_double]
        L.oracle_initialize.argtypes = [C.c_void_p]
        L.oracle_run.argtypes = [C.c_void_p, C.c_double, P(_abi.Record), C.c_int64, C.c_int32,
                                 P(_abi.Result)]
        L.oracle_result_free.argtypes = [P(_abi.Result)]
        L.oracle_read_voltage.argtypes = [C.c_void_p, P(C.c_double), C.c_int64]
        L.oracle_read_field.argtypes = [C.c_void_p, C.c_int, C.c_int, P(C.c_double), C.c_int64]
        L.oracle_node_count.argtypes = [C.c_void_p, P(C.c_int64), P(C.c_int64)]
        L.oracle_geometry.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_double), P(C.c_double),
                                      P(C.c_double), P(C.c_double)]
        L.oracle_location_node.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_double,
                                           C.c_int32, P(C.c_int64)]
        L.oracle_point_row.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32)]
        L.oracle_log_deliveries.argtypes = [C.c_void_p, C.c_int32]
        L.oracle_delivery_log.restype = C.c_int64
        L.oracle_delivery_log.argtypes = [C.c_void_p, P(C.c_double), P(C.c_int64), P(C.c_int64),
                                          P(C.c_int32), C.c_int64]
        L.oracle_philox.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
        L.oracle_uniform_stream.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int64,
                                            P(C.c_double)]
        L.oracle_hines.argtypes = [C.c_int64, P(C.c_int32), P(C.c_double), P(C.c_double),
                                   P(C.c_double)]
        L.oracle_exp.restype = C.c_double
        L.oracle_exp.argtypes = [C.c_double]
        L.oracle_register_mech.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, P(C.c_double),
                                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_log.restype = C.c_double
        L.oracle_log.argtypes = [C.c_double]
        L.oracle_last_error.restype = C.c_char_p
        _lib = L
    return _lib


_registered = {}  # runtime mechanism id -> its host build (kept loaded)

_WRAPPERS = """
extern "C" {{
void oracle_u{i}_current(double* data, int64_t rows, int64_t r, double v, double t, double dt, double celsius,
                        double* value, double* shifted, int32_t store, double* ish) {{
  mind::user::u{i}_current(mind::Cols{{data, rows}}, r, v, t, dt, *value, *shifted, store != 0, celsius, ish);
}}
void oracle_u{i}_advance(double* data, int64_t rows, int64_t r, double v, double t, double dt, double celsius) {{
  mind::user::u{i}_advance(mind::Cols{{data, rows}}, r, v, t, dt, celsius);
}}
void oracle_u{i}_initial(double* data, int64_t rows, int64_t r, double v, double celsius) {{
  mind::user::u{i}_initial(mind::Cols{{data, rows}}, r, v, celsius, nullptr, 0);
}}
void oracle_u{i}_receive(double* data, int64_t rows, int64_t r, double t, double weight) {{
  mind::user::u{i}_receive(mind::Cols{{data, rows}}, r, t, weight, 0.0, nullptr, 0);
}}
}}
"""


def _host_build(mech):
    """g++ build of libmind's lowering of a runtime-compiled mechanism (the text the
    device compiles with hipRTC), cached by content under the temp directory."""
    import hashlib
    import subprocess
    import tempfile

    from mind_amd import mechanisms

    csrc = os.path.join(os.path.dirname(HERE), "mind.jl_amd", "csrc")
    src = '#include "mech.h"\n' + mechanisms.lowering(mech) + _WRAPPERS.format(i=mech.id)
    key = hashlib.sha256(src.encode()).hexdigest()[:16]
    cache = os.path.join(tempfile.gettempdir(), f"mind_oracle_mechs_{os.getuid()}")
    os.makedirs(cache, exist_ok=True)
    so = os.path.join(cache, f"u{mech.id}_{key}.so")
    if not os.path.exists(so):
        cpp = os.path.join(cache, f"u{mech.id}_{key}.cpp")
        with open(cpp, "w") as f:
            f.write(src)
        tmp = f"{so}.{os.getpid()}.tmp"
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                        f"-I{csrc}", "-o", tmp, cpp], check=True, capture_output=True, text=True)
        os.replace(tmp, so)
    return C.CDLL(so)


def register_mechanisms(model):
    """Registers with the oracle every runtime-compiled mechanism the model uses."""
    from mind_amd import mechanisms
    from mind_amd.mechanisms import BY_ID

    ids = {mid for (_, _, mid, _) in model.density} | {p[0] for p in model.points}
    for mid in sorted(i for i in ids if i >= 10):
        if mid in _registered:
            continue
        mech = BY_ID[mid]
        h = _host_build(mech)
        fns = [C.cast(getattr(h, f"oracle_u{mid}_{ph}"), C.c_void_p)
               for ph in ("current", "advance", "initial", "receive")]
        n = len(mech.fields)
        defaults = (C.c_double * max(1, n))(*mech.defaults)
        rc = lib().oracle_register_mech(mid, n, 1 if mech.kind == "point" else 0, int(mech.has_state), defaults,
                                        *fns)
        if rc:
            raise RuntimeError(lib().oracle_last_error().decode())
        uses = mechanisms.ions(mech)
        is_state, in_recv = mechanisms.columns_info(mech)
        m = max(1, len(uses))
        names = (C.c_char_p * m)(*[u.name.encode() for u in uses])
        has_v = (C.c_int32 * m)(*[int(u.valence is not None) for u in uses])
        val = (C.c_double * m)(*[u.valence or 0.0 for u in uses])
        rd = (C.c_int32 * (4 * m))(*[c for u in uses for c in u.read])
        wr = (C.c_int32 * (4 * m))(*[c for u in uses for c in u.write])
        st = (C.c_uint8 * max(1, n))(*is_state)
        recv = [k for k in range(n) if in_recv[k]]
        rf = (C.c_int32 * max(1, len(recv)))(*recv)
        rc = lib().oracle_register_ions(mid, len(uses), names, has_v, val, rd, wr, st, len(recv), rf)
        if rc:
            raise RuntimeError(lib().oracle_last_error().decode())
        _registered[mid] = h


def default_threads():
    """Threads for the oracle's membrane phases: MIND_ORACLE_THREADS, else the
    CPUs this process may use, at most 16."""
    env = os.environ.get("MIND_ORACLE_THREADS")
    if env:
        return max(1, int(env))
    return max(1, min(16, len(os.sched_getaffinity(0))))


class OracleSim:
    """The oracle's view of one MicroModel (same descriptor as libmind)."""

    def __init__(self, model, dt=0.025, seed=1, secondorder=0, threads=None):
        register_mechanisms(model)
        desc, keep = model.descriptor()
        self.model = model
        self.dt = dt
        self.h = lib().oracle_build(C.byref(desc), float(dt), int(seed))
        del keep
        if not self.h:
            raise RuntimeError("oracle build failed")
        if secondorder and lib().oracle_set_secondorder(self.h, int(secondorder)):
            raise ValueError(lib().oracle_last_error().decode())
        self.set_threads(default_threads() if threads is None else threads)

    def set_threads(self, threads):
        """OpenMP threads of the membrane phases (results do not depend on it)."""
        if lib().oracle_set_threads(self.h, int(threads)):
            raise ValueError(lib().oracle_last_error().decode())
        return self

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_free(self.h)
            self.h = None

    def initialize(self):
        rc = lib().oracle_initialize(self.h)
        if rc:
            raise RuntimeError(f"oracle initialize failed [{rc}]: {lib().oracle_last_error()}")
        return self

    def run(self, stop, records=(), record_spikes=True):
        recs = (_abi.Record * max(1, len(records)))()
        for i, r in enumerate(records):
            recs[i] = self.model.record_struct(r.spec)
        res = _abi.Result()
        rc = lib().oracle_run(self.h, float(stop), recs, len(records), int(bool(record_spikes)),
                              C.byref(res))
        if rc:
            raise RuntimeError(f"oracle run failed [{rc}]: {lib().oracle_last_error().decode()}")
        try:
            time, traces, st, ss, secs = _abi.result_arrays(res)
        finally:
            lib().oracle_result_free(C.byref(res))
        from mind_amd.model import SimulationResult
        return SimulationResult(time, {r.name: traces[i] for i, r in enumerate(records)}, st, ss,
                                secs)

    def inject_event(self, connection, time):
        """inject_event! with the 1-based connection number connect() returned."""
        if lib().oracle_inject_event(self.h, int(connection) - 1, float(time)):
            raise ValueError(lib().oracle_last_error().decode())
        return self

    def inject_spike(self, source, time):
        if lib().oracle_inject_spike(self.h, int(source.index), float(time)):
            raise ValueError(lib().oracle_last_error().decode())
        return self

    def counts(self):
        n, s = C.c_int64(), C.c_int64()
        lib().oracle_node_count(self.h, C.byref(n), C.byref(s))
        return n.value, s.value

    def point_row(self, point):
        """Row of a point process in its mechanism's data."""
        out = C.c_int32()
        if lib().oracle_point_row(self.h, int(point.index), C.byref(out)):
            raise ValueError(lib().oracle_last_error().decode())
        return out.value

    def voltage(self):
        n, _ = self.counts()
        out = np.zeros(n)
        lib().oracle_read_voltage(self.h, out.ctypes.data_as(C.POINTER(C.c_double)), n)
        return out

    def rows(self, mech_id):
        n = C.c_int64()
        lib().oracle_row_count(self.h, int(mech_id), C.byref(n))
        return n.value

    def field(self, mech_id, field, rows=None):
        rows = self.rows(mech_id) if rows is None else rows
        out = np.zeros(rows)
        rc = lib().oracle_read_field(self.h, mech_id, field,
                                     out.ctypes.data_as(C.POINTER(C.c_double)), rows)
        if rc:
            raise RuntimeError(lib().oracle_last_error().decode())
        return out

    def geometry(self):
        n, _ = self.counts()
        parent = np.zeros(n, np.int32)
        area, cap, axial, adiag = (np.zeros(n) for _ in range(4))
        P = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
        lib().oracle_geometry(self.h, P(parent, C.c_int32), P(area, C.c_double),
                              P(cap, C.c_double), P(axial, C.c_double), P(adiag, C.c_double))
        return dict(parent=parent, area=area, cap=cap, axial=axial, adiag=adiag)

    def location_node(self, cell, section, x, range_=False):
        out = C.c_int64()
        lib().oracle_location_node(self.h, cell, section, float(x), int(range_), C.byref(out))
        return out.value

    def log_deliveries(self, on=True):
        lib().oracle_log_deliveries(self.h, int(on))

    def delivery_log(self, cap=1 << 22):
        """(time, ticket, edge, target point) of every applied event, in delivery order."""
        t = np.zeros(cap)
        tk = np.zeros(cap, np.int64)
        ed = np.zeros(cap, np.int64)
        tg = np.zeros(cap, np.int32)
        n = lib().oracle_delivery_log(self.h, t.ctypes.data_as(C.POINTER(C.c_double)),
                                      tk.ctypes.data_as(C.POINTER(C.c_int64)),
                                      ed.ctypes.data_as(C.POINTER(C.c_int64)),
                                      tg.ctypes.data_as(C.POINTER(C.c_int32)), cap)
        n = min(n, cap)
        return t[:n], tk[:n], ed[:n], tg[:n]


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().oracle_philox(c, k, o)
    return list(o)


def uniform_stream(seed, logical_id, digest, n):
    out = np.zeros(n)
    lib().oracle_uniform_stream(seed, logical_id, digest, n,
                                out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def hines(parent, g, d, b):
    parent = np.ascontiguousarray(parent, np.int32)
    g = np.ascontiguousarray(g, np.float64)
    d = np.array(d, np.float64)
    b = np.array(b, np.float64)
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
    rc = lib().oracle_hines(len(parent), P(parent, C.c_int32), P(g, C.c_double),
                            P(d, C.c_double), P(b, C.c_double))
    if rc:
        raise ZeroDivisionError("singular")
    return b


def nernst(ci, co, z, celsius):
    """The oracle's Nernst potential (mV)."""
    return lib().oracle_nernst(float(ci), float(co), float(z), float(celsius))


def exp(x):
    return lib().oracle_exp(float(x))


def log(x):
    return lib().oracle_log(float(x))


class OracleMacro:
    """The oracle's view of one MacroModel (same descriptors as libmind)."""

    def __init__(self, model, dt=0.025, dt_macro=0.1, seed=1):
        from mind_amd import macro as M

        L = _macro_lib()
        self.model = model
        register_mechanisms(model.micro)
        d, k1 = model.micro.descriptor()
        m, k2 = M.macro_descriptor(model, dt_macro)
        self.h = L.oracle_build_macro(C.byref(d), C.byref(m), float(dt), int(seed))
        del k1, k2

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_free_macro(self.h)
            self.h = None

    def initialize(self):
        rc = lib().oracle_initialize_macro(self.h)
        if rc:
            raise RuntimeError(lib().oracle_last_error().decode())
        return self

    def run(self, stop, record_spikes=True):
        from mind_amd import macro as M

        res = M.MacroResultC()
        rc = lib().oracle_run_macro(self.h, float(stop), int(bool(record_spikes)), C.byref(res))
        if rc:
            raise RuntimeError(lib().oracle_last_error().decode())
        try:
            return M.result_from_c(res, len(self.model.connectivity.labels))
        finally:
            lib().oracle_macro_result_free(C.byref(res))


def _macro_lib():
    from mind_amd import macro as M

    L = lib()
    if not getattr(L, "_macro", False):
        L.oracle_build_macro.restype = C.c_void_p
        L.oracle_build_macro.argtypes = [C.POINTER(_abi.ModelDesc), C.POINTER(M.MacroDesc),
                                         C.c_double, C.c_uint64]
        L.oracle_initialize_macro.argtypes = [C.c_void_p]
        L.oracle_run_macro.argtypes = [C.c_void_p, C.c_double, C.c_int32,
                                       C.POINTER(M.MacroResultC)]
        L.oracle_macro_result_free.argtypes = [C.POINTER(M.MacroResultC)]
        L.oracle_free_macro.argtypes = [C.c_void_p]
        L.oracle_wc_rk4.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.c_double, C.c_double]
        L._macro = True
    return L


def wc_rk4(params, e, i, u, h):
    p = (C.c_double * len(params))(*params)
    ce, ci = C.c_double(e), C.c_double(i)
    _macro_lib().oracle_wc_rk4(p, C.byref(ce), C.byref(ci), float(u), float(h))
    return ce.value, ci.value
