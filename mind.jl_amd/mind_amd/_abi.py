"""ctypes mirror of include/mind.h and the loader of the in-tree libmind.so.

The library is the product: every simulation call goes through it. If it is
missing the import of this module fails loudly; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PACKAGE_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("MIND_LIBRARY") or os.path.join(PACKAGE_DIR, "csrc", "libmind.so")

MIND_METHOD_PLAIN = 0
MIND_METHOD_BATCHED = 1

ERRORS = {
    1: "invalid argument",
    2: "HIP runtime failure",
    3: "device allocation failed",
    5: "invalid self-event delay or target",
    6: "singular or non-finite cable system",
    8: "non-finite numerical state",
    9: "event ring capacity exceeded",
    10: "model feature outside the device path",
}

_p = C.POINTER


class ModelDesc(C.Structure):
    _fields_ = [
        ("celsius", C.c_double),
        ("n_morphologies", C.c_int32),
        ("morph_offsets", _p(C.c_int64)),
        ("sec_parent", _p(C.c_int32)),
        ("sec_parent_x", _p(C.c_double)),
        ("sec_child_x", _p(C.c_double)),
        ("sec_nseg", _p(C.c_int32)),
        ("sec_length", _p(C.c_double)),
        ("sec_diameter", _p(C.c_double)),
        ("sec_Ra", _p(C.c_double)),
        ("sec_cm", _p(C.c_double)),
        ("sec_point_offsets", _p(C.c_int64)),
        ("points", _p(C.c_double)),
        ("n_cells", C.c_int32),
        ("cell_morphology", _p(C.c_int32)),
        ("cell_gid", _p(C.c_int64)),
        ("cell_voltage", _p(C.c_double)),
        ("n_density", C.c_int64),
        ("dens_cell", _p(C.c_int32)),
        ("dens_section", _p(C.c_int32)),
        ("dens_mech", _p(C.c_int32)),
        ("dens_param_offsets", _p(C.c_int64)),
        ("n_points", C.c_int64),
        ("pt_mech", _p(C.c_int32)),
        ("pt_cell", _p(C.c_int32)),
        ("pt_section", _p(C.c_int32)),
        ("pt_x", _p(C.c_double)),
        ("pt_param_offsets", _p(C.c_int64)),
        ("param_field", _p(C.c_int32)),
        ("param_value", _p(C.c_double)),
        ("n_sources", C.c_int64),
        ("src_kind", _p(C.c_int32)),
        ("src_cell", _p(C.c_int32)),
        ("src_section", _p(C.c_int32)),
        ("src_x", _p(C.c_double)),
        ("src_threshold", _p(C.c_double)),
        ("src_point", _p(C.c_int32)),
        ("src_sid", _p(C.c_int64)),
        ("n_connections", C.c_int64),
        ("con_source", _p(C.c_int32)),
        ("con_target", _p(C.c_int32)),
        ("con_weight", _p(C.c_double)),
        ("con_delay", _p(C.c_double)),
        ("pt_spike_offsets", _p(C.c_int64)),
        ("pt_spike_times", _p(C.c_double)),
        ("n_species", C.c_int32),
        ("species_name", _p(C.c_char_p)),
        ("species_valence", _p(C.c_double)),
        ("species_internal", _p(C.c_double)),
        ("species_external", _p(C.c_double)),
        ("species_reversal", _p(C.c_double)),
        ("species_nernst", _p(C.c_int32)),
        ("n_ion_settings", C.c_int64),
        ("ionset_cell", _p(C.c_int32)),
        ("ionset_section", _p(C.c_int32)),
        ("ionset_species", _p(C.c_int32)),
        ("ionset_what", _p(C.c_int32)),
        ("ionset_value", _p(C.c_double)),
    ]


class BuildOptions(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("seed", C.c_uint64),
        ("method", C.c_int32),
        ("cells_per_group", C.c_int32),
        ("window_steps", C.c_int32),
        ("event_capacity", C.c_int64),
        ("device", C.c_int32),
        ("secondorder", C.c_int32),
    ]


class Record(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("cell", C.c_int32),
        ("section", C.c_int32),
        ("x", C.c_double),
        ("point", C.c_int32),
        ("mech", C.c_int32),
        ("field", C.c_int32),
    ]


class Result(C.Structure):
    _fields_ = [
        ("n_samples", C.c_int64),
        ("time", _p(C.c_double)),
        ("n_records", C.c_int64),
        ("traces", _p(C.c_double)),
        ("n_spikes", C.c_int64),
        ("spike_times", _p(C.c_double)),
        ("spike_sids", _p(C.c_int64)),
        ("run_seconds", C.c_double),
    ]


class MechDesc(C.Structure):
    _fields_ = [
        ("name", C.c_char_p),
        ("filename", C.c_char_p),
        ("kind", C.c_int32),
        ("n_fields", C.c_int32),
        ("has_net_receive", C.c_int32),
        ("has_state", C.c_int32),
        ("compiled", C.c_int32),
        ("digest", C.c_uint64),
    ]


def as_ptr(array: np.ndarray, ctype):
    return array.ctypes.data_as(C.POINTER(ctype))


def result_arrays(res: Result):
    """Copy a Result into numpy arrays (time, traces[n_records, n_samples], spikes)."""
    ns, nr, nk = res.n_samples, res.n_records, res.n_spikes
    time = np.ctypeslib.as_array(res.time, (ns,)).copy() if ns else np.zeros(0)
    traces = (
        np.ctypeslib.as_array(res.traces, (nr * ns,)).reshape(nr, ns).copy()
        if nr and ns
        else np.zeros((nr, ns))
    )
    times = np.ctypeslib.as_array(res.spike_times, (nk,)).copy() if nk else np.zeros(0)
    sids = np.ctypeslib.as_array(res.spike_sids, (nk,)).copy() if nk else np.zeros(0, np.int64)
    return time, traces, times, sids, res.run_seconds


_lib = None


def lib():
    """Load libmind.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libmind.so is not built ({LIB_PATH}); run __graft_entry__.build() first"
        )
    L = C.CDLL(LIB_PATH)
    L.mind_build.argtypes = [_p(ModelDesc), _p(BuildOptions), _p(C.c_void_p)]
    L.mind_initialize.argtypes = [C.c_void_p]
    L.mind_run.argtypes = [C.c_void_p, C.c_double, _p(Record), C.c_int64, C.c_int32, _p(Result)]
    L.mind_step.argtypes = [C.c_void_p]
    L.mind_result_free.argtypes = [_p(Result)]
    L.mind_result_free.restype = None
    L.mind_state_destroy.argtypes = [C.c_void_p]
    L.mind_state_destroy.restype = None
    L.mind_node_count.argtypes = [C.c_void_p, _p(C.c_int64), _p(C.c_int64)]
    L.mind_read_voltage.argtypes = [C.c_void_p, _p(C.c_double), C.c_int64]
    L.mind_read_field.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _p(C.c_double), C.c_int64]
    L.mind_row_count.argtypes = [C.c_void_p, C.c_int32, _p(C.c_int64)]
    L.mind_time.argtypes = [C.c_void_p, _p(C.c_double), _p(C.c_int64)]
    L.mind_event_counts.argtypes = [C.c_void_p, _p(C.c_int64), _p(C.c_int64)]
    L.mind_read_value.argtypes = [C.c_void_p, _p(Record), _p(C.c_double)]
    L.mind_checkpoint.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, _p(C.c_size_t)]
    L.mind_restore.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t]
    L.mind_inject_event.argtypes = [C.c_void_p, C.c_int64, C.c_double]
    L.mind_inject_spike.argtypes = [C.c_void_p, C.c_int64, C.c_double]
    L.mind_connection_weights.argtypes = [C.c_void_p, C.c_int64, _p(C.c_double), C.c_int32, _p(C.c_int32)]
    L.mind_window_steps.argtypes = [C.c_void_p, _p(C.c_int32)]
    L.mind_batch_plan.argtypes = [C.c_void_p, _p(C.c_int64), _p(C.c_int32), _p(C.c_int32), _p(C.c_int64)]
    L.mind_batch_kernel.argtypes = [C.c_void_p, _p(C.c_char_p)]
    L.mind_set_batch_plan.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int32]
    L.mind_check_batch_plan.argtypes = [C.c_void_p, C.c_double, C.c_int64, C.c_int32, C.c_int32, _p(C.c_int64)]
    L.mind_pack_nodes.argtypes = [_p(ModelDesc), C.c_double, _p(C.c_int64), _p(C.c_int64),
                                  _p(C.c_int32), _p(C.c_double), _p(C.c_double), _p(C.c_double),
                                  _p(C.c_double)]
    L.mind_device_count.argtypes = [_p(C.c_int32)]
    L.mind_synchronize.argtypes = [C.c_void_p]
    L.mind_set_profiling.argtypes = [C.c_void_p, C.c_int32]
    L.mind_kernel_time.argtypes = [C.c_void_p, _p(C.c_int64), _p(C.c_int64), _p(C.c_double)]
    L.mind_batch_work.argtypes = [C.c_void_p, _p(C.c_int64)]
    L.mind_last_error.restype = C.c_char_p
    L.mind_build_info.restype = C.c_char_p
    L.mind_copy_bandwidth.argtypes = [C.c_int32, C.c_int64, C.c_int32, _p(C.c_double)]
    L.mind_compile_mod.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, _p(C.c_int32)]
    L.mind_mech_info.argtypes = [C.c_int32, _p(MechDesc)]
    L.mind_mech_fields.argtypes = [C.c_int32, _p(C.c_char_p), _p(C.c_double), C.c_int32, _p(C.c_int32)]
    L.mind_mech_lookup.argtypes = [C.c_char_p, _p(C.c_int32)]
    L.mind_mech_ions.argtypes = [C.c_int32, C.c_int32, _p(C.c_char_p), _p(C.c_int32), _p(C.c_double),
                                 _p(C.c_int32), _p(C.c_int32), _p(C.c_int32)]
    L.mind_mech_columns.argtypes = [C.c_int32, _p(C.c_uint8), _p(C.c_uint8), C.c_int32, _p(C.c_int32)]
    L.mind_mech_globals.argtypes = [C.c_int32, _p(C.c_uint8), C.c_int32, _p(C.c_int32)]
    L.mind_mech_lowering.argtypes = [C.c_int32, _p(C.c_char_p)]
    L.mind_rtc_check.argtypes = [_p(C.c_int32), C.c_int32, C.c_char_p]
    _lib = L
    return L


def build_info() -> str:
    """How the library's mechanism phases were built (NMODL lowering of which MOD files)."""
    return lib().mind_build_info().decode()


def device_count() -> int:
    n = C.c_int32(0)
    lib().mind_device_count(C.byref(n))
    return n.value


def check(code: int):
    if code != 0:
        message = lib().mind_last_error().decode(errors="replace")
        raise MindError(code, message)


class MindError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[{code}] {ERRORS.get(code, 'error')}: {message}")
        self.code = code
