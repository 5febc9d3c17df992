"""The C ABI library loads without a GPU and exports every function include/mind.h declares."""
import os
import re

import pytest

from mind_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "mind.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mind_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_interface():
    names = declared()
    for core in ("mind_build", "mind_initialize", "mind_run", "mind_step", "mind_build_macro",
                 "mind_run_macro", "mind_mlp_train", "mind_run_surrogate", "mind_last_error"):
        assert core in names
    assert len(names) >= 35


def test_library_exports_every_declared_symbol():
    lib = _abi.lib()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_device_count_without_gpu_is_safe():
    # counting devices never faults; 0 here, >= 1 on a GPU machine
    assert _abi.device_count() >= 0


def _c_structs():
    """Field (name, kind) lists of the typedef'd structs in include/mind.h."""
    import re
    text = open(os.path.join(ROOT, "include", "mind.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for name, body in re.findall(r"typedef struct (\w+) \{(.*?)\} \1;", text, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            m = re.match(r"(?:const )?(\w+)\s*(\*?)\s*(?:const\s*\*\s*)?(\w+(?:\s*,\s*\w+)*)$", decl)
            assert m, decl
            ctype, ptr, names = m.groups()
            kind = "p" if ptr else {"double": "f8", "int32_t": "i4", "int64_t": "i8", "uint64_t": "u8"}.get(ctype, "?")
            fields.extend((fname.strip(), kind) for fname in names.split(","))
        out[name] = fields
    return out


@pytest.mark.parametrize("cname,pyname,jlname", [
    ("mind_model_desc", "ModelDesc", "ModelDesc"), ("mind_build_options", "BuildOptions", "BuildOptions"),
    ("mind_record", "Record", "RecordSpec"), ("mind_result", "Result", "ResultC"),
    ("mind_mech_desc", "MechDesc", "MechDesc"), ("mind_macro_desc", "macro.MacroDesc", "MacroDesc"),
    ("mind_macro_result", "macro.MacroResultC", "MacroResultC"),
    ("mind_surrogate_desc", "surrogate.SurrogateDescC", "SurrogateDesc"),
])
def test_bindings_match_the_header_structs(cname, pyname, jlname):
    """The ctypes structures of mind_amd and the Julia structs of MINDAMD.jl lay out the
    header's structs field for field (name, width, pointer or not), in order."""
    import ctypes as C
    import re

    import importlib

    want = _c_structs()[cname]
    kinds = {C.c_double: "f8", C.c_int32: "i4", C.c_int64: "i8", C.c_uint64: "u8"}
    module, _, cls = pyname.rpartition(".")
    mod = importlib.import_module(f"mind_amd.{module or '_abi'}")
    py = [(n, kinds.get(t, "p" if issubclass(t, (C._Pointer, C.c_void_p, C.c_char_p)) else "?"))
          for n, t in getattr(mod, cls)._fields_]
    assert py == want
    jl = open(os.path.join(ROOT, "mind.jl_amd", "julia", "MINDAMD.jl")).read()
    body = re.search(rf"^struct {jlname}\n(.*?)^end", jl, flags=re.S | re.M).group(1)
    jkinds = {"Float64": "f8", "Int32": "i4", "Int64": "i8", "UInt64": "u8"}
    fields = [(n, "p" if t.startswith("Ptr") or t == "Cstring" else jkinds[t])
              for n, t in re.findall(r"^\s+(\w+)::(\S+)", body, flags=re.M)]
    assert fields == want
