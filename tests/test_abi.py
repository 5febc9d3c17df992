"""The C ABI library loads without a GPU and exports every function include/mind.h declares."""
import os
import re

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
