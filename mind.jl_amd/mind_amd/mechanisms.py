"""Mechanisms: the builtin table and runtime compilation of MOD files.

Mirrors the reference's `builtin(:name)`, `compile_mod` and `load_mod`
(src/compiler/frontend.jl:58-110). The ten builtins are the mechanisms of the
simulation step; `compile_mod` / `load_mod` lower a MOD text through libmind's
NMODL frontend (mind_compile_mod) into a new mechanism slot, whose device
phases are compiled with hipRTC when a model that uses it is built. Array
fields take one column per element, named name[1], name[2], ...; field order
is the column order of the packed mechanism data (include/mind.h).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field


@dataclass(frozen=True)
class Mechanism:
    name: str
    id: int
    kind: str  # "density", "point" or "artificial"
    fields: tuple
    has_net_receive: bool = False
    index: dict = field(default_factory=dict, compare=False, hash=False)
    compiled: bool = False    # a runtime slot (compile_mod / load_mod)
    has_state: bool = False   # compiled: BREAKPOINT has SOLVE statements
    defaults: tuple = ()      # compiled: column defaults (PARAMETER values, ion values)
    digest: int = 0           # compiled: SHA-256 prefix of the MOD text
    filename: str = ""

    def field_index(self, name: str, index: int = 1) -> int:
        """Column of a field; of element `index` (1-based, as the reference's
        `index` keyword) of an array field, whose columns are named name[1], name[2], ..."""
        if name in self.fields:
            if index != 1:
                raise ValueError(f"{self.name}.{name} is not an array field")
            return self.fields.index(name)
        if f"{name}[1]" in self.fields:
            if not 1 <= index <= self.extent(name):
                raise ValueError(f"index {index} is outside {self.name}.{name}")
            return self.fields.index(f"{name}[1]") + index - 1
        raise ValueError(f"mechanism {self.name} has no field {name}")

    def extent(self, name: str) -> int:
        """Element count of an array field (1 for a scalar field)."""
        return sum(1 for f in self.fields if f.startswith(f"{name}[")) or 1


_TABLE = {
    "hh": Mechanism(
        "hh", 0, "density",
        ("gnabar", "gkbar", "gl", "el", "m", "h", "n", "ena", "ek", "gna", "gk", "ina", "ik",
         "il", "minf", "hinf", "ninf", "mtau", "htau", "ntau"),
    ),
    "pas": Mechanism("pas", 1, "density", ("g", "e", "i")),
    "ExpSyn": Mechanism("ExpSyn", 2, "point", ("tau", "e", "g", "i"), True),
    "Exp2Syn": Mechanism(
        "Exp2Syn", 3, "point", ("tau1", "tau2", "e", "A", "B", "i", "g", "factor"), True
    ),
    "IClamp": Mechanism("IClamp", 4, "point", ("del", "dur", "amp", "i")),
    "NetStim": Mechanism(
        "NetStim", 5, "artificial",
        ("interval", "number", "start", "noise", "event", "on", "ispike"), True,
    ),
    "AlphaSynapse": Mechanism("AlphaSynapse", 6, "point", ("onset", "tau", "gmax", "e", "i", "g")),
    "PatternStim": Mechanism("PatternStim", 7, "artificial", ()),
    "VClamp": Mechanism(
        "VClamp", 9, "point",
        ("dur[1]", "dur[2]", "dur[3]", "amp[1]", "amp[2]", "amp[3]", "gain", "rstim", "tau1", "tau2", "e0",
         "vo0", "vi0", "fac", "e", "vo", "vi", "i", "stim", "tc"),
    ),
    "SEClamp": Mechanism(
        "SEClamp", 8, "point",
        ("rs", "dur1", "amp1", "dur2", "amp2", "dur3", "amp3", "i", "vc", "tc2", "tc3", "on"),
    ),
}

BY_ID = {m.id: m for m in _TABLE.values()}


_KINDS = {0: "density", 1: "point", 2: "artificial"}


def compile_mod(source: str, filename: str = "<mod>") -> Mechanism:
    """Lower MOD text into a runtime mechanism slot (compile_mod, frontend.jl:74-110).
    Compiling the same text and filename again returns the same mechanism.
    Malformed or unsupported input raises ValueError naming the file and line."""
    from . import _abi

    L = _abi.lib()
    raw = source.encode()
    mech = C.c_int32(-1)
    code = L.mind_compile_mod(raw, len(raw), str(filename).encode(), C.byref(mech))
    if code != 0:
        message = L.mind_last_error().decode(errors="replace")
        if code == 1:
            raise ValueError(message)
        raise _abi.MindError(code, message)
    if mech.value in BY_ID:
        return BY_ID[mech.value]
    desc = _abi.MechDesc()
    _abi.check(L.mind_mech_info(mech.value, C.byref(desc)))
    n = desc.n_fields
    names = (C.c_char_p * max(n, 1))()
    defaults = (C.c_double * max(n, 1))()
    count = C.c_int32(0)
    _abi.check(L.mind_mech_fields(mech.value, names, defaults, n, C.byref(count)))
    m = Mechanism(desc.name.decode(), mech.value, _KINDS[desc.kind], tuple(names[k].decode() for k in range(n)),
                  bool(desc.has_net_receive), compiled=True, has_state=bool(desc.has_state),
                  defaults=tuple(defaults[k] for k in range(n)), digest=int(desc.digest),
                  filename=desc.filename.decode())
    BY_ID[m.id] = m
    return m


def load_mod(path) -> Mechanism:
    """compile_mod of a MOD file (load_mod, frontend.jl:74-80)."""
    with open(path, encoding="utf-8", newline="") as f:
        return compile_mod(f.read(), os.fspath(path))


def lowering(mechanism: Mechanism) -> str:
    """C++ text of a compiled mechanism's phases (namespace mind::user), as the
    device compiles them."""
    from . import _abi

    text = C.c_char_p()
    _abi.check(_abi.lib().mind_mech_lowering(mechanism.id, C.byref(text)))
    return text.value.decode()


@dataclass(frozen=True)
class IonUse:
    """One USEION statement: the ion, its VALENCE (None if not declared) and per ion
    column (reversal, internal, external concentration, current) the field column
    read into / written from (-1 for none); reads include written concentrations."""

    name: str
    valence: float | None
    read: tuple
    write: tuple


_IONS: dict = {}


def ions(mechanism) -> tuple:
    """The USEION statements of a mechanism (hh: na and k), from libmind."""
    from . import _abi

    mech = resolve(mechanism)
    if mech.id not in _IONS:
        L = _abi.lib()
        out, count, k = [], C.c_int32(0), 0
        while True:
            name, has_v, val = C.c_char_p(), C.c_int32(0), C.c_double(0.0)
            rd, wr = (C.c_int32 * 4)(), (C.c_int32 * 4)()
            _abi.check(L.mind_mech_ions(mech.id, k, C.byref(name), C.byref(has_v), C.byref(val), rd, wr,
                                        C.byref(count)))
            if k >= count.value:
                break
            out.append(IonUse(name.value.decode(), val.value if has_v.value else None, tuple(rd), tuple(wr)))
            k += 1
        _IONS[mech.id] = tuple(out)
    return _IONS[mech.id]


def columns_info(mechanism):
    """Per column: (is a STATE variable, touched by NET_RECEIVE), from libmind."""
    from . import _abi

    mech = resolve(mechanism)
    n = len(mech.fields)
    st, rc, count = (C.c_uint8 * max(1, n))(), (C.c_uint8 * max(1, n))(), C.c_int32(0)
    _abi.check(_abi.lib().mind_mech_columns(mech.id, st, rc, n, C.byref(count)))
    return tuple(bool(st[k]) for k in range(n)), tuple(bool(rc[k]) for k in range(n))


def globals_of(mechanism) -> tuple:
    """Names of a mechanism's GLOBAL fields (one value every instance shares), from
    libmind; the builtins declare none."""
    from . import _abi

    mech = resolve(mechanism)
    if mech.id not in _GLOBALS:
        n = len(mech.fields)
        flags, count = (C.c_uint8 * max(1, n))(), C.c_int32(0)
        _abi.check(_abi.lib().mind_mech_globals(mech.id, flags, n, C.byref(count)))
        _GLOBALS[mech.id] = tuple(f for k, f in enumerate(mech.fields) if flags[k])
    return _GLOBALS[mech.id]


_GLOBALS: dict = {}


def builtin(name) -> Mechanism:
    """Select a packaged builtin mechanism by name (hh, pas, ExpSyn, Exp2Syn, IClamp, NetStim,
    AlphaSynapse, PatternStim, SEClamp, VClamp)."""
    key = str(name).lstrip(":")
    if key not in _TABLE:
        raise ValueError(f"unknown mechanism {key}; load its MOD with load_mod")
    return _TABLE[key]


def resolve(mechanism) -> Mechanism:
    return mechanism if isinstance(mechanism, Mechanism) else builtin(mechanism)
