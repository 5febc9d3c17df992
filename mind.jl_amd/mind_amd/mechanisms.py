"""Builtin mechanism table: names, ids and field columns.

Mirrors the reference's `builtin(:name)` (src/compiler/frontend.jl:58-72) for
the ten mechanisms of the simulation step (array fields take one column per
element, named name[1], name[2], ...). Field order is the column order of
the packed mechanism data (include/mind.h).
"""
from __future__ import annotations

from dataclasses import dataclass, field


@dataclass(frozen=True)
class Mechanism:
    name: str
    id: int
    kind: str  # "density", "point" or "artificial"
    fields: tuple
    has_net_receive: bool = False
    index: dict = field(default_factory=dict, compare=False, hash=False)

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


def builtin(name) -> Mechanism:
    """Select a packaged builtin mechanism by name (hh, pas, ExpSyn, Exp2Syn, IClamp, NetStim,
    AlphaSynapse, PatternStim, SEClamp, VClamp)."""
    key = str(name).lstrip(":")
    if key not in _TABLE:
        raise ValueError(f"unknown mechanism {key}; load its MOD with load_mod")
    return _TABLE[key]


def resolve(mechanism) -> Mechanism:
    return mechanism if isinstance(mechanism, Mechanism) else builtin(mechanism)
