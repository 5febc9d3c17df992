"""Model-building API of MIND.jl's micro layer, flattened into the C ABI.

Names and argument meanings follow the reference (Julia's `!` suffix dropped):
MicroModel (src/model/micro/model.jl:35-113), Section (morphology.jl:17-79),
add_population (model.jl:217-283), insert (model.jl:153-169), add_pointprocess
(model.jl:172-194), register_spike_source (model.jl:286-365, SID rules of
utils/model.jl:124-224), connect (model.jl:379-400), build (build.jl:120-220),
initialize (initialize.jl:256-267), run/step (step.jl:4-86) and the record
selectors (recording.jl:4-76). Model state lives in libmind; this module only
validates arguments and packs arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .mechanisms import Mechanism, resolve


class Section:
    """A cable section; `child_x` end (0 or 1) attaches at the parent's `parent_x`.

    Length and diameter in um, Ra in ohm*cm, cm in uF/cm2 (morphology.jl:31-79).
    """

    __slots__ = ("name", "label", "parent", "parent_x", "child_x", "nseg", "length",
                 "diameter", "points", "Ra", "cm")

    def __init__(self, name, label="dend", parent=None, parent_x=1.0, child_x=0.0, nseg=1,
                 length=100.0, diameter=1.0, points=(), Ra=100.0, cm=1.0):
        points = [tuple(map(float, p)) for p in points]
        if len(points) >= 2:
            length = float(sum(np.linalg.norm(np.subtract(points[i][:3], points[i - 1][:3]))
                               for i in range(1, len(points))))
            diameter = points[0][3]
        if not (isinstance(nseg, (int, np.integer)) and nseg > 0):
            raise ValueError("nseg must be a positive integer")
        if not 0.0 <= parent_x <= 1.0:
            raise ValueError("parent_x must be in [0, 1]")
        if child_x not in (0, 1):
            raise ValueError("child_x must be 0 or 1")
        if not all(np.isfinite([length, diameter, Ra, cm, parent_x])):
            raise ValueError("non-finite section parameter")
        if not (length > 0 and (diameter > 0 or (points and diameter == 0)) and Ra > 0
                and cm >= 0):
            raise ValueError("invalid section geometry or biophysics")
        self.name = str(name)
        self.label = str(label)
        self.parent = None if parent is None else str(parent)
        self.parent_x = float(parent_x)
        self.child_x = float(child_x)
        self.nseg = int(nseg)
        self.length = float(length)
        self.diameter = float(diameter)
        self.points = points
        self.Ra = float(Ra)
        self.cm = float(cm)


class Morphology:
    """Array form of a section list, shared by every cell of one population."""

    def __init__(self, names, labels, parent, parent_x, child_x, nseg, length, diameter, Ra, cm,
                 points=None):
        self.names = list(names)
        self.labels = list(labels)
        self.index = {n: i for i, n in enumerate(self.names)}
        if len(self.index) != len(self.names):
            raise ValueError("duplicate section name")
        self.parent = np.asarray(parent, np.int32)
        self.parent_x = np.asarray(parent_x, np.float64)
        self.child_x = np.asarray(child_x, np.float64)
        self.nseg = np.asarray(nseg, np.int32)
        self.length = np.asarray(length, np.float64)
        self.diameter = np.asarray(diameter, np.float64)
        self.Ra = np.asarray(Ra, np.float64)
        self.cm = np.asarray(cm, np.float64)
        self.points = points  # None or list of per-section point lists
        if np.any(self.nseg <= 0):
            raise ValueError("nseg must be a positive integer")

    @classmethod
    def from_sections(cls, sections):
        sections = list(sections)
        names = [s.name for s in sections]
        index = {n: i for i, n in enumerate(names)}
        parent = []
        for s in sections:
            if s.parent is None:
                parent.append(-1)
            elif s.parent not in index:
                raise ValueError(f"unknown parent section {s.parent}")
            else:
                parent.append(index[s.parent])
        points = [s.points for s in sections] if any(len(s.points) >= 2 for s in sections) else None
        return cls(names, [s.label for s in sections], parent, [s.parent_x for s in sections],
                   [s.child_x for s in sections], [s.nseg for s in sections],
                   [s.length for s in sections], [s.diameter for s in sections],
                   [s.Ra for s in sections], [s.cm for s in sections], points)

    def select(self, sections):
        """Section indices for :all, a label, a name or a list of names."""
        if isinstance(sections, str):
            key = sections.lstrip(":")
            if key == "all":
                return list(range(len(self.names)))
            if key in self.index:
                return [self.index[key]]
            picked = [i for i, label in enumerate(self.labels) if label == key]
            if not picked:
                raise ValueError(f"unknown section group {key}")
            return picked
        return [self.section_index(s) for s in sections]

    def section_index(self, name) -> int:
        key = str(name).lstrip(":")
        if key not in self.index:
            raise ValueError(f"unknown section {key}")
        return self.index[key]


@dataclass
class Cell:
    model: "MicroModel"
    index: int
    gid: int


@dataclass
class CellSection:
    """One section of one cell (reference CellSection, model.jl:73-76)."""
    cell: Cell
    name: str


def section(cell: Cell, name) -> CellSection:
    """Select one section by its exact name (model.jl:128-137)."""
    key = str(name).lstrip(":")
    cell.model._morph(cell).section_index(key)
    return CellSection(cell, key)


def section_group(cell: Cell, label) -> list:
    """Sections of one cell with a label (or all), in morphology order (model.jl:115-126)."""
    key = str(label).lstrip(":")
    morph = cell.model._morph(cell)
    names = [n for n, lab in zip(morph.names, morph.labels) if key == "all" or lab == key]
    if not names:
        raise ValueError(f"unknown section group {key}")
    return [CellSection(cell, n) for n in names]


@dataclass
class PointProcess:
    model: "MicroModel"
    index: int


@dataclass
class SpikeSource:
    model: "MicroModel"
    index: int
    sid: int


@dataclass
class SimulationResult:
    """SimulationResult (recording.jl:78-87) for micro runs."""

    time: np.ndarray
    traces: dict
    spike_times: np.ndarray
    spike_sources: np.ndarray
    run_seconds: float = 0.0


@dataclass(frozen=True)
class IonSpecies:
    """IonSpecies (src/model/micro/biophysics.jl:3-25): valence, concentrations (mM),
    reversal potential (mV) and whether the reversal follows the Nernst equation."""

    name: str
    valence: float
    internal: float
    external: float
    reversal: float = 0.0
    nernst: bool = False

    def __post_init__(self):
        if not np.isfinite(self.valence):
            raise ValueError("ion valence must be finite")
        if not all(np.isfinite(x) for x in (self.internal, self.external, self.reversal)):
            raise ValueError("non-finite ion parameter")
        if self.internal < 0 or self.external < 0:
            raise ValueError("ion concentrations must be nonnegative")


@dataclass(frozen=True)
class IonStyle:
    """IonStyle (biophysics.jl:27-52): c_style and e_style 0 (unused), 1 (parameter),
    2 (assigned), 3 (state); einit / eadvance request the Nernst reversal at
    initialization / every step; cinit resets concentrations to the species'."""

    c_style: int = 0
    e_style: int = 0
    einit: bool = False
    eadvance: bool = False
    cinit: bool = False

    def __post_init__(self):
        if not all(isinstance(v, (int, np.integer)) and 0 <= v <= 3 for v in (self.c_style, self.e_style)):
            raise ValueError("ion variable styles must be integers from 0 to 3")

    def bits(self) -> int:
        """ion_style_bits (utils/biophysics.jl:3-4)."""
        return (self.c_style + 4 * int(bool(self.cinit)) + 8 * self.e_style + 32 * int(bool(self.einit))
                + 64 * int(bool(self.eadvance)))


ion_style = IonStyle


def default_ions():
    """default_ions (utils/biophysics.jl:31-42)."""
    return {"na": IonSpecies("na", 1.0, 10.0, 140.0, 50.0),
            "k": IonSpecies("k", 1.0, 54.4, 2.5, -77.0),
            "ca": IonSpecies("ca", 2.0, 5.0e-5, 2.0, 132.4579341637009)}


_POLICIES = {"automatic": 0, "fixed": 1, "initial": 2, "advance": 3, "dynamic": 4}
# ion record fields -> column: e<ion>, <ion>i, <ion>o, i<ion>, di<ion>_dv_ (references.jl:161-172)
_ION_COLUMNS = ("e{}", "{}i", "{}o", "i{}", "di{}_dv_")


class MicroModel:
    def __init__(self, celsius=6.3, ions=()):
        if not (np.isfinite(celsius) and celsius > -273.15):
            raise ValueError("invalid temperature")
        self.celsius = float(celsius)
        self.ions = default_ions()
        for ion in (ions.values() if isinstance(ions, dict) else ions):
            if not isinstance(ion, IonSpecies):
                raise ValueError("ions takes IonSpecies values")
            self.ions[ion.name] = ion
        # section ion settings: (cell, section, ion name, what, value)
        self.ion_settings: list = []
        # GLOBAL values set for every instance of a mechanism: mech id -> {column: value}
        self.global_values: dict = {}
        self.morphologies: list[Morphology] = []
        self.cell_morph: list[int] = []
        self.cell_gid: list[int] = []
        self.cell_voltage: list[float] = []
        self.gids: set = set()
        # density insertions: (cell, section, mech id, params)
        self.density: list = []
        self._density_seen: set = set()
        # points: (mech, cell, section, x, params)
        self.points: list = []
        # sources: (kind, cell, section, x, threshold, point, sid)
        self.sources: list = []
        self.source_indices: dict = {}
        self.point_sources: dict = {}
        self.variable_sources: dict = {}
        self.next_source_id = 0
        # connections: (source, target, weight, delay)
        self.connections: list = []
        self.state = None
        self.options = None

    # -- cells ---------------------------------------------------------------
    def add_population(self, sections, count, first_gid=None, voltage=-65.0, ions=None):
        """Create `count` cells sharing one morphology (model.jl:217-283). ions maps a
        section selector to {ion: {internal, external, reversal, reversal_policy | style}}
        (cell_ion_values, utils/model.jl:280-320): values on every segment of the
        selected sections, the reversal policy (automatic, fixed, initial, advance,
        dynamic) or an IonStyle; any entry also places the ion there."""
        if count < 0:
            raise ValueError("population size must be nonnegative")
        morph = sections if isinstance(sections, Morphology) else Morphology.from_sections(sections)
        first_gid = len(self.cell_gid) if first_gid is None else int(first_gid)
        first_cell = len(self.cell_gid)
        if count == 0:
            return []
        self.morphologies.append(morph)
        m = len(self.morphologies) - 1
        cells = []
        for i in range(count):
            gid = first_gid + i
            if gid in self.gids:
                raise ValueError(f"duplicate cell gid {gid}")
            value = voltage(i + 1) if callable(voltage) else voltage
            if not np.isfinite(value):
                raise ValueError("non-finite initial voltage")
            self.gids.add(gid)
            self.cell_morph.append(m)
            self.cell_gid.append(gid)
            self.cell_voltage.append(float(value))
            cells.append(Cell(self, first_cell + i, gid))
        for selector, species in (ions or {}).items():
            secs = morph.select(selector)
            for ion, cfg in species.items():
                name = str(ion).lstrip(":")
                policy, style = cfg.get("reversal_policy"), cfg.get("style")
                if policy is not None and style is not None:
                    raise ValueError("choose either an ion style or a reversal policy")
                if policy is not None and str(policy).lstrip(":") not in _POLICIES:
                    raise ValueError("unknown ion reversal policy")
                if style is not None and not isinstance(style, IonStyle):
                    raise ValueError("style must be an IonStyle")
                for c in cells:
                    for sec in secs:
                        for what, key in ((0, "internal"), (1, "external"), (2, "reversal")):
                            if cfg.get(key) is None:
                                continue
                            v = float(cfg[key])
                            if not np.isfinite(v):
                                raise ValueError("non-finite ion value")
                            if what < 2 and v < 0:
                                raise ValueError("negative ion concentration")
                            self.ion_settings.append((c.index, sec, name, what, v))
                        if style is not None:
                            self.ion_settings.append((c.index, sec, name, 4, float(style.bits())))
                        else:
                            pol = _POLICIES[str(policy).lstrip(":")] if policy is not None else 0
                            self.ion_settings.append((c.index, sec, name, 3, float(pol)))
        return cells

    def _morph(self, cell: Cell) -> Morphology:
        return self.morphologies[self.cell_morph[cell.index]]

    # -- mechanisms ----------------------------------------------------------
    def _params(self, mech: Mechanism, params):
        """(column, value) overrides; an array field takes a sequence of its extent."""
        items = params.items() if isinstance(params, dict) else params
        out = []
        shared = _global_names(mech)
        for k, v in items:
            if str(k) in shared:
                raise ValueError(f"{k} is a GLOBAL of {mech.name}: set it for every instance with set_global")
            if isinstance(v, (list, tuple, np.ndarray)):
                if len(v) != mech.extent(str(k)) or mech.extent(str(k)) == 1 and str(k) in mech.fields:
                    raise ValueError(f"{mech.name}.{k} takes {mech.extent(str(k))} values")
                out.extend((mech.field_index(str(k), j + 1), float(x)) for j, x in enumerate(v))
            else:
                out.append((mech.field_index(str(k)), float(v)))
        return out

    def set_global(self, mechanism, fname, value, index=1):
        """The value of a mechanism's GLOBAL for every instance (the reference's global
        parameters, packed once per mechanism: pack_globals, src/model/build.jl:388)."""
        mech = resolve(mechanism)
        key = str(fname).lstrip(":")
        if key not in _global_names(mech):
            raise ValueError(f"{key} is not a shared GLOBAL of {mech.name}")
        if not np.isfinite(value):
            raise ValueError("a GLOBAL value must be finite")
        self.global_values.setdefault(mech.id, {})[mech.field_index(key, index)] = float(value)

    def _with_globals(self, mech_id, params):
        g = self.global_values.get(mech_id)
        return list(params) + sorted(g.items()) if g else params

    def insert(self, cell: Cell, mechanism, sections="all", params=None):
        """Insert a density mechanism where absent (model.jl:153-169)."""
        mech = resolve(mechanism)
        if mech.kind != "density":
            raise ValueError("cell insertion requires a distributed micro mechanism")
        pv = self._params(mech, params or {})
        for s in self._morph(cell).select(sections):
            key = (cell.index, s, mech.id)
            if key in self._density_seen:
                continue
            self._density_seen.add(key)
            self.density.append((cell.index, s, mech.id, pv))
        return cell

    def add_pointprocess(self, target, mechanism, section="soma", x=0.5, params=None, spike_times=()):
        """Create a point process on a cell, or an artificial cell on the model
        (create_pointprocess!, model/micro/utils/model.jl:72-105). `spike_times` are the
        spikes a PatternStim emits: finite, nonnegative and sorted."""
        mech = resolve(mechanism)
        times = [float(t) for t in spike_times]
        if not all(np.isfinite(t) and t >= 0.0 for t in times):
            raise ValueError("invalid spike times")
        if any(b < a for a, b in zip(times, times[1:])):
            raise ValueError("spike times must be sorted")
        if times and mech.name != "PatternStim":
            raise ValueError("spike_times applies to PatternStim only")
        if isinstance(target, MicroModel):
            if mech.kind != "artificial":
                raise ValueError(f"{mech.name} needs a cell location")
            cell, sec, x = -1, -1, 0.0
        else:
            if mech.kind == "density":
                raise ValueError("point creation requires a point process")
            cell = target.index
            sec = self._morph(target).section_index(section)
            if not 0.0 <= x <= 1.0:
                raise ValueError("section location must be in [0, 1]")
            if mech.kind == "artificial":  # an artificial cell reads no voltage: location unused
                cell, sec, x = -1, -1, 0.0
        self.points.append((mech.id, cell, sec, float(x), self._params(mech, params or {}), times))
        return PointProcess(self, len(self.points) - 1)

    # -- sources and connections --------------------------------------------
    def _new_source(self, entry, sid, key):
        identity = self.next_source_id if sid is None else int(sid)
        if identity < 0:
            raise ValueError("spike-source IDs must be nonnegative")
        if identity in self.source_indices:
            raise ValueError(f"duplicate spike-source ID {identity}")
        self.sources.append(entry + (identity,))
        index = len(self.sources) - 1
        self.source_indices[identity] = index
        if key is not None:
            self.variable_sources[key] = index
        self.next_source_id = max(self.next_source_id, identity + 1)
        return SpikeSource(self, index, identity)

    def register_spike_source(self, target, sid=None, section="soma", x=0.5, threshold=10.0):
        """Voltage detector on a cell, the net_event output of a point process, or (target
        the model itself) an external spike input driven by inject_spike (model.jl:286-365)."""
        if target is self:
            return self._new_source((2, -1, -1, 0.0, 0.0, -1), sid, None)
        if isinstance(target, PointProcess):
            existing = self.point_sources.get(target.index)
            if existing is not None:
                return SpikeSource(self, existing, self.sources[existing][-1])
            src = self._new_source((1, -1, -1, 0.0, 10.0, target.index), sid, None)
            self.point_sources[target.index] = src.index
            return src
        if not (0.0 <= x <= 1.0 and np.isfinite(threshold)):
            raise ValueError("invalid spike detector")
        morph = self._morph(target)
        s = morph.section_index(section)
        # endpoint aliases follow the parent attachment (utils/model.jl:171-190)
        name_idx, coord = s, float(x)
        while coord == morph.child_x[name_idx] and morph.parent[name_idx] >= 0:
            name_idx, coord = int(morph.parent[name_idx]), float(morph.parent_x[name_idx])
        nseg = int(morph.nseg[name_idx])
        segment = (0 if coord == 0.0 else nseg + 1) if coord in (0.0, 1.0) else \
            min(int(np.floor(coord * nseg)) + 1, nseg)
        key = ("voltage", target.index, name_idx, segment)
        existing = self.variable_sources.get(key)
        if existing is not None:
            return SpikeSource(self, existing, self.sources[existing][-1])
        return self._new_source((0, target.index, s, float(x), float(threshold), -1), sid, key)

    def connect(self, source, target: PointProcess, weight=1.0, delay=0.0):
        """One NetCon from a source (or a point's net_event) to a point process."""
        from .mechanisms import BY_ID
        mech = BY_ID[self.points[target.index][0]]
        if not mech.has_net_receive:
            raise ValueError("target has no NET_RECEIVE block")
        if not (np.isfinite(weight) and np.isfinite(delay) and delay >= 0):
            raise ValueError("connection weight and delay must be finite, delay nonnegative")
        if isinstance(source, PointProcess):
            source = self.register_spike_source(source)
        self.connections.append((source.index, target.index, float(weight), float(delay)))
        return len(self.connections)

    # -- flattening ----------------------------------------------------------
    def descriptor(self):
        """Flatten into a mind_model_desc; returns (desc, keepalive)."""
        keep = {}

        def arr(name, values, dtype):
            a = np.ascontiguousarray(np.asarray(values, dtype=dtype))
            if a.size == 0:
                a = np.zeros(1, dtype=dtype)
            keep[name] = a
            return a

        morphs = self.morphologies
        counts = [len(m.names) for m in morphs]
        offsets = np.zeros(len(morphs) + 1, np.int64)
        offsets[1:] = np.cumsum(counts) if counts else []

        def cat(attr, dtype):
            if not morphs:
                return np.zeros(0, dtype)
            return np.concatenate([getattr(m, attr) for m in morphs]).astype(dtype)

        d = _abi.ModelDesc()
        d.celsius = self.celsius
        d.n_morphologies = len(morphs)
        d.morph_offsets = _abi.as_ptr(arr("mo", offsets, np.int64), C.c_int64)
        d.sec_parent = _abi.as_ptr(arr("sp", cat("parent", np.int32), np.int32), C.c_int32)
        d.sec_parent_x = _abi.as_ptr(arr("spx", cat("parent_x", np.float64), np.float64), C.c_double)
        d.sec_child_x = _abi.as_ptr(arr("scx", cat("child_x", np.float64), np.float64), C.c_double)
        d.sec_nseg = _abi.as_ptr(arr("sn", cat("nseg", np.int32), np.int32), C.c_int32)
        d.sec_length = _abi.as_ptr(arr("sl", cat("length", np.float64), np.float64), C.c_double)
        d.sec_diameter = _abi.as_ptr(arr("sd", cat("diameter", np.float64), np.float64), C.c_double)
        d.sec_Ra = _abi.as_ptr(arr("sra", cat("Ra", np.float64), np.float64), C.c_double)
        d.sec_cm = _abi.as_ptr(arr("scm", cat("cm", np.float64), np.float64), C.c_double)
        if any(m.points is not None for m in morphs):
            pts, poff = [], [0]
            for m in morphs:
                for i in range(len(m.names)):
                    p = m.points[i] if m.points is not None else []
                    p = p if len(p) >= 2 else []
                    pts.extend(p)
                    poff.append(poff[-1] + len(p))
            d.sec_point_offsets = _abi.as_ptr(arr("po", poff, np.int64), C.c_int64)
            d.points = _abi.as_ptr(arr("pts", np.asarray(pts, np.float64).reshape(-1), np.float64),
                                   C.c_double)
        d.n_cells = len(self.cell_gid)
        d.cell_morphology = _abi.as_ptr(arr("cm_", self.cell_morph, np.int32), C.c_int32)
        d.cell_gid = _abi.as_ptr(arr("cg", self.cell_gid, np.int64), C.c_int64)
        d.cell_voltage = _abi.as_ptr(arr("cv", self.cell_voltage, np.float64), C.c_double)

        pf, pv = [], []

        def params_csr(entries, getter):
            off = [0]
            for e in entries:
                for f, v in getter(e):
                    pf.append(f)
                    pv.append(v)
                off.append(len(pf))
            return off

        d.n_density = len(self.density)
        d.dens_cell = _abi.as_ptr(arr("dc", [e[0] for e in self.density], np.int32), C.c_int32)
        d.dens_section = _abi.as_ptr(arr("ds", [e[1] for e in self.density], np.int32), C.c_int32)
        d.dens_mech = _abi.as_ptr(arr("dm", [e[2] for e in self.density], np.int32), C.c_int32)
        d.dens_param_offsets = _abi.as_ptr(
            arr("dpo", params_csr(self.density, lambda e: self._with_globals(e[2], e[3])), np.int64), C.c_int64)
        d.n_points = len(self.points)
        d.pt_mech = _abi.as_ptr(arr("pm", [e[0] for e in self.points], np.int32), C.c_int32)
        d.pt_cell = _abi.as_ptr(arr("pc", [e[1] for e in self.points], np.int32), C.c_int32)
        d.pt_section = _abi.as_ptr(arr("ps", [e[2] for e in self.points], np.int32), C.c_int32)
        d.pt_x = _abi.as_ptr(arr("px", [e[3] for e in self.points], np.float64), C.c_double)
        d.pt_param_offsets = _abi.as_ptr(
            arr("ppo", params_csr(self.points, lambda e: self._with_globals(e[0], e[4])), np.int64), C.c_int64)
        if any(e[5] for e in self.points):
            spike_off = np.cumsum([0] + [len(e[5]) for e in self.points])
            d.pt_spike_offsets = _abi.as_ptr(arr("pso", spike_off, np.int64), C.c_int64)
            d.pt_spike_times = _abi.as_ptr(
                arr("pst", [t for e in self.points for t in e[5]], np.float64), C.c_double)
        d.param_field = _abi.as_ptr(arr("pf", pf, np.int32), C.c_int32)
        d.param_value = _abi.as_ptr(arr("pv", pv, np.float64), C.c_double)
        S = self.sources
        d.n_sources = len(S)
        d.src_kind = _abi.as_ptr(arr("sk", [e[0] for e in S], np.int32), C.c_int32)
        d.src_cell = _abi.as_ptr(arr("sc", [e[1] for e in S], np.int32), C.c_int32)
        d.src_section = _abi.as_ptr(arr("ss", [e[2] for e in S], np.int32), C.c_int32)
        d.src_x = _abi.as_ptr(arr("sx", [e[3] for e in S], np.float64), C.c_double)
        d.src_threshold = _abi.as_ptr(arr("st", [e[4] for e in S], np.float64), C.c_double)
        d.src_point = _abi.as_ptr(arr("spt", [e[5] for e in S], np.int32), C.c_int32)
        d.src_sid = _abi.as_ptr(arr("ssid", [e[6] for e in S], np.int64), C.c_int64)
        K = self.connections
        d.n_connections = len(K)
        d.con_source = _abi.as_ptr(arr("ks", [e[0] for e in K], np.int32), C.c_int32)
        d.con_target = _abi.as_ptr(arr("kt", [e[1] for e in K], np.int32), C.c_int32)
        d.con_weight = _abi.as_ptr(arr("kw", [e[2] for e in K], np.float64), C.c_double)
        d.con_delay = _abi.as_ptr(arr("kd", [e[3] for e in K], np.float64), C.c_double)
        species = self.species_list()
        index = {s.name: k for k, s in enumerate(species)}
        names = (C.c_char_p * len(species))(*[s.name.encode() for s in species])
        keep["spn"] = names
        d.n_species = len(species)
        d.species_name = C.cast(names, C.POINTER(C.c_char_p))
        d.species_valence = _abi.as_ptr(arr("spv", [s.valence for s in species], np.float64), C.c_double)
        d.species_internal = _abi.as_ptr(arr("spi", [s.internal for s in species], np.float64), C.c_double)
        d.species_external = _abi.as_ptr(arr("spe", [s.external for s in species], np.float64), C.c_double)
        d.species_reversal = _abi.as_ptr(arr("spr", [s.reversal for s in species], np.float64), C.c_double)
        d.species_nernst = _abi.as_ptr(arr("spz", [int(s.nernst) for s in species], np.int32), C.c_int32)
        I = self.ion_settings
        for e in I:
            if e[2] not in index:
                raise ValueError(f"unknown ion {e[2]}")
        d.n_ion_settings = len(I)
        d.ionset_cell = _abi.as_ptr(arr("isc", [e[0] for e in I], np.int32), C.c_int32)
        d.ionset_section = _abi.as_ptr(arr("iss", [e[1] for e in I], np.int32), C.c_int32)
        d.ionset_species = _abi.as_ptr(arr("isp", [index[e[2]] for e in I], np.int32), C.c_int32)
        d.ionset_what = _abi.as_ptr(arr("isw", [e[3] for e in I], np.int32), C.c_int32)
        d.ionset_value = _abi.as_ptr(arr("isv", [e[4] for e in I], np.float64), C.c_double)
        return d, keep

    def species_list(self):
        """Species in descriptor order: the model's, then those only a used mechanism
        declares (with its VALENCE, 1 mM on both sides; model.jl:690-703)."""
        from .mechanisms import BY_ID, ions

        out = list(self.ions.values())
        names = {s.name for s in out}
        used = [e[2] for e in self.density] + [p[0] for p in self.points]
        for mid in dict.fromkeys(used):
            for u in ions(BY_ID[mid]):
                if u.name not in names:
                    if u.valence is None:
                        raise ValueError(f"USEION {u.name} needs a VALENCE declaration")
                    out.append(IonSpecies(u.name, u.valence, 1.0, 1.0, 0.0))
                    names.add(u.name)
        return out

    # -- records -------------------------------------------------------------
    def record_struct(self, spec):
        r = _abi.Record()
        r.kind, r.cell, r.section, r.x, r.point, r.mech, r.field = spec
        return r


# -- module-level API mirroring the reference's exported functions -----------
def add_population(model, sections, count, **kw):
    return model.add_population(sections, count, **kw)


def insert(cell, mechanism, sections="all", params=None):
    return cell.model.insert(cell, mechanism, sections, params)


def add_pointprocess(target, mechanism, section="soma", x=0.5, params=None, spike_times=()):
    model = target if isinstance(target, MicroModel) else target.model
    return model.add_pointprocess(target, mechanism, section, x, params, spike_times)


def register_spike_source(target, **kw):
    model = target if isinstance(target, MicroModel) else target.model
    return model.register_spike_source(target, **kw)


def connect(source, target, weight=1.0, delay=0.0):
    return target.model.connect(source, target, weight, delay)


@dataclass
class RecordSpec:
    name: str
    spec: tuple = field(default_factory=tuple)


def voltage_record(model, cell, section="soma", x=0.5, name=None):
    """voltage_record(model, cell; section, x) or voltage_record(model, cell_section; x)
    (recording.jl:4-9, 50-55)."""
    if isinstance(cell, CellSection):
        sec = cell
        s = model._morph(sec.cell).section_index(sec.name)
        return RecordSpec(name or f"{sec.cell.gid}.{sec.name}.v", (0, sec.cell.index, s, float(x), -1, -1, -1))
    s = model._morph(cell).section_index(section)
    return RecordSpec(name or f"v_{cell.gid}", (0, cell.index, s, float(x), -1, -1, -1))


def mechanism_record(model, target, *args, name=None, section="soma", x=0.5, index=1):
    """mechanism_record(model, point, field), mechanism_record(model, cell, mech, field)
    or mechanism_record(model, cell_section, mech, field) (recording.jl:10-30, 56-63);
    `index` selects an element of an array field (1-based, references.jl:44-69).
    ROI and MICRO2MACRO fields: mind_amd.macro.mechanism_record."""
    from .mechanisms import BY_ID
    if isinstance(target, CellSection):
        mechanism, fname = args
        mech = resolve(mechanism)
        s = model._morph(target.cell).section_index(target.name)
        return RecordSpec(name or f"{target.cell.gid}.{target.name}.{fname}",
                          (2, target.cell.index, s, float(x), -1, mech.id, mech.field_index(str(fname), index)))
    if isinstance(target, PointProcess):
        (fname,) = args
        mech = BY_ID[model.points[target.index][0]]
        return RecordSpec(name or str(fname),
                          (1, -1, -1, 0.0, target.index, mech.id, mech.field_index(str(fname), index)))
    mechanism, fname = args
    mech = resolve(mechanism)
    s = model._morph(target).section_index(section)
    return RecordSpec(name or f"{target.gid}.{fname}",
                      (2, target.index, s, float(x), -1, mech.id, mech.field_index(str(fname), index)))


def _global_names(mech) -> frozenset:
    """Field names (array fields by their base name) of a mechanism's GLOBALs."""
    if not mech.compiled:
        return frozenset()  # the builtins declare none
    from .mechanisms import globals_of
    return frozenset(f.split("[")[0] for f in globals_of(mech))


def set_global(model, mechanism, fname, value, index=1):
    return model.set_global(mechanism, fname, value, index)


def global_record(model, mechanism, fname, name=None, index=1):
    """global_record(model, mechanism, field) (recording.jl:32-38; global_ref,
    references.jl:147-154): a GLOBAL of a mechanism the model uses, one value every
    instance shares. Its phases never assign it (compile_mod refuses that), so every
    instance's column holds the value; the record reads the first instance's: the
    first point process of the mechanism, else its first density insertion."""
    mech = resolve(mechanism)
    key = str(fname).lstrip(":")
    if key not in _global_names(mech):
        raise ValueError(f"{key} is not a shared GLOBAL of {mech.name}")
    column = mech.field_index(key, index)
    for k, point in enumerate(model.points):
        if point[0] == mech.id:
            return RecordSpec(name or key, (1, -1, -1, 0.0, k, mech.id, column))
    for cell_index, s, mech_id, _ in model.density:
        if mech_id == mech.id:
            return RecordSpec(name or key, (2, cell_index, s, 0.5, -1, mech.id, column))
    raise ValueError(f"the model has no instance of {mech.name}")


def ion_record(model, target, fname, name=None, section="soma", x=0.5):
    """ion_record(model, cell, field; section, x) or ion_record(model, cell_section, field; x)
    (recording.jl:40-46, 64-70; ion_ref, references.jl:156-178): a reversal potential,
    concentration, current or current derivative such as ena, cai, cao, ica or
    dica_dv_ of the ion row at the location."""
    key = str(fname).lstrip(":")
    species = model.species_list()
    for k, sp in enumerate(species):
        for column, pattern in enumerate(_ION_COLUMNS):
            if key == pattern.format(sp.name):
                if isinstance(target, CellSection):
                    cell, s = target.cell, model._morph(target.cell).section_index(target.name)
                else:
                    cell, s = target, model._morph(target).section_index(section)
                return RecordSpec(name or key, (3, cell.index, s, float(x), -1, k, column))
    raise ValueError(f"unknown ion field {key}")


def read_value(model, spec):
    """read_value(state, selection) (state.jl:586-600): the current value of one record
    selector, read from the device."""
    rec = model.record_struct(spec.spec if isinstance(spec, RecordSpec) else spec)
    out = C.c_double()
    _abi.check(_abi.lib().mind_read_value(model.state, C.byref(rec), C.byref(out)))
    return out.value


def inject_event(model, connection, time):
    """inject_event!(state, connection, time) (state.jl:618-636): the connection's event
    (numbered as connect() returned it) arriving at absolute time `time`, without its delay."""
    _abi.check(_abi.lib().mind_inject_event(model.state, int(connection) - 1, float(time)))
    return model


def inject_spike(model, source, time):
    """inject_spike!(state, source, time) (state.jl:638-659): a spike of an external input
    (register_spike_source(model; sid)) at absolute time `time`, through all its connections."""
    if source.model is not model:
        raise ValueError("spike source belongs to another model")
    _abi.check(_abi.lib().mind_inject_spike(model.state, int(source.index), float(time)))
    return model


def checkpoint(model) -> bytes:
    """The state a run continues from, as bytes (restore() loads it into a state built
    from the same model, in this or another process)."""
    L = _abi.lib()
    n = C.c_size_t()
    _abi.check(L.mind_checkpoint(model.state, None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value)
    _abi.check(L.mind_checkpoint(model.state, buf, n.value, C.byref(n)))
    return buf.raw[:n.value]


def restore(model, data: bytes):
    """Load a checkpoint into the built state of the same model; the next run continues it."""
    _abi.check(_abi.lib().mind_restore(model.state, bytes(data), len(data)))
    return model


def connection_weights(model, connection):
    """connection_weights(state, connection) (state.jl:603-616): the NET_RECEIVE arguments of
    the connection numbered as connect() returned it (1-based)."""
    out = (C.c_double * 8)()
    n = C.c_int32()
    _abi.check(_abi.lib().mind_connection_weights(model.state, int(connection) - 1, out, 8, C.byref(n)))
    return np.array(out[:n.value])


def build(model, dt_micro=0.025, seed=1, method="batched", cells_per_group=0, window_steps=0,
          event_capacity=0, device=0, secondorder=0):
    """Pack the model and allocate device state (build.jl:120-220 + allocate_simulation).
    secondorder 1 or 2 selects Crank-Nicolson cable steps as the reference's keyword does."""
    if not (np.isfinite(dt_micro) and dt_micro > 0):
        raise ValueError("dt_micro must be positive")
    if secondorder not in (0, 1, 2):
        raise ValueError("secondorder must be 0, 1 or 2")
    L = _abi.lib()
    if model.state is not None:
        L.mind_state_destroy(model.state)
        model.state = None
    desc, keep = model.descriptor()
    opts = _abi.BuildOptions()
    opts.dt = float(dt_micro)
    opts.seed = int(seed)
    opts.method = {"plain": 0, "batched": 1}[method] if isinstance(method, str) else int(method)
    opts.cells_per_group = int(cells_per_group)
    opts.window_steps = int(window_steps)
    opts.event_capacity = int(event_capacity)
    opts.device = int(device)
    opts.secondorder = int(secondorder)
    handle = C.c_void_p()
    _abi.check(L.mind_build(C.byref(desc), C.byref(opts), C.byref(handle)))
    del keep
    model.state = handle
    model.options = opts
    return model


def initialize(model):
    _abi.check(_abi.lib().mind_initialize(model.state))
    return model


def run(model, stop, records=(), record_spikes=True):
    L = _abi.lib()
    recs = (_abi.Record * max(1, len(records)))()
    for i, r in enumerate(records):
        recs[i] = model.record_struct(r.spec)
    res = _abi.Result()
    _abi.check(L.mind_run(model.state, float(stop), recs, len(records), int(bool(record_spikes)),
                          C.byref(res)))
    try:
        time, traces, st, ss, secs = _abi.result_arrays(res)
    finally:
        L.mind_result_free(C.byref(res))
    return SimulationResult(time, {r.name: traces[i] for i, r in enumerate(records)}, st, ss, secs)


def step(model):
    _abi.check(_abi.lib().mind_step(model.state))
    return model


def read_voltage(model):
    L = _abi.lib()
    n, s = C.c_int64(), C.c_int64()
    _abi.check(L.mind_node_count(model.state, C.byref(n), C.byref(s)))
    out = np.zeros(n.value)
    _abi.check(L.mind_read_voltage(model.state, _abi.as_ptr(out, C.c_double), n.value))
    return out


def read_field(model, mechanism, fname):
    mech = resolve(mechanism)
    L = _abi.lib()
    rows = C.c_int64()
    _abi.check(L.mind_row_count(model.state, mech.id, C.byref(rows)))
    out = np.zeros(rows.value)
    _abi.check(L.mind_read_field(model.state, mech.id, mech.field_index(fname),
                                 _abi.as_ptr(out, C.c_double), rows.value))
    return out


def pack_nodes(model, dt_micro=0.025):
    """Packed cable geometry computed by libmind on the host (no device needed)."""
    L = _abi.lib()
    desc, keep = model.descriptor()
    n, s = C.c_int64(), C.c_int64()
    _abi.check(L.mind_pack_nodes(C.byref(desc), float(dt_micro), C.byref(n), C.byref(s),
                                 None, None, None, None, None))
    parent = np.zeros(n.value, np.int32)
    area, cap, axial, adiag = (np.zeros(n.value) for _ in range(4))
    _abi.check(L.mind_pack_nodes(C.byref(desc), float(dt_micro), C.byref(n), C.byref(s),
                                 _abi.as_ptr(parent, C.c_int32), _abi.as_ptr(area, C.c_double),
                                 _abi.as_ptr(cap, C.c_double), _abi.as_ptr(axial, C.c_double),
                                 _abi.as_ptr(adiag, C.c_double)))
    del keep
    return dict(nodes=n.value, compartments=s.value, parent=parent, area=area, cap=cap,
                axial=axial, adiag=adiag)


def check_batch_plan(model, budget, threads, layout, dt_micro=0.025):
    """Builds the batched plan of a model for (budget, threads, layout) on the host
    and checks every index its kernels derive from it (no device needed); returns
    the workgroup count, raises MindError on an inconsistent plan."""
    L = _abi.lib()
    desc, keep = model.descriptor()
    g = C.c_int64()
    _abi.check(L.mind_check_batch_plan(C.byref(desc), float(dt_micro), int(budget), int(threads), int(layout),
                                       C.byref(g)))
    del keep
    return g.value


def batch_plan(model):
    """The batched kernel's plan: node budget per workgroup, threads per workgroup,
    layout of those workgroups (0 all in LDS, 5 lean with v in LDS, 4 lean with v
    in HBM) and workgroup count."""
    L = _abi.lib()
    b, t, lay, g = C.c_int64(), C.c_int32(), C.c_int32(), C.c_int64()
    _abi.check(L.mind_batch_plan(model.state, C.byref(b), C.byref(t), C.byref(lay), C.byref(g)))
    return dict(budget=b.value, threads=t.value, layout=lay.value, groups=g.value)


def batch_kernel(model):
    """Name of the cell kernel the batched method launches for the workgroups within
    the node budget (k_window, k_window_lean or k_window_lean_hbm; with _core when
    every mechanism group of the model has an inlined path)."""
    name = C.c_char_p()
    _abi.check(_abi.lib().mind_batch_kernel(model.state, C.byref(name)))
    return name.value.decode()


def set_batch_plan(model, budget, threads=256, layout=0):
    """Replaces the autotuned plan; initialize(model) must follow."""
    _abi.check(_abi.lib().mind_set_batch_plan(model.state, int(budget), int(threads), int(layout)))


def synchronize(model):
    _abi.check(_abi.lib().mind_synchronize(model.state))


def set_profiling(model, on=True):
    """Time the dominant cell-kernel launch with HIP events on the library's stream."""
    _abi.check(_abi.lib().mind_set_profiling(model.state, int(bool(on))))


def kernel_time(model):
    """(launches, steps, total ms) of the timed cell-kernel launches."""
    n, s, ms = C.c_int64(), C.c_int64(), C.c_double()
    _abi.check(_abi.lib().mind_kernel_time(model.state, C.byref(n), C.byref(s), C.byref(ms)))
    return n.value, s.value, ms.value


def batch_work(model):
    """Per-step work of the dominant batched launch: nodes, pas, hh, Exp2Syn, IClamp, detectors."""
    out = (C.c_int64 * 6)()
    _abi.check(_abi.lib().mind_batch_work(model.state, out))
    keys = ("nodes", "pas", "hh", "exp2syn", "iclamp", "detectors")
    return dict(zip(keys, list(out)))


def event_counts(model):
    """(routed, delivered): connection events ticketed and gathered for delivery
    since initialize."""
    r, d = C.c_int64(), C.c_int64()
    _abi.check(_abi.lib().mind_event_counts(model.state, C.byref(r), C.byref(d)))
    return r.value, d.value


def window_steps(model):
    w = C.c_int32()
    _abi.check(_abi.lib().mind_window_steps(model.state, C.byref(w)))
    return w.value


def destroy(model):
    if model.state is not None:
        _abi.lib().mind_state_destroy(model.state)
        model.state = None
