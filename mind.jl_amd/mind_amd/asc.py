"""Neurolucida ASCII morphology import (load_asc, src/model/micro/morphology.jl:344-398;
parse_asc, asc_tree_label, asc_branches, asc_soma_points in
src/model/micro/utils/morphology.jl:124-298).

The file is a sequence of parenthesised trees. Within a tree:

- `(x y z d)` is a sample (d is a diameter: its magnitude is used, and a diameter of
  at most 1e-6 becomes 1);
- a parenthesised list that starts with a list is a split: its `|`-separated parts
  are the child branches of the current branch;
- other lists (`(Dendrite)`, `(Color Red)`, `(CellBody)`) and strings label or
  decorate; `<...>` markup, `;` comments and commas are ignored.

A tree's label is the last of CellBody/Axon/Dendrite/Apical (case and spaces
ignored) named by one of its strings or by the first item of one of its lists;
trees with any other label (markers, spines) are skipped. One CellBody contour at
most becomes the soma: a sphere-like section along the contour's principal axis
(see `soma_points`). Every axon/dend/apic branch becomes a section named
<label>_<ordinal>; a root branch attaches to the soma at 0.5 (or, without a soma,
is a root), and a child branch starts with a copy of its parent's last sample
(with the child's first diameter) unless it already starts there.
"""
from __future__ import annotations

import math
import re

import numpy as np

from .model import Section

_TOKEN = re.compile(r';[^\n]*|"[^"]*"|[()|<>,]|[^\s()|<>,;]+')
_NUMBER = re.compile(r"[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?$")
_NONFINITE = re.compile(r"[+-]?(inf|infinity|nan)$", re.IGNORECASE)
_LABELS = {"CELLBODY": "soma", "AXON": "axon", "DENDRITE": "dend", "APICAL": "apic"}


def parse(text, filename="<asc>"):
    """The top-level parenthesised expressions of `text` as nested lists of floats,
    strings and lists (parse_asc, utils/morphology.jl:132-193)."""
    tokens, lines = [], []
    line, last = 1, 0
    for m in _TOKEN.finditer(text):
        line += text.count("\n", last, m.start())
        last = m.start()
        tok = m.group(0)
        if tok.startswith(";"):
            continue
        tokens.append(tok)
        lines.append(line)
    pos = 0

    def expression():
        nonlocal pos
        items, start = [], pos
        pos += 1
        while pos < len(tokens):
            tok = tokens[pos]
            if tok == ")":
                pos += 1
                return items
            if tok == "(":
                items.append(expression())
            elif tok == "<":  # markup up to the closing '>'
                while pos < len(tokens) and tokens[pos] != ">":
                    pos += 1
                pos += 1
            else:
                if _NUMBER.match(tok):
                    items.append(float(tok))
                elif _NONFINITE.match(tok):
                    raise ValueError(f"{filename}:{lines[pos]}: non-finite ASC coordinate")
                elif tok.startswith('"'):
                    items.append(tok[1:-1])
                elif tok != ",":
                    items.append(tok)
                pos += 1
        raise ValueError(f"{filename}:{lines[start]}: unclosed ASC expression")

    roots = []
    while pos < len(tokens):
        if tokens[pos] == "(":
            roots.append(expression())
        else:
            pos += 1
    return roots


def _label(item):
    if not isinstance(item, str):
        return None
    return _LABELS.get(re.sub(r"\s+", "", item.upper()))


def tree_label(root):
    """asc_tree_label (utils/morphology.jl:202-212): the last label a tree names."""
    label = None
    for item in root:
        candidate = _label(item)
        if isinstance(item, list) and item:
            candidate = _label(item[0])
        if candidate is not None:
            label = candidate
    return label


class Branch:
    __slots__ = ("points", "children")

    def __init__(self):
        self.points, self.children = [], []


def branches(items):
    """Branches of a tree or split (asc_branches, utils/morphology.jl:214-238)."""
    out = [Branch()]
    for item in items:
        if isinstance(item, str) and item == "|":
            out.append(Branch())
        elif isinstance(item, list) and item:
            if isinstance(item[0], float):
                if len(item) < 3 or not all(isinstance(x, float) for x in item[:3]):
                    raise ValueError("ASC point requires three coordinates")
                d = abs(item[3]) if len(item) >= 4 and isinstance(item[3], float) else 1.0
                if d <= 1e-6:
                    d = 1.0
                out[-1].points.append((item[0], item[1], item[2], d))
            elif isinstance(item[0], list):
                out[-1].children.extend(branches(item))
    return [b for b in out if b.points or b.children]


def _interp(x, y, at):
    """Piecewise-linear y(at) over increasing x, clamped at the ends
    (interpolate_samples, utils/morphology.jl:240-248)."""
    x, y, at = np.asarray(x), np.asarray(y), np.asarray(at)
    out = np.empty(at.shape)
    for k, p in enumerate(at):
        if p <= x[0]:
            out[k] = y[0]
        elif p >= x[-1]:
            out[k] = y[-1]
        else:
            right = int(np.searchsorted(x, p, side="left"))
            left = right - 1
            frac = (p - x[left]) / (x[right] - x[left])
            out[k] = y[left] + frac * (y[right] - y[left])
    return out


def _monotone(x, y):
    keep = [0]
    for k in range(1, len(x)):
        if x[k] > x[keep[-1]]:
            keep.append(k)
    return x[keep], y[keep]


def soma_points(contour):
    """The soma of a CellBody contour (asc_soma_points, utils/morphology.jl:250-298):
    the contour resampled at 101 equal arc-length steps; its centroid and principal
    axis (largest eigenvector of the scatter matrix, its largest component made
    positive); the in-plane minor direction (middle eigenvector with z dropped); the
    outline split at its far ends along the major axis into two monotone halves; and
    21 points along the axis between the second-smallest and second-largest
    positions, each with the distance between the halves as its diameter (ends
    averaged with their neighbours, diameters of at most 1e-6 set to 1). Empty when
    the contour is too short or degenerate."""
    pts = np.asarray([p[:3] for p in contour], dtype=np.float64)
    if len(pts) < 3:
        return []
    arcs = np.zeros(len(pts))
    for k in range(1, len(pts)):
        arcs[k] = arcs[k - 1] + math.sqrt(float(np.sum((pts[k] - pts[k - 1]) ** 2)))
    if not arcs[-1] > 1e-12:
        return []
    at = np.linspace(0.0, arcs[-1], 101)
    sampled = np.stack([_interp(arcs, pts[:, j], at) for j in range(3)], axis=1)
    center = sampled.sum(axis=0) / len(sampled)
    centered = sampled - center
    _, vectors = np.linalg.eigh(centered.T @ centered)
    major = vectors[:, 2].copy()
    if major[np.argmax(np.abs(major))] < 0.0:
        major = -major
    minor = vectors[:, 1].copy()
    minor[2] = 0.0
    if np.linalg.norm(minor) < 1e-6:
        minor = np.array([-major[1], major[0], 0.0])
        if np.linalg.norm(minor) < 1e-12:
            minor = np.array([1.0, 0.0, 0.0])
    minor = minor / np.linalg.norm(minor)
    longitudinal, radial = centered @ major, centered @ minor
    first = int(np.argmax(longitudinal))
    longitudinal, radial = np.roll(longitudinal, -first), np.roll(radial, -first)
    split = int(np.argmin(longitudinal))
    if not 0 < split < len(longitudinal) - 1:
        return []
    x1, y1 = _monotone(longitudinal[:split][::-1], radial[:split][::-1])
    x2, y2 = _monotone(longitudinal[split:], radial[split:])
    if min(len(x1), len(x2)) < 2:
        return []
    merged = np.sort(np.concatenate([x1, x2]))
    if not merged[-2] > merged[1]:
        return []
    coords = np.linspace(merged[1], merged[-2], 21)
    diam = np.abs(_interp(x1, y1, coords) - _interp(x2, y2, coords))
    diam[0] = 0.5 * (diam[0] + diam[1])
    diam[-1] = 0.5 * (diam[-1] + diam[-2])
    out = []
    for c, d in zip(coords, diam):
        x, y, z = center + major * c
        out.append((float(x), float(y), float(z), float(d) if d > 1e-6 else 1.0))
    return out


def load_asc(path, nseg=1, Ra=100.0, cm=1.0):
    """Sections of the Neurolucida ASCII morphology in `path` (list of Section)."""
    with open(path) as f:
        roots = parse(f.read(), str(path))
    trees = [(tree_label(root), branches(root)) for root in roots]
    contours = [b.points for label, bs in trees if label == "soma" for b in bs if b.points]
    if len(contours) > 1:
        raise ValueError("ASC supports one CellBody contour")
    sections = []
    parent = None
    if contours:
        points = soma_points(contours[0])
        if len(points) >= 2:
            parent = "soma_0"
            sections.append(Section(parent, label="soma", points=points, nseg=nseg, Ra=Ra, cm=cm))
    counters = {"axon": 0, "dend": 0, "apic": 0}

    def emit(branch, label, parent, terminal, parent_x):
        points = list(branch.points)
        here = terminal if not points else points[-1]
        if terminal is not None and points and tuple(terminal[:3]) != tuple(points[0][:3]):
            points.insert(0, (terminal[0], terminal[1], terminal[2], points[0][3]))
        if points:
            if len(points) == 1:
                x, y, z, d = points[0]
                points = [(x - 0.5 * d, y, z, d), points[0], (x + 0.5 * d, y, z, d)]
            name = f"{label}_{counters[label]}"
            counters[label] += 1
            sections.append(Section(name, label=label, parent=parent, parent_x=parent_x, points=points,
                                    nseg=nseg, Ra=Ra, cm=cm))
            parent, parent_x = name, 1.0
        for child in branch.children:
            emit(child, label, parent, here, parent_x)

    for label, bs in trees:
        if label not in counters:
            continue
        for b in bs:
            emit(b, label, parent, None, 1.0 if parent is None else 0.5)
    return sections
