"""Neurolucida ASCII morphology import: the reference's `load_asc`
(src/model/micro/morphology.jl:344-398, helpers in src/model/micro/utils/morphology.jl:124-298).

The format is a sequence of parenthesised S-expressions:

* `(x y z d)` is a sample; d is a diameter whose magnitude is used (at most 1e-6
  becomes 1; a missing d is 1);
* a list whose first item is a list is a split: its parts, separated by `|`, are
  the child branches of the branch that contains it;
* other lists and strings label or decorate (`(Dendrite)`, `(Color Red)`,
  `"CellBody"`, words such as `Normal`); `<...>` markup, `;` comments to the end
  of a line and commas are ignored.

A top-level tree is classified by the last of CellBody / Axon / Dendrite / Apical
(case and blanks ignored) that it names in a string item or as the first word of
one of its lists; other trees (markers, spines) are skipped. At most one CellBody
contour is allowed; it becomes the soma section `soma_0`, a chain of 21 points
along the contour's principal axis with the contour's width there as diameter
(`soma_points`). Each axon/dend/apic branch becomes a section `<label>_<k>`: a
root branch hangs from the soma at 0.5 (or is a root without a soma); a child
branch hangs from the end of its parent and starts with a copy of the parent's
last sample (with the child's first diameter) unless it starts there already; a
single-sample branch becomes three points one diameter long along x.
"""
from __future__ import annotations

import math

import numpy as np

from .model import Section

KINDS = {"cellbody": "soma", "axon": "axon", "dendrite": "dend", "apical": "apic"}
SPLIT = "|"
_DIGITS = set("0123456789")


# ------------------------------------------------------------------ reading ----
def _is_number(word):
    """Decimal literal: optional sign, digits with at most one point (at least one
    digit), optional exponent."""
    body = word[1:] if word[:1] in "+-" else word
    mantissa, _, exponent = body.replace("E", "e").partition("e")
    if "e" in body.lower() and not exponent:
        return False
    if exponent:
        exp = exponent[1:] if exponent[:1] in "+-" else exponent
        if not exp or not set(exp) <= _DIGITS:
            return False
    whole, dot, frac = mantissa.partition(".")
    if "." in frac:
        return False
    digits = whole + frac
    return bool(digits) and set(digits) <= _DIGITS


def _nonfinite(word):
    return word.lower().lstrip("+-") in ("inf", "infinity", "nan")


def _tokens(text, filename):
    """(kind, value, line) tokens: '(' ')' '|' 'num' 'str'."""
    i, n, line = 0, len(text), 1
    stops = set("()|<>,;\" \t\r\n")
    while i < n:
        ch = text[i]
        if ch == "\n":
            line += 1
            i += 1
        elif ch in " \t\r,":
            i += 1
        elif ch == ";":  # comment to the end of the line
            j = text.find("\n", i)
            i = n if j < 0 else j
        elif ch == "<":  # markup to the next '>'
            j = text.find(">", i)
            j = n if j < 0 else j
            line += text.count("\n", i, j)
            i = j + 1
        elif ch in "()|":
            yield ch, ch, line
            i += 1
        elif ch == '"':
            j = text.find('"', i + 1)
            j = n if j < 0 else j
            yield "str", text[i + 1:j], line
            line += text.count("\n", i, j)
            i = j + 1
        elif ch == ">":
            i += 1
        else:
            j = i
            while j < n and text[j] not in stops:
                j += 1
            word = text[i:j]
            if _is_number(word):
                yield "num", float(word), line
            elif _nonfinite(word):
                raise ValueError(f"{filename}:{line}: ASC coordinate {word} is not finite")
            else:
                yield "str", word, line
            i = j


def parse(text, filename="<asc>"):
    """Top-level S-expressions of `text` as nested Python lists of floats, strings
    (the split marker is the string '|') and lists."""
    top, stack = [], []  # stack of (list, line where it opened)
    for kind, value, line in _tokens(text, filename):
        if kind == "(":
            stack.append(([], line))
        elif kind == ")":
            if stack:
                done, _ = stack.pop()
                (stack[-1][0] if stack else top).append(done)
        elif stack:
            stack[-1][0].append(value)
    if stack:
        raise ValueError(f"{filename}:{stack[0][1]}: ASC list opened here is never closed")
    return top


def _kind_of(word):
    return KINDS.get("".join(word.split()).lower()) if isinstance(word, str) else None


def classify(tree):
    """Label of a top-level tree: the last kind word among its string items and the
    first items of its lists (None when it names none)."""
    found = None
    for item in tree:
        word = item[0] if isinstance(item, list) and item else item
        kind = _kind_of(word)
        if kind is not None:
            found = kind
    return found


class Arbor:
    """A branch: its samples and the branches it splits into."""

    __slots__ = ("points", "children")

    def __init__(self, points=None, children=None):
        self.points = points if points is not None else []
        self.children = children if children is not None else []


def _sample(item):
    if len(item) < 3 or not all(isinstance(v, float) for v in item[:3]):
        raise ValueError("an ASC sample needs three numeric coordinates")
    d = abs(item[3]) if len(item) > 3 and isinstance(item[3], float) else 1.0
    return (item[0], item[1], item[2], d if d > 1e-6 else 1.0)


def arbors(items):
    """Branches of a tree's (or a split's) items: `|` starts the next branch;
    samples extend the current one; a split adds child branches to it. Branches
    with neither samples nor children are left out."""
    parts = [Arbor()]
    for item in items:
        if item == SPLIT:
            parts.append(Arbor())
        elif isinstance(item, list) and item:
            head = item[0]
            if isinstance(head, float):
                parts[-1].points.append(_sample(item))
            elif isinstance(head, list):
                parts[-1].children += arbors(item)
    return [a for a in parts if a.points or a.children]


# ------------------------------------------------------------------ soma -------
def _resample(curve, count):
    """`count` points at equal arc length along a polyline (ends kept)."""
    steps = np.linalg.norm(np.diff(curve, axis=0), axis=1)
    arc = np.concatenate([[0.0], np.cumsum(steps)])
    keep = np.concatenate([[True], steps > 0])  # coincident samples add no arc
    arc, curve = arc[keep], curve[keep]
    at = np.linspace(0.0, arc[-1], count)
    return np.column_stack([np.interp(at, arc, curve[:, j]) for j in range(3)])


def _rising(x, y):
    """The samples of (x, y) at which x exceeds every earlier x."""
    best, keep = -math.inf, []
    for k, v in enumerate(x):
        if v > best:
            best = v
            keep.append(k)
    return x[keep], y[keep]


def soma_points(contour, count=101, stations=21):
    """Soma section of a CellBody contour: the contour resampled at `count` equal
    arc-length steps; the centroid and principal axes of those samples (scatter
    matrix eigenvectors; the major one with its largest component positive; the
    middle one with its z component dropped as in-plane minor direction, or a
    perpendicular of the major one when that vanishes); the outline split at its
    far ends along the major axis into two halves, each made monotone along it;
    `stations` points on the major axis between the second-lowest and
    second-highest outline positions, with the distance between the halves as
    diameter (the end values averaged with their neighbours, at most 1e-6 set to
    1). Empty when the contour is too short or degenerate."""
    xyz = np.asarray([p[:3] for p in contour], np.float64)
    if len(xyz) < 3:
        return []
    if not np.sum(np.linalg.norm(np.diff(xyz, axis=0), axis=1)) > 1e-12:
        return []
    ring = _resample(xyz, count)
    center = ring.sum(axis=0) / count
    rel = ring - center
    _, axes = np.linalg.eigh(rel.T @ rel)  # ascending eigenvalues
    major = axes[:, 2].copy()
    if major[np.argmax(np.abs(major))] < 0.0:
        major = -major
    minor = np.array([axes[0, 1], axes[1, 1], 0.0])
    if np.linalg.norm(minor) < 1e-6:
        minor = np.array([-major[1], major[0], 0.0])
        if np.linalg.norm(minor) < 1e-12:
            minor = np.array([1.0, 0.0, 0.0])
    minor /= np.linalg.norm(minor)
    along, across = rel @ major, rel @ minor
    start = int(np.argmax(along))
    along, across = np.roll(along, -start), np.roll(across, -start)
    turn = int(np.argmin(along))  # from the far end back: one half, then the other
    if turn <= 0 or turn >= count - 1:
        return []
    x1, y1 = _rising(along[:turn][::-1], across[:turn][::-1])
    x2, y2 = _rising(along[turn:], across[turn:])
    if len(x1) < 2 or len(x2) < 2:
        return []
    pos = np.sort(np.concatenate([x1, x2]))
    lo, hi = pos[1], pos[-2]
    if not hi > lo:
        return []
    at = np.linspace(lo, hi, stations)
    width = np.abs(np.interp(at, x1, y1) - np.interp(at, x2, y2))
    width[0] = (width[0] + width[1]) / 2
    width[-1] = (width[-1] + width[-2]) / 2
    width[width <= 1e-6] = 1.0
    return [tuple(float(v) for v in center + major * c) + (float(w),) for c, w in zip(at, width)]


# ------------------------------------------------------------------ sections ---
def load_asc(path, nseg=1, Ra=100.0, cm=1.0):
    """Sections of the Neurolucida ASCII morphology in `path` (list of Section)."""
    with open(path) as f:
        trees = parse(f.read(), str(path))
    kinds = [(classify(t), arbors(t)) for t in trees]
    contours = [a.points for kind, arbs in kinds if kind == "soma" for a in arbs if a.points]
    if len(contours) > 1:
        raise ValueError(f"{path}: more than one CellBody contour (only one soma is supported)")
    out = []
    soma = None
    if contours:
        pts = soma_points(contours[0])
        if len(pts) >= 2:
            soma = "soma_0"
            out.append(Section(soma, label="soma", points=pts, nseg=nseg, Ra=Ra, cm=cm))
    serial = {"axon": 0, "dend": 0, "apic": 0}
    for kind, arbs in kinds:
        if kind not in serial:
            continue
        # (arbor, parent section, parent's end sample, attachment) in depth-first order
        todo = [(a, soma, None, 0.5 if soma else 1.0) for a in reversed(arbs)]
        while todo:
            arbor, parent, end, x = todo.pop()
            pts = list(arbor.points)
            if pts and end is not None and pts[0][:3] != end[:3]:
                pts.insert(0, end[:3] + (pts[0][3],))
            tail = pts[-1] if pts else end
            if pts:
                if len(pts) == 1:
                    px, py, pz, d = pts[0]
                    pts = [(px - d / 2, py, pz, d), pts[0], (px + d / 2, py, pz, d)]
                name = f"{kind}_{serial[kind]}"
                serial[kind] += 1
                out.append(Section(name, label=kind, parent=parent, parent_x=x, points=pts, nseg=nseg,
                                   Ra=Ra, cm=cm))
                parent, x = name, 1.0
            todo.extend((c, parent, tail, x) for c in reversed(arbor.children))
    return out
