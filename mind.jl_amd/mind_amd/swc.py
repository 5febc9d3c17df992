"""SWC morphology import: the reference's `load_swc` (src/model/micro/morphology.jl:130-338).

An SWC file lists samples, one per line: id, structure type (1 soma, 2 axon, 3
basal dendrite, 4 apical dendrite, others free), x, y, z, radius, parent id (-1
for a root); `#` starts a comment. This module turns the sample tree into
`Section`s with 3-D points (diameter = 2 x radius) by the section rules of
NEURON's Import3d, which the reference keeps:

* A section is a maximal run of samples, each continuing into one child:
  - a soma sample continues into its soma child if it has exactly one, and a root
    soma sample with several soma children continues into the first of them;
  - any other sample with exactly one child of its own type continues into it;
  - a root, or a sample hanging from a soma sample, that has several children
    continues into its first child of its own type; its other children start
    sections attached at the proximal end (x = 0) of its section.
* A root soma sample with exactly two soma children that are leaves, have its
  diameter and lie one diameter apart in total (NeuroMorpho's three-point soma)
  is a single one-point soma section; the two children are dropped.
* Attachment of a child section, whose first sample hangs from sample `p` of
  section `P`:
  - P is a one-point root soma and the child is a neurite: x = 0.5, and the
    child does not repeat `p` unless it is itself a single sample;
  - P is a root section and `p` its first sample: x = 0; a neurite child starts
    without repeating `p` only when `p` has several soma children;
  - P is a soma elsewhere: x = 0.5 if `p` is not P's last sample; a neurite child
    of more than one sample then does not repeat `p` (nor when `p` is P's last
    sample but has several soma children);
  - otherwise x = 1 and the child repeats `p`;
  - a proximal attachment (above) overrides x with 0.
  A repeated `p` takes the child's first diameter when P is a soma and the child
  is not.
* A child section of at most two points that starts and ends at one position is
  dropped; its children move to its parent at its attachment point. A remaining
  single-point section becomes three points one diameter long along x.
* Sections are formed depth first from the roots, children in sample order, and
  named `<label>_<k>` with label soma/axon/dend/apic by type (`sec_<type>_<k>`
  for other types), k counting per label in formation order.

Malformed input raises ValueError: too few columns, repeated ids, unknown
parents, samples that no root reaches, an empty file.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

from .model import Section

LABELS = {1: "soma", 2: "axon", 3: "dend", 4: "apic"}
SOMA = 1


@dataclass
class _Sample:
    kind: int
    point: tuple          # x, y, z, diameter
    up: int               # index of the parent sample, -1 for a root
    down: list = field(default_factory=list)  # child sample indices in file order


def _read(path):
    """Samples of the file in file order, parents resolved to indices."""
    ids, raw = {}, []
    with open(path) as f:
        for number, text in enumerate(f, 1):
            text = text.partition("#")[0]
            cols = text.split()
            if not cols:
                continue
            if len(cols) < 7:
                raise ValueError(f"{path}:{number}: an SWC sample line has 7 columns, found {len(cols)}")
            sid = int(cols[0])
            if sid in ids:
                raise ValueError(f"{path}:{number}: SWC sample id {sid} appears twice")
            ids[sid] = len(raw)
            x, y, z, r = (float(c) for c in cols[2:6])
            raw.append((int(cols[1]), (x, y, z, 2.0 * r), int(cols[6])))
    if not raw:
        raise ValueError(f"{path}: no SWC samples")
    samples = []
    for kind, point, parent in raw:
        if parent < 0:
            up = -1
        elif parent in ids:
            up = ids[parent]
        else:
            raise ValueError(f"{path}: SWC parent id {parent} is not a sample")
        samples.append(_Sample(kind, point, up))
    for i, s in enumerate(samples):
        if s.up >= 0:
            samples[s.up].down.append(i)
    return samples


class _Tree:
    """The sample tree with the per-sample facts the section rules ask for."""

    def __init__(self, samples):
        self.s = samples
        self.roots = [i for i, s in enumerate(samples) if s.up < 0]
        # every sample must be reachable from a root (parent links may form cycles)
        self.root_of = [-1] * len(samples)
        for r in self.roots:
            todo = [r]
            while todo:
                i = todo.pop()
                self.root_of[i] = r
                todo.extend(samples[i].down)
        if min(self.root_of, default=0) < 0:
            raise ValueError("SWC parent links form a cycle that no root reaches")
        self.soma_kids = [sum(samples[c].kind == SOMA for c in s.down) for s in samples]
        self.somas_in_tree = {r: 0 for r in self.roots}
        for i, s in enumerate(samples):
            self.somas_in_tree[self.root_of[i]] += s.kind == SOMA
        self.dropped = set()
        self.proximal = set()
        self.next = [self._successor(i) for i in range(len(samples))]
        for r in self.roots:
            if self._three_point_soma(r):
                self.next[r] = -1
                self.dropped.update(c for c in samples[r].down if samples[c].kind == SOMA)

    def _successor(self, i):
        s = self.s[i]
        kids = s.down
        if s.kind == SOMA:
            if self.soma_kids[i] == 1 or (s.up < 0 and self.soma_kids[i] > 1):
                return next(c for c in kids if self.s[c].kind == SOMA)
            return -1
        if len(kids) == 1 and self.s[kids[0]].kind == s.kind:
            return kids[0]
        if s.up < 0 or self.s[s.up].kind == SOMA:
            own = [c for c in kids if self.s[c].kind == s.kind]
            if own:
                self.proximal.update(c for c in kids if c != own[0])
                return own[0]
        return -1

    def _three_point_soma(self, r):
        root = self.s[r]
        if root.kind != SOMA or self.somas_in_tree[r] != 3 or self.soma_kids[r] != 2:
            return False
        ends = [self.s[c] for c in root.down if self.s[c].kind == SOMA]
        d = root.point[3]
        if any(e.down or e.point[3] != d for e in ends):
            return False
        reach = sum(math.dist(e.point[:3], root.point[:3]) for e in ends)
        return abs(reach / d - 1.0) < 0.01


@dataclass
class _Piece:
    """A section under construction."""
    first: int            # its first sample
    last: int             # its last sample
    kind: int
    points: list
    parent: int           # index of the parent piece, -1 for a root
    x: float              # attachment point on the parent


def _attach(tree, pieces, parent, hook, kind, points, proximal):
    """Attachment point on piece `parent` for a child whose first sample hangs from
    sample `hook`, and whether the child repeats the hook sample."""
    P = pieces[parent]
    soma = P.kind == SOMA
    neurite = kind != SOMA
    x, repeat = 1.0, True
    if P.parent < 0 and soma and neurite and len(P.points) == 1:
        x, repeat = 0.5, len(points) == 1
    elif P.parent < 0 and hook == P.first:
        x, repeat = 0.0, not (neurite and tree.soma_kids[hook] > 1)
    elif soma and hook != P.last:
        x, repeat = 0.5, not (neurite and len(points) > 1)
    elif soma and tree.soma_kids[hook] > 1:
        repeat = not (neurite and len(points) > 1)
    if proximal:
        x = 0.0
    return x, repeat


def _pieces(tree):
    """Sections in formation order (depth first from the roots)."""
    s = tree.s
    pieces = []
    home = {}  # sample -> piece holding it
    work = [(r, -1) for r in reversed(tree.roots)]
    while work:
        first, _ = work.pop()
        run = [first]
        while tree.next[run[-1]] >= 0:
            run.append(tree.next[run[-1]])
        kind = s[first].kind
        points = [s[i].point for i in run]
        hook = s[first].up
        parent, x = (-1, 1.0) if hook < 0 else (home[hook], 1.0)
        if parent >= 0:
            x, repeat = _attach(tree, pieces, parent, hook, kind, points, first in tree.proximal)
            if repeat:
                px, py, pz, pd = s[hook].point
                if pieces[parent].kind == SOMA and kind != SOMA:
                    pd = points[0][3]
                points.insert(0, (px, py, pz, pd))
        index = len(pieces)
        pieces.append(_Piece(first, run[-1], kind, points, parent, x))
        for i in run:
            home[i] = index
        branches = [c for i in run for c in s[i].down if c != tree.next[i] and c not in tree.dropped]
        work.extend((c, index) for c in reversed(branches))
    return pieces


def load_swc(path, nseg=1, Ra=100.0, cm=1.0):
    """Sections of the morphology in the SWC file `path` (list of Section)."""
    tree = _Tree(_read(path))
    pieces = _pieces(tree)
    gone = [False] * len(pieces)
    names = [None] * len(pieces)
    per_label = {}
    out = []
    for k, p in enumerate(pieces):
        while p.parent >= 0 and gone[p.parent]:  # hang from the dropped parent's own parent
            up = pieces[p.parent]
            p.parent, p.x = up.parent, up.x
        pts = p.points
        if p.parent >= 0 and len(pts) <= 2 and pts[0][:3] == pts[-1][:3]:
            gone[k] = True
            continue
        if len(pts) == 1:
            x, y, z, d = pts[0]
            pts = [(x - d / 2, y, z, d), (x, y, z, d), (x + d / 2, y, z, d)]
        label = LABELS.get(p.kind, str(p.kind))
        k_label = per_label.get(label, 0)
        per_label[label] = k_label + 1
        names[k] = f"{label}_{k_label}" if p.kind in LABELS else f"sec_{p.kind}_{k_label}"
        out.append(Section(names[k], label=label, parent=None if p.parent < 0 else names[p.parent],
                           parent_x=p.x, points=pts, nseg=nseg, Ra=Ra, cm=cm))
    return out
