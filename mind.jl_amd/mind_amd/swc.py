"""SWC morphology import (load_swc, src/model/micro/morphology.jl:130-338).

Reads the seven-column SWC format (id, type, x, y, z, radius, parent) into
`Section`s with 3-D points (diameter = 2 x radius), following the section rules
of NEURON's Import3d that the reference keeps:

1. samples are resolved to dense indices; every sample must descend from a root
   (parent -1);
2. a section is a maximal chain of samples: a soma sample continues to its soma
   child when it has exactly one (a root soma with several continues to its first);
   a neurite sample continues to its only child of the same type; below a root or
   a soma sample the first same-type child continues and its siblings attach
   proximally (x = 0);
3. NeuroMorpho's three-point soma (a root soma sample with two leaf soma children
   of the same diameter spanning one diameter) is one sphere-like section;
4. a child section starts from a copy of its parent sample (with the child's own
   diameter where it leaves a soma) unless it is wired to a one-point soma or to
   a soma's interior (then it attaches at x = 0.5); zero-length one- or two-point
   child sections are dropped, their children reattached through them; a
   one-point section becomes three points one diameter long along x.

Sections are named <label>_<ordinal> (soma, axon, dend, apic; other types
sec_<type>_<ordinal>) in the order they are formed (depth first from the roots).
"""
from __future__ import annotations

import math

from .model import Section

LABELS = {1: "soma", 2: "axon", 3: "dend", 4: "apic"}


def load_swc(path, nseg=1, Ra=100.0, cm=1.0):
    """Sections of the morphology in the SWC file `path` (list of Section)."""
    # 1. samples, dense indices, children, roots
    index_of, kinds, parents, samples = {}, [], [], []
    with open(path) as f:
        for line_number, line in enumerate(f, 1):
            content = line.split("#", 1)[0].strip()
            if not content:
                continue
            values = content.split()
            if len(values) < 7:
                raise ValueError(f"{path}:{line_number}: expected seven SWC columns")
            sid, kind = int(values[0]), int(values[1])
            if sid in index_of:
                raise ValueError(f"duplicate SWC sample {sid}")
            samples.append((float(values[2]), float(values[3]), float(values[4]), 2.0 * float(values[5])))
            kinds.append(kind)
            parents.append(int(values[6]))
            index_of[sid] = len(samples) - 1
    if not samples:
        raise ValueError("empty SWC morphology")
    n = len(samples)
    children = [[] for _ in range(n)]
    roots = []
    for i in range(n):
        p = parents[i]
        if p < 0:
            parents[i] = -1
            roots.append(i)
        else:
            if p not in index_of:
                raise ValueError(f"SWC has missing parent {p}")
            parents[i] = index_of[p]
            children[parents[i]].append(i)
    root_soma = [0] * len(roots)
    visited = 0
    for k, r in enumerate(roots):
        stack = [r]
        while stack:
            i = stack.pop()
            visited += 1
            root_soma[k] += kinds[i] == 1
            stack.extend(reversed(children[i]))
    if visited != n:
        raise ValueError("SWC contains an unrooted cycle")

    # 2. continuation of each sample's section, and proximal attachments
    soma_children = [sum(kinds[c] == 1 for c in children[i]) for i in range(n)]
    cont = [-1] * n
    proximal = [False] * n
    for i in range(n):
        branch, kind = children[i], kinds[i]
        cand = -1
        if kind == 1:
            if soma_children[i] == 1 or (parents[i] < 0 and soma_children[i] > 1):
                cand = next(c for c in branch if kinds[c] == 1)
        elif len(branch) == 1 and kinds[branch[0]] == kind:
            cand = branch[0]
        elif parents[i] < 0 or kinds[parents[i]] == 1:
            same = [c for c in branch if kinds[c] == kind]
            if same:
                cand = same[0]
                for c in branch:
                    if c != cand:
                        proximal[c] = True
        cont[i] = cand
    # three-point soma: one sphere, not three cables
    absorbed = [False] * n
    for k, r in enumerate(roots):
        if not (kinds[r] == 1 and root_soma[k] == 3 and soma_children[r] == 2):
            continue
        ends = [c for c in children[r] if kinds[c] == 1]
        cx, cy, cz, cd = samples[r]
        if not all(not children[c] and samples[c][3] == cd for c in ends):
            continue
        span = sum(math.sqrt((samples[c][0] - cx) ** 2 + (samples[c][1] - cy) ** 2 + (samples[c][2] - cz) ** 2)
                   for c in ends)
        if not abs(span / cd - 1.0) < 0.01:
            continue
        cont[r] = -1
        for c in ends:
            absorbed[c] = True

    # 3. sections and their attachments
    section_of = [-1] * n
    heads, tails, sec_parent, connection, geometry = [], [], [], [], []
    pending = list(reversed(roots))
    while pending:
        head = pending.pop()
        i = head
        kind, parent_sample = kinds[head], parents[head]
        points, branches = [], []
        while True:
            points.append(samples[i])
            section_of[i] = len(heads)
            for c in children[i]:
                if c != cont[i] and not absorbed[c]:
                    branches.append(c)
            if cont[i] < 0:
                break
            i = cont[i]
        parent = -1 if parent_sample < 0 else section_of[parent_sample]
        parent_x = 1.0
        if parent >= 0:
            soma = kinds[heads[parent]] == 1
            wire_first = False
            if sec_parent[parent] < 0 and soma and kind != 1 and len(geometry[parent]) == 1:
                parent_x = 0.5
                wire_first = len(points) > 1
            elif sec_parent[parent] < 0 and parent_sample == heads[parent]:
                parent_x = 0.0
                wire_first = kind != 1 and soma_children[parent_sample] > 1
            elif soma:
                if parent_sample != tails[parent]:
                    parent_x = 0.5
                    wire_first = kind != 1 and len(points) > 1
                elif soma_children[parent_sample] > 1:
                    wire_first = kind != 1 and len(points) > 1
            if proximal[head]:
                parent_x = 0.0
            if not wire_first:
                prev = samples[parent_sample]
                if soma and kind != 1:
                    prev = (prev[0], prev[1], prev[2], points[0][3])
                points.insert(0, prev)
        heads.append(head)
        tails.append(i)
        sec_parent.append(parent)
        connection.append(parent_x)
        geometry.append(points)
        pending.extend(reversed(branches))

    # 4. drop zero-length child sections, name and build the rest
    sections, names = [], [None] * len(heads)
    removed = [False] * len(heads)
    counts = {}
    for k in range(len(heads)):
        points = geometry[k]
        parent = sec_parent[k]
        if parent >= 0 and removed[parent]:
            connection[k] = connection[parent]
            parent = sec_parent[parent]
            sec_parent[k] = parent
        if parent >= 0 and len(points) <= 2 and points[0][:3] == points[-1][:3]:
            removed[k] = True
            continue
        if len(points) == 1:
            x, y, z, d = points[0]
            points = [(x - 0.5 * d, y, z, d), points[0], (x + 0.5 * d, y, z, d)]
        kind = kinds[heads[k]]
        label = LABELS.get(kind, str(kind))
        ordinal = counts.get(label, 0)
        counts[label] = ordinal + 1
        names[k] = f"{label}_{ordinal}" if kind in LABELS else f"sec_{kind}_{ordinal}"
        sections.append(Section(names[k], label=label, parent=None if parent < 0 else names[parent],
                                parent_x=connection[k], points=points, nseg=nseg, Ra=Ra, cm=cm))
    return sections
