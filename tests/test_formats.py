"""On-disk formats (no device): the connectome CSV reader (load_connectivity,
src/model/macro/model.jl:22-54), the SWC reader (load_swc, morphology.jl:130-338) and
the Neurolucida ASC reader (load_asc, morphology.jl:344-398) on files written here,
with closed-form geometry and their rejections."""
import numpy as np
import pytest

from mind_amd import macro as M


def _write(path, labels, W, D, quote=False):
    q = (lambda s: '"' + s.replace('"', '""') + '"') if quote else (lambda s: s)
    lines = ["matrix,weights", "," + ",".join(q(x) for x in labels)]
    lines += [q(labels[i]) + "," + ",".join(repr(float(v)) for v in W[i]) for i in range(len(labels))]
    lines += ["", "matrix,delays_ms", "," + ",".join(q(x) for x in labels)]
    lines += [q(labels[i]) + " , " + " , ".join(repr(float(v)) for v in D[i]) for i in range(len(labels))]
    path.write_text("\n".join(lines) + "\n")


def _matrices(n=5, seed=3):
    g = np.random.default_rng(seed)
    W = g.uniform(0.0, 0.2, (n, n))
    np.fill_diagonal(W, 0.0)
    D = g.uniform(2.0, 10.0, (n, n))
    np.fill_diagonal(D, 0.0)
    W[1, 3] = 0.0
    D[1, 3] = 0.0
    return W, D


@pytest.mark.parametrize("quote", [False, True])
def test_connectome_csv_round_trip(tmp_path, quote):
    W, D = _matrices()
    labels = ["V1", "lh, MT", 'say "hi"', "PFC", "A1"] if quote else ["V1", "V2", "MT", "PFC", "A1"]
    p = tmp_path / "c.csv"
    _write(p, labels, W, D, quote)
    c = M.load_connectivity(p)
    assert c.labels == labels
    assert np.array_equal(c.weights, W) and np.array_equal(c.delays, D)
    mm = M.MacroModel(c)
    assert mm.roi(labels[2]).index == 2


@pytest.mark.parametrize("defect,message", [
    ("diagonal", "diagonal"), ("delay", "delay must be zero"), ("labels", "different ROI labels"),
    ("missing", "no matrix,delays_ms"), ("short", "labels or size"), ("dup", "unique"),
])
def test_connectome_csv_rejections(tmp_path, defect, message):
    W, D = _matrices()
    labels = ["a", "b", "c", "d", "e"]
    dl = list(labels)
    if defect == "diagonal":
        W[2, 2] = 0.1
    elif defect == "delay":
        D[1, 3] = 4.0
    elif defect == "labels":
        dl[4] = "z"
    elif defect == "dup":
        labels[4] = "a"
        dl = list(labels)
    p = tmp_path / "c.csv"
    _write(p, labels, W, D)
    text = p.read_text()
    if defect == "labels":
        head, tail = text.split("matrix,delays_ms")
        tail = tail.replace(",a,b,c,d,e", ",a,b,c,d,z").replace("\ne ,", "\nz ,")
        text = head + "matrix,delays_ms" + tail
    if defect == "missing":
        text = text.split("matrix,delays_ms")[0]
    if defect == "short":
        text = text.replace("\nd,", "\nd,0.5,")
    p.write_text(text)
    with pytest.raises(ValueError, match=message):
        M.load_connectivity(p)


# -- SWC (load_swc, src/model/micro/morphology.jl:130-338) --------------------------
from mind_amd import MicroModel  # noqa: E402
from mind_amd.model import pack_nodes  # noqa: E402
from mind_amd.swc import load_swc  # noqa: E402


def _swc(tmp_path, text):
    p = tmp_path / "m.swc"
    p.write_text("# test morphology\n" + text)
    return p


def _by_name(sections):
    return {s.name: s for s in sections}


def test_swc_point_soma_and_chain(tmp_path):
    """A one-point soma becomes three points one diameter long; a neurite wired to it
    attaches at its middle and starts at its own first sample."""
    secs = load_swc(_swc(tmp_path, "1 1 0 0 0 5 -1\n2 3 0 5 0 1 1\n3 3 0 15 0 1 2\n4 3 0 30 0 1 3\n"))
    s = _by_name(secs)
    assert [x.name for x in secs] == ["soma_0", "dend_0"]
    assert s["soma_0"].points == [(-5.0, 0.0, 0.0, 10.0), (0.0, 0.0, 0.0, 10.0), (5.0, 0.0, 0.0, 10.0)]
    assert s["soma_0"].length == 10.0 and s["soma_0"].parent is None
    assert s["dend_0"].parent == "soma_0" and s["dend_0"].parent_x == 0.5
    assert s["dend_0"].length == 25.0 and s["dend_0"].label == "dend"
    assert s["dend_0"].points[0] == (0.0, 5.0, 0.0, 2.0)


def test_swc_three_point_soma_is_a_sphere(tmp_path):
    """NeuroMorpho's three-point soma: one section whose compartment has the sphere's area
    4 pi r^2."""
    text = ("1 1 0 0 0 5 -1\n2 1 0 -5 0 5 1\n3 1 0 5 0 5 1\n"
            "4 3 0 5 0 1 1\n5 3 0 20 0 1 4\n")
    secs = load_swc(_swc(tmp_path, text))
    s = _by_name(secs)
    assert [x.name for x in secs] == ["soma_0", "dend_0"]
    assert len(s["soma_0"].points) == 3 and s["soma_0"].length == 10.0
    assert s["dend_0"].parent_x == 0.5 and s["dend_0"].length == 15.0
    m = MicroModel()
    m.add_population([s["soma_0"]], 1)
    geo = pack_nodes(m)
    # nodes: the root (area 100 by convention), the soma's compartment, its end
    assert geo["nodes"] == 3
    assert geo["area"][1] == pytest.approx(4.0 * np.pi * 5.0 ** 2, rel=1e-14)


def test_swc_branches_and_zero_length_sections(tmp_path):
    """A bifurcation starts each child from a copy of the branch sample; a zero-length
    two-point child is dropped and its child reattached through it."""
    text = ("1 1 0 0 0 5 -1\n2 3 10 0 0 1 1\n3 3 20 0 0 1 2\n4 3 30 5 0 0.5 3\n"
            "5 3 30 -5 0 0.5 3\n6 3 20 0 0 0.5 3\n7 3 25 0 0 0.5 6\n")
    secs = load_swc(_swc(tmp_path, text))
    s = _by_name(secs)
    assert [x.name for x in secs] == ["soma_0", "dend_0", "dend_1", "dend_2", "dend_3"]
    assert s["dend_0"].parent == "soma_0" and s["dend_0"].length == 10.0
    for name, y in (("dend_1", 5.0), ("dend_2", -5.0)):
        assert s[name].parent == "dend_0" and s[name].parent_x == 1.0
        assert s[name].points[0] == (20.0, 0.0, 0.0, 2.0)
        assert s[name].length == pytest.approx(np.hypot(10.0, y), rel=1e-15)
    assert s["dend_3"].parent == "dend_0" and s["dend_3"].parent_x == 1.0
    assert s["dend_3"].length == 5.0
    m = MicroModel()
    (cell,) = m.add_population(secs, 1)
    assert m._morph(cell).section_index("dend_3") == 4


def test_swc_soma_chain_attachments(tmp_path):
    """A three-sample soma chain (not the three-point sphere: the ends continue the
    root): a neurite leaving its middle sample hangs at 0.5 without repeating the
    soma sample; one leaving its last sample hangs at 1 and starts with a copy of
    that sample carrying the neurite's own diameter."""
    text = ("1 1 0 0 0 4 -1\n2 1 0 6 0 4 1\n3 1 0 12 0 4 2\n"
            "4 3 5 6 0 1 2\n5 3 15 6 0 1 4\n6 2 0 20 0 0.5 3\n7 2 0 30 0 0.5 6\n")
    s = _by_name(load_swc(_swc(tmp_path, text)))
    assert sorted(s) == ["axon_0", "dend_0", "soma_0"]
    assert s["soma_0"].points == [(0.0, 0.0, 0.0, 8.0), (0.0, 6.0, 0.0, 8.0), (0.0, 12.0, 0.0, 8.0)]
    assert s["dend_0"].parent == "soma_0" and s["dend_0"].parent_x == 0.5
    assert s["dend_0"].points == [(5.0, 6.0, 0.0, 2.0), (15.0, 6.0, 0.0, 2.0)]
    assert s["axon_0"].parent == "soma_0" and s["axon_0"].parent_x == 1.0
    assert s["axon_0"].points == [(0.0, 12.0, 0.0, 1.0), (0.0, 20.0, 0.0, 1.0), (0.0, 30.0, 0.0, 1.0)]


def test_swc_neurite_root_with_proximal_sibling(tmp_path):
    """A root that is not soma with two children of its type: the first continues its
    section; the second starts a section at the proximal end (x = 0) that repeats
    the root sample; a child of another type under a continuing sample starts a
    section at x = 1; unknown types are named sec_<type>_<k>."""
    text = ("1 2 0 0 0 1 -1\n2 2 10 0 0 1 1\n3 2 0 10 0 1 1\n4 2 20 0 0 1 2\n"
            "5 7 20 10 0 0.5 4\n")
    secs = load_swc(_swc(tmp_path, text))
    s = _by_name(secs)
    assert [x.name for x in secs] == ["axon_0", "axon_1", "sec_7_0"]  # branches in sample order along the run
    assert s["axon_0"].parent is None
    assert s["axon_0"].points == [(0.0, 0.0, 0.0, 2.0), (10.0, 0.0, 0.0, 2.0), (20.0, 0.0, 0.0, 2.0)]
    assert s["axon_1"].parent == "axon_0" and s["axon_1"].parent_x == 0.0
    assert s["axon_1"].points == [(0.0, 0.0, 0.0, 2.0), (0.0, 10.0, 0.0, 2.0)]
    assert s["sec_7_0"].parent == "axon_0" and s["sec_7_0"].parent_x == 1.0
    assert s["sec_7_0"].label == "7" and s["sec_7_0"].points[0] == (20.0, 0.0, 0.0, 2.0)


@pytest.mark.parametrize("text,message", [
    ("1 1 0 0 0 5\n", "7 columns"),
    ("1 1 0 0 0 5 -1\n1 3 0 1 0 1 1\n", "appears twice"),
    ("1 1 0 0 0 5 -1\n2 3 0 1 0 1 9\n", "is not a sample"),
    ("1 3 0 0 0 1 2\n2 3 0 1 0 1 1\n", "cycle"),
    ("\n", "no SWC samples"),
])
def test_swc_rejections(tmp_path, text, message):
    with pytest.raises(ValueError, match=message):
        load_swc(_swc(tmp_path, text))


# --- Neurolucida ASC (load_asc, src/model/micro/morphology.jl:344-398) --------------
from mind_amd import MicroModel, load_asc  # noqa: E402
from mind_amd import asc as A  # noqa: E402
from mind_amd.model import pack_nodes  # noqa: E402


def _ellipse(a, b, center, n=64):
    cx, cy, cz = center
    return [(cx + a * np.cos(2 * np.pi * k / n), cy + b * np.sin(2 * np.pi * k / n), cz, 1.0)
            for k in range(n + 1)]


def _asc_text(soma=None, trees=()):
    parts = []
    if soma is not None:
        pts = " ".join(f"({float(x)!r} {float(y)!r} {float(z)!r} {float(d)!r})" for x, y, z, d in soma)
        parts.append(f'("CellBody"\n  (Color Red)\n  (CellBody)\n  {pts}\n)  ;  End of contour')
    parts.extend(trees)
    return "\n".join(parts) + "\n"


DEND = """( (Color Yellow)
  (Dendrite)
  (0 0 0 2) ; root sample
  (10, 0, 0, 2)
  (
    (20 5 0 1) <(comment markup)> (30 10 0 1) Normal
  |
    (10 0 0 1.5) (20 -5 0 -1.5)
    (
      (25 -5 0 1)
    |
      (20 -15 0 0)
    )
  )
)"""


def test_asc_branches_terminals_and_labels(tmp_path):
    """Branch structure without a soma: the root branch is a root section; a child that
    does not start at its parent's last sample gets a copy of it with the child's first
    diameter; one that does is not extended; single-sample branches become three points
    one diameter long along x; |d| is the diameter and d = 0 becomes 1; markup, comments,
    commas and words are ignored."""
    f = tmp_path / "tree.asc"
    f.write_text(_asc_text(trees=[DEND, DEND.replace("(Dendrite)", "(Apical)"),
                                  "((Marker) (1 2 3 1))"]))
    secs = {s.name: s for s in load_asc(f)}
    assert sorted(secs) == ["apic_0", "apic_1", "apic_2", "apic_3", "apic_4",
                            "dend_0", "dend_1", "dend_2", "dend_3", "dend_4"]
    d = secs
    assert d["dend_0"].parent is None and d["dend_0"].parent_x == 1.0
    assert d["dend_0"].points == [(0.0, 0.0, 0.0, 2.0), (10.0, 0.0, 0.0, 2.0)]
    assert d["dend_1"].parent == "dend_0" and d["dend_1"].parent_x == 1.0
    assert d["dend_1"].points == [(10.0, 0.0, 0.0, 1.0), (20.0, 5.0, 0.0, 1.0), (30.0, 10.0, 0.0, 1.0)]
    assert d["dend_2"].points == [(10.0, 0.0, 0.0, 1.5), (20.0, -5.0, 0.0, 1.5)]  # starts at the parent
    assert d["dend_3"].parent == "dend_2"
    assert d["dend_3"].points == [(20.0, -5.0, 0.0, 1.0), (25.0, -5.0, 0.0, 1.0)]
    assert d["dend_4"].points == [(20.0, -5.0, 0.0, 1.0), (20.0, -15.0, 0.0, 1.0)]
    one = A.arbors(A.parse("((Axon) (4 5 6 2))")[0])
    assert [b.points for b in one] == [[(4.0, 5.0, 6.0, 2.0)]]
    f2 = tmp_path / "one.asc"
    f2.write_text("((Axon) (4 5 6 2))\n")
    (s,) = load_asc(f2)
    assert s.name == "axon_0" and s.points == [(3.0, 5.0, 6.0, 2.0), (4.0, 5.0, 6.0, 2.0), (5.0, 5.0, 6.0, 2.0)]


def test_asc_soma_contour_is_resampled_along_its_major_axis(tmp_path):
    """An elliptic CellBody contour (semi-axes 12 along x, 5 along y, centre (1, 2, 3)):
    21 soma points on the major axis through the centre, spanning it between the
    second-extreme resampled positions, with diameters close to the chord 2b sqrt(1 -
    (x/a)^2); dendrites attach to soma_0 at 0.5."""
    a, b, c = 12.0, 5.0, (1.0, 2.0, 3.0)
    f = tmp_path / "soma.asc"
    f.write_text(_asc_text(soma=_ellipse(a, b, c), trees=[DEND]))
    secs = load_asc(f, nseg=3)
    soma = secs[0]
    assert soma.name == "soma_0" and soma.label == "soma" and soma.nseg == 3
    pts = np.array(soma.points)
    assert pts.shape == (21, 4)
    np.testing.assert_allclose(pts[:, 1], c[1], atol=0.05)  # on the major (x) axis
    np.testing.assert_allclose(pts[:, 2], c[2], atol=1e-9)
    x = pts[:, 0] - c[0]
    assert np.all(np.diff(x) > 0) and x[0] < -0.9 * a and x[-1] > 0.9 * a
    chord = 2 * b * np.sqrt(np.clip(1 - (x / a) ** 2, 0, None))
    np.testing.assert_allclose(pts[1:-1, 3], chord[1:-1], rtol=0.02, atol=0.05)
    assert all(s.parent == "soma_0" and s.parent_x == 0.5 for s in secs if s.name in ("dend_0",))
    # a soma section in a model packs like the oracle's
    m = MicroModel()
    m.add_population(secs, 2)
    import oracle
    ref = oracle.OracleSim(m).geometry()
    got = pack_nodes(m)
    for k in ("parent", "area", "cap", "axial", "adiag"):
        assert np.array_equal(got[k], ref[k]), k


def test_asc_soma_direction_and_degenerate_contours():
    """The major axis is oriented with its largest component positive; a contour with
    fewer than three samples, zero length, or a collinear outline gives no soma."""
    tilted = [(y, x, 0.0, 1.0) for x, y, _, _ in _ellipse(9.0, 3.0, (0.0, 0.0, 0.0))]  # major along y
    pts = np.array(A.soma_points(tilted))
    assert np.all(np.diff(pts[:, 1]) > 0) and np.allclose(pts[:, 0], 0.0, atol=0.05)
    assert A.soma_points([(0, 0, 0, 1), (1, 0, 0, 1)]) == []
    assert A.soma_points([(1, 1, 1, 1)] * 5) == []
    assert A.soma_points([(k, 0.0, 0.0, 1.0) for k in range(6)]) == []


@pytest.mark.parametrize("text,message", [
    ("((Dendrite) (0 0 0 1) (1 2 3 1)", "never closed"),
    ("((Dendrite) (0 0 inf 1))", "not finite"),
    ("((Dendrite) (0 0 1) (1 x 2 1))", "three numeric coordinates"),
    ('("CellBody" (0 0 0 1) (1 0 0 1) (0 1 0 1)) ("CellBody" (5 0 0 1) (6 0 0 1) (5 1 0 1))', "more than one CellBody"),
])
def test_asc_rejections(tmp_path, text, message):
    f = tmp_path / "bad.asc"
    f.write_text(text + "\n")
    with pytest.raises(ValueError, match=message):
        load_asc(f)
