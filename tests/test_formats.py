"""On-disk formats (no device): the connectome CSV reader (load_connectivity,
src/model/macro/model.jl:22-54) on files written here, including quoted labels,
and its rejections."""
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
    ("diagonal", "diagonal"), ("delay", "delay must be zero"), ("labels", "labels differ"),
    ("missing", "missing matrix,delays_ms"), ("short", "labels or dimensions"), ("dup", "unique"),
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


@pytest.mark.parametrize("text,message", [
    ("1 1 0 0 0 5\n", "seven SWC columns"),
    ("1 1 0 0 0 5 -1\n1 3 0 1 0 1 1\n", "duplicate"),
    ("1 1 0 0 0 5 -1\n2 3 0 1 0 1 9\n", "missing parent"),
    ("1 3 0 0 0 1 2\n2 3 0 1 0 1 1\n", "unrooted cycle"),
    ("\n", "empty"),
])
def test_swc_rejections(tmp_path, text, message):
    with pytest.raises(ValueError, match=message):
        load_swc(_swc(tmp_path, text))
