"""Class-tree selector (selector 2): the reference's offline-trained
decision trees retrained for this engine, one per application, mapping the
12 selector features to a variant class. Results stay bit-exact under the
committed trees and under one-leaf trees that pin a class; a pinned class is
what the inspected levels run; bad tables are rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def leaf(cls):
    return {"feat": [-1], "thr": [0.0], "left": [-1], "right": [-1], "cls": [cls]}


def record_class(gsw, r):
    from gswitch_amd import learn
    return learn.class_of_record(r)


@pytest.mark.parametrize("make", [lambda gsw: gsw.rmat(13, 16, seed=3),
                                  lambda gsw: gsw.grid2d(60, 60, 0.1, 0.01, 2),
                                  lambda gsw: gsw.barabasi_albert(5000, 4, 9)])
def test_committed_trees_bit_exact(gsw, orc, make):
    from gswitch_amd import learn
    gsw.selector_load_tree(learn.load_tree(learn.tree_path("bfs")), app="bfs")
    gsw.selector_load_tree(learn.load_tree(learn.tree_path("sssp")), app="sssp")
    h = make(gsw)
    g = gsw.Graph(h)
    root = h.max_degree_vertex()
    assert np.array_equal(g.bfs(root, selector=2), orc.bfs(h.row_ptr, h.col, root))
    g.close()


@pytest.mark.parametrize("make", [lambda gsw: gsw.rmat(12, 16, seed=2, weighted=True),
                                  lambda gsw: gsw.grid2d(50, 50, 0.1, 0.01, 3, weighted=True)])
def test_committed_sssp_tree_bit_exact(gsw, orc, make):
    from gswitch_amd import learn
    gsw.selector_load_tree(learn.load_tree(learn.tree_path("sssp")), app="sssp")
    h = make(gsw)
    g = gsw.Graph(h)
    root = h.max_degree_vertex()
    assert np.array_equal(g.sssp(root, selector=2), orc.spfa(h.row_ptr, h.col, h.w, root))
    g.close()


@pytest.mark.parametrize("cls", ["Push-Queue-CM-Standalone", "Pull-Bitmap-TM-Standalone",
                                 "Push-Bitmap-STRICT-Standalone"])
def test_one_leaf_tree_pins_the_class(gsw, orc, cls):
    """A tree that is one leaf picks the same class at every inspected
    level: the expanded standalone levels run that class."""
    c = gsw.MLP_CLASSES.index(cls)
    gsw.selector_load_tree(leaf(c), app="bfs")
    h = gsw.rmat(12, 16, seed=5)
    g = gsw.Graph(h)
    root = h.max_degree_vertex()
    assert np.array_equal(g.bfs(root, selector=2), orc.bfs(h.row_ptr, h.col, root))
    got = {record_class(gsw, r) for r in g.records() if not r["fused"]}
    assert got == {c}, (cls, [gsw.MLP_CLASSES[k] for k in got])
    g.close()


def test_bad_trees_are_rejected(gsw):
    good = leaf(3)
    for bad in [dict(good, cls=[30]), dict(good, cls=[-1]),
                {"feat": [12, -1, -1], "thr": [0.5, 0, 0], "left": [1, -1, -1],
                 "right": [2, -1, -1], "cls": [0, 1, 2]},      # feature out of range
                {"feat": [0, -1, -1], "thr": [0.5, 0, 0], "left": [0, -1, -1],
                 "right": [2, -1, -1], "cls": [0, 1, 2]},      # child not after parent
                {"feat": [0, -1, -1], "thr": [np.nan, 0, 0], "left": [1, -1, -1],
                 "right": [2, -1, -1], "cls": [0, 1, 2]}]:
        with pytest.raises(gsw.GswError):
            gsw.selector_load_tree(bad, app="bfs")
    with pytest.raises(ValueError):
        gsw.selector_load_tree(good, app="pagerank")
