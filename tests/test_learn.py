"""Host-side logic of the selector training pipeline (gswitch_amd/learn.py),
on the CPU: class numbering of engine records, label targets, regret, level
stride, seeded graph specs, model files, and the committed training sets."""
import json
import os

import numpy as np
import pytest

import gswitch_amd as gsw
from gswitch_amd import learn


def rec(d, f, lb, fused=False, sorted_=False):
    return {"dir": d, "fmt": f, "lb": lb, "fused": fused, "sorted": sorted_}


def test_class_numbering():
    # class = group * 5 + lb, lb in TM WM CM STRICT TWC
    assert learn.class_of_record(rec("Push", "Queue", "TM")) == 0
    assert learn.class_of_record(rec("Push", "Queue", "TWC", sorted_=True)) == 9
    assert learn.class_of_record(rec("Push", "Bitmap", "CM")) == 12
    assert learn.class_of_record(rec("Push", "Queue", "WM", fused=True)) == 16
    assert learn.class_of_record(rec("Pull", "Queue", "STRICT")) == 23
    assert learn.class_of_record(rec("Pull", "Bitmap", "TM")) == 25
    # edge-balanced relatives count as STRICT
    assert learn.class_of_record(rec("Pull", "Bitmap", "TWOD")) == 28
    # a fused level keeps its class whatever the queue's sortedness
    assert learn.class_of_record(rec("Push", "Queue", "CM", True, True)) == 17
    names = gsw.MLP_CLASSES
    assert len(names) == 30 and len(set(names)) == 30


def test_targets_and_regret():
    ms = np.array([[1.0, 2.0, 4.0, 1000.0], [3.0, 1.5, 1.5, 6.0]], np.float32)
    y, m = learn.targets(ms)
    # four classes fill the first four of the network's 30 outputs; the
    # rest are masked out of the loss
    assert y.shape == m.shape == (2, gsw.MLP_NOUT)
    np.testing.assert_allclose(y[0, :4], [0, 1, 2, learn.CLIP], atol=1e-6)
    np.testing.assert_allclose(y[1, :4], [1, 0, 0, 2], atol=1e-6)
    assert np.all(m[:, :4] == 1) and np.all(m[:, 4:] == 0) and np.all(y[:, 4:] == 0)
    best = ms.argmin(axis=1)
    assert learn.per_level_regret(ms, best) == pytest.approx(1.0)
    # choosing class 0 everywhere: (1 + 3) / (1 + 1.5)
    assert learn.per_level_regret(ms, np.zeros(2, int)) == pytest.approx(4 / 2.5)


def test_level_stride():
    assert learn.level_stride(5, 32) == 1
    assert learn.level_stride(32, 32) == 1
    assert learn.level_stride(33, 32) == 2
    assert learn.level_stride(1000, 32) == 32


def test_specs_are_seeded_and_prefix_stable():
    a, b = learn.training_specs(), learn.training_specs()
    assert a == b and len(a) == 300
    assert learn.training_specs(150) == a[:150]
    kinds = [s["kind"] for s in a]
    assert kinds.count("rmat") == 120 and kinds.count("grid") == 60
    for s in a:
        if s["kind"] == "rmat":
            assert 14 <= s["scale"] <= 22 and 0.45 <= s["a"] <= 0.65
            assert s["b"] == pytest.approx((1 - s["a"] - 0.05) / 2)
    e = learn.eval_specs()
    assert len(e) == 40 and e == learn.eval_specs()
    assert {s["scale"] for s in e if s["kind"] == "rmat"} == {20, 21, 22, 23, 24}
    # held-out graphs are not training graphs
    assert not ({s["seed"] for s in a} & {s["seed"] for s in e})


def test_model_file_roundtrip(tmp_path):
    w = learn.init_weights(seed=3)
    assert w.shape == (gsw.MLP_NPARAM,) and np.isfinite(w).all()
    np.testing.assert_array_equal(w, learn.init_weights(seed=3))
    p = str(tmp_path / "m.npz")
    learn.save_model(p, w, {"app": "sssp"})
    np.testing.assert_array_equal(learn.load_model(p), w)
    bad = w.copy()
    bad[5] = np.nan
    learn.save_model(p, bad, {})
    with pytest.raises(ValueError):
        learn.load_model(p)
    learn.save_model(p, w[:100], {})
    with pytest.raises(ValueError):
        learn.load_model(p)


@pytest.mark.parametrize("app", learn.APPS)
def test_committed_training_sets(orc, app):
    path = os.path.join(os.path.dirname(learn.MODEL_PATH), "selector_%s_train.npz" % app)
    if not os.path.exists(path):
        pytest.skip("no committed training set")
    with np.load(path, allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    n = len(d["ms"])
    k = {"pr": 8, "cc": 4}.get(app, 30)  # PageRank / CC: their own classes
    assert d["features"].shape == (n, 12) and d["ms"].shape == (n, k)
    assert np.all(d["ms"] > 0) and np.isfinite(d["features"]).all()
    assert np.all((d["rule"] >= 0) & (d["rule"] < k))
    assert len(np.unique(d["graph"])) == len(json.loads(str(d["specs"]))) == 300
    # the committed network's choices (FP64 oracle network) cost less than
    # the rule's over its own training levels
    w = learn.load_model(learn.model_path(app))
    choice = orc.mlp_forward(w, d["features"])[:, :k].argmin(axis=1)
    r_mlp = learn.per_level_regret(d["ms"], choice)
    r_rule = learn.per_level_regret(d["ms"], d["rule"])
    assert r_mlp < r_rule and r_mlp < 1.2, (app, r_mlp, r_rule)


def test_fit_tree_closed_forms():
    """Cost trees on hand-made levels: a feature that separates two regimes
    (class 1 cheap below 0.5, class 2 cheap above) gives one split at the
    midpoint with those leaves; a single cheapest class gives one leaf;
    tree_select follows the table with float32 comparisons."""
    from gswitch_amd import learn
    rng = np.random.default_rng(1)
    X = rng.uniform(0, 1, (200, 12)).astype(np.float32)
    ms = np.full((200, 30), 10.0)
    lo = X[:, 4] <= 0.5
    ms[lo, 1] = 1.0
    ms[~lo, 2] = 1.0
    t = learn.fit_tree(X, ms, depth=1, min_leaf=4)
    assert t["feat"][0] == 4 and len(t["feat"]) == 3
    xs = np.sort(X[:, 4])
    k = np.searchsorted(xs, 0.5, side="right")
    assert xs[k - 1] <= t["thr"][0] < xs[k]
    assert t["cls"][t["left"][0]] == 1 and t["cls"][t["right"][0]] == 2
    assert np.array_equal(learn.tree_select(t, X), np.where(lo, 1, 2))
    ms2 = np.full((50, 30), 5.0)
    ms2[:, 7] = 1.0
    t2 = learn.fit_tree(X[:50], ms2)
    assert t2["feat"] == [-1] and t2["cls"] == [7]


def test_pagerank_classes():
    """The PageRank classes are the from-all pulls of a bitmap frontier,
    each (variant, pull kernel) once; records map back to their class."""
    assert len(gsw.PR_CLASSES) == 8 and len(set(gsw.PR_CLASSES)) == 8
    assert gsw.PR_CLASS_MASK == 0xFF
    assert learn.MASK["pr"] == 0xFF and learn.MASK["bfs"] == (1 << 30) - 1
    assert learn.MASK["cc"] == 0xF and gsw.CC_CLASSES[0] is None

    def rec(lb, kernel):
        return dict(dir="Pull", fmt="Bitmap", lb=lb, kernel=kernel, fused=False,
                    sorted=False)
    for c, (cfg, pk) in enumerate(gsw.PR_CLASSES):
        lb = cfg.split("-")[2]
        assert gsw.pr_class_of_record(rec(lb, pk)) == c
    # TM/WM/CM do not depend on the pull kernel; STRICT under propagation
    # blocking keeps the merge-path walk
    assert gsw.pr_class_of_record(rec("WM", 0)) == 1
    assert gsw.pr_class_of_record(rec("STRICT", 2)) == 3
    assert gsw.pr_class_of_record(rec("ELB", 1)) == 4


def test_pagerank_features_are_the_first_level_inputs():
    """A PageRank run's first level has every vertex active: the inputs are
    the graph's, with level 0 and the whole edge count as push workload."""
    h = gsw.rmat(10, 8, seed=3)
    x = learn.pagerank_features(h)
    st = gsw.degree_stats(h.row_ptr)
    n, m = h.n, h.m
    assert x[0] == pytest.approx(np.log2(1 + n) / 24, rel=1e-6)
    assert x[1] == pytest.approx(np.log2(1 + m) / 32, rel=1e-6)
    assert x[2] == pytest.approx(np.log2(1 + st["avg"]) / 8, rel=1e-6)
    assert x[5] == 0 and x[6] == pytest.approx(1.0) and x[7] == 0
    assert x[8] == pytest.approx(np.log2(1 + m) / 32, rel=1e-6)
    assert x[9] == 0
