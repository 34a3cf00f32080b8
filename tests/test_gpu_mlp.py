"""GPU checks of the learned selector: the HIP forward / selection /
gradient kernels against the FP64 oracle network, reproducible training,
the engine's exhaustive per-level timing (its features against counts
recomputed from oracle BFS levels), and BFS under the learned policy staying
bit-exact whatever classes the network picks."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NP = 6942


def rand_problem(seed, B=50, scale=0.25):
    rng = np.random.default_rng(seed)
    w = rng.normal(0, scale, NP).astype(np.float32)
    x = rng.random((B, 12)).astype(np.float32)
    t = (rng.random((B, 30)) * 3).astype(np.float32)
    m = (rng.random((B, 30)) > 0.25).astype(np.float32)
    return w, x, t, m


def test_forward_matches_oracle(gsw, orc):
    w, x, _, _ = rand_problem(0, B=300)
    y = gsw.mlp_forward(w, x)
    ref = orc.mlp_forward(w, x)
    np.testing.assert_allclose(y, ref, rtol=1e-4, atol=1e-5 * np.abs(ref).max())


def test_select_is_masked_argmin(gsw, orc):
    w, x, _, _ = rand_problem(1, B=400)
    ref = orc.mlp_forward(w, x)
    for mask in [(1 << 30) - 1, 0b10101010101010101010101010101, 1 << 17]:
        cls = gsw.mlp_select(w, x, mask)
        allowed = np.array([(mask >> c) & 1 for c in range(30)], bool)
        r = np.where(allowed, ref, np.inf)
        assert allowed[cls].all()
        # the chosen class is a minimum up to FP32 rounding
        best = r.min(axis=1)
        assert np.all(r[np.arange(len(x)), cls] <= best + 1e-5 * (1 + np.abs(best)))


@pytest.mark.parametrize("seed,B", [(2, 1), (3, 37), (4, 513)])
def test_gradient_matches_oracle(gsw, orc, seed, B):
    w, x, t, m = rand_problem(seed, B=B)
    loss, g = gsw.mlp_grad(w, x, t, m)
    rl, rg = orc.mlp_grad(w, x, t, m)
    assert loss == pytest.approx(rl, rel=1e-4)
    np.testing.assert_allclose(g, rg, rtol=1e-3, atol=1e-5 * np.abs(rg).max())


def test_training_reproducible_and_learns(gsw, orc):
    rng = np.random.default_rng(5)
    x = rng.random((700, 12)).astype(np.float32)
    teacher = rng.normal(0, 0.3, NP)
    t = orc.mlp_forward(teacher, x).astype(np.float32)
    m = np.ones_like(t)
    w0 = (rng.normal(0, 1, NP) * 0.1).astype(np.float32)
    w1, c1 = gsw.mlp_train(w0, x, t, m, epochs=40, batch=64, lr=3e-3, seed=9)
    w2, c2 = gsw.mlp_train(w0, x, t, m, epochs=40, batch=64, lr=3e-3, seed=9)
    assert np.array_equal(w1, w2) and np.array_equal(c1, c2)
    assert c1[-1] < 0.3 * c1[0]
    # the reported epoch loss is the loss of the data under the weights
    l_end, _ = orc.mlp_grad(w1, x, t, m)
    assert l_end < c1[0]


def bfs_counts(h, lv, l):
    deg = h.degrees.astype(np.float64)
    act = lv == l
    inact = (lv < 0) | (lv > l)
    return (act.sum(), inact.sum(), deg[act].sum(), deg[inact].sum(),
            deg[act].max() if act.any() else 0)


def test_explore_bfs_features_and_times(gsw, orc):
    h = gsw.rmat(12, 16, seed=1)
    g = gsw.Graph(h)
    root = h.max_degree_vertex()
    lv = orc.bfs(h.row_ptr, h.col, root)
    ex = g.explore_bfs(root, max_levels=64, stride=1, reps=1)
    # levels the engine runs: a frontier exists and something is unvisited
    live = [l for l in range(lv.max() + 1) if bfs_counts(h, lv, l)[1] > 0]
    assert list(ex["level"]) == live and len(live) >= 3
    assert np.all(ex["ms"] > 0) and np.isfinite(ex["ms"]).all()
    assert np.all((ex["rule"] >= 0) & (ex["rule"] < 30))
    for k, l in enumerate(ex["level"]):
        a, i, pw, lw, md = bfs_counts(h, lv, l)
        c = (h.n, h.m, h.degrees.mean(), h.degrees.std(ddof=1),
             gsw.degree_stats(h.row_ptr)["GI"], l, a, i, pw, lw, md)
        np.testing.assert_allclose(ex["features"][k], gsw.mlp_features(c),
                                   rtol=1e-5, atol=1e-6)
    # the probes leave a finished BFS behind: levels still bit-exact after
    assert np.array_equal(g.bfs(root), lv)


def test_explore_stride(gsw, orc):
    h = gsw.grid2d(30, 30, 0.0, 0.0, 3)
    g = gsw.Graph(h)
    ex = g.explore_bfs(0, max_levels=8, stride=5, reps=1)
    assert list(ex["level"]) == [0, 5, 10, 15, 20, 25, 30, 35]


GRAPHS = [("rmat", lambda gsw: gsw.rmat(13, 16, seed=2)),
          ("grid", lambda gsw: gsw.grid2d(60, 60, 0.1, 0.01, 3)),
          ("ba", lambda gsw: gsw.barabasi_albert(4000, 4, 6)),
          ("er", lambda gsw: gsw.erdos_renyi(5000, 3.0, 5))]


@pytest.mark.parametrize("name,make", GRAPHS)
def test_learned_policy_bit_exact(gsw, orc, name, make):
    h = make(gsw)
    g = gsw.Graph(h)
    root = h.max_degree_vertex()
    exp = orc.bfs(h.row_ptr, h.col, root)
    seen = set()
    for seed in range(6):  # random networks visit many classes and switches
        w = np.random.default_rng(seed).normal(0, 0.5, NP).astype(np.float32)
        gsw.selector_load(w)
        assert np.array_equal(g.bfs(root, selector=1), exp), seed
        for r in g.records():
            seen.add((r["dir"], r["fmt"], r["lb"], r["fused"]))
    assert len(seen) >= 3
    # multi-root BFS under the learned policy
    roots = np.array([root, 0, h.n - 1], np.int32)
    ref = None
    for r in np.unique(roots):
        lv = orc.bfs(h.row_ptr, h.col, int(r))
        lv = np.where(lv < 0, 1 << 30, lv)
        ref = lv if ref is None else np.minimum(ref, lv)
    ref = np.where(ref == 1 << 30, -1, ref)
    assert np.array_equal(g.bfs_multi(roots, selector=1), ref)


def test_learned_policy_errors(gsw):
    h = gsw.rmat(10, 8, seed=3, weighted=True)
    g = gsw.Graph(h)
    with pytest.raises(gsw.GswError):  # no such selector
        g.cc(selector=4)
    with pytest.raises(gsw.GswError):  # no such selector
        g.pagerank(0.85, 3, selector=3)
    with pytest.raises(gsw.GswError):  # no such pull kernel
        g.pagerank(0.85, 3, pull_kernel=3)
    with pytest.raises(ValueError):
        gsw.selector_load(np.zeros(10, np.float32))
    with pytest.raises(ValueError):
        gsw.selector_load(np.zeros(NP, np.float32), app="pagerank")
    bad = np.zeros(NP, np.float32)
    assert gsw._lib.gsw_selector_load(bad.ctypes.data, 10) == 1  # GSW_E_INVALID
    assert gsw._lib.gsw_selector_load_app(4, bad.ctypes.data, NP) == 1
    bad[7] = np.nan
    assert gsw._lib.gsw_selector_load(bad.ctypes.data, NP) == 1


SSSP_GRAPHS = [("rmat", lambda gsw: gsw.rmat(12, 16, seed=2, weighted=True)),
               ("grid", lambda gsw: gsw.grid2d(50, 50, 0.1, 0.01, 3, weighted=True)),
               ("er", lambda gsw: gsw.erdos_renyi(4000, 4.0, 8, weighted=True))]


@pytest.mark.parametrize("name,make", SSSP_GRAPHS)
def test_learned_sssp_bit_exact(gsw, orc, name, make):
    h = make(gsw)
    g = gsw.Graph(h)
    root = h.max_degree_vertex()
    exp = orc.spfa(h.row_ptr, h.col, h.w, root)
    seen = set()
    for seed in range(6):  # random networks visit many classes and switches
        w = np.random.default_rng(100 + seed).normal(0, 0.5, NP).astype(np.float32)
        gsw.selector_load(w, app="sssp")
        assert np.array_equal(g.sssp(root, selector=1), exp), seed
        for r in g.records():
            seen.add((r["dir"], r["fmt"], r["lb"], r["fused"]))
    assert len(seen) >= 3


def test_explore_sssp(gsw, orc):
    h = gsw.rmat(11, 16, seed=2, weighted=True)
    g = gsw.Graph(h)
    root = h.max_degree_vertex()
    ex = g.explore_sssp(root, max_levels=64, stride=1, reps=1)
    L = len(ex["level"])
    assert L >= 3 and np.all(np.diff(ex["level"]) > 0) and ex["level"][0] == 0
    assert np.all(ex["ms"] > 0) and np.isfinite(ex["ms"]).all()
    assert np.all((ex["rule"] >= 0) & (ex["rule"] < 30))
    assert np.isfinite(ex["features"]).all()
    # level 0 is the root alone: active = 1, push workload = its degree
    c = (h.n, h.m, h.degrees.mean(), h.degrees.std(ddof=1),
         gsw.degree_stats(h.row_ptr)["GI"], 0, 1, h.n - 1, h.degrees[root],
         h.m - h.degrees[root], h.degrees[root])
    x0 = gsw.mlp_features(c)
    for j in (0, 1, 2, 3, 4, 5, 6, 8, 10, 11):
        np.testing.assert_allclose(ex["features"][0][j], x0[j], rtol=1e-5, atol=1e-6)
    # the probes leave the graph usable: a later run is still bit-exact
    assert np.array_equal(g.sssp(root), orc.spfa(h.row_ptr, h.col, h.w, root))


def test_trained_sssp_model_file(gsw, orc):
    from gswitch_amd import learn
    path = learn.model_path("sssp")
    if not os.path.exists(path):
        pytest.skip("no trained SSSP selector committed")
    gsw.selector_load(learn.load_model(path), app="sssp")
    for h in (gsw.rmat(14, 16, seed=11, weighted=True),
              gsw.grid2d(120, 120, 0.1, 0.01, 4, weighted=True)):
        g = gsw.Graph(h)
        root = h.max_degree_vertex()
        assert np.array_equal(g.sssp(root, selector=1),
                              orc.spfa(h.row_ptr, h.col, h.w, root))


def test_trained_model_file(gsw, orc):
    from gswitch_amd import learn
    if not os.path.exists(learn.MODEL_PATH):
        pytest.skip("no trained selector committed")
    w = learn.load_model()
    gsw.selector_load(w)
    h = gsw.rmat(14, 16, seed=11)
    g = gsw.Graph(h)
    root = h.max_degree_vertex()
    assert np.array_equal(g.bfs(root, selector=1), orc.bfs(h.row_ptr, h.col, root))


PR_GRAPHS = [("rmat", lambda gsw: gsw.rmat(12, 16, seed=5)),
             ("grid", lambda gsw: gsw.grid2d(60, 60, 0.1, 0.01, 3)),
             ("er", lambda gsw: gsw.erdos_renyi(5000, 6.0, 9))]


def pr_max_rel(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b) / np.abs(b)))


@pytest.mark.parametrize("name,make", PR_GRAPHS)
def test_learned_pagerank(gsw, orc, name, make):
    """PageRank under a learned network picks one from-all pull class for
    the whole run: the class the network gives the host-side features of
    the first level (every vertex active), and the ranks stay within 1e-5
    of the FP64 oracle whatever class that is."""
    from gswitch_amd import learn
    h = make(gsw)
    g = gsw.Graph(h)
    ref = orc.pagerank(h.row_ptr, h.col, 0.85, 30)
    x = learn.pagerank_features(h)
    seen = set()
    for seed in range(12):
        w = learn.init_weights(seed)
        w[4992:6912] *= 30.0  # spread the outputs so that seeds differ
        want = int(gsw.mlp_select(w, x, mask=gsw.PR_CLASS_MASK)[0])
        gsw.selector_load(w, app="pr")
        pr = g.pagerank(0.85, 30, selector=1)
        recs = g.records()
        got = {gsw.pr_class_of_record(r) for r in recs}
        assert got == {want} and len(recs) > 0, (seed, got, want)
        assert g.last_class() == want
        assert pr_max_rel(pr, ref) <= 1e-5, (seed, want)
        seen.add(want)
    assert len(seen) >= 2


@pytest.mark.parametrize("cls", range(8))
def test_pagerank_class_tree(gsw, orc, cls):
    """A one-leaf class tree pins every PageRank class (selector 2)."""
    h = gsw.rmat(11, 16, seed=6)
    g = gsw.Graph(h)
    gsw.selector_load_tree(dict(feat=[-1], thr=[0.0], left=[-1], right=[-1],
                                cls=[cls]), app="pr")
    pr = g.pagerank(0.85, 20, selector=2)
    assert {gsw.pr_class_of_record(r) for r in g.records()} == {cls}
    assert pr_max_rel(pr, orc.pagerank(h.row_ptr, h.col, 0.85, 20)) <= 1e-5


def test_trained_pagerank_model_file(gsw, orc):
    from gswitch_amd import learn
    path = learn.model_path("pr")
    if not os.path.exists(path):
        pytest.skip("no trained PageRank selector committed")
    w = learn.load_model(path)
    gsw.selector_load(w, app="pr")
    for h in (gsw.rmat(14, 16, seed=11), gsw.grid2d(120, 120, 0.1, 0.01, 4)):
        g = gsw.Graph(h)
        pr = g.pagerank(0.85, 50, selector=1)
        want = int(gsw.mlp_select(w, learn.pagerank_features(h), mask=gsw.PR_CLASS_MASK)[0])
        assert {gsw.pr_class_of_record(r) for r in g.records()} == {want}
        assert pr_max_rel(pr, orc.pagerank(h.row_ptr, h.col, 0.85, 50)) <= 1e-5


CC_GRAPHS = [("rmat", lambda gsw: gsw.rmat(12, 8, seed=7)),
             ("grid", lambda gsw: gsw.grid2d(70, 70, 0.3, 0.0, 5)),
             ("er", lambda gsw: gsw.erdos_renyi(6000, 1.5, 4))]


@pytest.mark.parametrize("name,make", CC_GRAPHS)
def test_learned_cc(gsw, orc, name, make):
    """CC under a learned network runs the class the network gives the
    first-level features (sampled hooking or one of the edge-centric
    loops); the labels are bit-exact against the oracle whatever it is."""
    from gswitch_amd import learn
    h = make(gsw)
    g = gsw.Graph(h)
    ref = orc.cc(h.row_ptr, h.col)
    x = learn.pagerank_features(h)
    seen = set()
    for seed in range(12):
        w = learn.init_weights(seed + 100)
        w[4992:6912] *= 30.0
        want = int(gsw.mlp_select(w, x, mask=gsw.CC_CLASS_MASK)[0])
        gsw.selector_load(w, app="cc")
        assert np.array_equal(g.cc(selector=1), ref), (seed, want)
        assert g.last_class() == want
        seen.add(want)
    assert len(seen) >= 2
    # forced configs take precedence over the selector
    assert np.array_equal(g.cc(configs="x-Queue-x-x", selector=1), ref)
    assert g.last_class() == -1


@pytest.mark.parametrize("cls", range(4))
def test_cc_class_tree(gsw, orc, cls):
    h = gsw.grid2d(40, 40, 0.25, 0.01, 8)
    g = gsw.Graph(h)
    gsw.selector_load_tree(dict(feat=[-1], thr=[0.0], left=[-1], right=[-1],
                                cls=[cls]), app="cc")
    assert np.array_equal(g.cc(selector=2), orc.cc(h.row_ptr, h.col))
    assert g.last_class() == cls


def test_trained_cc_model_file(gsw, orc):
    from gswitch_amd import learn
    path = learn.model_path("cc")
    if not os.path.exists(path):
        pytest.skip("no trained CC selector committed")
    w = learn.load_model(path)
    gsw.selector_load(w, app="cc")
    for h in (gsw.rmat(14, 16, seed=11), gsw.grid2d(120, 120, 0.1, 0.01, 4)):
        g = gsw.Graph(h)
        assert np.array_equal(g.cc(selector=1), orc.cc(h.row_ptr, h.col))
        want = int(gsw.mlp_select(w, learn.pagerank_features(h), mask=gsw.CC_CLASS_MASK)[0])
        assert g.last_class() == want
