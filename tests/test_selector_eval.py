"""What the trained selectors gain, from committed measurements (CPU only).

tests/golden/selector_eval_{bfs,sssp}.npz hold the exhaustive per-level
timings of all 30 variant classes on the 40 held-out graphs (scales 20-24)
that gswitch_amd.learn collected on an MI355X, with the class the rule trees
chose at each level in their own run. Each committed network
(gswitch_amd/model/selector_{bfs,sssp}.npz) is evaluated here with the FP64
oracle network: its choices must cost less than the rule trees' choices
summed over the held-out levels (the gain), and the regret against the
exhaustive optimum is reported in the assertion message."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module", params=["bfs", "sssp"])
def evald(request):
    path = os.path.join(HERE, "golden", "selector_eval_%s.npz" % request.param)
    if not os.path.exists(path):
        pytest.skip("no committed selector evaluation data")
    with np.load(path, allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    d["app"] = request.param
    return d


def test_eval_data_shape(evald):
    ms, f, r = evald["ms"], evald["features"], evald["rule"]
    assert ms.shape[1] == 30 and f.shape == (len(ms), 12) and len(r) == len(ms)
    assert np.all(ms > 0) and np.all((r >= 0) & (r < 30))
    assert len(np.unique(evald["graph"])) == 40


def test_learned_selector_beats_rule_on_held_out_levels(orc, evald):
    from gswitch_amd import learn
    w = learn.load_model(learn.model_path(evald["app"]))
    ms = evald["ms"].astype(np.float64)
    choice = orc.mlp_forward(w, evald["features"]).argmin(axis=1)
    best = ms.min(axis=1).sum()
    idx = np.arange(len(ms))
    t_mlp = ms[idx, choice].sum()
    t_rule = ms[idx, evald["rule"]].sum()
    msg = "%s per-level cost / optimum: learned %.3f, rule %.3f" % (
        evald["app"], t_mlp / best, t_rule / best)
    assert t_mlp < t_rule, msg
    assert t_mlp / best < 1.5, msg


def test_class_trees_beat_rule_on_held_out_levels(evald):
    """The committed gfx950 class trees (selector 2), evaluated with the
    same float32 comparisons as the engine: cheaper than the reference's
    rule trees over the held-out levels."""
    from gswitch_amd import learn
    tree = learn.load_tree(learn.tree_path(evald["app"]))
    ms = evald["ms"].astype(np.float64)
    idx = np.arange(len(ms))
    best = ms.min(axis=1).sum()
    t_tree = ms[idx, learn.tree_select(tree, evald["features"])].sum()
    t_rule = ms[idx, evald["rule"]].sum()
    msg = "%s per-level cost / optimum: class tree %.3f, rule %.3f" % (
        evald["app"], t_tree / best, t_rule / best)
    assert t_tree < t_rule, msg


def test_class_trees_refit_from_committed_data(evald):
    """The committed trees are what fit_tree gives on the committed training
    timings (the pipeline is deterministic)."""
    from gswitch_amd import learn
    path = os.path.join(os.path.dirname(learn.model_path()), "selector_%s_train.npz"
                        % evald["app"])
    with np.load(path, allow_pickle=False) as z:
        tree = learn.fit_tree(z["features"], z["ms"])
    committed = learn.load_tree(learn.tree_path(evald["app"]))
    for k in ("feat", "left", "right", "cls"):
        assert tree[k] == committed[k], k
    np.testing.assert_allclose(tree["thr"], committed["thr"], rtol=0, atol=0)
