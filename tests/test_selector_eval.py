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
