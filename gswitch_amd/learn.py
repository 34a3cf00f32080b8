"""Training pipeline of the learned variant selectors (BFS, SSSP and
PageRank, one network each).

1. `collect`: the engine times all 30 variant classes at every probed level
   of seeded synthetic graphs (gsw_explore_bfs / gsw_explore_sssp: snapshot,
   run each class, restore) and records the level's 12 features and the
   class the rule trees would pick.
2. `train`: targets are log2(time / fastest time) per class, clipped at
   CLIP; the 12 -> 64 -> 64 -> 30 network is fitted by the library's own
   HIP forward / backward / Adam kernels (gsw_mlp_train).
3. `evaluate`: on held-out graphs, per-level regret of the network's and
   the rule trees' choices against the exhaustive optimum, and end-to-end
   time under the rule trees, the network, and fixed strategies.

4. `fit-trees` (CPU): a class tree per application -- the reference's
   offline-trained decision trees (src/model/*.h, README.md:13) retrained
   for this engine -- fitted to the same per-level timings: every split
   and every leaf class is chosen to minimise the summed time of the
   levels it covers (selector 2 of the engine).

CC (--app cc) also chooses once per run, from the same first-level
features, among sampled hooking (its default) and the edge-centric level
loop with the rule's, bitmap or queue frontiers (gswitch_amd.CC_CLASSES);
its "rule" baseline is the default run.

PageRank (--app pr) pulls every vertex the same way in every iteration,
so it chooses once per run: a sample is one graph, its features are those
of the run's first level (everything active) and its timings are the
device time per iteration of each of the 8 from-all pull classes
(gswitch_amd.PR_CLASSES: TM, WM, CM, STRICT as a merge-path walk or a
segmented reduction, TWOD as length-binned slices, segmented slices or
propagation blocking); the network's other outputs are masked out.

Run as a module on a GPU machine (--app bfs | sssp | pr, default bfs):
    python -m gswitch_amd.learn collect --app A --set train --out DIR
    python -m gswitch_amd.learn train --app A --data DIR/train.npz --out DIR
    python -m gswitch_amd.learn collect --app A --set eval --out DIR
    python -m gswitch_amd.learn evaluate --app A --data DIR/eval.npz --model M --out DIR
    python -m gswitch_amd.learn evaluate --app pr|cc --model M --out DIR
        (PageRank, CC: the held-out graphs are timed per class during the
        end-to-end pass)
    python -m gswitch_amd.learn fit-trees --app A --data DIR/train.npz --out DIR
    python -m gswitch_amd.learn e2e --app A --out DIR      (held-out graphs end to end:
        rule trees, committed network, committed class tree)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

import gswitch_amd as gsw

HERE = os.path.dirname(os.path.abspath(__file__))
APPS = ("bfs", "sssp", "pr", "cc")
# classes each application's selector may pick
MASK = {"bfs": (1 << 30) - 1, "sssp": (1 << 30) - 1, "pr": gsw.PR_CLASS_MASK,
        "cc": gsw.CC_CLASS_MASK}
PR_ITERS = 3          # iterations per timed PageRank run while collecting
PR_SLOW = 50.0        # a class this many times slower than the best is timed once


def model_path(app="bfs"):
    return os.path.join(HERE, "model", "selector_%s.npz" % app)


def tree_path(app="bfs"):
    return os.path.join(HERE, "model", "trees_gfx950_%s.json" % app)


MODEL_PATH = model_path("bfs")
CLIP = 6.0
TRAIN_SEED, EVAL_SEED = 20240501, 20240777


def _u(rng, lo, hi):
    return float(rng.uniform(lo, hi))


def training_specs(count=300, seed=TRAIN_SEED):
    """R-MAT scales 14-22 (a in [0.45, 0.65], b = c = (1 - a - 0.05) / 2),
    Barabasi-Albert (m in [2, 16]), Erdos-Renyi and perturbed 2-D grids."""
    rng = np.random.default_rng(seed)
    kinds = ["rmat"] * 120 + ["ba"] * 60 + ["er"] * 60 + ["grid"] * 60
    kinds = [kinds[i] for i in rng.permutation(len(kinds))][:count]
    out = []
    for i, k in enumerate(kinds):
        s = dict(kind=k, seed=int(rng.integers(1, 1 << 30)),
                 root="max" if i % 2 == 0 else "random")
        if k == "rmat":
            a = _u(rng, 0.45, 0.65)
            s.update(scale=int(rng.integers(14, 23)), ef=int(rng.choice([8, 16])),
                     a=a, b=(1 - a - 0.05) / 2)
        elif k == "ba":
            s.update(n=int(2 ** rng.integers(14, 21)), m=int(rng.integers(2, 17)))
        elif k == "er":
            s.update(n=int(2 ** rng.integers(14, 22)), d=_u(rng, 2, 32))
        else:
            s.update(rows=int(rng.integers(64, 1025)), cols=int(rng.integers(64, 1025)),
                     p=_u(rng, 0.0, 0.3), lr=_u(rng, 0.0, 0.02))
        out.append(s)
    return out


def eval_specs(seed=EVAL_SEED):
    """40 held-out graphs at scales 20-24."""
    rng = np.random.default_rng(seed)
    out = []
    for scale in [20] * 4 + [21] * 4 + [22] * 4 + [23] * 2 + [24] * 2:
        a = _u(rng, 0.5, 0.62)
        out.append(dict(kind="rmat", scale=scale, ef=16, a=a, b=(1 - a - 0.05) / 2,
                        seed=int(rng.integers(1, 1 << 30)), root="max"))
    for _ in range(8):
        out.append(dict(kind="er", n=int(2 ** rng.integers(20, 23)), d=_u(rng, 4, 32),
                        seed=int(rng.integers(1, 1 << 30)), root="random"))
    for _ in range(8):
        out.append(dict(kind="ba", n=int(2 ** rng.integers(20, 22)),
                        m=int(rng.integers(4, 17)),
                        seed=int(rng.integers(1, 1 << 30)), root="max"))
    for _ in range(8):
        side = int(rng.integers(1024, 2049))
        out.append(dict(kind="grid", rows=side, cols=side, p=_u(rng, 0.0, 0.2),
                        lr=_u(rng, 0.0, 0.01), seed=int(rng.integers(1, 1 << 30)),
                        root="random"))
    return out


def build(spec, weighted=False):
    k = spec["kind"]
    if k == "rmat":
        return gsw.rmat(spec["scale"], spec["ef"], spec["a"], spec["b"], spec["b"],
                        seed=spec["seed"], weighted=weighted)
    if k == "ba":
        return gsw.barabasi_albert(spec["n"], spec["m"], seed=spec["seed"],
                                   weighted=weighted)
    if k == "er":
        return gsw.erdos_renyi(spec["n"], spec["d"], seed=spec["seed"],
                               weighted=weighted)
    return gsw.grid2d(spec["rows"], spec["cols"], spec["p"], spec["lr"],
                      seed=spec["seed"], weighted=weighted)


def run(g, app, root, **kw):
    """One traversal of `app` from `root` (levels or distances)."""
    return g.bfs(root, **kw) if app == "bfs" else g.sssp(root, **kw)


def pick_root(spec, h):
    if spec["root"] == "max":
        return h.max_degree_vertex()
    deg = h.degrees
    nz = np.flatnonzero(deg > 0)
    rng = np.random.default_rng(spec["seed"] ^ 0x5EED)
    return int(nz[rng.integers(len(nz))]) if len(nz) else 0


def level_stride(depth, max_levels):
    """Probe every level of shallow traversals, and about max_levels evenly
    spaced levels of deep ones (high-diameter grids)."""
    return max(1, (depth + max_levels - 1) // max_levels)


def class_of_record(r):
    """Variant class of one expand record of the engine."""
    fmt = r["fmt"]
    if r["dir"] == "Push" and fmt == "Queue" and r["sorted"] and not r["fused"]:
        fmt = "SortedQueue"
    lb = r["lb"] if r["lb"] in gsw.MLP_LBS else "STRICT"
    if r["fused"]:
        name = "Push-Queue-%s-Fused" % lb
    elif r["dir"] == "Pull":
        name = "Pull-%s-%s-Standalone" % ("Bitmap" if fmt == "Bitmap" else "Queue", lb)
    else:
        name = "Push-%s-%s-Standalone" % (fmt, lb)
    return gsw.MLP_CLASSES.index(name)


def rule_trajectory(g, root, app="bfs"):
    """The rule selector's actual class per level (fused levels included),
    keyed by the level of the expanded frontier."""
    run(g, app, root, out=False, profile=True)
    return {r["level"] - 1: class_of_record(r) for r in g.records()}


def pagerank_features(h):
    """The 12 inputs of a PageRank (or CC) run's first level (every vertex
    active, push workload m), as the engine computes them for its choice."""
    st = gsw.degree_stats(h.row_ptr)
    n, m = float(h.n), float(h.m)
    return gsw.mlp_features([n, m, st["avg"], st["std"], st["GI"], 0.0, n, 0.0, m,
                             0.0, m])


def time_pr_classes(g, reps=2, iters=PR_ITERS, slow=PR_SLOW):
    """Device ms per PageRank iteration of each class (PR_CLASSES order).
    Each class runs one warm-up iteration (which also builds its layout)
    and then `reps` runs of `iters` iterations; a class whose warm-up is
    `slow` times slower than the best time so far keeps the warm-up time."""
    t = np.full(len(gsw.PR_CLASSES), np.inf)
    for c in (7, 5, 6, 3, 4, 2, 1, 0):
        cfg, pk = gsw.PR_CLASSES[c]
        g.pagerank(0.85, 1, configs=cfg, pull_kernel=pk, out=False, profile=False)
        w1 = g.last_stats.device_ms
        best = t[np.isfinite(t)].min() if np.isfinite(t).any() else np.inf
        if w1 > slow * best:
            t[c] = w1
            continue
        for _ in range(reps):
            g.pagerank(0.85, iters, configs=cfg, pull_kernel=pk, out=False,
                       profile=False)
            t[c] = min(t[c], g.last_stats.device_ms / iters)
    return t


def collect_pr(specs, reps=2, log=print):
    rows = dict(features=[], ms=[], rule=[], level=[], graph=[])
    t0 = time.time()
    for gi, spec in enumerate(specs):
        h = build(spec)
        g = gsw.Graph(h)
        g.pagerank(0.85, 1, out=False)
        rule = gsw.pr_class_of_record(g.records()[0])
        ms = time_pr_classes(g, reps)
        g.close()
        rows["features"].append(pagerank_features(h)[None])
        rows["ms"].append(ms[None].astype(np.float32))
        rows["rule"].append(np.array([rule], np.int32))
        rows["level"].append(np.zeros(1, np.int32))
        rows["graph"].append(np.full(1, gi, np.int32))
        log("graph %d/%d %s n=%d m=%d rule %d best %d (%.3f ms/it) (%.0fs)" % (
            gi + 1, len(specs), spec["kind"], h.n, h.m, rule, int(np.argmin(ms)),
            float(ms.min()), time.time() - t0))
    return {k: np.concatenate(v) for k, v in rows.items()}


def time_cc_classes(g, reps=3):
    """Device ms of a CC run under each class (CC_CLASSES order), best of
    `reps` after a warm-up, and the labels of the last class (all equal)."""
    t = np.full(len(gsw.CC_CLASSES), np.inf)
    ref = None
    for c, cfg in enumerate(gsw.CC_CLASSES):
        lab = g.cc(configs=cfg, profile=False)
        if ref is None:
            ref = lab
        assert np.array_equal(lab, ref), c
        for _ in range(reps):
            g.cc(configs=cfg, out=False, profile=False)
            t[c] = min(t[c], g.last_stats.device_ms)
    return t, ref


def collect_cc(specs, reps=3, log=print):
    rows = dict(features=[], ms=[], rule=[], level=[], graph=[])
    t0 = time.time()
    for gi, spec in enumerate(specs):
        h = build(spec)
        g = gsw.Graph(h)
        ms, _ = time_cc_classes(g, reps)
        g.close()
        rows["features"].append(pagerank_features(h)[None])
        rows["ms"].append(ms[None].astype(np.float32))
        rows["rule"].append(np.zeros(1, np.int32))  # the default run
        rows["level"].append(np.zeros(1, np.int32))
        rows["graph"].append(np.full(1, gi, np.int32))
        log("graph %d/%d %s n=%d m=%d best %d (%s ms) (%.0fs)" % (
            gi + 1, len(specs), spec["kind"], h.n, h.m, int(np.argmin(ms)),
            " ".join("%.3f" % x for x in ms), time.time() - t0))
    return {k: np.concatenate(v) for k, v in rows.items()}


def collect(specs, reps=2, max_levels=32, log=print, app="bfs"):
    if app == "pr":
        return collect_pr(specs, reps, log)
    if app == "cc":
        return collect_cc(specs, max(reps, 3), log)
    rows = dict(features=[], ms=[], rule=[], level=[], graph=[])
    t0 = time.time()
    for gi, spec in enumerate(specs):
        h = build(spec, weighted=(app == "sssp"))
        g = gsw.Graph(h)
        root = pick_root(spec, h)
        if app == "bfs":
            depth = int(g.bfs(root).max()) + 1
        else:
            g.sssp(root, out=False)
            depth = int(g.last_stats.levels) + 1
        traj = rule_trajectory(g, root, app)
        explore = g.explore_bfs if app == "bfs" else g.explore_sssp
        ex = explore(root, max_levels=max_levels,
                     stride=level_stride(depth, max_levels), reps=reps)
        # the rule's class in its own run (fused state included) where known
        ex["rule"] = np.array([traj.get(int(l), int(c)) for l, c in
                               zip(ex["level"], ex["rule"])], np.int32)
        g.close()
        for key in ["features", "ms", "rule", "level"]:
            rows[key].append(ex[key])
        rows["graph"].append(np.full(len(ex["level"]), gi, np.int32))
        log("graph %d/%d %s n=%d m=%d levels=%d (%.0fs)" % (
            gi + 1, len(specs), spec["kind"], h.n, h.m, len(ex["level"]),
            time.time() - t0))
    return {k: np.concatenate(v) for k, v in rows.items()}


def targets(ms):
    """Regression targets and mask over the network's 30 outputs; an
    application with fewer classes leaves the rest masked out."""
    best = ms.min(axis=1, keepdims=True)
    y = np.minimum(np.log2(np.maximum(ms, 1e-6) / np.maximum(best, 1e-6)), CLIP)
    k = ms.shape[1]
    yt = np.zeros((len(ms), gsw.MLP_NOUT), np.float32)
    mk = np.zeros((len(ms), gsw.MLP_NOUT), np.float32)
    yt[:, :k] = y
    mk[:, :k] = 1.0
    return yt, mk


def init_weights(seed=1):
    """He-normal hidden layers, small output layer, zero biases."""
    rng = np.random.default_rng(seed)
    w = np.zeros(gsw.MLP_NPARAM, np.float32)
    w[0:768] = rng.normal(0, np.sqrt(2 / 12), 768)
    w[832:4928] = rng.normal(0, np.sqrt(2 / 64), 4096)
    w[4992:6912] = rng.normal(0, 0.1 * np.sqrt(1 / 64), 1920)
    w[6912:] = 1.0
    return w


def train(data, epochs=400, batch=128, lr=2e-3, seed=7):
    y, m = targets(data["ms"])
    w, curve = gsw.mlp_train(init_weights(), data["features"], y, m, epochs=epochs,
                             batch=batch, lr=lr, seed=seed)
    return w, curve


def per_level_regret(ms, choice):
    """sum over levels of the chosen class's time / the optimum's."""
    idx = np.arange(len(ms))
    return float(ms[idx, choice].sum() / ms.min(axis=1).sum())


# ------------------------------------------------------------ class trees
TREE_DEPTH = 3
TREE_MIN_LEAF = 16


def _leaf_class(ms, valid):
    tot = np.where(valid, ms.sum(axis=0), np.inf)
    c = int(np.argmin(tot))
    return c, float(tot[c])


def fit_tree(X, ms, valid=None, depth=TREE_DEPTH, min_leaf=TREE_MIN_LEAF):
    """Greedy cost tree: a leaf takes the class with the least summed time
    over its levels; a split (feature, threshold halfway between adjacent
    distinct values, >= min_leaf levels a side) is kept when it lowers the
    summed time of its two leaves. Returns the node table (fit order:
    children after their parent)."""
    X = np.asarray(X, np.float32)
    ms = np.asarray(ms, np.float64)
    if valid is None:
        valid = np.isfinite(ms).all(axis=0)
    nodes = []

    def grow(idx, d):
        i = len(nodes)
        c, cost = _leaf_class(ms[idx], valid)
        nodes.append([-1, 0.0, -1, -1, c])
        if d == 0 or len(idx) < 2 * min_leaf:
            return i
        best = None
        for f in range(X.shape[1]):
            o = idx[np.argsort(X[idx, f], kind="stable")]
            xs = X[o, f]
            pre = np.cumsum(ms[o], axis=0)
            k = np.arange(min_leaf, len(o) - min_leaf + 1)
            k = k[xs[k - 1] < xs[k]]
            if len(k) == 0:
                continue
            lc = np.where(valid, pre[k - 1], np.inf).min(axis=1)
            rc = np.where(valid, pre[-1] - pre[k - 1], np.inf).min(axis=1)
            j = int(np.argmin(lc + rc))
            if best is None or lc[j] + rc[j] < best[0]:
                thr = np.float32(0.5 * (np.float64(xs[k[j] - 1]) + np.float64(xs[k[j]])))
                # the threshold must separate the two values in float32
                if not (xs[k[j] - 1] <= thr < xs[k[j]]):
                    thr = xs[k[j] - 1]
                best = (float(lc[j] + rc[j]), f, float(thr))
        if best is None or best[0] >= cost * (1 - 1e-9):
            return i
        _, f, thr = best
        go_left = X[idx, f] <= np.float32(thr)
        nodes[i][0], nodes[i][1] = f, thr
        nodes[i][2] = grow(idx[go_left], d - 1)
        nodes[i][3] = grow(idx[~go_left], d - 1)
        return i

    grow(np.arange(len(ms)), depth)
    return {"feat": [int(r[0]) for r in nodes], "thr": [float(r[1]) for r in nodes],
            "left": [int(r[2]) for r in nodes], "right": [int(r[3]) for r in nodes],
            "cls": [int(r[4]) for r in nodes]}


def tree_select(tree, X):
    """Class of every row of X under a node table (float32 comparisons, as
    the engine evaluates it)."""
    X = np.asarray(X, np.float32)
    feat, thr = tree["feat"], np.asarray(tree["thr"], np.float32)
    out = np.empty(len(X), np.int32)
    for r, x in enumerate(X):
        i = 0
        while feat[i] >= 0:
            i = tree["left"][i] if x[feat[i]] <= thr[i] else tree["right"][i]
        out[r] = tree["cls"][i]
    return out


def save_tree(path, tree, meta):
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "w") as f:
        json.dump(dict(tree, meta=meta), f, indent=1)


def load_tree(path):
    with open(path) as f:
        t = json.load(f)
    return {k: t[k] for k in ("feat", "thr", "left", "right", "cls")}


def save_model(path, w, meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, w=np.asarray(w, np.float32), meta=json.dumps(meta))


def load_model(path=MODEL_PATH):
    with np.load(path, allow_pickle=False) as z:
        w = z["w"].astype(np.float32)
    if w.size != gsw.MLP_NPARAM or not np.isfinite(w).all():
        raise ValueError("bad selector model")
    return w


FIXED = ["Push-Queue-WM-Standalone", "Push-Queue-WM-Fused", None]


def end_to_end_cc(specs, w, reps=3, log=print, tree=None):
    """Median device time of CC per graph under the default run (sampled
    hooking), the network and the class tree (when given), and the
    exhaustive per-class optimum; labels bit-exact against the default
    run's."""
    gsw.selector_load(w, app="cc")
    if tree is not None:
        gsw.selector_load_tree(tree, app="cc")
    out = []
    for spec in specs:
        h = build(spec)
        g = gsw.Graph(h)
        ref = g.cc()
        rec = dict(kind=spec["kind"], n=h.n, m=h.m)
        runs = [("rule", {}), ("mlp", dict(selector=1))]
        if tree is not None:
            runs.append(("tree", dict(selector=2)))
        for name, kw in runs:
            assert np.array_equal(g.cc(**kw), ref), (name, spec)
            rec[name + "_class"] = max(g.last_class(), 0) if kw else 0
            ts = []
            for _ in range(reps):
                g.cc(out=False, profile=False, **kw)
                ts.append(g.last_stats.device_ms)
            rec[name] = float(np.median(ts))
        per, lab = time_cc_classes(g)
        assert np.array_equal(lab, ref)
        rec["classes_ms"] = [float(x) for x in per]
        rec["best_class"] = int(np.argmin(per))
        g.close()
        log("%s n=%d default %.3f mlp %.3f (class %d) best class %d: %.3f ms" % (
            spec["kind"], h.n, rec["rule"], rec["mlp"], rec["mlp_class"],
            rec["best_class"], per.min()))
        out.append(rec)
    return out


def end_to_end_pr(specs, w, iters=20, reps=3, log=print, tree=None):
    """Median device time of a PageRank run of `iters` iterations per graph
    under the rule trees, the network and the class tree (when given), and
    the exhaustive per-class optimum (results within 1e-5 relative of the
    rule's)."""
    gsw.selector_load(w, app="pr")
    if tree is not None:
        gsw.selector_load_tree(tree, app="pr")
    out = []
    for spec in specs:
        h = build(spec)
        g = gsw.Graph(h)
        ref = g.pagerank(0.85, iters).astype(np.float64)
        rec = dict(kind=spec["kind"], n=h.n, m=h.m)
        runs = [("rule", {}), ("mlp", dict(selector=1))]
        if tree is not None:
            runs.append(("tree", dict(selector=2)))
        for name, kw in runs:
            pr = g.pagerank(0.85, iters, **kw)
            err = float(np.max(np.abs(pr - ref) / np.abs(ref))) if h.n else 0.0
            assert err <= 1e-5, (name, spec, err)
            rec[name + "_class"] = gsw.pr_class_of_record(g.records()[0])
            ts = []
            for _ in range(reps):
                g.pagerank(0.85, iters, out=False, profile=False, **kw)
                ts.append(g.last_stats.device_ms)
            rec[name] = float(np.median(ts))
        per_it = time_pr_classes(g)
        rec["classes_ms_per_iter"] = [float(x) for x in per_it]
        rec["best_class"] = int(np.argmin(per_it))
        g.close()
        log("%s n=%d rule %.3f (class %d) mlp %.3f (class %d) best class %d: %.4f ms/it" % (
            spec["kind"], h.n, rec["rule"], rec["rule_class"], rec["mlp"],
            rec["mlp_class"], rec["best_class"], per_it.min()))
        out.append(rec)
    return out


def end_to_end(specs, w, reps=3, log=print, app="bfs", tree=None):
    """Median device time per graph under the rule trees, the network, the
    class tree (when given) and fixed strategies (results checked
    identical)."""
    gsw.selector_load(w, app=app)
    if tree is not None:
        gsw.selector_load_tree(tree, app=app)
    out = []
    for spec in specs:
        h = build(spec, weighted=(app == "sssp"))
        g = gsw.Graph(h)
        root = pick_root(spec, h)
        ref = run(g, app, root)
        rec = dict(kind=spec["kind"], n=h.n, m=h.m)
        runs = [("rule", {}), ("mlp", dict(selector=1)),
                ("fixed_wm", dict(configs=FIXED[0])),
                ("fixed_fused", dict(configs=FIXED[1]))]
        if tree is not None:
            runs.insert(2, ("tree", dict(selector=2)))
        for name, kw in runs:
            ts = []
            for _ in range(reps):
                res = run(g, app, root, profile=False, **kw)
                assert np.array_equal(res, ref), (name, spec)
                ts.append(g.last_stats.device_ms)
            rec[name] = float(np.median(ts))
        g.close()
        log("%s n=%d rule %.3f mlp %.3f fixed %.3f / %.3f ms" % (
            spec["kind"], h.n, rec["rule"], rec["mlp"], rec["fixed_wm"],
            rec["fixed_fused"]))
        out.append(rec)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["collect", "train", "evaluate", "fit-trees", "e2e"])
    ap.add_argument("--app", default="bfs", choices=APPS)
    ap.add_argument("--set", default="train", choices=["train", "eval"])
    ap.add_argument("--count", type=int, default=300)
    ap.add_argument("--start", type=int, default=0,
                    help="collect specs[start:start+count] of the set")
    ap.add_argument("--data")
    ap.add_argument("--model", default=None,
                    help="model file (default: the committed one of --app)")
    ap.add_argument("--out", default="selector_out")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--epochs", type=int, default=400)
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    log = lambda s: print(s, flush=True)
    if a.cmd == "fit-trees":
        path = a.data or os.path.join(HERE, "model", "selector_%s_train.npz" % a.app)
        with np.load(path, allow_pickle=False) as z:
            d = {k: z[k] for k in z.files}
        tree = fit_tree(d["features"], d["ms"])
        ch = tree_select(tree, d["features"])
        log("train regret: tree %.3f rule %.3f (%d nodes)" % (
            per_level_regret(d["ms"], ch), per_level_regret(d["ms"], d["rule"]),
            len(tree["feat"])))
        save_tree(os.path.join(a.out, "trees_gfx950_%s.json" % a.app), tree,
                  dict(app=a.app, depth=TREE_DEPTH, min_leaf=TREE_MIN_LEAF,
                       samples=int(len(d["ms"])), data=os.path.basename(path)))
        return 0
    if a.cmd == "e2e":
        # held-out graphs end to end under the rule trees, the committed
        # network and the committed class tree
        w = load_model(a.model or model_path(a.app))
        tree = load_tree(tree_path(a.app))
        specs = eval_specs()
        if a.app == "pr":
            e2e = end_to_end_pr(specs, w, log=log, tree=tree)
        elif a.app == "cc":
            e2e = end_to_end_cc(specs, w, log=log, tree=tree)
        else:
            e2e = end_to_end(specs, w, log=log, app=a.app, tree=tree)
        gm = lambda k: float(np.exp(np.mean([np.log(r["rule"] / r[k]) for r in e2e])))
        res = dict(graphs=len(e2e), speedup_mlp_over_rule=gm("mlp"),
                   speedup_tree_over_rule=gm("tree"), end_to_end=e2e)
        if a.app in ("pr", "cc"):
            key = "classes_ms_per_iter" if a.app == "pr" else "classes_ms"
            ms = np.array([r[key] for r in e2e])
            for k in ("rule", "mlp", "tree"):
                res["regret_" + k] = per_level_regret(
                    ms, np.array([r[k + "_class"] for r in e2e]))
        log(json.dumps({k: v for k, v in res.items() if k != "end_to_end"}))
        with open(os.path.join(a.out, "e2e_%s.json" % a.app), "w") as f:
            json.dump(res, f, indent=1)
        return 0
    if a.cmd == "collect":
        specs = training_specs() if a.set == "train" else eval_specs()
        specs = specs[a.start:a.start + a.count]
        d = collect(specs, reps=a.reps, log=log, app=a.app)
        name = a.set if a.start == 0 else "%s_%d" % (a.set, a.start)
        np.savez(os.path.join(a.out, name + ".npz"), specs=json.dumps(specs), **d)
    elif a.cmd == "train":
        with np.load(a.data, allow_pickle=False) as z:
            d = {k: z[k] for k in z.files}
        w, curve = train(d, epochs=a.epochs)
        log("loss %.4f -> %.4f" % (curve[0], curve[-1]))
        cls = gsw.mlp_select(w, d["features"], mask=MASK[a.app])
        log("train regret: mlp %.3f rule %.3f" % (per_level_regret(d["ms"], cls),
                                                  per_level_regret(d["ms"], d["rule"])))
        save_model(os.path.join(a.out, "selector_%s.npz" % a.app), w,
                   dict(app=a.app, samples=int(len(d["ms"])), epochs=a.epochs,
                        loss_first=float(curve[0]), loss_last=float(curve[-1])))
    elif a.app in ("pr", "cc") and not a.data:
        # one choice per graph: the end-to-end pass times every class as well
        w = load_model(a.model or model_path(a.app))
        if a.app == "pr":
            e2e = end_to_end_pr(eval_specs(), w, log=log)
            ms = np.array([r["classes_ms_per_iter"] for r in e2e])
        else:
            e2e = end_to_end_cc(eval_specs(), w, log=log)
            ms = np.array([r["classes_ms"] for r in e2e])
        res = dict(graphs=len(e2e),
                   regret_mlp=per_level_regret(ms, np.array([r["mlp_class"] for r in e2e])),
                   regret_rule=per_level_regret(ms, np.array([r["rule_class"] for r in e2e])),
                   end_to_end=e2e)
        log("eval per-graph regret: mlp %.3f rule %.3f" % (res["regret_mlp"],
                                                           res["regret_rule"]))
        with open(os.path.join(a.out, "eval.json"), "w") as f:
            json.dump(res, f, indent=1)
    else:
        with np.load(a.data, allow_pickle=False) as z:
            d = {k: z[k] for k in z.files}
        w = load_model(a.model or model_path(a.app))
        cls = gsw.mlp_select(w, d["features"], mask=MASK[a.app])
        res = dict(levels=int(len(d["ms"])),
                   regret_mlp=per_level_regret(d["ms"], cls),
                   regret_rule=per_level_regret(d["ms"], d["rule"]))
        log("eval per-level regret: mlp %.3f rule %.3f" % (res["regret_mlp"],
                                                           res["regret_rule"]))
        specs = json.loads(str(d["specs"]))
        if a.app == "pr":
            res["end_to_end"] = end_to_end_pr(specs, w, log=log)
        else:
            res["end_to_end"] = end_to_end(specs, w, log=log, app=a.app)
        with open(os.path.join(a.out, "eval.json"), "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
