"""GPU parity at the full sizes of the benchmark configurations.

Each test builds the configuration's seeded graph and compares the HIP path
with the CPU oracle on it:
  * SSSP + 50-iteration PageRank on R-MAT scale 22, EF 16, seed 2 (uint32
    weights in [1, 64]): distances bit-exact, ranks within 1e-5 max
    relative error of the FP64 oracle;
  * the 4096 x 4096 grid (10 % of edges deleted, 1 % long-range, seed 3):
    BFS levels and CC labels bit-exact;
  * 64-source bitmask BFS on R-MAT scale 22 seed 4 from the 64
    highest-degree vertices: masks and per-source level sums bit-exact
    against 64 oracle BFS runs;
  * one whole 1000-instance piece of the 20,000-graph Erdos-Renyi batch
    (seed 5) through the autotuned captured BFS and the captured CC of
    gsw_batch_bfs_cc, every instance against its own oracle runs, plus
    seeded samples of instances from the rest of the batch;
  * held-out selector evaluation graphs (R-MAT scale 20, Erdos-Renyi, grid,
    Barabasi-Albert): BFS, SSSP, 50-iteration PageRank and CC under the rule
    trees and under the committed learned selectors.
The oracle runs of one test are spread over threads (the ctypes calls
release the GIL); each is still the single-threaded restated routine.
"""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

PR_TOL = 1e-5
FIXED = "Push-Queue-WM-Standalone"
WORKERS = 8


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def pmap(fn, items):
    with ThreadPoolExecutor(WORKERS) as ex:
        return list(ex.map(fn, items))


def max_rel(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b) / np.abs(b)))


def load_selectors(gsw):
    from gswitch_amd import learn
    have = {}
    for app in ("bfs", "sssp", "pr", "cc"):
        p = learn.model_path(app)
        if os.path.exists(p):
            gsw.selector_load(learn.load_model(p), app=app)
            have[app] = True
    return have


# ------------------------------------------------------------ C2: R-MAT 22
@pytest.fixture(scope="module")
def rmat22w(gsw):
    say("C2: building R-MAT 22 EF16 seed 2 (weighted)")
    h = gsw.rmat(22, 16, 0.57, 0.19, 0.19, seed=2, weighted=True)
    g = gsw.Graph(h)
    yield h, g
    g.close()


def test_c2_sssp_rmat22(gsw, orc, rmat22w):
    h, g = rmat22w
    root = h.max_degree_vertex()
    ref = orc.spfa(h.row_ptr, h.col, h.w, root)
    assert (ref != np.uint32(0xFFFFFFFF)).sum() > h.n // 2
    sel = load_selectors(gsw)
    runs = [dict(), dict(configs=FIXED), dict(configs="Push-Queue-STRICT-Fused"),
            dict(configs="Pull-Bitmap-STRICT-Standalone")]
    if sel.get("sssp"):
        runs.append(dict(selector=1))
    for kw in runs:
        say("C2 SSSP", kw)
        assert np.array_equal(g.sssp(root, **kw), ref), kw


def test_c2_pagerank50_rmat22(gsw, orc, rmat22w):
    h, g = rmat22w
    say("C2: FP64 PageRank oracle, 50 iterations")
    ref = orc.pagerank(h.row_ptr, h.col, 0.85, 50)
    for kw in [dict(), dict(pull_kernel=1), dict(pull_kernel=2),
               dict(configs="Pull-Bitmap-STRICT-Standalone"),
               dict(configs="Pull-Bitmap-TM-Standalone")]:
        say("C2 PageRank", kw)
        pr = g.pagerank(0.85, 50, **kw)
        err = max_rel(pr, ref)
        assert err <= PR_TOL, (kw, err)


# ------------------------------------------------------------ C3: grid 4096^2
def test_c3_grid4096_bfs_cc(gsw, orc):
    say("C3: building the 4096 x 4096 grid, seed 3")
    h = gsw.grid2d(4096, 4096, 0.1, 0.01, 3)
    g = gsw.Graph(h)
    ref = orc.bfs(h.row_ptr, h.col, 0)
    assert ref.max() > 40  # high diameter: tens of levels despite the long-range edges
    sel = load_selectors(gsw)
    runs = [dict(), dict(configs=FIXED), dict(configs="Push-Queue-WM-Fused"),
            dict(configs="Push-SortedQueue-TWC-Standalone")]
    if sel.get("bfs"):
        runs.append(dict(selector=1))
    for kw in runs:
        say("C3 BFS", kw)
        assert np.array_equal(g.bfs(0, **kw), ref), kw
    say("C3 CC")
    assert np.array_equal(g.cc(), orc.cc(h.row_ptr, h.col))
    g.close()


# ------------------------------------------------------------ C4a: 64 sources
def test_c4a_msbfs64_rmat22(gsw, orc):
    say("C4a: building R-MAT 22 EF16 seed 4")
    h = gsw.rmat(22, 16, 0.57, 0.19, 0.19, seed=4)
    g = gsw.Graph(h)
    src = np.argsort(-h.degrees, kind="stable")[:64].astype(np.int32)
    say("C4a: 64 oracle BFS runs")
    levels = pmap(lambda r: orc.bfs(h.row_ptr, h.col, int(r)), src)
    em = np.zeros(h.n, np.uint64)
    es = np.zeros(64, np.uint64)
    for s, lv in enumerate(levels):
        em[lv >= 0] |= np.uint64(1) << np.uint64(s)
        es[s] = lv[lv >= 0].astype(np.int64).sum()
    del levels
    for direction in [None, "Push-x-x-x", "Pull-x-x-x"]:
        say("C4a msbfs64", direction)
        masks, sums = g.msbfs64(src, configs=direction)
        assert np.array_equal(masks, em), direction
        assert np.array_equal(sums, es), direction
    g.close()


# ------------------------------------------------------------ C4b: ER batch
def instance(h, off, i):
    lo, hi = int(off[i]), int(off[i + 1])
    rp = h.row_ptr[lo:hi + 1] - h.row_ptr[lo]
    col = (h.col[h.row_ptr[lo]:h.row_ptr[hi]] - lo).astype(np.int32)
    return rp, col


def check_instances(orc, h, off, lv, lab):
    def one(i):
        rp, col = instance(h, off, i)
        lo, hi = int(off[i]), int(off[i + 1])
        ok_bfs = np.array_equal(lv[lo:hi], orc.bfs(rp, col, 0))
        ok_cc = np.array_equal(lab[lo:hi] - lo, orc.cc(rp, col))
        return ok_bfs, ok_cc
    res = pmap(one, range(len(off) - 1))
    bad = [i for i, (a, b) in enumerate(res) if not (a and b)]
    assert not bad, "instances differ from the oracle: %s" % bad[:10]


def run_batch(g, off, tune=None, direction_optimising=False):
    roots = off[:-1].astype(np.int32)
    if direction_optimising:
        lv, lab = g.batch_bfs_cc(off, capture=8)
        return lv, lab, tune
    if tune is None:
        tune = g.autotune_bfs(roots, capture=8)
    lv, lab = g.batch_bfs_cc(off, configs="Push-x-%s-x" % tune["lb"], capture=8,
                             grid=tune["grid"], twc_small=tune["twc_small"],
                             twc_large=tune["twc_large"])
    assert g.last_stats.levels == lv.max() + 1
    return lv, lab, tune


def test_c4b_er_batch_full_piece(gsw, orc):
    say("C4b: building instances 0..999 of the seed-5 ER batch")
    h, off = gsw.erdos_renyi_batch(1000, 1000, 200000, 2.0, 32.0, seed=5, first=0)
    sizes = np.diff(off)
    assert sizes.min() >= 1000 and sizes.max() <= 200000
    g = gsw.Graph(h)
    lv, lab, tune = run_batch(g, off)
    say("C4b piece 0: autotuned", tune, "- checking 1000 instances")
    check_instances(orc, h, off, lv, lab)
    lv2, lab2, _ = run_batch(g, off, direction_optimising=True)
    say("C4b piece 0: direction-optimising loop")
    assert np.array_equal(lv2, lv) and np.array_equal(lab2, lab)
    g.close()
    del h, g
    # seeded samples of consecutive instances from the other 19 pieces
    rng = np.random.default_rng(5)
    for first in sorted(rng.choice(np.arange(1000, 20000 - 16), 4, replace=False)):
        hs, offs = gsw.erdos_renyi_batch(16, 1000, 200000, 2.0, 32.0, seed=5,
                                         first=int(first))
        gs = gsw.Graph(hs)
        say("C4b sample: instances %d..%d" % (first, first + 15))
        lv, lab, _ = run_batch(gs, offs, tune)
        check_instances(orc, hs, offs, lv, lab)
        lv2, lab2, _ = run_batch(gs, offs, direction_optimising=True)
        assert np.array_equal(lv2, lv) and np.array_equal(lab2, lab)
        gs.close()


def test_c4b_instances_do_not_depend_on_the_piece(gsw):
    """Instance i of the batch is the same graph whether it is built inside
    a piece starting at 0 or in a piece of its own."""
    h, off = gsw.erdos_renyi_batch(12, 1000, 200000, 2.0, 32.0, seed=5, first=0)
    h2, off2 = gsw.erdos_renyi_batch(4, 1000, 200000, 2.0, 32.0, seed=5, first=8)
    for k in range(4):
        a, b = instance(h, off, 8 + k), instance(h2, off2, k)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# ------------------------------------------------------------ C5: held-out graphs
def heldout_specs():
    from gswitch_amd import learn
    specs = learn.eval_specs()
    pick = [next(s for s in specs if s["kind"] == "rmat" and s["scale"] == 20),
            next(s for s in specs if s["kind"] == "er" and s["n"] <= 1 << 21),
            next(s for s in specs if s["kind"] == "grid"),
            next(s for s in specs if s["kind"] == "ba" and s["n"] <= 1 << 20)]
    return pick


@pytest.mark.parametrize("k", range(4))
def test_c5_heldout_selectors(gsw, orc, k):
    from gswitch_amd import learn
    spec = heldout_specs()[k]
    say("C5 held-out graph", spec)
    h = learn.build(spec, weighted=True)
    g = gsw.Graph(h)
    root = learn.pick_root(spec, h)
    sel = load_selectors(gsw)
    assert all(sel.get(a) for a in ("bfs", "sssp", "pr", "cc")), "committed networks missing"
    ref_l = orc.bfs(h.row_ptr, h.col, root)
    ref_d = orc.spfa(h.row_ptr, h.col, h.w, root)
    ref_p = orc.pagerank(h.row_ptr, h.col, 0.85, 50)
    ref_c = orc.cc(h.row_ptr, h.col)
    for s in (0, 1):
        assert np.array_equal(g.bfs(root, selector=s), ref_l), ("bfs", s)
        assert np.array_equal(g.sssp(root, selector=s), ref_d), ("sssp", s)
        assert max_rel(g.pagerank(0.85, 50, selector=s), ref_p) <= PR_TOL, ("pr", s)
        assert np.array_equal(g.cc(selector=s), ref_c), ("cc", s)
    g.close()
