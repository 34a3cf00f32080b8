"""GPU parity of batched execution: a batch of Erdos-Renyi instances stored
as one disjoint union, BFS from every instance's vertex 0 in one multi-root
run (per-level engine loop, and the hipGraph-captured fused loop with WM or
TWC expands), CC on the union, and multi-root BFS on one graph. Every
instance's result must equal the oracle run on that instance alone
(instance isolation), bit-exact."""
import numpy as np
import pytest

from conftest import csr_from_edges

pytestmark = pytest.mark.gpu


def instance(h, off, i):
    lo, hi = int(off[i]), int(off[i + 1])
    rp = h.row_ptr[lo:hi + 1] - h.row_ptr[lo]
    col = h.col[h.row_ptr[lo]:h.row_ptr[hi]] - lo
    return rp, col.astype(np.int32)


@pytest.fixture(scope="module")
def batch(gsw):
    h, off = gsw.erdos_renyi_batch(40, 200, 20000, 2.0, 32.0, seed=5)
    return h, off, gsw.Graph(h)


def check_batched_levels(orc, h, off, lv):
    for i in range(len(off) - 1):
        rp, col = instance(h, off, i)
        exp = orc.bfs(rp, col, 0)
        got = lv[off[i]:off[i + 1]]
        assert np.array_equal(got, exp), "instance %d" % i


@pytest.mark.parametrize("configs,capture,grid", [
    (None, 0, 0),                 # per-level engine loop, rule selector
    ("Push-Queue-WM-Fused", 0, 0),
    ("Push-x-WM-x", 1, 0),        # captured: 2 levels per replay
    ("Push-x-WM-x", 8, 7),        # captured, tiny fixed grid
    ("Push-x-TWC-x", 4, 0),
    ("Push-x-TWC-x", 8, 3),
])
def test_batched_bfs(gsw, orc, batch, configs, capture, grid):
    h, off, g = batch
    roots = off[:-1].astype(np.int32)
    lv = g.bfs_multi(roots, configs=configs, capture=capture, grid=grid)
    check_batched_levels(orc, h, off, lv)
    if capture:
        # the device level counter ends at the number of non-empty frontiers
        assert g.last_stats.levels == lv.max() + 1


def test_batched_bfs_twc_thresholds(gsw, orc, batch):
    h, off, g = batch
    roots = off[:-1].astype(np.int32)
    for small, large in [(1, 2), (4, 16), (128, 8192)]:
        lv = g.bfs_multi(roots, configs="Push-x-TWC-x", capture=2,
                         twc_small=small, twc_large=large)
        check_batched_levels(orc, h, off, lv)


def test_batched_cc(gsw, orc, batch):
    h, off, g = batch
    lab = g.cc()
    for i in range(len(off) - 1):
        rp, col = instance(h, off, i)
        assert np.array_equal(lab[off[i]:off[i + 1]] - off[i], orc.cc(rp, col)), i


@pytest.mark.parametrize("configs,capture,grid", [
    ("Push-x-WM-x", 1, 0), ("Push-x-WM-x", 8, 7), ("Push-x-TWC-x", 4, 0),
    (None, 8, 0), (None, 3, 5), ("x-x-WM-x", 1, 0)])
def test_batch_bfs_cc(gsw, orc, batch, configs, capture, grid):
    """BFS from every instance's first vertex and CC in one call (captured
    push loop when the direction is forced, else the device-resident
    direction-optimising loop from all the roots; captured CC sequence):
    every instance equals its own oracle runs, also on repeated calls (the
    captured sequences are replayed)."""
    h, off, g = batch
    for rep in range(2):
        lv, lab = g.batch_bfs_cc(off, configs=configs, capture=capture, grid=grid)
        check_batched_levels(orc, h, off, lv)
        for i in range(len(off) - 1):
            rp, col = instance(h, off, i)
            assert np.array_equal(lab[off[i]:off[i + 1]] - off[i], orc.cc(rp, col)), \
                (rep, i)
    assert g.last_stats.levels == lv.max() + 1


def test_batch_bfs_cc_instances_sharing_edges(gsw, orc):
    """Offsets that do not cut the graph into separate pieces: edges between
    "instances" are linked from both ends, so the labels are still the
    whole graph's components; the levels are the multi-root BFS from the
    instances' first vertices."""
    h = gsw.grid2d(30, 40, 0.3, 0.01, 11)
    g = gsw.Graph(h)
    off = np.array([0, 1, 300, 301, 640, 1199, 1200], np.int64)
    lv, lab = g.batch_bfs_cc(off, configs="Push-x-WM-x", capture=2)
    assert np.array_equal(lab, orc.cc(h.row_ptr, h.col))
    big = np.iinfo(np.int32).max
    exp = np.full(h.n, big)
    for r in off[:-1]:
        x = orc.bfs(h.row_ptr, h.col, int(r))
        exp = np.minimum(exp, np.where(x < 0, big, x))
    assert np.array_equal(lv, np.where(exp == big, -1, exp).astype(np.int32))
    with pytest.raises(gsw.GswError):  # the direction-optimising loop pulls too
        g.batch_bfs_cc(np.array([0, 700, h.n], np.int64), configs="Pull-x-x-x")
    with pytest.raises(gsw.GswError):
        g.batch_bfs_cc(np.array([0, 5], np.int64))  # does not end at n
    with pytest.raises(gsw.GswError):
        g.batch_bfs_cc(np.array([0, 700, 600, h.n], np.int64))  # decreasing
    g.close()


def test_captured_replays_repeat(gsw, orc, batch):
    """The captured graph is reused: repeated runs and alternating shapes
    give the same levels (state is re-initialised outside the graph)."""
    h, off, g = batch
    roots = off[:-1].astype(np.int32)
    first = g.bfs_multi(roots, configs="Push-x-WM-x", capture=2)
    for cfg in ["Push-x-TWC-x", "Push-x-WM-x", "Push-x-WM-x"]:
        assert np.array_equal(g.bfs_multi(roots, configs=cfg, capture=2), first)
    assert np.array_equal(g.bfs_multi(roots), first)


def test_multi_root_single_graph(gsw, orc):
    h = gsw.rmat(12, 16, seed=9)
    g = gsw.Graph(h)
    roots = np.array([5, 77, 77, 1000, 4095, 2048], np.int32)
    exp = None
    for r in np.unique(roots):
        lv = orc.bfs(h.row_ptr, h.col, int(r))
        lv = np.where(lv < 0, np.iinfo(np.int32).max, lv)
        exp = lv if exp is None else np.minimum(exp, lv)
    exp = np.where(exp == np.iinfo(np.int32).max, -1, exp).astype(np.int32)
    for configs, capture in [(None, 0), ("Pull-Bitmap-x-x", 0),
                             ("Push-x-WM-x", 3), ("Push-x-TWC-x", 1)]:
        assert np.array_equal(g.bfs_multi(roots, configs=configs, capture=capture),
                              exp), configs


def test_multi_root_errors(gsw):
    rp, col = csr_from_edges(4, [(0, 1), (2, 3)])
    g = gsw.Graph(row_ptr=rp, col=col)
    with pytest.raises(gsw.GswError):
        g.bfs_multi(np.array([4], np.int32))
    with pytest.raises(gsw.GswError):
        g.bfs_multi(np.array([], np.int32))
    with pytest.raises(gsw.GswError):
        g.bfs_multi(np.array([0], np.int32), configs="Push-x-STRICT-x", capture=1)
    assert list(g.bfs_multi(np.array([0, 3], np.int32), capture=1)) == [0, 1, 1, 0]


def test_autotune(gsw, orc, batch):
    h, off, g = batch
    roots = off[:-1].astype(np.int32)
    best = g.autotune_bfs(roots, capture=4)
    assert best["lb"] in ("WM", "TWC") and best["ms"] > 0 and best["grid"] >= 1
    lv = g.bfs_multi(roots, configs="Push-x-%s-x" % best["lb"], capture=4,
                     twc_small=best["twc_small"], twc_large=best["twc_large"],
                     grid=best["grid"])
    check_batched_levels(orc, h, off, lv)
