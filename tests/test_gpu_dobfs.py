"""GPU parity of the device-resident single-query BFS (gsw_bfs with
opts.capture > 0): the direction is chosen on the device and levels replay
from a captured hipGraph, so these tests check that the level array is
bit-exact against the oracle's BFS whatever the replay length, the grid,
the switching thresholds and the heavy-vertex handling (listed and split
over the grid, or walked by one wavefront when the list overflows).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class env:
    """Set environment variables for the duration of a block."""

    def __init__(self, **kv):
        self.kv = {k: str(v) for k, v in kv.items()}

    def __enter__(self):
        self.old = {k: os.environ.get(k) for k in self.kv}
        os.environ.update(self.kv)

    def __exit__(self, *exc):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.fixture(scope="module")
def rmat14(gsw):
    h = gsw.rmat(14, 16, seed=1)
    return h, gsw.Graph(h)


@pytest.mark.parametrize("capture", [1, 3, 8])
@pytest.mark.parametrize("grid", [0, 1, 7])
def test_device_bfs_rmat(gsw, orc, rmat14, capture, grid):
    h, g = rmat14
    root = h.max_degree_vertex()
    lv = g.bfs(root, capture=capture, grid=grid)
    assert np.array_equal(lv, orc.bfs(h.row_ptr, h.col, root))
    assert g.last_stats.levels == int(lv.max()) + 1


@pytest.mark.parametrize("spec", ["grid", "er", "ba"])
def test_device_bfs_other_graphs(gsw, orc, spec):
    h = {"grid": lambda: gsw.grid2d(300, 200, 0.1, 0.01, 3),
         "er": lambda: gsw.erdos_renyi(20000, 3.0, 5),
         "ba": lambda: gsw.barabasi_albert(20000, 4, 6)}[spec]()
    g = gsw.Graph(h)
    for root in (0, h.max_degree_vertex(), h.n - 1):
        lv = g.bfs(root, capture=8)
        ref = orc.bfs(h.row_ptr, h.col, root)
        assert np.array_equal(lv, ref), (spec, root)
        assert g.last_stats.levels == int(ref.max()) + 1
    g.close()


@pytest.mark.parametrize("alpha,beta", [(1e-6, 24), (1e9, 1e-6), (1e9, 1e9), (14, 1e9), (0.5, 2)])
def test_device_bfs_switching_thresholds(gsw, orc, rmat14, alpha, beta):
    """Top-down to bottom-up when the frontier's edges exceed the unexplored
    edges / alpha: never (alpha tiny), whenever the frontier grows (alpha
    huge) with an immediate switch back (beta tiny) or none (beta huge),
    and in between: the levels do not depend on the directions."""
    h, g = rmat14
    root = h.max_degree_vertex()
    with env(GSW_DO_ALPHA=alpha, GSW_DO_BETA=beta):
        lv = g.bfs(root, capture=4)
    assert np.array_equal(lv, orc.bfs(h.row_ptr, h.col, root))


@pytest.mark.parametrize("heavy", [64, 100, 100000])
def test_device_bfs_heavy_vertices(gsw, orc, heavy):
    """Top-down only (alpha tiny) with a low heavy degree: the listed heavy
    vertices' edges are split over the grid; on R-MAT 19 levels 1 and 2
    hold 25416 and 18320 vertices of degree > 64, more than the list's
    16384 entries, so the overflow path (one wavefront per heavy vertex)
    runs as well."""
    h = gsw.rmat(19, 16, seed=3)
    g = gsw.Graph(h)
    deg = np.diff(h.row_ptr)
    root = h.max_degree_vertex()
    ref = orc.bfs(h.row_ptr, h.col, root)
    assert ((ref == 1) & (deg > 64)).sum() > 16384
    with env(GSW_DO_ALPHA=1e-9, GSW_DO_HEAVY=heavy):
        for cap in (1, 5):
            assert np.array_equal(g.bfs(root, capture=cap), ref)
    g.close()


def test_device_bfs_directed(gsw, orc):
    """One-way edges: top-down walks out-edges, bottom-up in-edges."""
    rng = np.random.default_rng(7)
    n = 5000
    h = gsw.from_edges(n, rng.integers(0, n, 15000), rng.integers(0, n, 15000),
                       directed=True)
    g = gsw.Graph(h)
    root = h.max_degree_vertex()
    ref = orc.bfs(h.row_ptr, h.col, root)
    for a in (1e-6, 14, 1e9):
        with env(GSW_DO_ALPHA=a):
            assert np.array_equal(g.bfs(root, capture=4), ref)
    g.close()


def test_device_bfs_edge_cases(gsw, orc):
    # isolated root, a graph without edges, one vertex
    h = gsw.from_edges(10, [1, 2], [2, 3])
    g = gsw.Graph(h)
    lv = g.bfs(0, capture=2)
    assert np.array_equal(lv, orc.bfs(h.row_ptr, h.col, 0))
    assert g.last_stats.levels == 1
    g.close()
    h = gsw.from_edges(1, [], [])
    g = gsw.Graph(h)
    assert np.array_equal(g.bfs(0, capture=1), np.array([0], np.int32))
    g.close()
    # a path longer than one replay: 130 vertices, 129 levels
    n = 130
    h = gsw.from_edges(n, np.arange(n - 1), np.arange(1, n))
    g = gsw.Graph(h)
    lv = g.bfs(0, capture=16)
    assert np.array_equal(lv, np.arange(n, dtype=np.int32))
    assert g.last_stats.levels == n
    g.close()


def test_device_bfs_alternates_with_engine(gsw, orc, rmat14):
    """The device loop and the engine loop share the level array: runs in
    any order, from different roots, stay exact."""
    h, g = rmat14
    roots = [h.max_degree_vertex(), 5, 777]
    for r in roots:
        ref = orc.bfs(h.row_ptr, h.col, r)
        assert np.array_equal(g.bfs(r, capture=8), ref)
        assert np.array_equal(g.bfs(r), ref)
        assert np.array_equal(g.bfs(r, capture=2), ref)


def test_device_bfs_rejects_configs(gsw, rmat14):
    h, g = rmat14
    with pytest.raises(gsw.GswError):
        g.bfs(0, capture=4, configs="Push-x-x-x")
    with pytest.raises(gsw.GswError):
        g.bfs(0, capture=4, selector=1)


@pytest.mark.parametrize("alpha", [1e-6, 14, 1e9])
def test_device_bfs_level_records(gsw, orc, rmat14, alpha):
    """The device history gives one record per level: the frontier is the
    level's vertex count, a top-down level walks every out-edge of its
    frontier, a bottom-up level examines every vertex still unvisited."""
    h, g = rmat14
    root = h.max_degree_vertex()
    ref = orc.bfs(h.row_ptr, h.col, root)
    deg = np.diff(h.row_ptr)
    with env(GSW_DO_ALPHA=alpha):
        g.bfs(root, capture=4, out=False)
    recs = g.records()
    assert len(recs) == int(ref.max()) + 1 == g.last_stats.levels
    for L, r in enumerate(recs):
        assert r["level"] == L
        assert r["frontier"] == int((ref == L).sum())
        if r["dir"] == "Push":
            assert r["verts"] == r["frontier"]
            assert r["edges"] == int(deg[ref == L].sum())
        else:
            assert r["verts"] == int(((ref == -1) | (ref > L)).sum())
    dirs = [r["dir"] for r in recs]
    assert dirs[0] == "Push"
    if alpha == 1e-6:
        assert set(dirs) == {"Push"}
    if alpha == 1e9:
        assert "Pull" in dirs
