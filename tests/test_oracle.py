"""Pin the CPU oracle (oracle/oracle.c) before trusting it.

Every oracle function is checked against closed forms, against an
independent algorithm (scipy.sparse.csgraph, union-find, Dijkstra,
Bellman-Ford, brute-force path enumeration) or both, so that a dropped term,
a wrong index or an off-by-one in the restatement fails a test.
"""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

from conftest import csr_from_edges


def path_graph(n):
    return csr_from_edges(n, [(i, i + 1) for i in range(n - 1)])


def grid_graph(r, c):
    e = []
    for i in range(r):
        for j in range(c):
            if j + 1 < c:
                e.append((i * c + j, i * c + j + 1))
            if i + 1 < r:
                e.append((i * c + j, (i + 1) * c + j))
    return csr_from_edges(r * c, e)


def random_graph(n, m, seed, weights=False):
    rng = np.random.default_rng(seed)
    e = rng.integers(0, n, size=(m, 2))
    w = rng.integers(1, 65, size=m) if weights else None
    return csr_from_edges(n, e, w)


def to_scipy(rp, col, w=None):
    n = len(rp) - 1
    data = np.ones(len(col)) if w is None else w.astype(np.float64)
    return sp.csr_matrix((data, col, rp), shape=(n, n))


# ---------------------------------------------------------------- BFS
def test_bfs_path_closed_form(orc):
    rp, col = path_graph(50)
    assert np.array_equal(orc.bfs(rp, col, 0), np.arange(50))
    lv = orc.bfs(rp, col, 20)
    assert np.array_equal(lv, np.abs(np.arange(50) - 20))


def test_bfs_grid_is_manhattan(orc):
    rp, col = grid_graph(13, 17)
    lv = orc.bfs(rp, col, 0)
    r, c = np.divmod(np.arange(13 * 17), 17)
    assert np.array_equal(lv, r + c)


def test_bfs_star_and_disconnected(orc):
    rp, col = csr_from_edges(8, [(0, i) for i in range(1, 6)] + [(6, 7)])
    assert list(orc.bfs(rp, col, 0)) == [0, 1, 1, 1, 1, 1, -1, -1]
    assert list(orc.bfs(rp, col, 3)) == [1, 2, 2, 0, 2, 2, -1, -1]
    assert list(orc.bfs(rp, col, 7)) == [-1] * 6 + [1, 0]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_bfs_matches_scipy(orc, seed):
    rp, col = random_graph(3000, 6000, seed)
    d = cg.shortest_path(to_scipy(rp, col), unweighted=True, indices=7)
    ref = np.where(np.isinf(d), -1, d).astype(np.int32)
    assert np.array_equal(orc.bfs(rp, col, 7), ref)


def test_bfs_level_invariants(orc, gsw):
    h = gsw.rmat(12, seed=9)
    root = h.max_degree_vertex()
    lv = orc.bfs(h.row_ptr, h.col, root)
    src = np.repeat(np.arange(h.n), np.diff(h.row_ptr))
    a, b = lv[src], lv[h.col]
    both = (a >= 0) & (b >= 0)
    assert np.all(np.abs(a[both] - b[both]) <= 1)
    assert np.all((a >= 0) == (b >= 0))  # reachability is symmetric per edge
    # every reached non-root vertex has a parent one level up
    par = np.zeros(h.n, bool)
    par[h.col[(a >= 0) & (b == a + 1)]] = True
    reached = lv > 0
    assert np.all(par[reached])


# ---------------------------------------------------------------- SSSP
def test_sssp_weighted_path(orc):
    n = 40
    w = np.arange(1, n, dtype=np.uint32)
    rp, col, ww = csr_from_edges(n, [(i, i + 1) for i in range(n - 1)], w)
    d = orc.spfa(rp, col, ww, 0)
    assert np.array_equal(d, np.concatenate([[0], np.cumsum(w)]).astype(np.uint32))


@pytest.mark.parametrize("seed", [4, 5])
def test_sssp_three_algorithms_and_scipy(orc, seed):
    rp, col, w = random_graph(2000, 5000, seed, weights=True)
    a = orc.spfa(rp, col, w, 3)
    assert np.array_equal(a, orc.dijkstra(rp, col, w, 3))
    d = cg.dijkstra(to_scipy(rp, col, w), indices=3)
    ref = np.where(np.isinf(d), np.iinfo(np.uint32).max, d).astype(np.uint32)
    assert np.array_equal(a, ref)


def test_sssp_bellman_ford_tiny(orc):
    for seed in range(10):
        rp, col, w = random_graph(12, 20, 100 + seed, weights=True)
        assert np.array_equal(orc.spfa(rp, col, w, 0), orc.bellman_ford(rp, col, w, 0))


def test_sssp_unreached_sentinel(orc):
    rp, col, w = csr_from_edges(4, [(0, 1), (2, 3)], [5, 7])
    assert list(orc.spfa(rp, col, w, 0)) == [0, 5, 2**32 - 1, 2**32 - 1]


def test_sssp_saturates_long_paths(orc):
    """Path lengths past 2^32 - 1 saturate to the unreached sentinel rather
    than wrapping: 0-1-2 costs 2^31 + 2^31 = 2^32 (wraps to 0), 2-3 adds 5
    (wraps to 5, shorter than the direct 0-3 edge of 100 only if it wraps).
    scipy's float64 Dijkstra gives the exact lengths to compare with."""
    big = 2**31
    rp, col, w = csr_from_edges(5, [(0, 1), (1, 2), (2, 3), (0, 3), (3, 4)],
                                [big, big, 5, 100, 2**32 - 1 - 100])
    exact = cg.dijkstra(to_scipy(rp, col, w), indices=0)
    mx = 2**32 - 1
    ref = np.where(exact >= mx, mx, exact).astype(np.uint64)
    for fn in (orc.spfa, orc.dijkstra, orc.bellman_ford):
        d = fn(rp, col, w, 0)
        assert list(d) == list(ref), fn
        assert list(d) == [0, big, 100 + 5, 100, mx], fn


# ---------------------------------------------------------------- CC
@pytest.mark.parametrize("seed", [6, 7])
def test_cc_min_label_three_ways(orc, seed):
    rp, col = random_graph(4000, 3000, seed)
    a = orc.cc(rp, col)
    assert np.array_equal(a, orc.cc_unionfind(rp, col))
    k, lab = cg.connected_components(to_scipy(rp, col), directed=False)
    mins = np.full(k, np.iinfo(np.int64).max)
    np.minimum.at(mins, lab, np.arange(len(lab)))
    assert np.array_equal(a, mins[lab].astype(np.int32))


def test_cc_closed_form(orc):
    rp, col = csr_from_edges(7, [(5, 3), (3, 1), (6, 4)])
    assert list(orc.cc(rp, col)) == [0, 1, 2, 1, 4, 1, 4]


# ---------------------------------------------------------------- PageRank
def test_pagerank_complete_graph_closed_form(orc):
    n = 9
    rp, col = csr_from_edges(n, list(itertools.combinations(range(n), 2)))
    for k in (0, 1, 7, 50):
        pr = orc.pagerank(rp, col, 0.85, k)
        assert np.allclose(pr, 1 - 0.85 ** (k + 1), rtol=1e-14)


def test_pagerank_matches_sparse_power_iteration(orc):
    rp, col = random_graph(3000, 9000, 11)
    n = len(rp) - 1
    A = to_scipy(rp, col)  # symmetric adjacency
    od = np.diff(rp).astype(np.float64)
    inv = np.where(od > 0, 1.0 / np.maximum(od, 1), 0.0)
    pr = np.full(n, 0.15)
    for _ in range(50):
        pr = 0.15 + 0.85 * (A.T @ (pr * inv))
    assert np.allclose(orc.pagerank(rp, col, 0.85, 50), pr, rtol=1e-12)


def test_pagerank_isolated_vertex(orc):
    rp, col = csr_from_edges(3, [(0, 1)])
    pr = orc.pagerank(rp, col, 0.85, 5)
    assert pr[2] == pytest.approx(0.15)


# ---------------------------------------------------------------- BC
def brute_bc(n, edges, root):
    adj = {i: set() for i in range(n)}
    for a, b in edges:
        if a != b:
            adj[a].add(b)
            adj[b].add(a)
    # all shortest paths from root by BFS layering, then enumeration
    dist = {root: 0}
    frontier = [root]
    while frontier:
        nxt = []
        for v in frontier:
            for u in adj[v]:
                if u not in dist:
                    dist[u] = dist[v] + 1
                    nxt.append(u)
        frontier = nxt

    def paths(t):
        if t == root:
            return [[root]]
        out = []
        for p in adj[t]:
            if p in dist and dist[p] == dist[t] - 1:
                out += [q + [t] for q in paths(p)]
        return out

    nsp = np.zeros(n)
    dep = np.zeros(n)
    allp = {t: paths(t) for t in dist}
    for t in dist:
        nsp[t] = len(allp[t])
    for v in dist:
        s = 0.0
        for t in dist:
            if t == v or dist[t] <= dist[v]:
                continue
            s += sum(1 for p in allp[t] if v in p) / len(allp[t])
        dep[v] = s
    return nsp, dep


@pytest.mark.parametrize("seed", range(6))
def test_bc_brute_force(orc, seed):
    rng = np.random.default_rng(50 + seed)
    n = 9
    edges = [tuple(e) for e in rng.integers(0, n, size=(14, 2))]
    rp, col = csr_from_edges(n, edges)
    lvl, nsp, dep = orc.bc(rp, col, 0)
    bn, bd = brute_bc(n, edges, 0)
    assert np.array_equal(lvl, orc.bfs(rp, col, 0))
    assert np.allclose(nsp, bn)
    assert np.allclose(dep, bd)


# ---------------------------------------------------------------- msbfs
def test_msbfs_masks(orc):
    rp, col = random_graph(2500, 2200, 13)
    srcs = np.arange(0, 2500, 40)[:64]
    mask = orc.msbfs64(rp, col, srcs)
    k, lab = cg.connected_components(to_scipy(rp, col), directed=False)
    for s, v in enumerate(srcs):
        expect = lab == lab[v]
        got = ((mask >> np.uint64(s)) & np.uint64(1)).astype(bool)
        assert np.array_equal(got, expect)


def test_traversed_edges(orc):
    rp, col = csr_from_edges(6, [(0, 1), (1, 2), (3, 4)])
    lv = orc.bfs(rp, col, 0)
    assert orc.traversed_edges(rp, lv) == 2


# ---------------------------------------------------------------- colouring
def gc_recurrence(n, row_ptr, col):
    """Independent form of id-priority Jones-Plassmann: a vertex is coloured
    in the round after all its higher-id neighbours, so colour(v) = 1 + the
    largest colour among neighbours with a larger id (0 if none); computed
    in decreasing id order."""
    color = np.zeros(n, np.int64)
    for v in range(n - 1, -1, -1):
        nb = col[row_ptr[v]:row_ptr[v + 1]]
        hi = nb[nb > v]
        color[v] = 1 + (color[hi].max() if len(hi) else 0)
    return color


def test_gc_closed_forms(orc):
    # path 0-1-...-(n-1): only the last vertex beats its neighbours, then the
    # one before it, ...: colour(v) = n - v
    n = 9
    rp, col = csr_from_edges(n, [(i, i + 1) for i in range(n - 1)])
    assert list(orc.gc(rp, col)) == [n - v for v in range(n)]
    # star centred on 0: the leaves all colour in round 1, the centre in 2
    rp, col = csr_from_edges(6, [(0, i) for i in range(1, 6)])
    assert list(orc.gc(rp, col)) == [2, 1, 1, 1, 1, 1]
    # complete graph: one vertex per round, highest id first
    k = 5
    rp, col = csr_from_edges(k, [(i, j) for i in range(k) for j in range(i + 1, k)])
    assert list(orc.gc(rp, col)) == [k - v for v in range(k)]
    # isolated vertices colour in round 1
    rp, col = csr_from_edges(3, [])
    assert list(orc.gc(rp, col)) == [1, 1, 1]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_gc_matches_recurrence_and_is_proper(orc, seed):
    rng = np.random.default_rng(seed)
    n = 300
    e = rng.integers(0, n, size=(900, 2))
    rp, col = csr_from_edges(n, [tuple(x) for x in e if x[0] != x[1]])
    c = orc.gc(rp, col)
    assert np.array_equal(c, gc_recurrence(n, rp, col))
    u = np.repeat(np.arange(n), np.diff(rp))
    assert np.all(c >= 1) and np.all(c[u] != c[col])


def test_complete_bipartite_closed_forms(orc):
    """Closed forms used at full size past 2^31 entries (K_{a,a} with
    a = 33000, tests/test_gpu_large.py), pinned here on small K_{a,a} built
    by the same helper."""
    from test_gpu_large import complete_bipartite, expected_levels
    for a in (1, 2, 7, 50):
        rp, col = complete_bipartite(a)
        ref_rp, ref_col = csr_from_edges(2 * a, [(i, a + j) for i in range(a)
                                                 for j in range(a)])
        assert np.array_equal(rp, ref_rp) and np.array_equal(col, ref_col)
        for root in (0, a, 2 * a - 1):
            assert np.array_equal(orc.bfs(rp, col, root), expected_levels(a, root))
        assert not orc.cc(rp, col).any()
        pr = orc.pagerank(rp, col, 0.85, 10)
        assert np.allclose(pr, 1 - 0.85 ** 11, rtol=1e-12)
