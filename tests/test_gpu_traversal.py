"""GPU parity: every traversal variant against the CPU oracle.

BFS levels, SSSP distances and CC labels must be bit-exact whatever variant
the selector picks or the test forces; PageRank (FP32 output, FP64
accumulation) must be within 1e-5 max relative error of the FP64 oracle.
Variants are forced with the reference's --configs syntax
(Dir-Fmt-LB-Fusion, src/utils/cmdline.cuh:51-94).
"""
import itertools

import numpy as np
import pytest

from conftest import csr_from_edges

pytestmark = pytest.mark.gpu

DIRS = ["Push", "Pull"]
FMTS = ["Queue", "SortedQueue", "Bitmap"]
LBS = ["TM", "WM", "CM", "STRICT", "TWC"]
FUS = ["Standalone", "Fused"]
ALL_VARIANTS = ["-".join(v) for v in itertools.product(DIRS, FMTS, LBS, FUS)]
PUSH_VARIANTS = [v for v in ALL_VARIANTS if v.startswith("Push")]

PR_TOL = 1e-5


def graphs(gsw):
    out = {}
    out["rmat12"] = gsw.rmat(12, 16, seed=1)
    out["grid"] = gsw.grid2d(40, 50, 0.1, 0.01, 3)
    out["er"] = gsw.erdos_renyi(3000, 3.0, 5)
    out["ba"] = gsw.barabasi_albert(2000, 3, 6)
    return out


@pytest.fixture(scope="module")
def gset(gsw):
    hs = graphs(gsw)
    return {k: (h, gsw.Graph(h)) for k, h in hs.items()}


@pytest.fixture(scope="module")
def wset(gsw):
    out = {}
    for k, h in {"rmat12w": gsw.rmat(12, 16, seed=2, weighted=True),
                 "gridw": gsw.grid2d(30, 40, 0.1, 0.01, 3, weighted=True),
                 "erw": gsw.erdos_renyi(2000, 4.0, 8, weighted=True)}.items():
        out[k] = (h, gsw.Graph(h))
    return out


# ---------------------------------------------------------------- primitives
def test_wave_scan64(gsw):
    rng = np.random.default_rng(0)
    x = rng.integers(-1000, 1000, 64).astype(np.int32)
    assert np.array_equal(gsw.wave_scan64(x), np.cumsum(x).astype(np.int32))


@pytest.mark.parametrize("n", [0, 1, 2047, 2048, 2049, 100000, 3000001])
def test_device_scan(gsw, n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 1 << 20, n).astype(np.int64)
    out = gsw.scan_exclusive_i64(x)
    ref = np.concatenate([[0], np.cumsum(x)])
    assert np.array_equal(out, ref)


# ---------------------------------------------------------------- BFS
@pytest.mark.parametrize("gname", ["rmat12", "grid", "er", "ba"])
def test_bfs_default_selector(gsw, orc, gset, gname):
    h, g = gset[gname]
    for root in [h.max_degree_vertex(), 0, h.n - 1]:
        got = g.bfs(root)
        assert np.array_equal(got, orc.bfs(h.row_ptr, h.col, root)), root


@pytest.mark.parametrize("variant", ALL_VARIANTS)
def test_bfs_forced_variant(gsw, orc, gset, variant):
    for gname in ["rmat12", "grid"]:
        h, g = gset[gname]
        root = h.max_degree_vertex()
        got = g.bfs(root, configs=variant)
        ref = orc.bfs(h.row_ptr, h.col, root)
        assert np.array_equal(got, ref), "%s on %s" % (variant, gname)


def test_bfs_records_and_edge_cases(gsw, orc):
    rp, col = csr_from_edges(6, [(0, 1), (1, 2)])
    g = gsw.Graph(row_ptr=rp, col=col)
    assert list(g.bfs(0)) == [0, 1, 2, -1, -1, -1]
    assert list(g.bfs(4)) == [-1, -1, -1, -1, 0, -1]  # isolated root
    recs = g.records()
    assert all(r["frontier"] >= 0 for r in recs)
    with pytest.raises(gsw.GswError):
        g.bfs(6)
    g.close()


def test_bfs_c1_rmat16(gsw, orc):
    """R-MAT scale 16, EF 16, seed 1, source = highest-degree vertex."""
    h = gsw.rmat(16, 16, 0.57, 0.19, 0.19, seed=1)
    g = gsw.Graph(h)
    root = h.max_degree_vertex()
    ref = orc.bfs(h.row_ptr, h.col, root)
    assert np.array_equal(g.bfs(root), ref)
    for v in ["Push-Queue-WM-Fused", "Pull-Bitmap-TM-Standalone",
              "Push-SortedQueue-STRICT-Standalone", "Push-Queue-TWC-Standalone"]:
        assert np.array_equal(g.bfs(root, configs=v), ref), v


# ---------------------------------------------------------------- SSSP
@pytest.mark.parametrize("gname", ["rmat12w", "gridw", "erw"])
def test_sssp_default(gsw, orc, wset, gname):
    h, g = wset[gname]
    for root in [h.max_degree_vertex(), 0]:
        assert np.array_equal(g.sssp(root), orc.spfa(h.row_ptr, h.col, h.w, root))


def test_sssp_saturates_long_paths(gsw, orc):
    """uint32 path lengths saturate to the unreached sentinel instead of
    wrapping (a wrapped 0-1-2 path would cost 0 and undercut 0-3-2)."""
    big = 2**31
    rp, col, w = csr_from_edges(5, [(0, 1), (1, 2), (2, 3), (0, 3), (3, 4)],
                                [big, big, 5, 100, 2**32 - 1 - 100])
    g = gsw.Graph(row_ptr=rp, col=col, w=w)
    ref = orc.spfa(rp, col, w, 0)
    assert list(ref) == [0, big, 105, 100, 2**32 - 1]
    for v in [None, "Push-Queue-WM-Standalone", "Pull-Bitmap-TM-Standalone",
              "Push-Queue-STRICT-Fused", "Pull-Bitmap-STRICT-Standalone"]:
        assert np.array_equal(g.sssp(0, configs=v), ref), v
    g.close()


@pytest.mark.parametrize("variant", ALL_VARIANTS)
def test_sssp_forced_variant(gsw, orc, wset, variant):
    h, g = wset["rmat12w"]
    root = h.max_degree_vertex()
    assert np.array_equal(g.sssp(root, configs=variant),
                          orc.spfa(h.row_ptr, h.col, h.w, root)), variant


@pytest.mark.parametrize("gname", ["rmat12w", "gridw"])
def test_sssp_window_scales(gsw, orc, wset, gname):
    """The stepping window only shapes the work per level: every scale of
    the initial window, and the autotuned one, give the oracle's distances
    (the graph keeps its scale until changed)."""
    h, g = wset[gname]
    root = h.max_degree_vertex()
    ref = orc.spfa(h.row_ptr, h.col, h.w, root)
    try:
        for k in (4.0, 1.0, 0.3, 0.0625, 0.001):
            assert g.sssp_window(k) == k
            for v in (None, "Push-Queue-WM-Standalone", "Push-Queue-CM-Fused"):
                assert np.array_equal(g.sssp(root, configs=v), ref), (k, v)
        k, ms = g.autotune_sssp([root, 0, h.n // 2])
        assert k in (1.0, 0.5, 0.25, 0.125, 0.0625) and ms > 0
        assert g.sssp_window() == k
        assert np.array_equal(g.sssp(root), ref)
        for bad in (0.0, -1.0):
            with pytest.raises(gsw.GswError):
                g.sssp_window(bad)
        with pytest.raises(gsw.GswError):
            g.autotune_sssp([h.n])
    finally:
        g.sssp_window(1.0)


def test_sssp_requires_weights(gsw, gset):
    h, g = gset["er"]
    with pytest.raises(gsw.GswError):
        g.sssp(0)


# ---------------------------------------------------------------- CC
@pytest.mark.parametrize("gname", ["rmat12", "grid", "er", "ba"])
@pytest.mark.parametrize("variant", [None, "x-Queue-x-x", "x-Bitmap-x-x"])
def test_cc(gsw, orc, gset, gname, variant):
    h, g = gset[gname]
    assert np.array_equal(g.cc(configs=variant), orc.cc(h.row_ptr, h.col))


def test_cc_sampled_hooking_shapes(gsw, orc):
    """The default CC (sampled hooking): many small components, isolated
    vertices, a star (one hub holding every edge, walked by whole
    wavefronts) and a directed graph (weak components, no skipped set)."""
    rng = np.random.default_rng(4)
    h = gsw.grid2d(50, 50, 0.45, 0.0, 9)
    g = gsw.Graph(h)
    assert np.array_equal(g.cc(), orc.cc(h.row_ptr, h.col))
    g.close()
    rp, col = csr_from_edges(3000, [(0, v) for v in range(1, 3000, 2)] +
                             [(v, v + 1) for v in range(2, 2990, 7)])
    g = gsw.Graph(row_ptr=rp, col=col)
    assert np.array_equal(g.cc(), orc.cc(rp, col))
    g.close()
    n = 2000
    src, dst = rng.integers(0, n, 1500), rng.integers(0, n, 1500)
    hd = gsw.from_edges(n, src, dst, directed=True)
    g = gsw.Graph(hd)
    srp, scol = csr_from_edges(n, np.stack([src, dst], 1))
    assert np.array_equal(g.cc(), orc.cc(srp, scol))
    g.close()


# ---------------------------------------------------------------- PageRank
def max_rel(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b) / np.abs(b)))


@pytest.mark.parametrize("variant", [None] + ["%s-Bitmap-%s-Standalone" % (d, l)
                                              for d in DIRS for l in LBS + ["TWOD"]])
@pytest.mark.parametrize("pull_kernel", [0, 1, 2])
def test_pagerank(gsw, orc, gset, variant, pull_kernel):
    for gname in ["rmat12", "ba", "grid"]:
        h, g = gset[gname]
        pr = g.pagerank(0.85, 50, configs=variant, pull_kernel=pull_kernel)
        ref = orc.pagerank(h.row_ptr, h.col, 0.85, 50)
        assert max_rel(pr, ref) <= PR_TOL, (variant, gname, max_rel(pr, ref))


@pytest.mark.parametrize("slice_kb", [1, 4, 64])
def test_pagerank_twod_slices(gsw, orc, monkeypatch, slice_kb):
    """TWOD with several source slices (one launch each): R-MAT 12 has
    n * 4 B = 16 KB of source state, so 1 KB slices give 16 launches per
    iteration; its degrees reach 1345, so rows of every length bin and rows
    cut into several 512-edge chunks (combined with compAtomic) occur."""
    monkeypatch.setenv("GSW_TWOD_SLICE_KB", str(slice_kb))
    h = gsw.rmat(12, 16, seed=1)
    deg = np.diff(h.row_ptr)
    assert deg.max() > 2 * 512 and (deg <= 8).any() and ((deg > 8) & (deg <= 64)).any()
    ref = orc.pagerank(h.row_ptr, h.col, 0.85, 50)
    for pull_kernel in (0, 1):
        g = gsw.Graph(h)  # slices are built on the graph's first TWOD pull
        pr = g.pagerank(0.85, 50, configs="Pull-Bitmap-TWOD-Standalone",
                        pull_kernel=pull_kernel)
        g.close()
        assert max_rel(pr, ref) <= PR_TOL, (slice_kb, pull_kernel, max_rel(pr, ref))


@pytest.mark.parametrize("ecap,groups", [(0, 0), (300, 0), (0, 3), (300, 5), (0, 64)])
def test_pagerank_propagation_blocking(gsw, orc, monkeypatch, ecap, groups):
    """Propagation-blocked pulls (pull_kernel 2). R-MAT 15 has 32 768
    vertices, so its sources fall in two 16 384-id chunks and its targets
    in bins of at most 8 192; with pieces of 300 slots (GSW_PB_ECAP) the
    hubs (degrees in the thousands) become one-target bins split over
    several workgroups that add their parts atomically, and chunks are cut
    into several phase-1 pieces. Also a grid (no hubs) and a graph with
    isolated vertices. With GSW_PB_GROUPS the bins form groups whose phase 1
    and phase 2 run group by group (a hub's one-target bin can hold more
    than a group's share, so some groups are empty)."""
    if ecap:
        monkeypatch.setenv("GSW_PB_ECAP", str(ecap))
    if groups:
        monkeypatch.setenv("GSW_PB_GROUPS", str(groups))
    h = gsw.rmat(15, 16, seed=7)
    deg = np.diff(h.row_ptr)
    assert deg.max() > 3 * 300 and (deg == 0).any()
    for hh in (h, gsw.grid2d(70, 90, 0.2, 0.01, 5)):
        ref = orc.pagerank(hh.row_ptr, hh.col, 0.85, 30)
        g = gsw.Graph(hh)  # the layout is built on the graph's first such pull
        for rep in range(2):
            pr = g.pagerank(0.85, 30, configs="Pull-Bitmap-TWOD-Standalone",
                            pull_kernel=2)
            assert max_rel(pr, ref) <= PR_TOL, (ecap, groups, rep, max_rel(pr, ref))
        g.close()


def test_unsorted_rows_are_sorted_on_upload(gsw, orc, monkeypatch):
    """Rows handed over in any order: the TWOD slices split rows by source
    range (1 KB slices: 16 per iteration on R-MAT 12) and the bottom-up pull
    takes each row's first entry, so both see the rows sorted on upload."""
    monkeypatch.setenv("GSW_TWOD_SLICE_KB", "1")
    h = gsw.rmat(12, 16, seed=1, weighted=True)
    rng = np.random.default_rng(3)
    perm = np.arange(h.m)
    for v in range(h.n):
        lo, hi = h.row_ptr[v], h.row_ptr[v + 1]
        perm[lo:hi] = lo + rng.permutation(hi - lo)
    col, w = h.col[perm], h.w[perm]
    assert not all(np.all(np.diff(col[h.row_ptr[v]:h.row_ptr[v + 1]]) > 0)
                   for v in range(h.n))
    g = gsw.Graph(row_ptr=h.row_ptr, col=col, w=w)
    pr = g.pagerank(0.85, 50, configs="Pull-Bitmap-TWOD-Standalone")
    ref = orc.pagerank(h.row_ptr, h.col, 0.85, 50)
    assert max_rel(pr, ref) <= PR_TOL
    root = h.max_degree_vertex()
    assert np.array_equal(g.bfs(root, configs="Pull-Bitmap-TM-Standalone"),
                          orc.bfs(h.row_ptr, h.col, root))
    assert np.array_equal(g.sssp(root), orc.spfa(h.row_ptr, h.col, h.w, root))
    g.close()


def test_pagerank_iteration_counts(gsw, orc, gset):
    h, g = gset["er"]
    for k in (0, 1, 2, 5):
        pr = g.pagerank(0.85, k)
        assert max_rel(pr, orc.pagerank(h.row_ptr, h.col, 0.85, k)) <= PR_TOL, k


# ---------------------------------------------------------------- degenerate
@pytest.mark.parametrize("n", [1, 5])
def test_graphs_without_edges(gsw, orc, n):
    """Vertices and no edges: every traversal ends at once with the root
    alone reached, whatever the variant."""
    rp = np.zeros(n + 1, np.int64)
    col = np.zeros(0, np.int32)
    g = gsw.Graph(row_ptr=rp, col=col)
    exp = np.full(n, -1, np.int32)
    exp[n - 1] = 0
    assert np.array_equal(g.bfs(n - 1), exp)
    for v in ["Push-Queue-STRICT-Standalone", "Pull-Bitmap-TM-Standalone",
              "Push-Queue-WM-Fused", "Push-Queue-TWC-Standalone"]:
        assert np.array_equal(g.bfs(n - 1, configs=v), exp), v
    assert np.array_equal(g.cc(), np.arange(n, dtype=np.int32))
    pr = g.pagerank(0.85, 10)
    assert max_rel(pr, orc.pagerank(rp, col, 0.85, 10)) <= PR_TOL
    for pk in (1, 2):
        pr2 = g.pagerank(0.85, 10, configs="Pull-Bitmap-TWOD-Standalone", pull_kernel=pk)
        assert max_rel(pr2, orc.pagerank(rp, col, 0.85, 10)) <= PR_TOL, pk
    lv, nsp, dep = g.bc(0)
    rl, rn, rd = orc.bc(rp, col, 0)
    assert np.array_equal(lv, rl) and np.allclose(nsp, rn) and np.allclose(dep, rd)
    masks, _ = g.msbfs64(np.arange(min(n, 64), dtype=np.int32))
    assert np.array_equal(masks, (np.uint64(1) << np.arange(n, dtype=np.uint64)))
    multi = g.bfs_multi(np.array([0], np.int32), configs="Push-x-WM-x", capture=2)
    exp0 = np.full(n, -1, np.int32)
    exp0[0] = 0
    assert np.array_equal(multi, exp0)
    ex = g.explore_bfs(0, max_levels=4, stride=1, reps=1)
    assert len(ex["level"]) <= 1
    g.close()


def test_traversal_on_a_matrix_market_file(gsw, orc, tmp_path):
    """A graph read from a Matrix Market file (rinsed, as the reference
    loads it) runs BFS and CC bit-exact against the oracle."""
    from test_host_graph import write_mtx
    h0 = gsw.rmat(12, 16, seed=5)
    p = str(tmp_path / "g.mtx")
    write_mtx(p, h0)
    h = gsw.read_mtx(p, header=True)
    assert h.n < h0.n  # R-MAT 12 has isolated vertices, removed by rinse
    g = gsw.Graph(h)
    root = h.max_degree_vertex()
    assert np.array_equal(g.bfs(root), orc.bfs(h.row_ptr, h.col, root))
    assert np.array_equal(g.cc(), orc.cc(h.row_ptr, h.col))
    g.close()


@pytest.mark.parametrize("app", ["bfs", "sssp", "cc", "pagerank", "bc", "gc"])
def test_command_line_driver(gsw, tmp_path, capsys, app):
    """The command-line driver runs each application on a Matrix Market
    file and its --validation passes."""
    from test_host_graph import write_mtx
    from gswitch_amd import cli
    p = str(tmp_path / "g.mtx")
    write_mtx(p, gsw.rmat(11, 16, seed=6))
    assert cli.main([app, p, "-H", "-V"]) == 0
    out = capsys.readouterr().out
    assert ("validation: PASS" in out) == (app != "bc")


def test_training_records(gsw, gset, tmp_path):
    """gsw_record_features appends one reference-format record per expanded
    level (keys "0".."22": the 21 features, then filter and expand times),
    at most 100 per setting."""
    h, g = gset["rmat12"]
    root = h.max_degree_vertex()
    p = str(tmp_path / "rec.json")
    gsw.record_features(p)
    try:
        g.bfs(root)
        levels = g.records()
        recs = gsw.read_records(p)
        assert len(recs) == len(levels) > 0
        for r, lv in zip(recs, levels):
            assert sorted(r, key=int) == [str(i) for i in range(23)]
            assert r["0"] == h.n and r["1"] == h.m // 2
            if not lv["fused"]:
                assert r["7"] == lv["active"], (r["7"], lv)
        for _ in range(100 // len(levels) + 1):
            g.bfs(root, out=False)
        assert len(gsw.read_records(p)) == 100
    finally:
        gsw.record_features(None)
    g.bfs(root, out=False)
    assert len(gsw.read_records(p)) == 100


# ---------------------------------------------------------------- colouring
def test_gc_rejects_directed(gsw):
    h = gsw.from_edges(4, [0, 1, 2], [1, 2, 3], directed=True)
    g = gsw.Graph(h)
    with pytest.raises(gsw.GswError):
        g.gc()
    g.close()


@pytest.mark.parametrize("variant", [None, "Pull-Queue-TM-Standalone",
                                     "Pull-Bitmap-TWOD-Standalone",
                                     "Pull-Bitmap-STRICT-Standalone",
                                     "Pull-Queue-WM-Standalone",
                                     "Push-Queue-WM-Standalone",
                                     "Push-Bitmap-CM-Standalone"])
def test_gc(gsw, orc, gset, variant):
    """Graph colouring (Jones-Plassmann rounds, id priorities) bit-exact
    against the oracle, whatever the variant."""
    for gname in ["rmat12", "grid", "er", "ba"]:
        h, g = gset[gname]
        assert np.array_equal(g.gc(configs=variant), orc.gc(h.row_ptr, h.col)), \
            (variant, gname)


# ---------------------------------------------------------------- Delta-PageRank
@pytest.mark.parametrize("variant", [None, "Push-Queue-WM-Standalone",
                                     "Push-Bitmap-STRICT-Standalone",
                                     "Pull-Bitmap-TM-Standalone",
                                     "Pull-Queue-CM-Standalone"])
def test_delta_pagerank(gsw, orc, gset, variant):
    """Delta-PageRank reaches the PageRank fixed point: within 1e-5 max
    relative error of the FP64 oracle iterated to convergence (300
    iterations: 0.85^300 ~ 1e-21), in fewer rounds than that."""
    for gname in ["rmat12", "ba", "grid", "er"]:
        h, g = gset[gname]
        pr = g.delta_pagerank(0.85, 1e-10, configs=variant)
        ref = orc.pagerank(h.row_ptr, h.col, 0.85, 300)
        assert max_rel(pr, ref) <= PR_TOL, (variant, gname, max_rel(pr, ref))
        assert 0 < g.last_stats.levels < 300, (variant, gname, g.last_stats.levels)
    with pytest.raises(gsw.GswError):
        g.delta_pagerank(0.85, 0.0)
    with pytest.raises(gsw.GswError):  # its filter applies the changes: never fused
        g.delta_pagerank(0.85, 1e-10, configs="Push-Queue-WM-Fused")
