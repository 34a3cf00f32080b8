"""GPU parity of the 64-source bitmask BFS against 64 independent oracle
BFS runs: reachability masks bit-exact, and per-source sums of BFS levels
equal (a checksum of every source's level array)."""
import numpy as np
import pytest

from conftest import csr_from_edges

pytestmark = pytest.mark.gpu


def oracle_masks_and_sums(orc, h, sources):
    masks = np.zeros(h.n, np.uint64)
    sums = np.zeros(64, np.uint64)
    for s, r in enumerate(sources):
        lv = orc.bfs(h.row_ptr, h.col, int(r))
        masks[lv >= 0] |= np.uint64(1) << np.uint64(s)
        sums[s] = lv[lv >= 0].astype(np.int64).sum()
    return masks, sums


def top_degree(h, k):
    return np.argsort(-h.degrees, kind="stable")[:k].astype(np.int32)


@pytest.mark.parametrize("direction", [None, "Push-x-x-x", "Pull-x-x-x"])
def test_msbfs_rmat(gsw, orc, direction):
    h = gsw.rmat(13, 16, seed=4)
    g = gsw.Graph(h)
    src = top_degree(h, 64)
    masks, sums = g.msbfs64(src, configs=direction)
    em, es = oracle_masks_and_sums(orc, h, src)
    assert np.array_equal(masks, em)
    assert np.array_equal(sums, es)
    assert np.array_equal(masks, orc.msbfs64(h.row_ptr, h.col, src))


def test_msbfs_few_and_duplicate_sources(gsw, orc):
    h = gsw.grid2d(30, 30, 0.3, 0.0, 7)  # many small components
    g = gsw.Graph(h)
    src = np.array([0, 5, 5, 899, 450], np.int32)
    masks, sums = g.msbfs64(src)
    em, es = oracle_masks_and_sums(orc, h, src)
    assert np.array_equal(masks, em)
    assert np.array_equal(sums[:5], es[:5])
    assert not np.any(sums[5:])


def test_msbfs_isolated_and_errors(gsw, orc):
    rp, col = csr_from_edges(5, [(0, 1), (1, 2)])
    g = gsw.Graph(row_ptr=rp, col=col)
    masks, sums = g.msbfs64(np.array([4, 0], np.int32))
    assert list(masks) == [2, 2, 2, 0, 1]
    assert list(sums[:2]) == [0, 3]
    with pytest.raises(gsw.GswError):
        g.msbfs64(np.array([5], np.int32))
    with pytest.raises(gsw.GswError):
        g.msbfs64(np.zeros(65, np.int32))


def test_msbfs_c4a_shape_rmat16(gsw, orc):
    """64 highest-degree sources on R-MAT 16 (the C4a setup at small scale)."""
    h = gsw.rmat(16, 16, seed=4)
    g = gsw.Graph(h)
    src = top_degree(h, 64)
    masks, sums = g.msbfs64(src)
    em, es = oracle_masks_and_sums(orc, h, src)
    assert np.array_equal(masks, em)
    assert np.array_equal(sums, es)


@pytest.mark.parametrize("direction", [None, "Push-x-x-x", "Pull-x-x-x"])
def test_msbfs_directed(gsw, orc, direction):
    """One-way edges: bit s of v is set iff v is reachable from source s along
    the edges' directions, whether the levels run top-down (out-edges) or
    bottom-up (in-edges). Reachability against the edges differs on this
    graph, so a pull over the wrong adjacency fails."""
    rng = np.random.default_rng(11)
    n = 3000
    h = gsw.from_edges(n, rng.integers(0, n, 6000), rng.integers(0, n, 6000),
                       directed=True)
    g = gsw.Graph(h)
    src = top_degree(h, 64)
    masks, sums = g.msbfs64(src, configs=direction)
    em, es = oracle_masks_and_sums(orc, h, src)
    assert np.array_equal(masks, em)
    assert np.array_equal(sums, es)
    # reachability against the edge directions, from the transposed CSR
    e_src = np.repeat(np.arange(n), np.diff(h.row_ptr))
    order = np.lexsort((e_src, h.col))
    rrp = np.concatenate([[0], np.cumsum(np.bincount(h.col, minlength=n))])
    rcol = e_src[order].astype(np.int32)
    rev = np.zeros(n, np.uint64)
    for s, r in enumerate(src):
        rev[orc.bfs(rrp, rcol, int(r)) >= 0] |= np.uint64(1) << np.uint64(s)
    assert not np.array_equal(rev, em)
    g.close()
