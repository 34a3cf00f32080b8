"""A graph with more CSR entries than int32 can index.

K_{a,a} with a = 33000 holds 2 * a^2 = 2,178,000,000 directed entries, past
2^31 = 2,147,483,648: row offsets, edge positions, the STRICT slot
prefixes, the TWC bins and the TWOD slices all run on int64 there, and the
second BFS level scans a^2 = 1.09e9 edges at once. Results are checked
against closed forms that the oracle reproduces on small K_{a,a}
(tests/test_oracle.py::test_complete_bipartite_closed_forms):
  BFS from vertex 0: level 1 on the other side, 2 on its own side;
  CC: one component, label 0;
  PageRank: every vertex has degree a, so pr_k = 1 - 0.85^(k+1).
"""
import sys

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

A = 33000


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def complete_bipartite(a):
    """K_{a,a}: side 0 = ids 0..a-1, side 1 = a..2a-1; each row is one
    contiguous ascending id range."""
    n = 2 * a
    row_ptr = np.arange(n + 1, dtype=np.int64) * a
    col = np.empty(n * a, np.int32)
    col[: a * a].reshape(a, a)[:] = np.arange(a, 2 * a, dtype=np.int32)
    col[a * a:].reshape(a, a)[:] = np.arange(0, a, dtype=np.int32)
    return row_ptr, col


def expected_levels(a, root):
    lv = np.full(2 * a, 2, np.int32)
    other = slice(a, 2 * a) if root < a else slice(0, a)
    lv[other] = 1
    lv[root] = 0
    return lv


@pytest.fixture(scope="module")
def big(gsw):
    say("building K_{%d,%d}" % (A, A))
    rp, col = complete_bipartite(A)
    assert len(col) > 2**31
    g = gsw.Graph(row_ptr=rp, col=col)
    del col
    yield g
    g.close()


@pytest.mark.parametrize("variant", [None, "Push-Queue-STRICT-Standalone",
                                     "Push-Queue-TWC-Standalone",
                                     "Push-Bitmap-CM-Standalone",
                                     "Push-Queue-WM-Fused",
                                     "Push-Queue-STRICT-Fused",
                                     "Pull-Bitmap-TM-Standalone",
                                     "Pull-Bitmap-STRICT-Standalone"])
def test_bfs_past_int32_entries(big, variant):
    say("K_{a,a} BFS", variant)
    exp = expected_levels(A, 0)
    assert np.array_equal(big.bfs(0, configs=variant), exp)
    recs = big.records()
    # the second level scans every edge of one side (pushes) or probes one
    # in-edge per target (pulls)
    assert max(r["edges"] for r in recs) >= (A * A if variant and
                                            variant.startswith("Push") else A)


def test_cc_and_pagerank_past_int32_entries(big):
    say("K_{a,a} CC")
    assert not big.cc().any()
    for kw in [dict(), dict(configs="Pull-Bitmap-STRICT-Standalone")]:
        say("K_{a,a} PageRank", kw)
        pr = big.pagerank(0.85, 10, **kw)
        exact = 1.0 - 0.85 ** 11
        assert np.max(np.abs(pr - exact) / exact) <= 1e-5, kw
