"""GPU parity of betweenness centrality (forward path counting + backward
dependency accumulation) against the oracle's Brandes pass (restating
bc_cpu, application/bc.cu:162-214). Levels are bit-exact, path counts are
exact integers held in FP64 (bit-exact below 2^53), dependencies agree to
1e-9 relative (FP64 atomics add in a different order)."""
import itertools

import numpy as np
import pytest

from conftest import csr_from_edges

pytestmark = pytest.mark.gpu

VARIANTS = [None] + ["-".join(v) for v in itertools.product(
    ["Push", "Pull"], ["Queue", "SortedQueue", "Bitmap"],
    ["TM", "WM", "CM", "STRICT", "TWC"], ["Standalone", "Fused"])]


def check(got, exp):
    lv, nsp, dep = got
    el, en, ed = exp
    assert np.array_equal(lv, el)
    assert np.array_equal(nsp, en)
    np.testing.assert_allclose(dep, ed, rtol=1e-9, atol=1e-12)


@pytest.fixture(scope="module")
def rmat(gsw):
    h = gsw.rmat(12, 16, seed=1)
    return h, gsw.Graph(h)


@pytest.mark.parametrize("configs", VARIANTS)
def test_bc_rmat(gsw, orc, rmat, configs):
    h, g = rmat
    root = h.max_degree_vertex()
    check(g.bc(root, configs=configs), orc.bc(h.row_ptr, h.col, root))


@pytest.mark.parametrize("configs", [None, "Push-Queue-WM-Fused",
                                     "Pull-Bitmap-WM-Standalone"])
def test_bc_grid_many_levels(gsw, orc, configs):
    h = gsw.grid2d(40, 40, 0.1, 0.0, 3)
    g = gsw.Graph(h)
    check(g.bc(0, configs=configs), orc.bc(h.row_ptr, h.col, 0))


def test_bc_small_and_disconnected(gsw, orc):
    edges = [(0, 1), (0, 2), (1, 3), (2, 3), (3, 4), (5, 6)]
    rp, col = csr_from_edges(8, edges)
    g = gsw.Graph(row_ptr=rp, col=col)
    lv, nsp, dep = g.bc(0)
    assert list(lv) == [0, 1, 1, 2, 3, -1, -1, -1]
    assert list(nsp) == [1, 1, 1, 2, 2, 0, 0, 0]
    # dep(v) = sum over targets t of sigma_st(v)/sigma_st (root excluded as v)
    np.testing.assert_allclose(dep[1:5], [1.0, 1.0, 1.0, 0.0])
    check((lv, nsp, dep), orc.bc(rp, col, 0))
    # an isolated root
    lv, nsp, dep = g.bc(7)
    assert list(lv) == [-1] * 7 + [0] and nsp[7] == 1 and not dep.any()


def test_bc_repeat_and_errors(gsw, orc, rmat):
    h, g = rmat
    a = g.bc(3)
    b = g.bc(3, configs="Push-Queue-WM-Fused")
    check(a, orc.bc(h.row_ptr, h.col, 3))
    check(b, orc.bc(h.row_ptr, h.col, 3))
    with pytest.raises(gsw.GswError):
        g.bc(h.n)
