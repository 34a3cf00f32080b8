"""The command-line driver's validation checks (gswitch_amd/cli.py), on the
CPU: each accepts the oracle's result and rejects a copy with one planted
mistake."""
import numpy as np
import pytest

from gswitch_amd import cli


@pytest.fixture(scope="module")
def wgraph(gsw):
    return gsw.rmat(11, 8, seed=4, weighted=True)


def test_check_bfs(orc, wgraph):
    h = wgraph
    root = h.max_degree_vertex()
    lv = orc.bfs(h.row_ptr, h.col, root)
    assert cli.check_bfs(h, root, lv)
    far = int(np.argmax(lv))
    bad = lv.copy()
    bad[far] += 1  # one level too deep
    assert not cli.check_bfs(h, root, bad)
    bad = lv.copy()
    bad[far] = -1  # a reached vertex dropped
    assert not cli.check_bfs(h, root, bad)
    bad = lv.copy()
    bad[far] -= 1  # too shallow: no neighbour one level up
    assert not cli.check_bfs(h, root, bad)


def test_check_sssp(orc, wgraph):
    h = wgraph
    root = h.max_degree_vertex()
    d = orc.spfa(h.row_ptr, h.col, h.w, root)
    assert cli.check_sssp(h, h.w, root, d)
    v = int(np.argmax(np.where(d == np.uint32(0xFFFFFFFF), 0, d)))
    for delta in (1, -1):
        bad = d.copy()
        bad[v] = np.uint32(int(bad[v]) + delta)
        assert not cli.check_sssp(h, h.w, root, bad), delta


def test_check_cc(gsw, orc):
    h = gsw.erdos_renyi(3000, 1.5, 7)  # many components
    lab = orc.cc(h.row_ptr, h.col)
    assert len(np.unique(lab)) > 10 and cli.check_cc(h, lab)
    big = np.bincount(lab).argmax()
    members = np.flatnonzero(lab == big)
    bad = lab.copy()
    bad[members[-1]] = members[-1]  # one member split off
    assert not cli.check_cc(h, bad)
    bad = lab.copy()
    bad[members] = members[1]  # a label that is not the smallest id
    assert not cli.check_cc(h, bad)


def test_pagerank_residual(orc, wgraph):
    h = wgraph
    assert cli.pagerank_residual(h, orc.pagerank(h.row_ptr, h.col, 0.85, 50), 0.85) < 0.05
    assert cli.pagerank_residual(h, orc.pagerank(h.row_ptr, h.col, 0.85, 2), 0.85) > 0.05


def test_pair_weights_are_symmetric(wgraph):
    h = wgraph
    w = cli.pair_weights(h)
    assert w.min() >= 1 and w.max() <= 64
    u = np.repeat(np.arange(h.n), np.diff(h.row_ptr))
    back = {(int(a), int(b)): int(x) for a, b, x in zip(u, h.col, w)}
    assert all(back[(b, a)] == x for (a, b), x in back.items())


def test_cli_reads_the_file_before_any_gpu_call(gsw, tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("1 2\n2 3\n")
    with pytest.raises(gsw.GswError):
        cli.main(["bfs", str(tmp_path / "missing.txt")])
    if gsw.device_count() < 1:
        with pytest.raises(gsw.GswError):  # no GPU here: fails loudly
            cli.main(["bfs", str(p)])


def test_check_gc(gsw, orc):
    h = gsw.erdos_renyi(2000, 6.0, 9)
    c = orc.gc(h.row_ptr, h.col)
    assert cli.check_gc(h, c)
    u = np.repeat(np.arange(h.n), np.diff(h.row_ptr))
    e = int(np.flatnonzero(u < h.col)[0])
    bad = c.copy()
    bad[h.col[e]] = bad[u[e]]  # two neighbours share a colour
    assert not cli.check_gc(h, bad)
    bad = c.copy()
    bad[5] = -1  # uncoloured
    assert not cli.check_gc(h, bad)


def test_cli_cache_file(gsw, tmp_path):
    """--cache writes .csr_<name> beside the input before any GPU call."""
    p = tmp_path / "g.txt"
    p.write_text("1 2\n2 3\n3 1\n")
    try:
        cli.main(["cc", str(p), "--cache"])
    except gsw.GswError:
        pass  # no GPU in this process
    c = tmp_path / ".csr_g.txt"
    assert c.exists()
    h = gsw.load_graph(str(c))
    assert h.n == 3 and list(h.col) == [1, 2, 0, 2, 0, 1]
