"""Applications written the way GSWITCH applications are written.

tests/apps/{bfs,sssp,cc}_app.hip use only the template API of gswitch.h in
the reference applications' call shape (global actors, a Functor,
parse_cmd / cmd_opt, build_graph from a file or an edge list, the host
level loop, p_jump, f.data.sync_wa, -V CPU check). They are compiled by
build() and run here as programs on Matrix Market files; their results
must equal the oracle's on the same file (vertex numbering as the
reference's loader gives it: rinsed for unweighted undirected BFS, kept for
SSSP and CC).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from test_host_graph import write_mtx

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "tests", "apps", "build")


def app(name):
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        pytest.fail("%s is not built (run __graft_entry__.build())" % p)
    return p


def run(name, args, out_file, timeout=120):
    env = dict(os.environ, GSW_APP_OUT=str(out_file))
    print("running", name, *args, file=sys.stderr, flush=True)
    r = subprocess.run([app(name)] + [str(a) for a in args], env=env,
                       capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    return r.stdout


def test_apps_print_usage_without_a_graph():
    """argv[1] missing: the usage text, exit status 0 (parse_cmd), before any
    GPU call."""
    for name in ("bfs_app", "sssp_app", "cc_app"):
        r = subprocess.run([app(name)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0
        assert "--configs" in r.stdout and "--validation" in r.stdout


@pytest.fixture(scope="module")
def mtx(gsw, tmp_path_factory):
    d = tmp_path_factory.mktemp("apps")
    files = {}
    h = gsw.rmat(12, 16, seed=5)
    files["rmat"] = str(d / "rmat.mtx")
    write_mtx(files["rmat"], h)
    g = gsw.grid2d(60, 70, 0.15, 0.01, 3)
    files["grid"] = str(d / "grid.mtx")
    write_mtx(files["grid"], g)
    # weighted file: third column = weight
    hw = gsw.rmat(11, 8, seed=7, weighted=True)
    u = np.repeat(np.arange(hw.n), np.diff(hw.row_ptr))
    keep = u < hw.col
    files["weighted"] = str(d / "weighted.mtx")
    with open(files["weighted"], "w") as f:
        f.write("%d %d %d\n" % (hw.n, hw.n, int(keep.sum())))
        np.savetxt(f, np.stack([u[keep] + 1, hw.col[keep] + 1, hw.w[keep]], 1),
                   fmt="%d")
    files["dir"] = d
    return files


@pytest.mark.gpu
@pytest.mark.parametrize("graph", ["rmat", "grid"])
@pytest.mark.parametrize("configs", [None, "Push-Queue-WM-Fused",
                                     "Pull-Bitmap-TM-Standalone",
                                     "Push-SortedQueue-STRICT-Standalone"])
def test_bfs_application(gsw, orc, mtx, graph, configs):
    h = gsw.read_mtx(mtx[graph], header=True)  # rinsed, as build_graph<VC> loads it
    root = h.max_degree_vertex()
    out = mtx["dir"] / ("bfs_%s.bin" % graph)
    args = [mtx[graph], "-H", "-V", "-r", root] + (["-c", configs] if configs else [])
    stdout = run("bfs_app", args, out)
    assert "passed" in stdout
    got = np.fromfile(out, np.int32)
    assert got[0] == root
    assert np.array_equal(got[1:], orc.bfs(h.row_ptr, h.col, root))


@pytest.mark.gpu
def test_bfs_application_random_root(gsw, orc, mtx):
    """No -r: random_root() picks a vertex with edges; the result is that
    root's BFS."""
    h = gsw.read_mtx(mtx["rmat"], header=True)
    out = mtx["dir"] / "bfs_rand.bin"
    assert "passed" in run("bfs_app", [mtx["rmat"], "-H", "-V"], out)
    got = np.fromfile(out, np.int32)
    root = int(got[0])
    assert h.degrees[root] > 0
    assert np.array_equal(got[1:], orc.bfs(h.row_ptr, h.col, root))


@pytest.mark.gpu
@pytest.mark.parametrize("configs", [None, "Push-Queue-WM-Standalone",
                                     "Pull-Bitmap-STRICT-Standalone",
                                     "Push-Queue-TWC-Fused"])
def test_sssp_application(gsw, orc, mtx, configs):
    h = gsw.read_mtx(mtx["weighted"], weighted=True, header=True)
    root = h.max_degree_vertex()
    out = mtx["dir"] / "sssp.bin"
    args = [mtx["weighted"], "-H", "-W", "-V", "-r", root] + \
        (["-c", configs] if configs else [])
    stdout = run("sssp_app", args, out)
    assert "passed" in stdout
    got = np.fromfile(out, np.int32)
    ref = orc.spfa(h.row_ptr, h.col, h.w, root).astype(np.int64)
    ref[ref == 2**32 - 1] = 2**31 - 1  # the application's unreached value
    assert got[0] == root
    assert np.array_equal(got[1:].astype(np.int64), ref)


@pytest.mark.gpu
def test_sssp_application_generated_weights(mtx):
    """An unweighted file: gen_weight(64) draws the weights; the
    application's own CPU check must agree with the GPU result."""
    out = mtx["dir"] / "sssp_gen.bin"
    assert "passed" in run("sssp_app", [mtx["grid"], "-H", "-V", "-r", 0], out)


@pytest.mark.gpu
@pytest.mark.parametrize("graph", ["rmat", "grid"])
def test_cc_application(gsw, orc, mtx, graph):
    h = gsw.read_mtx(mtx[graph], header=True, rinse=False)  # COO keeps the ids
    out = mtx["dir"] / ("cc_%s.bin" % graph)
    stdout = run("cc_app", [mtx[graph], "-H", "-V"], out)
    assert "passed" in stdout
    got = np.fromfile(out, np.int32)
    assert np.array_equal(got, orc.cc(h.row_ptr, h.col))
