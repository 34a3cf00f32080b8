"""Command-line driver with the reference applications' options
(src/utils/cmdline.cuh:141-232; application/{bfs,sssp,cc,pagerank,bc}.cu):

    python -m gswitch_amd.cli bfs graph.mtx [-r SRC] [-H] [-W] [-d] [-c CONFIGS] [-v] [-V]

    -r, --src          root vertex (default: the highest-degree vertex, where
                       the reference draws one at random)
    -H, --with-header  the first numeric line is "rows cols nnz" and is ignored
    -W, --with-weight  the file holds integer edge weights (SSSP)
    -d, --directed     directed graph (default undirected)
    -c, --configs      forced variant, e.g. Push-Queue-CM-Fused
    -v, --verbose      per-level records
    -V, --validation   check the result's defining properties on the host
    -D, --device       accepted for compatibility; the library uses GPU 0
    -j, --json         append the per-level training records (the reference's
                       feature_t::record format) to this file
        --cache        keep the parsed CSR in .csr_<file name> next to the file
                       and load it from there on later runs (the reference's
                       store_to_binfile / load_from_binfile naming)
        --iters        PageRank iterations (default 50), --damping (0.85)

The graph is read as the reference reads it (gsw_read_mtx: comments and the
optional size line skipped, self-loops dropped, ids shifted to 0, an
undirected unweighted graph renumbered without its degree-zero vertices).
SSSP on a file without weights draws symmetric weights in [1, 64] per edge
from a fixed seed, where the reference draws them with rand()
(src/utils/fileIO.cuh:148-166).

--validation checks properties that define a correct result, on the host:
BFS levels grow by at most one along every edge out of a reached vertex and
every reached vertex has an in-neighbour one level up; SSSP distances satisfy every edge's triangle
inequality and every reached vertex has a tight in-edge; CC labels agree
across every edge, no label exceeds its vertex id and every label is a
vertex labelled by itself; GC colours every vertex and no edge joins two vertices of
one colour; PageRank's residual (the change one more
iteration would make) is small after 30 or more iterations. These are
necessary conditions, not a second computation.
"""
import argparse
import json
import os
import sys

import numpy as np

import gswitch_amd as gsw

APPS = ("bfs", "sssp", "cc", "pagerank", "bc", "gc")


def edge_sources(h):
    return np.repeat(np.arange(h.n, dtype=np.int64), np.diff(h.row_ptr))


def pair_weights(h, seed=1, lim=64):
    """Symmetric weights in [1, lim] from a hash of the unordered pair."""
    u = edge_sources(h).astype(np.uint64)
    v = h.col.astype(np.uint64)
    lo, hi = np.minimum(u, v), np.maximum(u, v)
    z = (lo << np.uint64(32)) | hi
    z = z + np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    return (np.uint64(1) + z % np.uint64(lim)).astype(np.uint32)


def check_bfs(h, root, lv):
    u, v = edge_sources(h), h.col
    if lv[root] != 0 or np.any(lv[np.arange(h.n) != root] == 0):
        return False
    a, b = lv[u], lv[v]
    # an edge u -> v from a reached u: v reached, at most one level below
    # (both directions of an undirected edge: levels differ by at most one)
    if np.any((a >= 0) & ((b < 0) | (b > a + 1))):
        return False
    has_parent = np.zeros(h.n, bool)
    has_parent[v[(a >= 0) & (a == b - 1)]] = True
    reached = lv > 0
    return bool(np.all(has_parent[reached]))


def check_sssp(h, w, root, d):
    inf = np.uint32(0xFFFFFFFF)
    u, v = edge_sources(h), h.col
    if d[root] != 0:
        return False
    du = d[u].astype(np.int64)
    dv = d[v].astype(np.int64)
    fin = d[u] != inf
    if np.any(d[v][fin] == inf) or np.any(dv[fin] > du[fin] + w[fin]):
        return False
    tight = np.zeros(h.n, bool)
    tight[v[fin & (dv == du + w)]] = True
    reached = (d != inf)
    reached[root] = False
    return bool(np.all(tight[reached]))


def check_cc(h, lab):
    u, v = edge_sources(h), h.col
    ids = np.arange(h.n)
    if np.any(lab[u] != lab[v]) or np.any(lab > ids) or np.any(lab[lab] != lab):
        return False
    return True


def check_gc(h, color):
    """Every vertex coloured and no edge joins two vertices of one colour
    (the reference's validation, application/gc.cu:45-74)."""
    u, v = edge_sources(h), h.col
    return bool(np.all(color >= 1) and np.all(color[u] != color[v]))


def pagerank_residual(h, pr, damping):
    """Largest relative change one more iteration would make (the residual
    the reference's check looks at, application/pagerank.cu:44-79): value
    v <- (1 - d) + d * sum over v's row of value[u] / degree[u]."""
    deg = np.diff(h.row_ptr).astype(np.float64)
    p = pr.astype(np.float64)
    contrib = np.where(deg > 0, p / np.maximum(deg, 1), 0.0)
    acc = np.bincount(edge_sources(h), weights=contrib[h.col], minlength=h.n)
    nxt = (1 - damping) + damping * acc
    return float(np.max(np.abs(nxt - p) / np.maximum(np.abs(p), 1e-30)))


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m gswitch_amd.cli")
    ap.add_argument("app", choices=APPS)
    ap.add_argument("path")
    ap.add_argument("-r", "--src", type=int, default=-1)
    ap.add_argument("-H", "--with-header", action="store_true")
    ap.add_argument("-W", "--with-weight", action="store_true")
    ap.add_argument("-d", "--directed", action="store_true")
    ap.add_argument("-c", "--configs", default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("-V", "--validation", action="store_true")
    ap.add_argument("-D", "--device", type=int, default=0)
    ap.add_argument("-j", "--json", default=None)
    ap.add_argument("--cache", action="store_true")
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--damping", type=float, default=0.85)
    a = ap.parse_args(argv)
    if a.device != 0:
        print("note: the library runs on GPU 0", file=sys.stderr)
    cache = os.path.join(os.path.dirname(os.path.abspath(a.path)),
                         ".csr_" + os.path.basename(a.path))
    if a.cache and os.path.exists(cache):
        h = gsw.load_graph(cache)
    else:
        h = gsw.read_mtx(a.path, directed=a.directed, weighted=a.with_weight,
                         header=a.with_header)
        if a.cache:
            h.save(cache)
    print("graph: n=%d m=%d (directed CSR entries)%s" % (
        h.n, h.m, " directed" if h.directed else ""))
    w = None
    if a.app == "sssp":
        if h.directed:
            raise SystemExit("sssp: weighted directed graphs are not supported")
        w = h.w if h.weighted else pair_weights(h)
        g = gsw.Graph(row_ptr=h.row_ptr, col=h.col, w=w)
    else:
        g = gsw.Graph(h)
    if a.json:
        gsw.record_features(a.json)
    root = a.src if a.src >= 0 else h.max_degree_vertex()
    if a.app in ("bfs", "sssp", "bc") and not 0 <= root < h.n:
        raise SystemExit("root %d out of range [0, %d)" % (root, h.n))
    ok = None
    if a.app == "bfs":
        res = g.bfs(root, configs=a.configs)
        if a.validation:
            ok = check_bfs(h, root, res)
    elif a.app == "sssp":
        res = g.sssp(root, configs=a.configs)
        if a.validation:
            ok = check_sssp(h, w, root, res)
    elif a.app == "cc":
        res = g.cc(configs=a.configs)
        if a.validation:
            ok = check_cc(h, res)
    elif a.app == "pagerank":
        res = g.pagerank(a.damping, a.iters, configs=a.configs)
        if a.validation:
            r = pagerank_residual(h, res, a.damping)
            print("pagerank residual (max relative change of one more iteration): %.3g" % r)
            if a.iters >= 30:
                # each iteration shrinks the change by about the damping
                # factor: 0.85^30 ~ 8e-3
                ok = r < 0.05
    elif a.app == "gc":
        res = g.gc(configs=a.configs)
        if a.validation:
            ok = check_gc(h, res)
    else:
        res = g.bc(root, configs=a.configs)
    st = g.last_stats
    print("%s: %.3f ms on the GPU, %d levels" % (a.app, st.device_ms, st.levels))
    if a.verbose:
        for r in g.records():
            print(json.dumps(r))
    if ok is not None:
        print("validation: %s" % ("PASS" if ok else "FAIL"))
    g.close()
    if a.json:
        gsw.record_features(None)
    return 0 if ok in (None, True) else 1


if __name__ == "__main__":
    sys.exit(main())
