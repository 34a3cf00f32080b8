"""CPU oracle — TEST INFRASTRUCTURE ONLY.

ctypes wrapper of liborc.so (oracle.c): plain-C restatements of GSWITCH's
CPU validation routines. Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, and only as the checker.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")


def _load():
    if not os.path.exists(LIB_PATH):
        import subprocess
        subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    return C.CDLL(LIB_PATH)


_lib = _load()
_P = C.c_void_p
for name, args, res in [
    ("orc_bfs", [C.c_int64, _P, _P, C.c_int32, _P], None),
    ("orc_spfa", [C.c_int64, _P, _P, _P, C.c_int32, _P], None),
    ("orc_dijkstra", [C.c_int64, _P, _P, _P, C.c_int32, _P], None),
    ("orc_bellman_ford", [C.c_int64, _P, _P, _P, C.c_int32, _P], None),
    ("orc_cc", [C.c_int64, _P, _P, _P], None),
    ("orc_cc_unionfind", [C.c_int64, _P, _P, _P], None),
    ("orc_gc", [C.c_int64, _P, _P, _P], None),
    ("orc_pagerank", [C.c_int64, _P, _P, C.c_double, C.c_int, _P], None),
    ("orc_bc", [C.c_int64, _P, _P, C.c_int32, _P, _P, _P], None),
    ("orc_msbfs64", [C.c_int64, _P, _P, _P, C.c_int, _P], None),
    ("orc_traversed_edges", [C.c_int64, _P, _P], C.c_int64),
]:
    f = getattr(_lib, name)
    f.argtypes = args
    f.restype = res


def _p(a):
    return a.ctypes.data_as(_P)


def _csr(row_ptr, col):
    return (np.ascontiguousarray(row_ptr, np.int64),
            np.ascontiguousarray(col, np.int32))


def bfs(row_ptr, col, root):
    rp, cl = _csr(row_ptr, col)
    out = np.empty(len(rp) - 1, np.int32)
    _lib.orc_bfs(len(rp) - 1, _p(rp), _p(cl), root, _p(out))
    return out


def _sssp(fn, row_ptr, col, w, root):
    rp, cl = _csr(row_ptr, col)
    ww = np.ascontiguousarray(w, np.uint32)
    out = np.empty(len(rp) - 1, np.uint32)
    fn(len(rp) - 1, _p(rp), _p(cl), _p(ww), root, _p(out))
    return out


def spfa(row_ptr, col, w, root):
    return _sssp(_lib.orc_spfa, row_ptr, col, w, root)


def dijkstra(row_ptr, col, w, root):
    return _sssp(_lib.orc_dijkstra, row_ptr, col, w, root)


def bellman_ford(row_ptr, col, w, root):
    return _sssp(_lib.orc_bellman_ford, row_ptr, col, w, root)


def cc(row_ptr, col):
    rp, cl = _csr(row_ptr, col)
    out = np.empty(len(rp) - 1, np.int32)
    _lib.orc_cc(len(rp) - 1, _p(rp), _p(cl), _p(out))
    return out


def gc(row_ptr, col):
    """Jones-Plassmann colouring rounds with id priorities (colours from 1)."""
    rp, cl = _csr(row_ptr, col)
    out = np.empty(len(rp) - 1, np.int32)
    _lib.orc_gc(len(rp) - 1, _p(rp), _p(cl), _p(out))
    return out


def cc_unionfind(row_ptr, col):
    rp, cl = _csr(row_ptr, col)
    out = np.empty(len(rp) - 1, np.int32)
    _lib.orc_cc_unionfind(len(rp) - 1, _p(rp), _p(cl), _p(out))
    return out


def pagerank(row_ptr, col, damping=0.85, iters=50):
    rp, cl = _csr(row_ptr, col)
    out = np.empty(len(rp) - 1, np.float64)
    _lib.orc_pagerank(len(rp) - 1, _p(rp), _p(cl), damping, iters, _p(out))
    return out


def bc(row_ptr, col, root):
    rp, cl = _csr(row_ptr, col)
    n = len(rp) - 1
    lvl = np.empty(n, np.int32)
    nsp = np.empty(n, np.float64)
    dep = np.empty(n, np.float64)
    _lib.orc_bc(n, _p(rp), _p(cl), root, _p(lvl), _p(nsp), _p(dep))
    return lvl, nsp, dep


def msbfs64(row_ptr, col, sources):
    rp, cl = _csr(row_ptr, col)
    src = np.ascontiguousarray(sources, np.int32)
    out = np.empty(len(rp) - 1, np.uint64)
    _lib.orc_msbfs64(len(rp) - 1, _p(rp), _p(cl), _p(src), len(src), _p(out))
    return out


def traversed_edges(row_ptr, levels):
    rp = np.ascontiguousarray(row_ptr, np.int64)
    lv = np.ascontiguousarray(levels, np.int32)
    return _lib.orc_traversed_edges(len(rp) - 1, _p(rp), _p(lv))


_lib.orc_mlp_forward.argtypes = [_P, _P, C.c_int64, _P]
_lib.orc_mlp_forward.restype = None
_lib.orc_mlp_grad.argtypes = [_P, _P, _P, _P, C.c_int64, _P]
_lib.orc_mlp_grad.restype = C.c_double


def mlp_forward(w, x):
    w = np.ascontiguousarray(w, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 12)
    y = np.empty((len(x), 30), np.float64)
    _lib.orc_mlp_forward(_p(w), _p(x), len(x), _p(y))
    return y


def mlp_grad(w, x, t, mask):
    """(loss, grad) of sum(mask*(y-t)^2)/sum(mask) in FP64."""
    w = np.ascontiguousarray(w, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 12)
    t = np.ascontiguousarray(t, np.float64).reshape(-1, 30)
    mask = np.ascontiguousarray(mask, np.float64).reshape(-1, 30)
    g = np.empty(6942, np.float64)
    loss = _lib.orc_mlp_grad(_p(w), _p(x), _p(t), _p(mask), len(x), _p(g))
    return loss, g
