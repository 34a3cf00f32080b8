"""Python binding of the MI355X traversal engine (ctypes over the C-ABI).

The engine itself is C++/HIP (gswitch_amd/csrc, C-ABI in
include/gswitch_amd.h); this module only loads libgswitch_amd.so and turns
numpy arrays into pointers. There is no fallback: if the library is missing
the import fails.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgswitch_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "gswitch_amd.h")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "libgswitch_amd.so is not built; run `python -m gswitch_amd.build` "
        "(or __graft_entry__.build())")

_lib = C.CDLL(LIB_PATH)

OK = 0
STATUS = {0: "OK", 1: "INVALID", 2: "HIP", 3: "ALLOC", 4: "OVERFLOW",
          5: "UNSUPPORTED", 6: "INTERNAL"}
LB_NAMES = ["TM", "WM", "CM", "STRICT", "TWOD", "ELB", "TWC"]


class GswError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (STATUS.get(code, code), msg))
        self.code = code


class RunOpts(C.Structure):
    _fields_ = [("configs", C.c_char_p), ("max_levels", C.c_int32),
                ("verbose", C.c_int32), ("twc_small", C.c_int32),
                ("twc_large", C.c_int32), ("profile", C.c_int32),
                ("pull_kernel", C.c_int32), ("capture", C.c_int32),
                ("grid", C.c_int32), ("selector", C.c_int32)]


class RunStats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("levels", C.c_int32),
                ("n_records", C.c_int32), ("reached", C.c_int64)]


class LevelRecord(C.Structure):
    _fields_ = [("level", C.c_int32), ("dir", C.c_int8), ("fmt", C.c_int8),
                ("lb", C.c_int8), ("fused", C.c_int8), ("sorted", C.c_int8),
                ("kernel", C.c_int8), ("pad", C.c_int8 * 2), ("frontier", C.c_int64),
                ("active", C.c_int64), ("inactive", C.c_int64),
                ("edges", C.c_int64), ("verts", C.c_int64), ("ms", C.c_float),
                ("pad2", C.c_float)]


_P = C.c_void_p
_SIGS = {
    "gsw_version": (C.c_char_p, []),
    "gsw_last_error": (C.c_char_p, []),
    "gsw_device_count": (C.c_int32, []),
    "gsw_run_opts_init": (None, [C.POINTER(RunOpts)]),
    "gsw_host_graph_from_edges": (C.c_int32, [C.c_int64, C.c_int64, _P, _P, _P,
                                              C.c_int32, C.POINTER(_P)]),
    "gsw_gen_rmat": (C.c_int32, [C.c_int32, C.c_int32, C.c_double, C.c_double,
                                 C.c_double, C.c_uint64, C.c_int32, C.POINTER(_P)]),
    "gsw_gen_grid2d": (C.c_int32, [C.c_int64, C.c_int64, C.c_double, C.c_double,
                                   C.c_uint64, C.c_int32, C.POINTER(_P)]),
    "gsw_gen_er": (C.c_int32, [C.c_int64, C.c_double, C.c_uint64, C.c_int32,
                               C.POINTER(_P)]),
    "gsw_gen_ba": (C.c_int32, [C.c_int64, C.c_int32, C.c_uint64, C.c_int32,
                               C.POINTER(_P)]),
    "gsw_record_features": (C.c_int32, [C.c_char_p]),
    "gsw_host_graph_save": (C.c_int32, [_P, C.c_char_p]),
    "gsw_host_graph_load": (C.c_int32, [C.c_char_p, C.POINTER(_P)]),
    "gsw_read_mtx": (C.c_int32, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                 C.POINTER(_P)]),
    "gsw_gen_er_batch": (C.c_int32, [C.c_int32, C.c_int64, C.c_int64, C.c_double,
                                     C.c_double, C.c_uint64, C.c_int32,
                                     C.POINTER(_P), _P]),
    "gsw_gen_er_batch_range": (C.c_int32, [C.c_int64, C.c_int32, C.c_int64, C.c_int64,
                                           C.c_double, C.c_double, C.c_uint64,
                                           C.c_int32, C.POINTER(_P), _P]),
    "gsw_host_graph_info": (C.c_int32, [_P, C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int32)]),
    "gsw_host_graph_export": (C.c_int32, [_P, _P, _P, _P]),
    "gsw_host_graph_destroy": (None, [_P]),
    "gsw_degree_stats": (C.c_int32, [_P, C.c_int64, _P]),
    "gsw_graph_create_csr": (C.c_int32, [_P, _P, _P, C.c_int64, C.c_int64,
                                         C.c_int32, C.POINTER(_P)]),
    "gsw_graph_from_host": (C.c_int32, [_P, C.POINTER(_P)]),
    "gsw_graph_destroy": (C.c_int32, [_P]),
    "gsw_bfs": (C.c_int32, [_P, C.c_int32, C.POINTER(RunOpts), _P,
                            C.POINTER(RunStats)]),
    "gsw_sssp": (C.c_int32, [_P, C.c_int32, C.POINTER(RunOpts), _P,
                             C.POINTER(RunStats)]),
    "gsw_cc": (C.c_int32, [_P, C.POINTER(RunOpts), _P, C.POINTER(RunStats)]),
    "gsw_cc_noalloc": (C.c_int32, [_P, C.c_int32, _P, _P]),
    "gsw_pagerank": (C.c_int32, [_P, C.c_float, C.c_int32, C.c_float,
                                 C.POINTER(RunOpts), _P, C.POINTER(RunStats)]),
    "gsw_gc": (C.c_int32, [_P, C.POINTER(RunOpts), _P, C.POINTER(RunStats)]),
    "gsw_delta_pagerank": (C.c_int32, [_P, C.c_float, C.c_double, C.c_int32,
                                       C.POINTER(RunOpts), _P, C.POINTER(RunStats)]),
    "gsw_bc": (C.c_int32, [_P, C.c_int32, C.POINTER(RunOpts), _P, _P, _P,
                           C.POINTER(RunStats)]),
    "gsw_last_records": (C.c_int32, [_P, _P, C.c_int32, C.POINTER(C.c_int32)]),
    "gsw_msbfs64": (C.c_int32, [_P, _P, C.c_int32, C.POINTER(RunOpts), _P, _P,
                                C.POINTER(RunStats)]),
    "gsw_bfs_multi": (C.c_int32, [_P, _P, C.c_int32, C.POINTER(RunOpts), _P,
                                  C.POINTER(RunStats)]),
    "gsw_batch_bfs_cc": (C.c_int32, [_P, _P, C.c_int32, C.POINTER(RunOpts), _P, _P,
                                     C.POINTER(RunStats)]),
    "gsw_autotune_bfs": (C.c_int32, [_P, _P, C.c_int32, C.c_int32,
                                     C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_double)]),
    "gsw_selector_load": (C.c_int32, [_P, C.c_int32]),
    "gsw_selector_load_app": (C.c_int32, [C.c_int32, _P, C.c_int32]),
    "gsw_last_class": (C.c_int32, [_P, _P]),
    "gsw_sssp_set_window": (C.c_int32, [_P, C.c_double]),
    "gsw_sssp_get_window": (C.c_int32, [_P, _P]),
    "gsw_autotune_sssp": (C.c_int32, [_P, _P, C.c_int32, _P, _P, _P]),
    "gsw_selector_load_tree": (C.c_int32, [C.c_int32, C.c_int32, _P, _P, _P, _P, _P]),
    "gsw_mlp_features": (C.c_int32, [_P, _P]),
    "gsw_mlp_forward": (C.c_int32, [_P, _P, C.c_int64, _P]),
    "gsw_mlp_select": (C.c_int32, [_P, _P, C.c_int64, C.c_uint32, _P]),
    "gsw_mlp_grad": (C.c_int32, [_P, _P, _P, _P, C.c_int64, C.POINTER(C.c_float),
                                 _P]),
    "gsw_mlp_train": (C.c_int32, [_P, _P, _P, C.c_int64, C.c_int32, C.c_int32,
                                  C.c_float, C.c_uint64, _P, _P]),
    "gsw_explore_bfs": (C.c_int32, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                    _P, _P, _P, _P, C.POINTER(C.c_int32)]),
    "gsw_explore_sssp": (C.c_int32, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                     _P, _P, _P, _P, C.POINTER(C.c_int32)]),
    "gsw_tree_eval": (C.c_int32, [C.c_char_p, _P, C.c_int32,
                                  C.POINTER(C.c_int32)]),
    "gsw_scan_exclusive_i64": (C.c_int32, [_P, C.c_int64, _P]),
    "gsw_wave_scan64": (C.c_int32, [_P, _P]),
    "gsw_stream_copy": (C.c_int32, [C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
}
for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(_lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def exported_symbols():
    return sorted(_SIGS)


def _check(rc):
    if rc != OK:
        raise GswError(rc, _lib.gsw_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def version():
    return _lib.gsw_version().decode()


def device_count():
    return _lib.gsw_device_count()


# --------------------------------------------------------------- host graphs
class HostGraph:
    """Host CSR built by the library (generators, edge-list ingest)."""

    def __init__(self, handle):
        self._h = handle
        n, m = C.c_int64(), C.c_int64()
        d, w = C.c_int32(), C.c_int32()
        _check(_lib.gsw_host_graph_info(self._h, C.byref(n), C.byref(m),
                                        C.byref(d), C.byref(w)))
        self.n, self.m = n.value, m.value
        self.directed, self.weighted = bool(d.value), bool(w.value)
        self.row_ptr = np.empty(self.n + 1, np.int64)
        self.col = np.empty(max(self.m, 1), np.int32)[: self.m]
        self.w = np.empty(max(self.m, 1), np.uint32)[: self.m] if self.weighted else None
        _check(_lib.gsw_host_graph_export(self._h, _ptr(self.row_ptr),
                                          _ptr(self.col), _ptr(self.w)))

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.gsw_host_graph_destroy(self._h)
            self._h = None

    @property
    def degrees(self):
        return np.diff(self.row_ptr)

    def save(self, path):
        """Write this CSR to a binary cache file (load it with load_graph)."""
        _check(_lib.gsw_host_graph_save(self._h, os.fsencode(path)))

    def max_degree_vertex(self):
        """Highest-degree vertex, ties broken by the smallest id."""
        return int(np.argmax(self.degrees))


def _gen(fn, *args):
    h = _P()
    _check(fn(*args, C.byref(h)))
    return HostGraph(h)


def rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, weighted=False):
    return _gen(_lib.gsw_gen_rmat, scale, edge_factor, a, b, c, seed, int(weighted))


def grid2d(rows, cols, p_delete=0.0, long_range_frac=0.0, seed=3, weighted=False):
    return _gen(_lib.gsw_gen_grid2d, rows, cols, p_delete, long_range_frac, seed,
                int(weighted))


def erdos_renyi(n, mean_degree, seed=5, weighted=False):
    return _gen(_lib.gsw_gen_er, n, mean_degree, seed, int(weighted))


def barabasi_albert(n, m, seed=6, weighted=False):
    return _gen(_lib.gsw_gen_ba, n, m, seed, int(weighted))


def load_graph(path):
    """A HostGraph from a binary CSR cache file written by HostGraph.save."""
    h = _P()
    _check(_lib.gsw_host_graph_load(os.fsencode(path), C.byref(h)))
    return HostGraph(h)


def record_features(path):
    """Append the reference's per-level training records (JSON objects with
    keys "0".."22") to `path` from now on, at most 100; None turns it off."""
    _check(_lib.gsw_record_features(None if path is None else os.fsencode(path)))


def read_records(path):
    """The JSON objects a record file holds, in order."""
    import json
    txt, dec, out, i = open(path).read(), json.JSONDecoder(), [], 0
    while True:
        while i < len(txt) and txt[i].isspace():
            i += 1
        if i >= len(txt):
            return out
        obj, i = dec.raw_decode(txt, i)
        out.append(obj)


def read_mtx(path, directed=False, weighted=False, header=False, rinse=None):
    """Matrix Market / edge-list file -> HostGraph (edgelist_t::read_mtx +
    host_graph_t<CSR>::build, src/utils/fileIO.cuh:73-332). rinse=None does
    what the reference does: drop degree-zero vertices and renumber, for
    undirected unweighted graphs only."""
    h = _P()
    r = -1 if rinse is None else int(bool(rinse))
    _check(_lib.gsw_read_mtx(os.fsencode(path), int(directed), int(weighted), int(header),
                             r, C.byref(h)))
    return HostGraph(h)


def erdos_renyi_batch(count, nmin=1000, nmax=200000, dmin=2.0, dmax=32.0, seed=5,
                      weighted=False, first=0):
    """Instances first .. first+count-1 of an ER batch as one disjoint union
    numbered from vertex 0: (HostGraph, offsets[count+1])."""
    off = np.empty(max(count, 0) + 1, np.int64)
    h = _P()
    _check(_lib.gsw_gen_er_batch_range(first, count, nmin, nmax, dmin, dmax, seed,
                                       int(weighted), C.byref(h), _ptr(off)))
    return HostGraph(h), off


def from_edges(n, src, dst, w=None, directed=False):
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    wv = None if w is None else np.ascontiguousarray(w, np.uint32)
    h = _P()
    _check(_lib.gsw_host_graph_from_edges(n, len(src), _ptr(src), _ptr(dst),
                                          _ptr(wv), int(directed), C.byref(h)))
    return HostGraph(h)


def degree_stats(row_ptr):
    row_ptr = np.ascontiguousarray(row_ptr, np.int64)
    out = np.zeros(6)
    _check(_lib.gsw_degree_stats(_ptr(row_ptr), len(row_ptr) - 1, _ptr(out)))
    return dict(zip(["avg", "std", "max", "range", "GI", "Her"], out))


def tree_eval(name, fv):
    fv = np.ascontiguousarray(fv, np.float64)
    cls = C.c_int32()
    _check(_lib.gsw_tree_eval(name.encode(), _ptr(fv), len(fv), C.byref(cls)))
    return cls.value


# --------------------------------------------------------------- learned selector
MLP_NPARAM, MLP_NIN, MLP_NOUT = 6942, 12, 30
MLP_GROUPS = ["Push-Queue-%s-Standalone", "Push-SortedQueue-%s-Standalone",
              "Push-Bitmap-%s-Standalone", "Push-Queue-%s-Fused",
              "Pull-Queue-%s-Standalone", "Pull-Bitmap-%s-Standalone"]
MLP_LBS = ["TM", "WM", "CM", "STRICT", "TWC"]
# class id -> the forced-variant string of that class
MLP_CLASSES = [g % lb for g in MLP_GROUPS for lb in MLP_LBS]


def _f32(a, n=None):
    a = np.ascontiguousarray(a, np.float32)
    if n is not None and a.size != n:
        raise ValueError("expected %d values" % n)
    return a


SELECTOR_APPS = {"bfs": 0, "sssp": 1, "pr": 2, "cc": 3}
# PageRank's learned classes (the network's first 8 outputs): one from-all
# pull for the whole run, as (forced variant, pull_kernel)
PR_CLASSES = [("Pull-Bitmap-TM-Standalone", 2), ("Pull-Bitmap-WM-Standalone", 2),
              ("Pull-Bitmap-CM-Standalone", 2), ("Pull-Bitmap-STRICT-Standalone", 0),
              ("Pull-Bitmap-STRICT-Standalone", 1), ("Pull-Bitmap-TWOD-Standalone", 0),
              ("Pull-Bitmap-TWOD-Standalone", 1), ("Pull-Bitmap-TWOD-Standalone", 2)]
PR_CLASS_MASK = (1 << len(PR_CLASSES)) - 1
# CC's learned classes (the first 4 outputs): sampled hooking (the default
# run), or the edge-centric level loop with the rule's frontier formats, a
# bitmap or an edge queue, as forced configs
CC_CLASSES = [None, "x-x-x-x", "x-Bitmap-x-x", "x-Queue-x-x"]
CC_CLASS_MASK = (1 << len(CC_CLASSES)) - 1


def pr_class_of_record(r):
    """PageRank class of one expand record (a from-all pull)."""
    lb = r["lb"] if r["lb"] in ("TM", "WM", "CM", "STRICT", "TWOD") else "STRICT"
    name = "Pull-Bitmap-%s-Standalone" % lb
    if lb in ("TM", "WM", "CM"):
        return [c for c, _ in PR_CLASSES].index(name)
    pk = 0 if r["kernel"] == 2 and lb == "STRICT" else r["kernel"]
    return PR_CLASSES.index((name, pk))


def selector_load(w, app="bfs"):
    """Install an application's learned selector weights, used by its runs
    with selector=1 ("bfs", "sssp", "pr" or "cc")."""
    if app not in SELECTOR_APPS:
        raise ValueError("no learned selector for %r" % (app,))
    w = _f32(w, MLP_NPARAM)
    _check(_lib.gsw_selector_load_app(SELECTOR_APPS[app], _ptr(w), MLP_NPARAM))


def selector_load_tree(tree, app="bfs"):
    """Install an application's class tree, used by its runs with
    selector=2. `tree` is a dict of equal-length node arrays feat, thr,
    left, right, cls (gswitch_amd.learn.fit_tree / load_tree)."""
    if app not in SELECTOR_APPS:
        raise ValueError("no class-tree selector for %r" % (app,))
    feat = np.ascontiguousarray(tree["feat"], np.int32)
    n = feat.size
    thr = np.ascontiguousarray(tree["thr"], np.float32)
    left = np.ascontiguousarray(tree["left"], np.int32)
    right = np.ascontiguousarray(tree["right"], np.int32)
    cls = np.ascontiguousarray(tree["cls"], np.int32)
    if not (thr.size == left.size == right.size == cls.size == n):
        raise ValueError("tree arrays differ in length")
    _check(_lib.gsw_selector_load_tree(SELECTOR_APPS[app], n, _ptr(feat), _ptr(thr),
                                       _ptr(left), _ptr(right), _ptr(cls)))


def mlp_features(counts):
    """12 MLP inputs from (n, m, avg, std, gini, level, active, inactive,
    push workload, pull workload, max active degree)."""
    c = np.ascontiguousarray(counts, np.float64)
    if c.size != 11:
        raise ValueError("expected 11 counts")
    x = np.empty(MLP_NIN, np.float32)
    _check(_lib.gsw_mlp_features(_ptr(c), _ptr(x)))
    return x


def mlp_forward(w, x):
    w, x = _f32(w, MLP_NPARAM), _f32(x).reshape(-1, MLP_NIN)
    y = np.empty((len(x), MLP_NOUT), np.float32)
    _check(_lib.gsw_mlp_forward(_ptr(w), _ptr(x), len(x), _ptr(y)))
    return y


def mlp_select(w, x, mask=(1 << 30) - 1):
    w, x = _f32(w, MLP_NPARAM), _f32(x).reshape(-1, MLP_NIN)
    cls = np.empty(len(x), np.int32)
    _check(_lib.gsw_mlp_select(_ptr(w), _ptr(x), len(x), mask, _ptr(cls)))
    return cls


def mlp_grad(w, x, t, mask):
    """(loss, grad[6942]) of sum(mask*(y-t)^2)/sum(mask), on the device."""
    w, x = _f32(w, MLP_NPARAM), _f32(x).reshape(-1, MLP_NIN)
    t = _f32(t).reshape(-1, MLP_NOUT)
    mask = _f32(mask).reshape(-1, MLP_NOUT)
    g = np.empty(MLP_NPARAM, np.float32)
    loss = C.c_float()
    _check(_lib.gsw_mlp_grad(_ptr(w), _ptr(x), _ptr(t), _ptr(mask), len(x),
                             C.byref(loss), _ptr(g)))
    return loss.value, g


def mlp_train(w0, x, t, mask, epochs=200, batch=256, lr=1e-3, seed=7):
    """Adam training on the device; returns (weights, loss per epoch)."""
    w = _f32(w0, MLP_NPARAM).copy()
    x = _f32(x).reshape(-1, MLP_NIN)
    t = _f32(t).reshape(-1, MLP_NOUT)
    mask = _f32(mask).reshape(-1, MLP_NOUT)
    curve = np.zeros(max(epochs, 1), np.float32)
    _check(_lib.gsw_mlp_train(_ptr(x), _ptr(t), _ptr(mask), len(x), epochs, batch,
                              lr, seed, _ptr(w), _ptr(curve)))
    return w, curve[:epochs]


# --------------------------------------------------------------- device graphs
def run_opts(configs=None, max_levels=0, verbose=False, twc_small=0,
             twc_large=0, profile=True, pull_kernel=2, capture=0, grid=0,
             selector=0):
    o = RunOpts()
    _lib.gsw_run_opts_init(C.byref(o))
    o.configs = configs.encode() if configs else None
    if max_levels:
        o.max_levels = max_levels
    o.verbose = int(verbose)
    if twc_small:
        o.twc_small = twc_small
    if twc_large:
        o.twc_large = twc_large
    o.profile = int(profile)
    o.pull_kernel = int(pull_kernel)
    o.capture = int(capture)
    o.grid = int(grid)
    o.selector = int(selector)
    return o


class Graph:
    """Device-resident graph; runs reuse its frontier and vertex arrays."""

    def __init__(self, host=None, row_ptr=None, col=None, w=None):
        h = _P()
        if host is not None:
            _check(_lib.gsw_graph_from_host(host._h, C.byref(h)))
            self.n, self.m = host.n, host.m
        else:
            row_ptr = np.ascontiguousarray(row_ptr, np.int64)
            col = np.ascontiguousarray(col, np.int32)
            wv = None if w is None else np.ascontiguousarray(w, np.uint32)
            self.n, self.m = len(row_ptr) - 1, len(col)
            _check(_lib.gsw_graph_create_csr(_ptr(row_ptr), _ptr(col), _ptr(wv),
                                             self.n, self.m, 0, C.byref(h)))
        self._h = h
        self.last_stats = None

    def close(self):
        if getattr(self, "_h", None):
            _check(_lib.gsw_graph_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _opts(self, configs, kw):
        return run_opts(configs=configs, **kw)

    def bfs(self, root, configs=None, out=True, **kw):
        o = self._opts(configs, kw)
        st = RunStats()
        lv = np.empty(self.n, np.int32) if out else None
        _check(_lib.gsw_bfs(self._h, int(root), C.byref(o), _ptr(lv), C.byref(st)))
        self.last_stats = st
        return lv

    def sssp(self, root, configs=None, out=True, **kw):
        o = self._opts(configs, kw)
        st = RunStats()
        d = np.empty(self.n, np.uint32) if out else None
        _check(_lib.gsw_sssp(self._h, int(root), C.byref(o), _ptr(d), C.byref(st)))
        self.last_stats = st
        return d

    def cc(self, configs=None, out=True, **kw):
        o = self._opts(configs, kw)
        st = RunStats()
        lab = np.empty(self.n, np.int32) if out else None
        _check(_lib.gsw_cc(self._h, C.byref(o), _ptr(lab), C.byref(st)))
        self.last_stats = st
        return lab

    def cc_noalloc(self, labels_ptr, samples=1024, stream=None):
        """Allocation-free CC into device memory at `labels_ptr` (n int32),
        queued on `stream` (a hipStream_t as an int; None: the null
        stream); returns before the kernels finish."""
        _check(_lib.gsw_cc_noalloc(self._h, int(samples), C.c_void_p(int(labels_ptr)),
                                   C.c_void_p(stream or 0)))

    def pagerank(self, damping=0.85, iters=50, eps=0.0, configs=None, out=True, **kw):
        o = self._opts(configs, kw)
        st = RunStats()
        pr = np.empty(self.n, np.float32) if out else None
        _check(_lib.gsw_pagerank(self._h, damping, iters, eps, C.byref(o), _ptr(pr),
                                 C.byref(st)))
        self.last_stats = st
        return pr

    def delta_pagerank(self, damping=0.85, eps=1e-9, max_iters=1000, configs=None,
                       out=True, **kw):
        """Delta-PageRank: the converged PageRank values (same form as
        pagerank()), sending only changes larger than eps."""
        o = self._opts(configs, kw)
        st = RunStats()
        pr = np.empty(self.n, np.float32) if out else None
        _check(_lib.gsw_delta_pagerank(self._h, damping, eps, max_iters, C.byref(o),
                                       _ptr(pr), C.byref(st)))
        self.last_stats = st
        return pr

    def gc(self, configs=None, out=True, **kw):
        """Graph colouring (Jones-Plassmann, id priorities): colour per vertex."""
        o = self._opts(configs, kw)
        st = RunStats()
        col = np.empty(self.n, np.int32) if out else None
        _check(_lib.gsw_gc(self._h, C.byref(o), _ptr(col), C.byref(st)))
        self.last_stats = st
        return col

    def bc(self, root, configs=None, out=True, **kw):
        """Betweenness dependencies from `root`: (levels, nsp, dep)."""
        o = self._opts(configs, kw)
        st = RunStats()
        lv = np.empty(self.n, np.int32) if out else None
        nsp = np.empty(self.n, np.float64) if out else None
        dep = np.empty(self.n, np.float64) if out else None
        _check(_lib.gsw_bc(self._h, int(root), C.byref(o), _ptr(lv), _ptr(nsp),
                           _ptr(dep), C.byref(st)))
        self.last_stats = st
        return lv, nsp, dep

    def msbfs64(self, sources, out=True, **kw):
        """64-source bitmask BFS: (masks uint64[n], per-source level sums)."""
        src = np.ascontiguousarray(sources, np.int32)
        o = run_opts(**kw)
        st = RunStats()
        masks = np.empty(self.n, np.uint64) if out else None
        sums = np.zeros(64, np.uint64)
        _check(_lib.gsw_msbfs64(self._h, _ptr(src), len(src), C.byref(o),
                                _ptr(masks), _ptr(sums), C.byref(st)))
        self.last_stats = st
        return masks, sums

    def bfs_multi(self, roots, configs=None, out=True, **kw):
        """BFS levels from the nearest of `roots`; capture=L replays a captured
        loop of 2L fused levels per launch."""
        r = np.ascontiguousarray(roots, np.int32)
        o = self._opts(configs, kw)
        st = RunStats()
        lv = np.empty(self.n, np.int32) if out else None
        _check(_lib.gsw_bfs_multi(self._h, _ptr(r), len(r), C.byref(o), _ptr(lv),
                                  C.byref(st)))
        self.last_stats = st
        return lv

    def batch_bfs_cc(self, offsets, configs=None, out=True, **kw):
        """BFS from every instance's first vertex and CC of a disjoint union
        (instance offsets[i] .. offsets[i+1]-1): (levels, labels)."""
        off = np.ascontiguousarray(offsets, np.int64)
        kw.setdefault("capture", 8)
        o = self._opts(configs, kw)
        st = RunStats()
        lv = np.empty(self.n, np.int32) if out else None
        lab = np.empty(self.n, np.int32) if out else None
        _check(_lib.gsw_batch_bfs_cc(self._h, _ptr(off), len(off) - 1, C.byref(o),
                                     _ptr(lv), _ptr(lab), C.byref(st)))
        self.last_stats = st
        return lv, lab

    def explore_bfs(self, root, max_levels=64, stride=1, reps=2):
        """Exhaustive per-level timing of the 30 variant classes for BFS from
        `root`: dict(features[L,12], ms[L,30], rule[L], level[L])."""
        return self._explore(_lib.gsw_explore_bfs, root, max_levels, stride, reps)

    def explore_sssp(self, root, max_levels=64, stride=1, reps=2):
        """The same for SSSP from `root` (weighted graphs)."""
        return self._explore(_lib.gsw_explore_sssp, root, max_levels, stride, reps)

    def _explore(self, fn, root, max_levels, stride, reps):
        feat = np.zeros((max_levels, MLP_NIN), np.float32)
        ms = np.zeros((max_levels, MLP_NOUT), np.float32)
        rule = np.zeros(max_levels, np.int32)
        lvl = np.zeros(max_levels, np.int32)
        k = C.c_int32()
        _check(fn(self._h, int(root), max_levels, stride, reps, _ptr(feat), _ptr(ms),
                  _ptr(rule), _ptr(lvl), C.byref(k)))
        k = k.value
        return dict(features=feat[:k], ms=ms[:k], rule=rule[:k], level=lvl[:k])

    def autotune_bfs(self, roots, capture=8):
        """Fastest captured-loop shape for these roots:
        dict(lb, twc_small, twc_large, grid, ms)."""
        r = np.ascontiguousarray(roots, np.int32)
        lb, s, l, g = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        ms = C.c_double()
        _check(_lib.gsw_autotune_bfs(self._h, _ptr(r), len(r), capture, C.byref(lb),
                                     C.byref(s), C.byref(l), C.byref(g),
                                     C.byref(ms)))
        return dict(lb=LB_NAMES[lb.value], twc_small=s.value, twc_large=l.value,
                    grid=g.value, ms=ms.value)

    def sssp_window(self, scale=None):
        """The graph's SSSP window scale; with `scale`, set it first."""
        if scale is not None:
            _check(_lib.gsw_sssp_set_window(self._h, float(scale)))
        k = C.c_double()
        _check(_lib.gsw_sssp_get_window(self._h, C.byref(k)))
        return k.value

    def autotune_sssp(self, roots, configs=None, **kw):
        """Pick (and keep) the SSSP window scale that is fastest from
        `roots`; returns (scale, total ms)."""
        r = np.ascontiguousarray(roots, np.int32)
        o = self._opts(configs, kw)
        k, ms = C.c_double(), C.c_double()
        _check(_lib.gsw_autotune_sssp(self._h, _ptr(r), len(r), C.byref(o), C.byref(k),
                                      C.byref(ms)))
        return k.value, ms.value

    def last_class(self):
        """Class chosen by the learned selector in the latest PageRank or CC
        run (PR_CLASSES / CC_CLASSES index), -1 if none."""
        c = C.c_int32()
        _check(_lib.gsw_last_class(self._h, C.byref(c)))
        return c.value

    def records(self):
        cap = max(1, self.last_stats.n_records if self.last_stats else 1 << 16)
        buf = (LevelRecord * cap)()
        n = C.c_int32()
        _check(_lib.gsw_last_records(self._h, buf, cap, C.byref(n)))
        out = []
        for r in buf[: n.value]:
            out.append(dict(level=r.level, dir="Pull" if r.dir else "Push",
                            fmt="Bitmap" if r.fmt else "Queue",
                            lb=LB_NAMES[r.lb], fused=bool(r.fused),
                            sorted=bool(r.sorted), kernel=int(r.kernel),
                            frontier=r.frontier,
                            active=r.active, inactive=r.inactive, edges=r.edges,
                            verts=r.verts, ms=r.ms))
        return out


def scan_exclusive_i64(x):
    x = np.ascontiguousarray(x, np.int64)
    out = np.empty(len(x) + 1, np.int64)
    _check(_lib.gsw_scan_exclusive_i64(_ptr(x), len(x), _ptr(out)))
    return out


def wave_scan64(x):
    x = np.ascontiguousarray(x, np.int32)
    assert len(x) == 64
    out = np.empty(64, np.int32)
    _check(_lib.gsw_wave_scan64(_ptr(x), _ptr(out)))
    return out


def stream_copy(nbytes=1 << 30, reps=10):
    """Best (read + write) GB/s of a plain device copy of `nbytes`."""
    g = C.c_double()
    _check(_lib.gsw_stream_copy(int(nbytes), int(reps), C.byref(g)))
    return g.value
