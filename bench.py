#!/usr/bin/env python
"""Benchmark: GTEPS and fraction of the HBM roofline on one MI355X.

Workload (one "step"): SSSP from the highest-degree vertex plus a
50-iteration PageRank (damping 0.85) on R-MAT scale 22, edge factor 16,
(a, b, c, d) = (0.57, 0.19, 0.19, 0.05), seed 2, symmetric int32 CSR with
uint32 weights in [1, 64] -- the SSSP / PageRank configuration of the
engine's headline metric. SSSP and PageRank pick their variants with the
learned selectors (the committed networks; --rule-only for the rule
trees). The graph is resident in HBM before the timed region (PageRank's
propagation-blocked layout is built during the warm-up); runs keep their
results on the device.

GTEPS (Graph500 convention): SSSP counts sum(deg(v), v reached)/2,
PageRank counts m_undirected per iteration; value = total / wall time.
roofline: the expand variant with the largest total time; achieved =
algorithmic bytes per launch / mean launch time (hipEvent per expand).
cpu_baseline: the oracle's SPFA and FP64 PageRank on one host core
(5 of the 50 PageRank iterations timed and scaled by 10).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 22]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "GTEPS + % HBM peak; batched+learned-selector vs plain fixed-strategy, 1 GPU"
HBM_PEAK_GBS = 8000.0
PR_ITERS = 50
DAMPING = 0.85


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ------------------------------------------------------------ byte accounting
def pr_bytes(n, m):
    """One PageRank iteration over the whole graph: per directed edge a
    column index and a source contribution; per vertex the offset pair,
    the out-degree and the written rank."""
    return 8.0 * m + 24.0 * n


def sssp_level_bytes(rec, next_active, n):
    if rec["dir"] == "Push":
        return 20.0 * rec["frontier"] + 12.0 * rec["edges"] + 8.0 * next_active
    # pull: two bitmaps + per target offsets/state + per inspected in-edge
    # (column, weight, source probe) + found vertices
    return n / 4.0 + 20.0 * rec["frontier"] + 12.0 * rec["edges"] + 4.0 * next_active


def kernel_key(app, r):
    """Variant of an expand record; pulls name their from-all pull kernel."""
    k = "%s:%s-%s-%s-%s" % (app, r["dir"], "SortedQueue" if r["sorted"] and
                            r["fmt"] == "Queue" else r["fmt"], r["lb"],
                            "Fused" if r["fused"] else "Standalone")
    return k + ("/pull_kernel=%d" % r["kernel"] if r["dir"] == "Pull" else "")


def accumulate(acc, app, recs, n, m):
    for i, r in enumerate(recs):
        if app == "pagerank":
            b = pr_bytes(n, m)
        else:
            nxt = recs[i + 1]["active"] if i + 1 < len(recs) else 0
            b = sssp_level_bytes(r, nxt, n)
        k = kernel_key(app, r)
        e = acc.setdefault(k, {"launches": 0, "ms": 0.0, "bytes": 0.0})
        e["launches"] += 1
        e["ms"] += r["ms"]
        e["bytes"] += b


# ------------------------------------------------------------ CPU baseline
def host_cpu():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


class one_core:
    """Pin the calling thread to one host core while timing the oracle
    (the reference's CPU validation path is single-threaded)."""

    def __enter__(self):
        self.prev = None
        try:
            self.prev = os.sched_getaffinity(0)
            os.sched_setaffinity(0, {sorted(self.prev)[0]})
        except (AttributeError, OSError):
            self.prev = None
        return self

    def __exit__(self, *exc):
        if self.prev:
            os.sched_setaffinity(0, self.prev)


def timed(fn, reps=3):
    """One warm-up call, then the median of `reps` timed calls (seconds) and
    the last result."""
    out = fn()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)), out


def cpu_baseline(h, root, sssp_edges, pr_edges):
    """The oracle's SPFA and FP64 PageRank on one core: one warm-up and the
    median of three runs each; PageRank times 5 of the 50 iterations (the
    iterations cost the same) and scales by 10."""
    import oracle
    with one_core():
        t_sssp, dist = timed(lambda: oracle.spfa(h.row_ptr, h.col, h.w, root))
        t_pr5, pr5 = timed(lambda: oracle.pagerank(h.row_ptr, h.col, DAMPING, 5))
    t_total = t_sssp + t_pr5 * (PR_ITERS / 5.0)
    model, ncpu = host_cpu()
    return {
        "value": (sssp_edges + pr_edges) / t_total / 1e9,
        "unit": "GTEPS",
        "cores": 1,
        "kind": "port",
        "sample": "SPFA from the same root (whole run) + 5 of the 50 FP64 PageRank "
                  "iterations (x10) on the same R-MAT 22 graph; one warm-up, median "
                  "of 3; one pinned core",
        "host_cpu": model, "nproc": ncpu,
        "seconds": {"spfa": round(t_sssp, 3), "pagerank_5it": round(t_pr5, 3)},
    }, dist, pr5


def cpu_leg(fn, edges, reps=3, scale=1.0, sample=None):
    """GTEPS of one oracle routine on one core (warm-up + median of reps);
    `scale` multiplies the timed sample up to the whole workload."""
    with one_core():
        t, out = timed(fn, reps)
    leg = {"seconds": round(t * scale, 4), "gteps": round(edges / (t * scale) / 1e9, 4)}
    if sample:
        leg["sample"] = sample
    return leg, out


def bfs_level_bytes(rec, next_active, n):
    """Minimal bytes of one BFS level (no weights): push reads the queue
    and offsets of the frontier, a column and the target state per edge,
    writes state + queue entry per discovered vertex; pull reads the two
    frontier bitmaps, offsets + state per target, a column and a frontier
    word per inspected in-edge, and writes found vertices."""
    if rec["dir"] == "Push":
        return 20.0 * rec["frontier"] + 8.0 * rec["edges"] + 8.0 * next_active
    return n / 4.0 + 20.0 * rec["verts"] + 8.0 * rec["edges"] + 4.0 * next_active


def graph500_edges(h, levels):
    return int(np.diff(h.row_ptr)[levels >= 0].sum()) // 2


def bfs_single(gsw, h, root, reps, g=None, **kw):
    """Best-of-reps device time of one BFS, its records and byte model."""
    own = g is None
    if own:
        g = gsw.Graph(h)
    ref = g.bfs(root, **kw)
    best, recs = None, None
    for _ in range(reps):
        g.bfs(root, out=False, **kw)
        if best is None or g.last_stats.device_ms < best:
            best, recs = g.last_stats.device_ms, g.records()
    if own:
        g.close()
    nbytes = 0.0
    for i, r in enumerate(recs):
        nxt = recs[i + 1]["active"] if i + 1 < len(recs) else 0
        nbytes += bfs_level_bytes(r, nxt, h.n)
    return ref, best, recs, nbytes


PQWM = "Push-Queue-WM-Standalone"
DEVICE_LEVELS = 8        # device-resident BFS: levels per hipGraph replay
DEVICE_LEVELS_GRID = 16


def fixed_sweep(run, ref, reps=1):
    """Every one of the 30 variant classes as a fixed strategy: run(configs)
    -> (result, device ms). Returns the best class, its time, the
    Push-Queue-WM time and whether every class gave `ref`."""
    import gswitch_amd as gsw
    times, exact = {}, True
    for c in gsw.MLP_CLASSES:
        best = None
        for _ in range(reps):
            out, ms = run(c)
            exact &= bool(np.array_equal(out, ref))
            best = ms if best is None else min(best, ms)
        times[c] = best
    b = min(times, key=times.get)
    return {"best": b, "best_ms": round(times[b], 3),
            "push_queue_wm_ms": round(times[PQWM], 3),
            "all_classes_bit_exact": exact}


def bfs_runner(g, root):
    def run(c):
        lv = g.bfs(root, configs=c, profile=False)
        return lv, g.last_stats.device_ms
    return run


def extended(gsw, args):
    """Single-query BFS at R-MAT 16 and 24, BFS and CC on the 4096 x 4096 grid
    (10 % of edges deleted, 1 % long-range, seed 3), 64-source BFS (R-MAT 22
    seed 4), the 20,000-graph ER batch (seed 5, built and run in pieces), the
    learned selector against the rule trees, every fixed strategy, and the
    oracle's CPU time per configuration. Every comparison checks that the
    results agree."""
    import oracle
    out = {}
    cpu = {}
    from gswitch_amd import learn
    have_model = os.path.exists(learn.MODEL_PATH)
    if have_model:
        gsw.selector_load(learn.load_model())
    # --- single query, R-MAT 16 (the CPU-runnable configuration)
    h = gsw.rmat(16, 16, 0.57, 0.19, 0.19, seed=1)
    root = h.max_degree_vertex()
    g = gsw.Graph(h)
    lv16, ms16, _, _ = bfs_single(gsw, h, root, 5, g=g)
    sweep16 = fixed_sweep(bfs_runner(g, root), lv16, reps=3)
    c1 = {"n": h.n, "m_undirected": h.m // 2, "ms": round(ms16, 3),
          "gteps": round(graph500_edges(h, lv16) / (ms16 * 1e-3) / 1e9, 2),
          "fixed_strategies": sweep16}
    if have_model:
        lv16_l, ms16_l, _, _ = bfs_single(gsw, h, root, 5, g=g, selector=1)
        c1["learned_ms"] = round(ms16_l, 3)
        c1["learned_bit_exact"] = bool(np.array_equal(lv16, lv16_l))
    g.close()
    out["bfs_rmat16"] = c1
    cpu["c1_bfs_rmat16"], ref = cpu_leg(lambda: oracle.bfs(h.row_ptr, h.col, root),
                                        graph500_edges(h, lv16))
    c1["cpu_bit_exact"] = bool(np.array_equal(ref, lv16))
    del h
    # --- single query, R-MAT 24
    t = time.perf_counter()
    h = gsw.rmat(args.bfs_scale, 16, 0.57, 0.19, 0.19, seed=1)
    root = h.max_degree_vertex()
    log("bfs graph: scale %d n=%d m=%d built in %.1fs" % (args.bfs_scale, h.n, h.m,
                                                          time.perf_counter() - t))
    g = gsw.Graph(h)
    lv, ms, recs, nbytes = bfs_single(gsw, h, root, 7, g=g)
    e = graph500_edges(h, lv)
    sweep24 = fixed_sweep(bfs_runner(g, root), lv)
    single = {"scale": args.bfs_scale, "n": h.n, "m_undirected": h.m // 2,
              "levels": len(recs), "ms": round(ms, 3),
              "gteps": round(e / (ms * 1e-3) / 1e9, 2),
              "algorithmic_GBps": round(nbytes / (ms * 1e-3) / 1e9, 1),
              "hbm_frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
              "expand_ms": round(sum(r["ms"] for r in recs), 3),
              "fixed_strategies": sweep24,
              "variants": ["%s-%s-%s-%s" % (r["dir"], r["fmt"], r["lb"],
                                            "F" if r["fused"] else "S") for r in recs]}
    learned = None
    if have_model:
        lv2, ms_l, recs_l, _ = bfs_single(gsw, h, root, 7, g=g, selector=1)
        single["learned_ms"] = round(ms_l, 3)
        single["learned_bit_exact"] = bool(np.array_equal(lv2, lv))
        learned = ms_l
    # the device-resident loop (direction chosen on the device, levels
    # replayed from a captured hipGraph)
    lvd, ms_d, recs_d, nbytes_d = bfs_single(gsw, h, root, 7, g=g, capture=DEVICE_LEVELS)
    single["device_loop"] = {
        "ms": round(ms_d, 3), "gteps": round(e / (ms_d * 1e-3) / 1e9, 2),
        "algorithmic_GBps": round(nbytes_d / (ms_d * 1e-3) / 1e9, 1),
        "hbm_frac": round(nbytes_d / (ms_d * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "levels": len(recs_d), "bit_exact": bool(np.array_equal(lvd, lv)),
        "directions": "".join("P" if r["dir"] == "Pull" else "p" for r in recs_d)}
    g.close()
    out["bfs_single"] = single
    del h
    # --- high-diameter grid: 4096 x 4096, 10 % of edges deleted, 1 % long-range
    t = time.perf_counter()
    h = gsw.grid2d(args.grid, args.grid, 0.1, 0.01, 3)
    log("grid graph: %dx%d n=%d m=%d built in %.1fs" % (args.grid, args.grid, h.n, h.m,
                                                      time.perf_counter() - t))
    g = gsw.Graph(h)
    lvg, ms_g, recs_g, _ = bfs_single(gsw, h, 0, 7, g=g)
    sweepg = fixed_sweep(bfs_runner(g, 0), lvg)
    grid = {"n": h.n, "m_undirected": h.m // 2, "levels": len(recs_g),
            "bfs_ms": round(ms_g, 3),
            "bfs_gteps": round(graph500_edges(h, lvg) / (ms_g * 1e-3) / 1e9, 2),
            "bfs_fixed_strategies": sweepg}
    lvg_d, ms_gd, recs_gd, _ = bfs_single(gsw, h, 0, 7, g=g, capture=DEVICE_LEVELS_GRID)
    grid["bfs_device_loop_ms"] = round(ms_gd, 3)
    grid["bfs_device_loop_bit_exact"] = bool(np.array_equal(lvg, lvg_d))
    ms_g_learned = ms_g
    if learned is not None:
        lvg_l, ms_g_learned, _, _ = bfs_single(gsw, h, 0, 7, g=g, selector=1)
        grid["bfs_learned_ms"] = round(ms_g_learned, 3)
        grid["bfs_learned_bit_exact"] = bool(np.array_equal(lvg, lvg_l))
    lab = g.cc()
    t_cc = []
    for _ in range(5):
        g.cc(out=False)
        t_cc.append(g.last_stats.device_ms)
    lab_ec = g.cc(configs="x-Bitmap-x-x")
    t_ec = []
    for _ in range(3):
        g.cc(configs="x-Bitmap-x-x", out=False)
        t_ec.append(g.last_stats.device_ms)
    grid["cc_ms"] = round(min(t_cc), 3)
    grid["cc_gteps"] = round((h.m // 2) / (min(t_cc) * 1e-3) / 1e9, 2)
    grid["cc_edge_centric_ms"] = round(min(t_ec), 3)
    grid["cc_edge_centric_bit_exact"] = bool(np.array_equal(lab, lab_ec))
    ms_cc_learned = min(t_cc)
    if os.path.exists(learn.model_path("cc")):
        gsw.selector_load(learn.load_model(learn.model_path("cc")), app="cc")
        lab_l = g.cc(selector=1)
        t_l = []
        for _ in range(5):
            g.cc(selector=1, out=False)
            t_l.append(g.last_stats.device_ms)
        ms_cc_learned = min(t_l)
        grid["cc_learned_ms"] = round(ms_cc_learned, 3)
        grid["cc_learned_class"] = gsw.CC_CLASSES[g.last_class()] or "sampled"
        grid["cc_learned_bit_exact"] = bool(np.array_equal(lab, lab_l))
    g.close()
    cpu["c3_bfs_grid"], ref_l = cpu_leg(lambda: oracle.bfs(h.row_ptr, h.col, 0),
                                        graph500_edges(h, lvg))
    cpu["c3_cc_grid"], ref_c = cpu_leg(lambda: oracle.cc(h.row_ptr, h.col), h.m // 2)
    grid["bfs_cpu_bit_exact"] = bool(np.array_equal(ref_l, lvg))
    grid["cc_cpu_bit_exact"] = bool(np.array_equal(ref_c, lab))
    out["grid"] = grid
    del h, ref_l, ref_c
    # --- 64-source BFS, R-MAT 22 seed 4
    h = gsw.rmat(22, 16, 0.57, 0.19, 0.19, seed=4)
    g = gsw.Graph(h)
    src = np.argsort(-np.diff(h.row_ptr), kind="stable")[:64].astype(np.int32)
    masks, _ = g.msbfs64(src)
    t_ms = []
    for _ in range(3):
        g.msbfs64(src, out=False)
        t_ms.append(g.last_stats.device_ms)
    ms64 = min(t_ms)
    lv0 = g.bfs(int(src[0]))
    sweep64 = fixed_sweep(bfs_runner(g, int(src[0])), lv0)
    best64 = sweep64["best"]
    t_seq = t_pqwm = 0.0
    edges64 = 0
    ref = np.zeros(h.n, np.uint64)
    for k, s in enumerate(src):
        lv = g.bfs(int(s), configs=best64, profile=False)
        t_seq += g.last_stats.device_ms
        g.bfs(int(s), configs=PQWM, out=False, profile=False)
        t_pqwm += g.last_stats.device_ms
        edges64 += graph500_edges(h, lv)
        ref[lv >= 0] |= np.uint64(1) << np.uint64(k)
    out["msbfs64"] = {"ms": round(ms64, 3), "gteps": round(edges64 / (ms64 * 1e-3) / 1e9, 1),
                      "fixed_64_single_ms": round(t_seq, 3),
                      "fixed_strategy": best64,
                      "push_queue_wm_64_single_ms": round(t_pqwm, 3),
                      "speedup": round(t_seq / ms64, 2),
                      "bit_exact_vs_64_bfs": bool(np.array_equal(masks, ref))}
    g.close()
    k4 = 4
    cpu["c4a_msbfs64"], _ = cpu_leg(
        lambda: [oracle.bfs(h.row_ptr, h.col, int(s)) for s in src[:k4]], edges64,
        scale=64 / k4, sample="%d of the 64 sources, scaled x%d" % (k4, 64 // k4))
    del h
    # --- ER batch: instances 0 .. batch-1 of the seed-5 batch, built and run
    # as consecutive pieces of batch_piece instances (one disjoint union
    # each); the captured BFS shape is autotuned on the first piece
    tune, kw = None, None
    ms_batch = t_loop = t_gen = 0.0
    e_bfs = e_cc = n_tot = m_tot = pieces = 0
    same = True
    fixed_b = fixed_c = None
    for first in range(0, args.batch, args.batch_piece):
        k = min(args.batch_piece, args.batch - first)
        t = time.perf_counter()
        h, off = gsw.erdos_renyi_batch(k, 1000, 200000, 2.0, 32.0, seed=5, first=first)
        t_gen += time.perf_counter() - t
        g = gsw.Graph(h)
        roots = off[:-1].astype(np.int32)
        if tune is None:
            # the autotuned captured push loop, or the direction-optimising
            # device loop from all the roots, whichever is faster here
            tune = g.autotune_bfs(roots, capture=8)
            kw = dict(configs="Push-x-%s-x" % tune["lb"], capture=8, grid=tune["grid"],
                      twc_small=tune["twc_small"], twc_large=tune["twc_large"])
            t_do = []
            for _ in range(3):
                g.batch_bfs_cc(off, out=False, profile=False, capture=8)
                t_do.append(g.last_stats.device_ms)
            t_push = []
            for _ in range(3):
                g.batch_bfs_cc(off, out=False, profile=False, **kw)
                t_push.append(g.last_stats.device_ms)
            tune["bfs_cc_ms_push"] = round(min(t_push), 3)
            tune["bfs_cc_ms_direction_optimising"] = round(min(t_do), 3)
            if min(t_do) < min(t_push):
                kw = dict(capture=8)
                tune["chosen"] = "direction-optimising device loop"
            else:
                tune["chosen"] = "captured push loop"
        lvb, labb = g.batch_bfs_cc(off, **kw)
        bms = []
        for _ in range(3):
            g.batch_bfs_cc(off, out=False, profile=False, **kw)
            bms.append(g.last_stats.device_ms)
        ms_piece = min(bms)
        ms_batch += ms_piece
        e_bfs += graph500_edges(h, lvb)
        e_cc += h.m // 2
        n_tot += h.n
        m_tot += h.m // 2

        def inst(i):
            lo, hi = int(off[i]), int(off[i + 1])
            rp = h.row_ptr[lo:hi + 1] - h.row_ptr[lo]
            return lo, hi, gsw.Graph(row_ptr=rp, col=h.col[h.row_ptr[lo]:h.row_ptr[hi]] - lo)
        if fixed_b is None:
            # the fixed strategies of the per-graph loop: the fastest BFS
            # class and the faster edge-centric CC format over a sample
            tb = {c: 0.0 for c in gsw.MLP_CLASSES}
            tc = {"x-Queue-x-x": 0.0, "x-Bitmap-x-x": 0.0}
            for i in range(min(k, 40)):
                lo, hi, gi = inst(i)
                for c in tb:
                    gi.bfs(0, configs=c, out=False, profile=False)
                    tb[c] += gi.last_stats.device_ms
                for c in tc:
                    gi.cc(configs=c, out=False, profile=False)
                    tc[c] += gi.last_stats.device_ms
                gi.close()
            fixed_b, fixed_c = min(tb, key=tb.get), min(tc, key=tc.get)
        # per-graph launch loop with the fixed strategies (upload outside
        # the timing)
        t_piece = 0.0
        for i in range(k):
            lo, hi, gi = inst(i)
            lvi = gi.bfs(0, configs=fixed_b, profile=False)
            t_piece += gi.last_stats.device_ms
            labi = gi.cc(configs=fixed_c, profile=False)
            t_piece += gi.last_stats.device_ms
            same &= bool(np.array_equal(lvi, lvb[lo:hi])) and \
                bool(np.array_equal(labi, labb[lo:hi] - lo))
            if first == 0 and i < 20:
                rp = h.row_ptr[lo:hi + 1] - h.row_ptr[lo]
                col = h.col[h.row_ptr[lo]:h.row_ptr[hi]] - lo
                cpu.setdefault("_c4b", []).append((rp, col))
            gi.close()
        t_loop += t_piece
        pieces += 1
        g.close()
        del h, g
        log("batch piece %d: instances %d..%d n=%d batched %.2f ms, loop %.1f ms, "
            "bit-exact so far %s" % (pieces, first, first + k - 1, n_tot, ms_piece,
                                     t_piece, same))
    log("batch: %d graphs in %d pieces, host build %.1fs" % (args.batch, pieces, t_gen))
    out["er_batch"] = {"graphs": args.batch, "pieces": pieces, "n": n_tot,
                       "m_undirected": m_tot, "autotune": tune,
                       "captured": "BFS level loop (%s) + CC launch sequence "
                                   "(gsw_batch_bfs_cc)" % tune["chosen"],
                       "batched_ms": round(ms_batch, 3),
                       "batched_gteps": round((e_bfs + e_cc) / (ms_batch * 1e-3) / 1e9, 2),
                       "per_graph_loop_ms": round(t_loop, 3),
                       "per_graph_fixed": [fixed_b, "edge-centric CC " + fixed_c],
                       "speedup": round(t_loop / ms_batch, 2),
                       "bit_exact_vs_per_graph": same}
    samp = cpu.pop("_c4b", [])
    if samp:
        def c4b_sample():
            for rp, col in samp:
                oracle.bfs(rp, col, 0)
                oracle.cc(rp, col)
        e_s = sum(int(len(col)) // 2 * 2 for rp, col in samp)
        cpu["c4b_er_batch"], _ = cpu_leg(
            c4b_sample, e_s, sample="BFS + CC of instances 0-%d (of %d); GTEPS over "
                                    "their undirected edges, counted for BFS and CC"
                                    % (len(samp) - 1, args.batch))
    out["cpu_legs"] = cpu
    # --- aggregate: batched + learned vs plain per-query fixed strategy
    t_new = ms_batch + ms64 + (learned if learned is not None else ms) + ms_g_learned \
        + ms_cc_learned
    t_old = t_loop + t_seq + sweep24["best_ms"] + sweepg["best_ms"] + \
        min(grid["cc_ms"], grid["cc_edge_centric_ms"])
    out["batched_learned_vs_fixed"] = {
        "speedup": round(t_old / t_new, 2),
        "note": "same work (ER batch BFS+CC, 64 BFS, one R-MAT 24 BFS, one grid BFS, "
                "one grid CC); device ms; fixed = the fastest of the 30 variant classes "
                "per workload (per-graph loop: fastest on a 40-instance sample; grid CC: "
                "the faster of sampled hooking and edge-centric bitmap)"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rule-only", action="store_true",
                    help="headline SSSP under the rule trees even with a trained network")
    ap.add_argument("--configs-sssp", default=None)
    ap.add_argument("--configs-pr", default=None)
    ap.add_argument("--quick", action="store_true",
                    help="headline workload only (no BFS / batch / selector lines)")
    ap.add_argument("--bfs-scale", type=int, default=24)
    ap.add_argument("--grid", type=int, default=4096,
                    help="side of the high-diameter grid")
    ap.add_argument("--batch", type=int, default=20000,
                    help="ER batch size (instances 0..batch-1 of the seed-5 batch)")
    ap.add_argument("--batch-piece", type=int, default=5000,
                    help="instances per union built and run at once")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU "
              "path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import gswitch_amd as gsw
    if gsw.device_count() < 1:
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)

    t = time.perf_counter()
    h = gsw.rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=args.seed, weighted=True)
    root = h.max_degree_vertex()
    log("graph: n=%d m=%d (directed CSR entries) root=%d built in %.1fs"
        % (h.n, h.m, root, time.perf_counter() - t))
    g = gsw.Graph(h)
    n, m = h.n, h.m
    copy_gbs = gsw.stream_copy(1 << 30, 10)
    log("stream copy: %.0f GB/s" % copy_gbs)

    # SSSP and PageRank run under their learned selectors when trained
    # networks are committed (and no variant is forced)
    from gswitch_amd import learn
    sel, sel_pr = {}, {}
    if os.path.exists(learn.model_path("sssp")) and args.configs_sssp is None \
            and not args.rule_only:
        gsw.selector_load(learn.load_model(learn.model_path("sssp")), app="sssp")
        sel = {"selector": 1}
    if os.path.exists(learn.model_path("pr")) and args.configs_pr is None \
            and not args.rule_only:
        gsw.selector_load(learn.load_model(learn.model_path("pr")), app="pr")
        sel_pr = {"selector": 1}

    def step():
        g.sssp(root, configs=args.configs_sssp, out=False, **sel)
        s_ms = g.last_stats.device_ms
        s_rec = g.records()
        g.pagerank(DAMPING, PR_ITERS, configs=args.configs_pr, out=False, **sel_pr)
        return s_ms, s_rec, g.last_stats.device_ms, g.records()

    for _ in range(args.warmup):
        step()

    acc = {}
    wall = 0.0
    dev_sssp = dev_pr = 0.0
    for _ in range(args.steps):
        t0 = time.perf_counter()
        g.sssp(root, configs=args.configs_sssp, out=False, **sel)
        t1 = time.perf_counter()
        s_ms, s_rec = g.last_stats.device_ms, g.records()
        t2 = time.perf_counter()
        g.pagerank(DAMPING, PR_ITERS, configs=args.configs_pr, out=False, **sel_pr)
        t3 = time.perf_counter()
        p_ms, p_rec = g.last_stats.device_ms, g.records()
        wall += (t1 - t0) + (t3 - t2)
        dev_sssp += s_ms
        dev_pr += p_ms
        accumulate(acc, "sssp", s_rec, n, m)
        accumulate(acc, "pagerank", p_rec, n, m)

    # traversed edges (results fetched outside the timed region)
    dist = g.sssp(root, configs=args.configs_sssp, **sel)
    reached = dist != np.uint32(0xFFFFFFFF)
    sssp_edges = int(np.diff(h.row_ptr)[reached].sum()) // 2
    pr_edges = (m // 2) * PR_ITERS
    step_edges = sssp_edges + pr_edges
    value = step_edges * args.steps / wall / 1e9

    dom_key = max(acc, key=lambda k: acc[k]["ms"])
    dom = acc[dom_key]
    per_launch_ms = dom["ms"] / dom["launches"]
    per_launch_bytes = dom["bytes"] / dom["launches"]
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9
    # HBM bytes per launch of this kernel, measured by the PMC runs recorded
    # in profiles/ (FETCH_SIZE / WRITE_SIZE, calibrated with the copy kernel)
    traffic = None
    tfile = os.path.join(REPO, "profiles", "traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get(dom_key, {}).get("traffic_bytes")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": dom_key, "launch_ms": round(per_launch_ms, 4),
                "algorithmic_bytes_per_launch": per_launch_bytes,
                "stream_copy_gbs": round(copy_gbs, 1)}

    detail = {
        "sssp_gteps": round(sssp_edges * args.steps / (dev_sssp * 1e-3) / 1e9, 3),
        "pagerank_gteps": round(pr_edges * args.steps / (dev_pr * 1e-3) / 1e9, 3),
        "sssp_ms": round(dev_sssp / args.steps, 3),
        "pagerank_ms": round(dev_pr / args.steps, 3),
        "sssp_levels": len(s_rec),
        "kernels": {k: {"launches": v["launches"] // args.steps,
                        "ms_per_step": round(v["ms"] / args.steps, 4),
                        "GBps": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)}
                    for k, v in sorted(acc.items(), key=lambda kv: -kv[1]["ms"])},
    }

    cpu = None
    if not args.no_cpu:
        cpu, cdist, cpr5 = cpu_baseline(h, root, sssp_edges, pr_edges)
        gpr5 = g.pagerank(DAMPING, 5, configs=args.configs_pr, **sel_pr)
        detail["parity"] = {
            "sssp_bit_exact": bool(np.array_equal(dist, cdist)),
            "pagerank5_max_rel_err": float(np.max(np.abs(gpr5 - cpr5) / np.abs(cpr5))),
        }

    if not args.quick:
        # the same SSSP and PageRank under every fixed strategy: SSSP's 30
        # variant classes, PageRank's 8 from-all pulls (no selector)
        def sssp_run(c):
            d = g.sssp(root, configs=c, profile=False)
            return d, g.last_stats.device_ms
        sweep_s = fixed_sweep(sssp_run, dist)
        pr_ref = g.pagerank(DAMPING, PR_ITERS, configs=args.configs_pr, **sel_pr)
        pr_times, pr_err = {}, 0.0
        for cfg, pk in gsw.PR_CLASSES:
            ts = []
            for _ in range(2):
                pr_f = g.pagerank(DAMPING, PR_ITERS, configs=cfg, pull_kernel=pk,
                                  profile=False)
                ts.append(g.last_stats.device_ms)
            pr_times["%s/pull_kernel=%d" % (cfg, pk)] = min(ts)
            pr_err = max(pr_err, float(np.max(np.abs(pr_f.astype(np.float64) - pr_ref)
                                              / np.abs(pr_ref))))
        best_p = min(pr_times, key=pr_times.get)
        detail["fixed_strategy"] = {
            "sssp": sweep_s,
            "pagerank": {"best": best_p, "best_ms": round(pr_times[best_p], 3),
                         "classes_ms": {k: round(v, 3) for k, v in pr_times.items()},
                         "max_rel_vs_headline": pr_err},
            "speedup_selected_vs_best_fixed": round(
                (sweep_s["best_ms"] + pr_times[best_p]) /
                ((dev_sssp + dev_pr) / args.steps), 3)}
        # the same SSSP under other stepping windows (the reference's initial
        # window times each scale), and the scale the window autotuner picks
        # from 4 seeded sample roots other than the query's
        wt, wexact = {}, True
        for k in (1.0, 0.5, 0.25, 0.125, 0.0625):
            g.sssp_window(k)
            ts = []
            for _ in range(3):
                dist_w = g.sssp(root, configs=args.configs_sssp, **sel)
                ts.append(g.last_stats.device_ms)
                wexact &= bool(np.array_equal(dist_w, dist))
            wt[str(k)] = round(min(ts), 3)
        rng = np.random.default_rng(7)
        wsample = rng.choice(np.flatnonzero(np.diff(h.row_ptr) > 0), 4,
                             replace=False).astype(np.int32)
        wscale, _ = g.autotune_sssp(wsample, configs=args.configs_sssp, **sel)
        g.sssp_window(1.0)
        detail["sssp_window"] = {"query_ms_by_scale": wt, "bit_exact": wexact,
                                 "autotuned_scale": wscale,
                                 "autotune_roots": [int(x) for x in wsample]}
        # the same SSSP and PageRank under the rule trees
        t_r, t_pr = [], []
        for _ in range(args.steps):
            dist_r = g.sssp(root, configs=args.configs_sssp)
            t_r.append(g.last_stats.device_ms)
            pr_r = g.pagerank(DAMPING, PR_ITERS, configs=args.configs_pr)
            t_pr.append(g.last_stats.device_ms)
        detail["selectors"] = {
            "headline": "learned" if sel else "rule",
            "sssp_headline_ms": round(dev_sssp / args.steps, 3),
            "sssp_rule_ms": round(min(t_r), 3),
            "sssp_rule_bit_exact": bool(np.array_equal(dist_r, dist)),
            "pagerank_headline_ms": round(dev_pr / args.steps, 3),
            "pagerank_rule_ms": round(min(t_pr), 3),
            "pagerank_rule_max_rel": float(np.max(np.abs(pr_r.astype(np.float64) - pr_ref)
                                                  / np.abs(pr_ref)))}

    g.close()
    if not args.quick:
        detail.update(extended(gsw, args))

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/f64",
        "data": "synthetic",
        "config": {"workload": "SSSP + PageRank(50 it) on R-MAT scale %d EF16 seed %d"
                               % (args.scale, args.seed),
                   "scale": args.scale, "edge_factor": 16, "n": n,
                   "m_undirected": m // 2, "root": root,
                   "selector": ("learned (SSSP, PageRank)" if sel and sel_pr else
                                "learned (SSSP)" if sel else "rule"),
                   "parallelism": "single-gpu"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "detail": detail,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
