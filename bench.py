#!/usr/bin/env python
"""Benchmark: GTEPS and fraction of the HBM roofline on one MI355X.

Workload (one "step"): SSSP from the highest-degree vertex plus a
50-iteration PageRank (damping 0.85) on R-MAT scale 22, edge factor 16,
(a, b, c, d) = (0.57, 0.19, 0.19, 0.05), seed 2, symmetric int32 CSR with
uint32 weights in [1, 64] -- the SSSP / PageRank configuration of the
engine's headline metric. SSSP picks its variants with the learned selector
(the committed network; --rule-only for the rule trees), PageRank with the
rule trees. The graph is resident in HBM before the timed region; runs keep
their results on the device.

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
    return "%s:%s-%s-%s-%s" % (app, r["dir"], "SortedQueue" if r["sorted"] and
                               r["fmt"] == "Queue" else r["fmt"], r["lb"],
                               "Fused" if r["fused"] else "Standalone")


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
def cpu_baseline(h, root, sssp_edges, pr_edges):
    import oracle
    cores = 1
    prev = None
    try:
        prev = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {sorted(prev)[0]})
    except (AttributeError, OSError):
        prev = None
    try:
        t = time.perf_counter()
        dist = oracle.spfa(h.row_ptr, h.col, h.w, root)
        t_sssp = time.perf_counter() - t
        t = time.perf_counter()
        pr5 = oracle.pagerank(h.row_ptr, h.col, DAMPING, 5)
        t_pr5 = time.perf_counter() - t
    finally:
        if prev:
            os.sched_setaffinity(0, prev)
    t_total = t_sssp + t_pr5 * (PR_ITERS / 5.0)
    return {
        "value": (sssp_edges + pr_edges) / t_total / 1e9,
        "unit": "GTEPS",
        "cores": cores,
        "kind": "port",
        "sample": "full SPFA from the same root + 5 of 50 FP64 PageRank "
                  "iterations (x10) on the same R-MAT 22 graph; single core",
        "seconds": {"spfa": round(t_sssp, 3), "pagerank_5it": round(t_pr5, 3)},
    }, dist, pr5


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


def bfs_single(gsw, h, root, reps, **kw):
    """Best-of-reps device time of one BFS, its records and byte model."""
    g = gsw.Graph(h)
    ref = g.bfs(root, **kw)
    best, recs = None, None
    for _ in range(reps):
        g.bfs(root, out=False, **kw)
        if best is None or g.last_stats.device_ms < best:
            best, recs = g.last_stats.device_ms, g.records()
    g.close()
    nbytes = 0.0
    for i, r in enumerate(recs):
        nxt = recs[i + 1]["active"] if i + 1 < len(recs) else 0
        nbytes += bfs_level_bytes(r, nxt, h.n)
    return ref, best, recs, nbytes


def extended(gsw, args):
    """Single-query BFS at R-MAT 16 and 24, BFS and CC on the 4096 x 4096 grid
    (10 % of edges deleted, 1 % long-range, seed 3), 64-source BFS (R-MAT 22
    seed 4), the 20,000-graph ER batch (seed 5, built and run in pieces), and the learned
    selector against the rule trees and a fixed strategy. Every comparison
    checks that the results agree."""
    out = {}
    fixed = "Push-Queue-WM-Standalone"
    from gswitch_amd import learn
    have_model = os.path.exists(learn.MODEL_PATH)
    if have_model:
        gsw.selector_load(learn.load_model())
    # --- single query, R-MAT 16 (the CPU-runnable configuration)
    h = gsw.rmat(16, 16, 0.57, 0.19, 0.19, seed=1)
    root = h.max_degree_vertex()
    lv16, ms16, _, _ = bfs_single(gsw, h, root, 5)
    lv16_f, ms16_f, _, _ = bfs_single(gsw, h, root, 5, configs=fixed)
    c1 = {"n": h.n, "m_undirected": h.m // 2, "ms": round(ms16, 3),
          "gteps": round(graph500_edges(h, lv16) / (ms16 * 1e-3) / 1e9, 2),
          "fixed_strategy_ms": round(ms16_f, 3),
          "fixed_bit_exact": bool(np.array_equal(lv16, lv16_f))}
    if have_model:
        lv16_l, ms16_l, _, _ = bfs_single(gsw, h, root, 5, selector=1)
        c1["learned_ms"] = round(ms16_l, 3)
        c1["learned_bit_exact"] = bool(np.array_equal(lv16, lv16_l))
    out["bfs_rmat16"] = c1
    del h
    # --- single query, R-MAT 24
    t = time.perf_counter()
    h = gsw.rmat(args.bfs_scale, 16, 0.57, 0.19, 0.19, seed=1)
    root = h.max_degree_vertex()
    log("bfs graph: scale %d n=%d m=%d built in %.1fs" % (args.bfs_scale, h.n, h.m,
                                                          time.perf_counter() - t))
    lv, ms, recs, nbytes = bfs_single(gsw, h, root, 5)
    e = graph500_edges(h, lv)
    _, ms_fixed, _, _ = bfs_single(gsw, h, root, 3, configs=fixed)
    single = {"scale": args.bfs_scale, "n": h.n, "m_undirected": h.m // 2,
              "levels": len(recs), "ms": round(ms, 3),
              "gteps": round(e / (ms * 1e-3) / 1e9, 2),
              "algorithmic_GBps": round(nbytes / (ms * 1e-3) / 1e9, 1),
              "hbm_frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
              "expand_ms": round(sum(r["ms"] for r in recs), 3),
              "fixed_strategy_ms": round(ms_fixed, 3),
              "variants": ["%s-%s-%s-%s" % (r["dir"], r["fmt"], r["lb"],
                                            "F" if r["fused"] else "S") for r in recs]}
    learned = None
    if have_model:
        lv2, ms_l, recs_l, _ = bfs_single(gsw, h, root, 5, selector=1)
        single["learned_ms"] = round(ms_l, 3)
        single["learned_bit_exact"] = bool(np.array_equal(lv2, lv))
        learned = ms_l
    out["bfs_single"] = single
    del h
    # --- high-diameter grid: 4096 x 4096, 10 % of edges deleted, 1 % long-range
    t = time.perf_counter()
    h = gsw.grid2d(args.grid, args.grid, 0.1, 0.01, 3)
    log("grid graph: %dx%d n=%d m=%d built in %.1fs" % (args.grid, args.grid, h.n, h.m,
                                                      time.perf_counter() - t))
    lvg, ms_g, recs_g, _ = bfs_single(gsw, h, 0, 7)
    lvg_f, ms_g_fixed, _, _ = bfs_single(gsw, h, 0, 3, configs=fixed)
    grid = {"n": h.n, "m_undirected": h.m // 2, "levels": len(recs_g),
            "bfs_ms": round(ms_g, 3),
            "bfs_gteps": round(graph500_edges(h, lvg) / (ms_g * 1e-3) / 1e9, 2),
            "bfs_fixed_strategy_ms": round(ms_g_fixed, 3),
            "bfs_fixed_bit_exact": bool(np.array_equal(lvg, lvg_f))}
    ms_g_learned = ms_g
    if learned is not None:
        lvg_l, ms_g_learned, _, _ = bfs_single(gsw, h, 0, 7, selector=1)
        grid["bfs_learned_ms"] = round(ms_g_learned, 3)
        grid["bfs_learned_bit_exact"] = bool(np.array_equal(lvg, lvg_l))
    g = gsw.Graph(h)
    lab = g.cc()
    t_cc = []
    for _ in range(3):
        g.cc(out=False)
        t_cc.append(g.last_stats.device_ms)
    g.close()
    grid["cc_ms"] = round(min(t_cc), 3)
    grid["cc_gteps"] = round((h.m // 2) / (min(t_cc) * 1e-3) / 1e9, 2)
    # size-independent checks: BFS levels differ by at most one across every
    # edge and every reached non-root vertex has a parent one level up; CC
    # labels agree across every edge and each label is its component's
    # smallest id
    src = np.repeat(np.arange(h.n, dtype=np.int64), np.diff(h.row_ptr))
    dst = h.col.astype(np.int64)
    a, b = lvg[src], lvg[dst]
    bfs_ok = bool(lvg[0] == 0 and np.all((a < 0) == (b < 0)) and
                  np.all(np.abs(a - b)[a >= 0] <= 1))
    has_parent = np.zeros(h.n, bool)
    has_parent[dst[(a >= 0) & (a == b - 1)]] = True
    bfs_ok &= bool(np.all(has_parent[(lvg > 0)]))
    cc_ok = bool(np.all(lab[src] == lab[dst]) and np.all(lab <= np.arange(h.n)) and
                 np.all(lab[lab] == lab))
    grid["bfs_invariants"] = bfs_ok
    grid["cc_invariants"] = cc_ok
    out["grid"] = grid
    del h, src, dst, a, b
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
    t_seq = 0.0
    edges64 = 0
    ref = np.zeros(h.n, np.uint64)
    for k, s in enumerate(src):
        lv = g.bfs(int(s), configs=fixed)
        t_seq += g.last_stats.device_ms
        edges64 += graph500_edges(h, lv)
        ref[lv >= 0] |= np.uint64(1) << np.uint64(k)
    out["msbfs64"] = {"ms": round(ms64, 3), "gteps": round(edges64 / (ms64 * 1e-3) / 1e9, 1),
                      "fixed_64_single_ms": round(t_seq, 3),
                      "speedup": round(t_seq / ms64, 2),
                      "bit_exact_vs_64_bfs": bool(np.array_equal(masks, ref))}
    g.close()
    del h
    # --- ER batch: instances 0 .. batch-1 of the seed-5 batch, built and run
    # as consecutive pieces of batch_piece instances (one union each, far
    # below 2^31 directed entries); the bins are autotuned on the first piece
    tune, kw = None, None
    ms_batch = t_loop = t_gen = 0.0
    e_bfs = e_cc = n_tot = m_tot = pieces = 0
    same = True
    for first in range(0, args.batch, args.batch_piece):
        k = min(args.batch_piece, args.batch - first)
        t = time.perf_counter()
        h, off = gsw.erdos_renyi_batch(k, 1000, 200000, 2.0, 32.0, seed=5, first=first)
        t_gen += time.perf_counter() - t
        g = gsw.Graph(h)
        roots = off[:-1].astype(np.int32)
        if tune is None:
            tune = g.autotune_bfs(roots, capture=8)
            kw = dict(configs="Push-x-%s-x" % tune["lb"], capture=8, grid=tune["grid"],
                      twc_small=tune["twc_small"], twc_large=tune["twc_large"])
        lvb = g.bfs_multi(roots, **kw)
        bms = []
        for _ in range(3):
            g.bfs_multi(roots, out=False, profile=False, **kw)
            t_b = g.last_stats.device_ms
            g.cc(out=False, profile=False)
            bms.append(t_b + g.last_stats.device_ms)
        ms_piece = min(bms)
        ms_batch += ms_piece
        labb = g.cc()
        e_bfs += graph500_edges(h, lvb)
        e_cc += h.m // 2
        n_tot += h.n
        m_tot += h.m // 2
        # per-graph launch loop with a fixed strategy (upload outside the timing)
        t_piece = 0.0
        for i in range(k):
            lo, hi = int(off[i]), int(off[i + 1])
            rp = h.row_ptr[lo:hi + 1] - h.row_ptr[lo]
            col = h.col[h.row_ptr[lo]:h.row_ptr[hi]] - lo
            gi = gsw.Graph(row_ptr=rp, col=col)
            lvi = gi.bfs(0, configs=fixed, profile=False)
            t_piece += gi.last_stats.device_ms
            labi = gi.cc(profile=False)
            t_piece += gi.last_stats.device_ms
            same &= bool(np.array_equal(lvi, lvb[lo:hi])) and \
                bool(np.array_equal(labi, labb[lo:hi] - lo))
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
                       "batched_ms": round(ms_batch, 3),
                       "batched_gteps": round((e_bfs + e_cc) / (ms_batch * 1e-3) / 1e9, 2),
                       "per_graph_loop_ms": round(t_loop, 3),
                       "speedup": round(t_loop / ms_batch, 2),
                       "bit_exact_vs_per_graph": same}
    # --- aggregate: batched + learned vs plain per-query fixed strategy
    t_new = ms_batch + ms64 + (learned if learned is not None else ms) + ms_g_learned
    t_old = t_loop + t_seq + ms_fixed + ms_g_fixed
    out["batched_learned_vs_fixed"] = {
        "speedup": round(t_old / t_new, 2),
        "note": "same work (ER batch BFS+CC, 64 BFS, one R-MAT BFS, one grid BFS); device ms"}
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
    ap.add_argument("--batch-piece", type=int, default=1000,
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

    # SSSP runs under its learned selector when a trained network is
    # committed (and no variant is forced); PageRank keeps the rule trees
    from gswitch_amd import learn
    sel = {}
    if os.path.exists(learn.model_path("sssp")) and args.configs_sssp is None \
            and not args.rule_only:
        gsw.selector_load(learn.load_model(learn.model_path("sssp")), app="sssp")
        sel = {"selector": 1}

    def step():
        g.sssp(root, configs=args.configs_sssp, out=False, **sel)
        s_ms = g.last_stats.device_ms
        s_rec = g.records()
        g.pagerank(DAMPING, PR_ITERS, configs=args.configs_pr, out=False)
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
        g.pagerank(DAMPING, PR_ITERS, configs=args.configs_pr, out=False)
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
        gpr5 = g.pagerank(DAMPING, 5, configs=args.configs_pr)
        detail["parity"] = {
            "sssp_bit_exact": bool(np.array_equal(dist, cdist)),
            "pagerank5_max_rel_err": float(np.max(np.abs(gpr5 - cpr5) / np.abs(cpr5))),
        }

    if not args.quick:
        # the same SSSP and PageRank under one fixed strategy (no selector)
        fixed_s, fixed_p = "Push-Queue-WM-Standalone", "Push-Bitmap-WM-Standalone"
        dist_f = g.sssp(root, configs=fixed_s)
        s_fixed = g.last_stats.device_ms
        pr_rule = g.pagerank(DAMPING, PR_ITERS, configs=args.configs_pr)
        pr_f = g.pagerank(DAMPING, PR_ITERS, configs=fixed_p)
        p_fixed = g.last_stats.device_ms
        detail["fixed_strategy"] = {
            "configs": [fixed_s, fixed_p], "sssp_ms": round(s_fixed, 3),
            "sssp_bit_exact": bool(np.array_equal(dist_f, dist)),
            "pagerank_ms": round(p_fixed, 3),
            "pagerank_max_rel_vs_rule": float(np.max(np.abs(pr_f.astype(np.float64) - pr_rule)
                                                     / np.abs(pr_rule))),
            "speedup_rule_vs_fixed": round((s_fixed + p_fixed) /
                                           ((dev_sssp + dev_pr) / args.steps), 2)}
        # the same SSSP under the rule trees
        t_r = []
        for _ in range(args.steps):
            dist_r = g.sssp(root, configs=args.configs_sssp)
            t_r.append(g.last_stats.device_ms)
        detail["sssp_selectors"] = {"headline": "learned" if sel else "rule",
                                    "headline_ms": round(dev_sssp / args.steps, 3),
                                    "rule_ms": round(min(t_r), 3),
                                    "rule_bit_exact": bool(np.array_equal(dist_r, dist))}

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
                   "selector": ("learned (SSSP), rule (PageRank)" if sel else "rule"),
                   "parallelism": "single-gpu"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "detail": detail,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
