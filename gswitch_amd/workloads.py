"""Single-workload driver for profiling (rocprofv3) and quick timing.

    python -m gswitch_amd.workloads pr   --scale 22 --iters 5  [--configs ...]
    python -m gswitch_amd.workloads sssp --scale 22
    python -m gswitch_amd.workloads bfs  --scale 24
    python -m gswitch_amd.workloads cc   --er-batch 1000
    python -m gswitch_amd.workloads bfs  --grid 4096
    python -m gswitch_amd.workloads copy --mib 1024
    python -m gswitch_amd.workloads msbfs --scale 22 --seed 4
    python -m gswitch_amd.workloads batch --er-batch 5000

Graphs are the seeded R-MAT (a, b, c) = (0.57, 0.19, 0.19), edge factor 16;
the root is the highest-degree vertex. `--er-batch K` replaces the R-MAT graph
by instances 0..K-1 of the seed-5 Erdos-Renyi batch (n log-uniform in
[1000, 200000], mean degree uniform in [2, 32]) as one union; `--grid S`
by the S x S grid with 10 % of its edges deleted and 1 % long-range edges
(seed 3; root vertex 0). `msbfs` is the 64-source bitmask BFS from the 64
highest-degree vertices; `batch` the batched BFS from every instance's
first vertex + CC of an ER batch union (gsw_batch_bfs_cc, the
direction-optimising device loop). `copy` runs the plain device copy
kernel used to calibrate hardware byte counters. Each run prints one JSON
line with the device time and the per-expand records.
"""
import argparse
import json
import sys
import time

import gswitch_amd as gsw


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("app", choices=["bfs", "sssp", "pr", "cc", "copy", "msbfs", "batch"])
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--repeat", type=int, default=1)
    ap.add_argument("--configs", default=None)
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--records", action="store_true")
    ap.add_argument("--pull-kernel", type=int, default=2)
    ap.add_argument("--er-batch", type=int, default=0)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--selector", type=int, default=0)
    ap.add_argument("--capture", type=int, default=0,
                    help="bfs: device-resident loop, levels per replay")
    ap.add_argument("--no-profile", action="store_true",
                    help="no per-expand hipEvents (records keep no times)")
    a = ap.parse_args(argv)
    if a.app == "copy":
        g = gsw.stream_copy(a.mib << 20, a.repeat)
        print(json.dumps({"app": "copy", "bytes": a.mib << 20, "GBps": g}))
        return
    t = time.perf_counter()
    off = None
    if a.er_batch > 0:
        h, off = gsw.erdos_renyi_batch(a.er_batch, 1000, 200000, 2.0, 32.0, seed=5,
                                       weighted=(a.app == "sssp"))
    elif a.grid > 0:
        h = gsw.grid2d(a.grid, a.grid, 0.1, 0.01, 3, weighted=(a.app == "sssp"))
    else:
        h = gsw.rmat(a.scale, 16, 0.57, 0.19, 0.19, seed=a.seed,
                     weighted=(a.app == "sssp"))
    root = 0 if a.grid > 0 else h.max_degree_vertex()
    g = gsw.Graph(h)
    if a.selector:
        from gswitch_amd import learn
        for app in ("bfs", "sssp"):
            gsw.selector_load(learn.load_model(learn.model_path(app)), app=app)
    kw = dict(configs=a.configs, out=False, profile=not a.no_profile)
    if a.selector and a.app in ("bfs", "sssp"):
        kw["selector"] = 1
    if a.capture and a.app == "bfs":
        kw["capture"] = a.capture
    print("graph built in %.1fs: n=%d m=%d" % (time.perf_counter() - t, h.n, h.m),
          file=sys.stderr, flush=True)
    if a.app == "batch" and off is None:
        raise SystemExit("batch needs --er-batch K")
    src = None
    if a.app == "msbfs":
        import numpy as np
        src = np.argsort(-h.degrees, kind="stable")[:64].astype(np.int32)
    for _ in range(a.repeat):
        if a.app == "msbfs":
            g.msbfs64(src, out=False)
        elif a.app == "batch":
            g.batch_bfs_cc(off, out=False, profile=not a.no_profile, capture=8)
        elif a.app == "bfs":
            g.bfs(root, **kw)
        elif a.app == "sssp":
            g.sssp(root, **kw)
        elif a.app == "cc":
            g.cc(**kw)
        else:
            g.pagerank(0.85, a.iters, pull_kernel=a.pull_kernel, **kw)
        st = g.last_stats
        name = ("er_batch %d" % a.er_batch) if a.er_batch > 0 else \
            ("grid %d" % a.grid) if a.grid > 0 else ("rmat %d" % a.scale)
        out = {"app": a.app, "graph": name, "device_ms": st.device_ms,
               "levels": st.levels}
        if a.records:
            out["records"] = g.records()
        print(json.dumps(out), flush=True)
    g.close()


if __name__ == "__main__":
    main()
