"""Single-workload driver for profiling (rocprofv3) and quick timing.

    python -m gswitch_amd.workloads pr   --scale 22 --iters 5  [--configs ...]
    python -m gswitch_amd.workloads sssp --scale 22
    python -m gswitch_amd.workloads bfs  --scale 24
    python -m gswitch_amd.workloads cc   --er-batch 1000
    python -m gswitch_amd.workloads copy --mib 1024

Graphs are the seeded R-MAT (a, b, c) = (0.57, 0.19, 0.19), edge factor 16;
the root is the highest-degree vertex. `--er-batch K` replaces the R-MAT graph
by instances 0..K-1 of the seed-5 Erdos-Renyi batch (n log-uniform in
[1000, 200000], mean degree uniform in [2, 32]) as one union. `copy` runs the plain device copy
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
    ap.add_argument("app", choices=["bfs", "sssp", "pr", "cc", "copy"])
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--repeat", type=int, default=1)
    ap.add_argument("--configs", default=None)
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--records", action="store_true")
    ap.add_argument("--pull-kernel", type=int, default=0)
    ap.add_argument("--er-batch", type=int, default=0)
    a = ap.parse_args(argv)
    if a.app == "copy":
        g = gsw.stream_copy(a.mib << 20, a.repeat)
        print(json.dumps({"app": "copy", "bytes": a.mib << 20, "GBps": g}))
        return
    t = time.perf_counter()
    if a.er_batch > 0:
        h, _ = gsw.erdos_renyi_batch(a.er_batch, 1000, 200000, 2.0, 32.0, seed=5,
                                     weighted=(a.app == "sssp"))
    else:
        h = gsw.rmat(a.scale, 16, 0.57, 0.19, 0.19, seed=a.seed,
                     weighted=(a.app == "sssp"))
    root = h.max_degree_vertex()
    g = gsw.Graph(h)
    print("graph built in %.1fs: n=%d m=%d" % (time.perf_counter() - t, h.n, h.m),
          file=sys.stderr, flush=True)
    for _ in range(a.repeat):
        if a.app == "bfs":
            g.bfs(root, configs=a.configs, out=False)
        elif a.app == "sssp":
            g.sssp(root, configs=a.configs, out=False)
        elif a.app == "cc":
            g.cc(configs=a.configs, out=False)
        else:
            g.pagerank(0.85, a.iters, configs=a.configs, out=False,
                       pull_kernel=a.pull_kernel)
        st = g.last_stats
        out = {"app": a.app, "graph": ("er_batch %d" % a.er_batch) if a.er_batch > 0
               else ("rmat %d" % a.scale), "device_ms": st.device_ms, "levels": st.levels}
        if a.records:
            out["records"] = g.records()
        print(json.dumps(out), flush=True)
    g.close()


if __name__ == "__main__":
    main()
