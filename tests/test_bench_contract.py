"""bench.py's contract on the CPU: one GPU only, and the bookkeeping that
turns per-expand records into the roofline line."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def rec(d, fmt, lb, fused=False, sorted_=False, frontier=0, edges=0, verts=0,
        active=0, ms=1.0, kernel=2):
    return {"dir": d, "fmt": fmt, "lb": lb, "fused": fused, "sorted": sorted_,
            "frontier": frontier, "edges": edges, "verts": verts, "active": active,
            "ms": ms, "kernel": kernel}


def test_kernel_keys():
    assert bench.kernel_key("sssp", rec("Push", "Queue", "WM", sorted_=True)) == \
        "sssp:Push-SortedQueue-WM-Standalone"
    # sortedness only names queue frontiers
    assert bench.kernel_key("pagerank", rec("Pull", "Bitmap", "TWOD", sorted_=True)) == \
        "pagerank:Pull-Bitmap-TWOD-Standalone/pull_kernel=2"
    # pulls name their from-all pull kernel, pushes do not
    assert bench.kernel_key("pagerank", rec("Pull", "Bitmap", "TWOD", kernel=0)) == \
        "pagerank:Pull-Bitmap-TWOD-Standalone/pull_kernel=0"
    assert bench.kernel_key("sssp", rec("Push", "Queue", "WM", kernel=0)) == \
        "sssp:Push-Queue-WM-Standalone"
    assert bench.kernel_key("sssp", rec("Push", "Queue", "CM", fused=True)) == \
        "sssp:Push-Queue-CM-Fused"


def test_accumulate_groups_launches_by_variant():
    acc = {}
    n, m = 1000, 16000
    recs = [rec("Pull", "Bitmap", "TWOD", ms=2.0) for _ in range(3)]
    bench.accumulate(acc, "pagerank", recs, n, m)
    e = acc["pagerank:Pull-Bitmap-TWOD-Standalone/pull_kernel=2"]
    assert e["launches"] == 3 and e["ms"] == pytest.approx(6.0)
    assert e["bytes"] == pytest.approx(3 * bench.pr_bytes(n, m))
    # SSSP levels: each charged with the next level's active count
    s = [rec("Push", "Queue", "WM", frontier=10, edges=100, active=10, ms=1.0),
         rec("Push", "Queue", "WM", frontier=30, edges=300, active=30, ms=1.5)]
    bench.accumulate(acc, "sssp", s, n, m)
    e = acc["sssp:Push-Queue-WM-Standalone"]
    assert e["launches"] == 2 and e["ms"] == pytest.approx(2.5)
    assert e["bytes"] == pytest.approx(bench.sssp_level_bytes(s[0], 30, n) +
                                       bench.sssp_level_bytes(s[1], 0, n))


def test_bfs_level_bytes_use_the_right_counts():
    # a pull level is charged per inspected target (verts), not per frontier
    # vertex; a push level per frontier vertex
    pull = rec("Pull", "Bitmap", "TM", frontier=5, verts=1000, edges=4000)
    push = rec("Push", "Queue", "WM", frontier=5, verts=1000, edges=4000)
    n = 4096
    assert bench.bfs_level_bytes(pull, 0, n) - bench.bfs_level_bytes(
        dict(pull, verts=0), 0, n) == pytest.approx(20.0 * 1000)
    assert bench.bfs_level_bytes(push, 0, n) - bench.bfs_level_bytes(
        dict(push, frontier=0), 0, n) == pytest.approx(20.0 * 5)
    # bitmaps are read only by pulls
    assert bench.bfs_level_bytes(dict(pull, verts=0, edges=0), 0, n) == pytest.approx(n / 4)
    assert bench.bfs_level_bytes(dict(push, frontier=0, edges=0), 0, n) == 0
