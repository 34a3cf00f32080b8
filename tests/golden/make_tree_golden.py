"""Write golden (feature vector -> class) tables for the rule selector's trees.

The classes come from the reference's own decision trees (src/model/*.h),
compiled where they lie into oracle/_ref/dump_trees (``make -C oracle ref``).
Inputs mix (a) vectors whose features sit exactly on, just below and just
above thresholds the tree tests, so most branches are taken, and (b) seeded
log-uniform vectors. Output: tests/golden/trees_golden.npz.

Run in the build container (needs /root/reference):
    make -C oracle ref && python tests/golden/make_tree_golden.py
"""
import os
import re
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "gswitch_amd", "csrc", "gswitch", "model"))
import gen_trees  # noqa: E402

MODEL_DIR = "/root/reference/src/model"
# Length of the feature vector each call site builds (src/abstraction/
# selector.cuh, src/gswitch.h, src/data_structures/window.cuh).
NFEAT = {
    "select_dir_idem": 8, "select_dir_asso": 8, "select_dir_asso_backup": 7,
    "select_fmt": 1, "select_fusion": 5, "select_lb_idem": 13,
    "select_lb_asso_forward": 13, "select_lb_asso_backward": 13,
    "select_lb_asso_forward_backup": 7, "select_lb_asso_backward_backup": 7,
    "select_stepping": 2,
}
PER_TREE = 600


def vectors(nodes, k, rng):
    thr = {}
    for f, t, _, _, _ in nodes:
        if f >= 0:
            thr.setdefault(f, []).append(t)
    out = []
    for i in range(PER_TREE):
        v = np.empty(k)
        for f in range(k):
            if i % 2 == 0 and f in thr:
                t = thr[f][rng.integers(len(thr[f]))]
                mode = rng.integers(3)
                v[f] = t if mode == 0 else (t - abs(t) * 1e-3 - 1e-6 if mode == 1
                                            else t + abs(t) * 1e-3 + 1e-6)
            else:
                v[f] = 10 ** rng.uniform(-4, 8) * (1 if rng.random() < 0.95 else 0)
        out.append(v)
    return np.array(out)


def main():
    exe = os.path.join(REPO, "oracle", "_ref", "dump_trees")
    rng = np.random.default_rng(20240601)
    payload = {}
    for name in gen_trees.TREES:
        nodes, _ = gen_trees.parse_tree(os.path.join(MODEL_DIR, name + ".h"), name)
        k = NFEAT[name]
        X = vectors(nodes, k, rng)
        text = "".join("%s %d %s\n" % (name, k, " ".join(repr(float(x)) for x in row))
                       for row in X)
        res = subprocess.run([exe], input=text, capture_output=True, text=True, check=True)
        y = np.array([int(s) for s in res.stdout.split()], dtype=np.int32)
        assert len(y) == len(X)
        payload[name + "_x"] = X
        payload[name + "_y"] = y
    np.savez_compressed(os.path.join(HERE, "trees_golden.npz"), **payload)
    print("wrote", sum(len(v) for k, v in payload.items() if k.endswith("_y")), "rows")


if __name__ == "__main__":
    main()
