"""Rule selector trees: the re-encoded node tables evaluated through the
C-ABI must reproduce the reference's compiled decision trees
(src/model/*.h) on the golden vectors in tests/golden/trees_golden.npz
(written by tests/golden/make_tree_golden.py from oracle/_ref/dump_trees)."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "trees_golden.npz")
TREES = ["select_dir_idem", "select_dir_asso", "select_dir_asso_backup",
         "select_fmt", "select_fusion", "select_lb_idem",
         "select_lb_asso_forward", "select_lb_asso_backward",
         "select_lb_asso_forward_backup", "select_lb_asso_backward_backup",
         "select_stepping"]


@pytest.mark.parametrize("name", TREES)
def test_tree_matches_reference_golden(gsw, name):
    d = np.load(GOLDEN)
    X, y = d[name + "_x"], d[name + "_y"]
    got = np.array([gsw.tree_eval(name, x) for x in X])
    assert np.array_equal(got, y), "%d mismatches" % (got != y).sum()


def test_tree_closed_forms(gsw):
    # select_fmt: Queue (0) iff active ratio <= 0.06 (src/model/select_fmt.h)
    assert gsw.tree_eval("select_fmt", [0.06]) == 0
    assert gsw.tree_eval("select_fmt", [0.0600001]) == 1
    with pytest.raises(gsw.GswError):
        gsw.tree_eval("select_lb_idem", [1.0, 2.0])
    with pytest.raises(gsw.GswError):
        gsw.tree_eval("no_such_tree", [1.0])
