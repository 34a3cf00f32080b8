import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: large inputs")


def csr_from_edges(n, edges, weights=None):
    """Symmetric, deduplicated, sorted CSR built with numpy (test helper)."""
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    keep = e[:, 0] != e[:, 1]
    e = e[keep]
    w = None if weights is None else np.asarray(weights, np.int64)[keep]
    both = np.concatenate([e, e[:, ::-1]])
    wb = None if w is None else np.concatenate([w, w])
    key = both[:, 0] * n + both[:, 1]
    if wb is None:
        key = np.unique(key)
    else:
        order = np.lexsort((wb, key))
        key, wb = key[order], wb[order]
        first = np.ones(len(key), bool)
        first[1:] = key[1:] != key[:-1]
        key, wb = key[first], wb[first]
    src, dst = key // n, key % n
    row_ptr = np.zeros(n + 1, np.int64)
    np.add.at(row_ptr, src + 1, 1)
    row_ptr = np.cumsum(row_ptr)
    out = (row_ptr, dst.astype(np.int32))
    if wb is not None:
        out = out + (wb.astype(np.uint32),)
    return out


@pytest.fixture(scope="session")
def gsw():
    import gswitch_amd
    return gswitch_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


def have_gpu():
    try:
        import gswitch_amd
        return gswitch_amd.device_count() > 0
    except Exception:
        return False
