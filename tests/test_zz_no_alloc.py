"""gsw_cc_noalloc: connected components that allocate nothing.

The entry point writes the labels into a device buffer the caller owns and
keeps its own scratch (the vertex sample and the chosen root) in a static
device array. Its labels must equal gsw_cc's on the same graph, on two
calls in a row whose second one is smaller (fewer vertices and fewer
samples), with the scratch left over from the first call; and both must
equal the oracle. The device buffers here come from the HIP runtime
through ctypes (the test allocates; the entry point must not).
"""
import ctypes as C

import numpy as np
import pytest

from conftest import csr_from_edges

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip(gsw):
    lib = C.CDLL("libamdhip64.so.7")  # the runtime the engine already loaded
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    lib.hipStreamSynchronize.argtypes = [C.c_void_p]
    lib.hipStreamDestroy.argtypes = [C.c_void_p]
    return lib


def dev_labels(hip, n, fill=-7):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), 4 * n) == 0
    h = np.full(n, fill, np.int32)
    assert hip.hipMemcpy(p, h.ctypes.data, 4 * n, H2D) == 0
    return p


def read_labels(hip, p, n):
    h = np.empty(n, np.int32)
    assert hip.hipMemcpy(h.ctypes.data, p, 4 * n, D2H) == 0
    return h


def test_noalloc_matches_cc_twice_second_smaller(gsw, orc, hip):
    big = gsw.grid2d(300, 300, 0.45, 0.0, 9)      # many components
    small = gsw.rmat(12, 4, 0.57, 0.19, 0.19, seed=3)
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    for h, samples, st in [(big, 4096, None), (small, 100, stream.value)]:
        g = gsw.Graph(h)
        ref = g.cc()
        assert np.array_equal(ref, orc.cc(h.row_ptr, h.col))
        p = dev_labels(hip, h.n)
        g.cc_noalloc(p.value, samples=samples, stream=st)
        assert hip.hipStreamSynchronize(C.c_void_p(st or 0)) == 0
        got = read_labels(hip, p, h.n)
        hip.hipFree(p)
        g.close()
        assert np.array_equal(got, ref), (h.n, samples)
    hip.hipStreamDestroy(stream)


def test_noalloc_directed_and_repeat(gsw, orc, hip):
    """Directed graphs (weak components, no skipped set) and a repeated call
    on the same buffer give the same labels."""
    rng = np.random.default_rng(4)
    n = 3000
    src, dst = rng.integers(0, n, 2500), rng.integers(0, n, 2500)
    hd = gsw.from_edges(n, src, dst, directed=True)
    g = gsw.Graph(hd)
    srp, scol = csr_from_edges(n, np.stack([src, dst], 1))
    p = dev_labels(hip, n)
    for _ in range(2):
        g.cc_noalloc(p.value, samples=1)
        assert hip.hipStreamSynchronize(C.c_void_p(0)) == 0
        assert np.array_equal(read_labels(hip, p, n), orc.cc(srp, scol))
    hip.hipFree(p)
    g.close()


def test_noalloc_rejects_bad_arguments(gsw, hip):
    h = gsw.grid2d(20, 20, 0.1, 0.0, 1)
    g = gsw.Graph(h)
    p = dev_labels(hip, h.n)
    for bad in (0, 4097):
        with pytest.raises(gsw.GswError, match="samples"):
            g.cc_noalloc(p.value, samples=bad)
    host = np.zeros(h.n, np.int32)
    with pytest.raises(gsw.GswError, match="device memory"):
        g.cc_noalloc(host.ctypes.data)
    # a rejected call leaves the library usable
    g.cc_noalloc(p.value)
    assert hip.hipStreamSynchronize(C.c_void_p(0)) == 0
    assert np.array_equal(read_labels(hip, p, h.n), g.cc())
    hip.hipFree(p)
    g.close()
