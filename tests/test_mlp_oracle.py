"""Learned-selector pieces that run without a GPU: the FP64 network in the
oracle pinned by central finite differences and closed forms, the class
table, and the product's feature function (C, host side) against a numpy
restatement of its definition."""
import numpy as np
import pytest

NP = 6942
OW1, OB1, OW2, OB2, OW3, OB3 = 0, 768, 832, 4928, 4992, 6912


def rand_problem(seed, B=6):
    rng = np.random.default_rng(seed)
    w = rng.normal(0, 0.25, NP)
    x = rng.random((B, 12))
    t = rng.random((B, 30)) * 3
    m = (rng.random((B, 30)) > 0.25).astype(np.float64)
    return w, x, t, m


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_grad_finite_differences(orc, seed):
    w, x, t, m = rand_problem(seed)
    loss, g = orc.mlp_grad(w, x, t, m)
    rng = np.random.default_rng(100 + seed)
    # parameters from every block, plus the block edges
    idx = np.concatenate([rng.integers(lo, hi, 12) for lo, hi in
                          [(OW1, OB1), (OB1, OW2), (OW2, OB2), (OB2, OW3),
                           (OW3, OB3), (OB3, NP)]] + [[0, OB1 - 1, NP - 1]])
    for i in idx:
        e = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[i] += e
        wm[i] -= e
        fd = (orc.mlp_grad(wp, x, t, m)[0] - orc.mlp_grad(wm, x, t, m)[0]) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_oracle_forward_closed_forms(orc):
    rng = np.random.default_rng(3)
    x = rng.random((4, 12))
    w = np.zeros(NP)
    w[OB3:] = np.arange(30)
    assert np.array_equal(orc.mlp_forward(w, x), np.tile(np.arange(30.0), (4, 1)))
    # one path: x0 -> h1[5] -> h2[7] -> y[2], all weights 2, ReLU passes
    w = np.zeros(NP)
    w[OW1 + 0 * 64 + 5] = 2.0
    w[OW2 + 5 * 64 + 7] = 2.0
    w[OW3 + 7 * 30 + 2] = 2.0
    y = orc.mlp_forward(w, x)
    assert np.allclose(y[:, 2], 8 * x[:, 0]) and not y[:, [0, 1, 3]].any()
    w[OW1 + 0 * 64 + 5] = -2.0  # ReLU blocks the path
    assert not orc.mlp_forward(w, x).any()
    # loss of a perfect fit is 0; an empty mask gives 0 and a zero gradient
    w, x, t, m = rand_problem(4)
    y = orc.mlp_forward(w, x)
    assert orc.mlp_grad(w, x, y, m)[0] == pytest.approx(0, abs=1e-20)
    loss, g = orc.mlp_grad(w, x, t, np.zeros_like(m))
    assert loss == 0 and not g.any()


def test_class_table(gsw):
    assert len(gsw.MLP_CLASSES) == 30 and len(set(gsw.MLP_CLASSES)) == 30
    assert gsw.MLP_CLASSES[0] == "Push-Queue-TM-Standalone"
    assert gsw.MLP_CLASSES[3 * 5 + 1] == "Push-Queue-WM-Fused"
    assert gsw.MLP_CLASSES[5 * 5 + 3] == "Pull-Bitmap-STRICT-Standalone"


def features_np(c):
    n, m, avg, std, gini, level, act, inact, pw, lw, md = c
    f = np.float32
    sd = lambda a, b: f(a) / f(b) if b > 0 else f(0)
    return np.array([
        np.log2(f(1) + f(n)) / f(24), np.log2(f(1) + f(m)) / f(32),
        np.log2(f(1) + f(avg)) / f(8), np.log2(f(1) + sd(std, avg)) / f(4),
        f(gini), np.log2(f(1) + max(f(level), f(0))) / f(8),
        sd(act, n), sd(inact, n),
        np.log2(f(1) + f(pw)) / f(32), np.log2(f(1) + f(lw)) / f(32),
        np.log2(f(1) + sd(pw, max(f(act), f(1)))) / f(16),
        np.log2(f(1) + f(md)) / f(24)], np.float32)


@pytest.mark.parametrize("counts", [
    (65536, 2.1e6, 32.0, 110.5, 0.61, 0, 1, 65535, 3000, 2.1e6 - 3000, 3000),
    (1e6, 8e6, 8.0, 2.5, 0.2, 5, 123456, 400000, 987654, 3.2e6, 55),
    (4096, 0, 0.0, 0.0, 0.0, 3, 0, 0, 0, 0, 0),
])
def test_features_restated(gsw, counts):
    np.testing.assert_allclose(gsw.mlp_features(counts), features_np(counts),
                               rtol=2e-6, atol=1e-7)
