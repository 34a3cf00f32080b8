"""Host ingest: generators and the CSR builder, restated independently in
numpy from their documented definitions (counter-based SplitMix64 draws,
R-MAT quadrant thresholds, grid edge numbering), plus the degree features
restated from graph_attr_t::fill (src/utils/fileIO.cuh:35-67)."""
import numpy as np
import pytest

from conftest import csr_from_edges

U64 = np.uint64


def mix64(z):
    with np.errstate(over="ignore"):
        z = z.astype(U64)
        z = (z ^ (z >> U64(30))) * U64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> U64(27))) * U64(0x94D049BB133111EB)
        return z ^ (z >> U64(31))


def stream_key(seed, stream):
    with np.errstate(over="ignore"):
        return mix64(mix64(np.array([seed], U64)) + U64(stream) * U64(0xD1B54A32D192ED03))[0]


def draw(key, ctr):
    with np.errstate(over="ignore"):
        return mix64(U64(key) + (np.asarray(ctr, U64) + U64(1)) * U64(0x9E3779B97F4A7C15))


def rmat_np(scale, ef, a, b, c, seed):
    ne = ef << scale
    key = stream_key(seed, 1)
    t1 = np.uint64(int(np.floor(a * 2**32)))
    t2 = np.uint64(int(np.floor((a + b) * 2**32)))
    t3 = np.uint64(int(np.floor((a + b + c) * 2**32)))
    per = (scale + 1) // 2
    i = np.arange(ne, dtype=U64)
    u = np.zeros(ne, U64)
    v = np.zeros(ne, U64)
    for k in range(scale):
        r = draw(key, i * U64(per) + U64(k >> 1))
        x = (r >> U64(32)) if (k & 1) else (r & U64(0xFFFFFFFF))
        ub = (x >= t2).astype(U64)
        vb = (((x >= t1) & (x < t2)) | (x >= t3)).astype(U64)
        u = (u << U64(1)) | ub
        v = (v << U64(1)) | vb
    return u.astype(np.int64), v.astype(np.int64)


def pair_weight_np(wkey, lo, hi, lim=64):
    r = draw(wkey, (lo.astype(U64) << U64(32)) | hi.astype(U64))
    return (1 + (((r >> U64(32)) * U64(lim)) >> U64(32))).astype(np.uint32)


@pytest.mark.parametrize("scale,seed", [(8, 1), (11, 7)])
def test_rmat_matches_numpy_restatement(gsw, scale, seed):
    h = gsw.rmat(scale, 16, 0.57, 0.19, 0.19, seed)
    u, v = rmat_np(scale, 16, 0.57, 0.19, 0.19, seed)
    rp, col = csr_from_edges(1 << scale, np.stack([u, v], 1))
    assert np.array_equal(h.row_ptr, rp)
    assert np.array_equal(h.col, col)


def test_rmat_weights_symmetric_and_restated(gsw):
    h = gsw.rmat(9, 8, 0.57, 0.19, 0.19, 3, weighted=True)
    src = np.repeat(np.arange(h.n), np.diff(h.row_ptr))
    lo, hi = np.minimum(src, h.col), np.maximum(src, h.col)
    assert np.array_equal(h.w, pair_weight_np(stream_key(3, 2), lo, hi))
    assert h.w.min() >= 1 and h.w.max() <= 64


def test_grid_structure(gsw):
    R, Cc = 20, 30
    full = gsw.grid2d(R, Cc, 0.0, 0.0, 3)
    assert full.m == 2 * (R * (Cc - 1) + (R - 1) * Cc)
    # deletions follow the documented edge numbering
    h = gsw.grid2d(R, Cc, 0.1, 0.0, 3)
    key = stream_key(3, 3)
    thr = int(0.1 * 2**64)
    kept = []
    for r in range(R):
        for c in range(Cc - 1):
            if int(draw(key, r * (Cc - 1) + c)) >= thr:
                kept.append((r * Cc + c, r * Cc + c + 1))
    for r in range(R - 1):
        for c in range(Cc):
            if int(draw(key, R * (Cc - 1) + r * Cc + c)) >= thr:
                kept.append((r * Cc + c, (r + 1) * Cc + c))
    rp, col = csr_from_edges(R * Cc, kept)
    assert np.array_equal(h.row_ptr, rp) and np.array_equal(h.col, col)
    lr = gsw.grid2d(R, Cc, 0.1, 0.05, 3)
    assert lr.m >= h.m


def test_er_and_ba_properties(gsw):
    h = gsw.erdos_renyi(5000, 8.0, 5)
    assert abs(h.m / h.n - 8.0) < 0.3
    ba = gsw.barabasi_albert(3000, 4, 6)
    deg = ba.degrees
    assert deg.min() >= 4 and deg.max() > 40  # preferential attachment hubs
    assert ba.m == 2 * (10 + (3000 - 5) * 4)  # clique(5) + 4 per new vertex
    again = gsw.barabasi_albert(3000, 4, 6)
    assert np.array_equal(ba.col, again.col)


def test_csr_builder_semantics(gsw):
    rng = np.random.default_rng(0)
    e = rng.integers(0, 300, size=(4000, 2))
    h = gsw.from_edges(300, e[:, 0], e[:, 1])
    rp, col = csr_from_edges(300, e)
    assert np.array_equal(h.row_ptr, rp) and np.array_equal(h.col, col)
    # weighted duplicates keep the smallest weight
    w = rng.integers(1, 100, size=4000)
    hw = gsw.from_edges(300, e[:, 0], e[:, 1], w)
    rp2, col2, w2 = csr_from_edges(300, e, w)
    assert np.array_equal(hw.col, col2) and np.array_equal(hw.w, w2)


def test_csr_builder_directed(gsw):
    h = gsw.from_edges(4, [0, 0, 2, 2], [1, 2, 3, 3], directed=True)
    assert list(h.row_ptr) == [0, 2, 2, 3, 3]
    assert list(h.col) == [1, 2, 3]


def test_csr_builder_rejects_bad_ids(gsw):
    with pytest.raises(gsw.GswError):
        gsw.from_edges(3, [0, 5], [1, 1])


def test_degree_stats_restated(gsw):
    h = gsw.rmat(10, 8, seed=4)
    d = np.sort(np.diff(h.row_ptr)).astype(np.float64)
    n, s = len(d), d.sum()
    i = np.arange(1, n + 1)
    gi = 1 - 2 * np.sum((n - i + 0.5) * d / n / s)
    nz = d > 0
    her = np.sum(d[nz] * (np.log(s) - np.log(d[nz])) / s) / np.log(n)
    st = gsw.degree_stats(h.row_ptr)
    assert st["avg"] == pytest.approx(d.mean())
    assert st["std"] == pytest.approx(d.std(ddof=1))
    assert st["max"] == d.max() and st["range"] == d.max() - d.min()
    assert st["GI"] == pytest.approx(gi)
    assert st["Her"] == pytest.approx(her)
    # closed forms: a regular graph has Gini 0 and entropy 1
    ring = csr_from_edges(50, [(i, (i + 1) % 50) for i in range(50)])
    st = gsw.degree_stats(ring[0])
    assert st["GI"] == pytest.approx(0.0, abs=1e-12)
    assert st["Her"] == pytest.approx(1.0)


def er_batch_params(count, nmin, nmax, dmin, dmax, seed):
    """Instance sizes, mean degrees and seeds of an ER batch, restated."""
    key = stream_key(seed, 7)
    ctr = np.arange(2 * count, dtype=U64)
    u = (draw(key, ctr) >> U64(11)).astype(np.float64) * 2.0 ** -53
    u1, u2 = u[0::2], u[1::2]
    ln0, ln1 = np.log(nmin), np.log(nmax)
    n = np.floor(np.exp(ln0 + u1 * (ln1 - ln0))).astype(np.int64)
    n = np.clip(n, 1, nmax)
    d = dmin + u2 * (dmax - dmin)
    seeds = [int(stream_key(seed, 1000 + i)) for i in range(count)]
    return n, d, seeds


def test_er_batch_is_union_of_instances(gsw):
    count, nmin, nmax = 12, 50, 4000
    h, off = gsw.erdos_renyi_batch(count, nmin, nmax, 2.0, 32.0, seed=5)
    n, d, seeds = er_batch_params(count, nmin, nmax, 2.0, 32.0, 5)
    assert np.array_equal(np.diff(off), n)
    assert off[0] == 0 and off[-1] == h.n
    for i in range(count):
        # instance isolation: the union's block equals the instance built alone
        one = gsw.erdos_renyi(int(n[i]), float(d[i]), seeds[i])
        lo, hi = int(off[i]), int(off[i + 1])
        rp = h.row_ptr[lo:hi + 1] - h.row_ptr[lo]
        cols = h.col[h.row_ptr[lo]:h.row_ptr[hi]]
        assert np.array_equal(rp, one.row_ptr)
        assert np.array_equal(cols - lo, one.col)
        assert cols.min(initial=lo) >= lo and cols.max(initial=lo) < hi
    again, off2 = gsw.erdos_renyi_batch(count, nmin, nmax, 2.0, 32.0, seed=5)
    assert np.array_equal(again.col, h.col) and np.array_equal(off, off2)
    with pytest.raises(gsw.GswError):
        gsw.erdos_renyi_batch(0)
    with pytest.raises(gsw.GswError):
        gsw.erdos_renyi_batch(3, nmin=10, nmax=5)
    with pytest.raises(gsw.GswError):
        gsw.erdos_renyi_batch(3, first=-1)


def test_er_batch_ranges_cut_the_batch(gsw):
    """Consecutive instance ranges are the whole batch cut into pieces."""
    count, nmin, nmax = 10, 50, 3000
    h, off = gsw.erdos_renyi_batch(count, nmin, nmax, 2.0, 32.0, seed=5)
    for first, k in [(0, 4), (4, 1), (5, 5)]:
        p, po = gsw.erdos_renyi_batch(k, nmin, nmax, 2.0, 32.0, seed=5, first=first)
        lo, hi = int(off[first]), int(off[first + k])
        assert np.array_equal(po, off[first:first + k + 1] - lo)
        assert p.n == hi - lo
        assert np.array_equal(p.row_ptr, h.row_ptr[lo:hi + 1] - h.row_ptr[lo])
        assert np.array_equal(p.col, h.col[h.row_ptr[lo]:h.row_ptr[hi]] - lo)


# ---------------------------------------------------------------- file ingest
def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


def test_read_mtx_matrix_market(gsw, tmp_path):
    """Banner, comments and the size line skipped (with header); ids shifted
    to 0; both directions stored; rows sorted; duplicates and self-loops
    dropped."""
    p = _write(tmp_path, "a.mtx", "%%MatrixMarket matrix coordinate pattern symmetric\n"
               "% a comment\n4 4 6\n1 2\n1 3\n3 2\n2 1\n4 4\n3 4\n")
    h = gsw.read_mtx(p, header=True)
    assert h.n == 4 and not h.directed and not h.weighted
    assert list(h.row_ptr) == [0, 2, 4, 7, 8]
    assert list(h.col) == [1, 2, 0, 2, 0, 1, 3, 2]


def test_read_mtx_header_is_an_edge_without_the_flag(gsw, tmp_path):
    # without with_header the size line "3 3 1" is a self-loop on id 3:
    # dropped, but it widens the id range to 1..3
    p = _write(tmp_path, "b.mtx", "3 3 1\n1 2\n")
    h = gsw.read_mtx(p, rinse=False)
    assert h.n == 3 and list(h.row_ptr) == [0, 1, 2, 2] and list(h.col) == [1, 0]


def test_read_mtx_rinse_and_id_shift(gsw, tmp_path):
    # ids 10..15; 12 and 14 have no edge (14 only in a self-loop)
    p = _write(tmp_path, "c.txt", "10 11\n11 13\n14 14\n13 15\n15 10\n")
    keep = gsw.read_mtx(p, rinse=False)
    assert keep.n == 6 and list(np.diff(keep.row_ptr)) == [2, 2, 0, 2, 0, 2]
    r = gsw.read_mtx(p)  # the reference's default for an undirected unweighted graph
    assert r.n == 4
    # kept ids 10, 11, 13, 15 -> 0, 1, 2, 3
    assert list(r.row_ptr) == [0, 2, 4, 6, 8]
    assert list(r.col) == [1, 3, 0, 2, 1, 3, 0, 2]


def test_read_mtx_last_line_without_newline(gsw, tmp_path):
    """The reference's stream loop drops a final record not ended by a
    newline; an unweighted record with extra tokens still counts its ids."""
    h = gsw.read_mtx(_write(tmp_path, "d.txt", "1 2\n2 3"), rinse=False)
    assert h.n == 2 and list(h.col) == [1, 0]
    h = gsw.read_mtx(_write(tmp_path, "e.txt", "1 2\n2 5 9"), rinse=False)
    assert h.n == 5 and list(h.col) == [1, 0]
    h = gsw.read_mtx(_write(tmp_path, "f.txt", "1 2\n2 3\n"), rinse=False)
    assert h.n == 3 and list(h.col) == [1, 0, 2, 1]


def test_read_mtx_weights_and_directed(gsw, tmp_path):
    p = _write(tmp_path, "g.txt", "0 1 7\n1 2 3\n0 1 5\n2 0 9\n")
    h = gsw.read_mtx(p, weighted=True)
    assert h.weighted and h.n == 3
    assert list(h.row_ptr) == [0, 2, 4, 6]
    assert list(h.col) == [1, 2, 0, 2, 0, 1]
    # duplicate 0-1 keeps its smaller weight (shortest paths are unchanged)
    assert list(h.w) == [5, 9, 5, 3, 9, 3]
    d = gsw.read_mtx(p, directed=True)
    assert d.directed and d.n == 3
    assert list(d.row_ptr) == [0, 1, 2, 3] and list(d.col) == [1, 2, 0]
    with pytest.raises(gsw.GswError):
        gsw.read_mtx(p, directed=True, rinse=True)


def test_read_mtx_errors(gsw, tmp_path):
    with pytest.raises(gsw.GswError):
        gsw.read_mtx(str(tmp_path / "missing.mtx"))
    with pytest.raises(gsw.GswError):
        gsw.read_mtx(_write(tmp_path, "h.txt", "% only a comment\n"))
    with pytest.raises(gsw.GswError):
        gsw.read_mtx(_write(tmp_path, "i.txt", "1 1\n2 2\n"))  # self-loops only
    with pytest.raises(gsw.GswError):
        gsw.read_mtx(_write(tmp_path, "j.txt", "1 2 0.5\n2 3 1\n"), weighted=True)


def write_mtx(path, h, base=1):
    """Each undirected edge once (u < v), Matrix Market coordinate layout."""
    u = np.repeat(np.arange(h.n), np.diff(h.row_ptr))
    keep = u < h.col
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern symmetric\n")
        f.write("%d %d %d\n" % (h.n, h.n, int(keep.sum())))
        np.savetxt(f, np.stack([u[keep] + base, h.col[keep] + base], 1), fmt="%d")


def test_read_mtx_round_trip(gsw, tmp_path):
    h = gsw.rmat(10, 8, seed=3)
    p = str(tmp_path / "r.mtx")
    write_mtx(p, h)
    # vertex 0 of R-MAT has edges, so the id shift is 1 (1-based file); the
    # id range ends at the largest id in an edge (n = vmax - vmin + 1), so
    # keeping every id gives back the generator's CSR up to that id
    live = np.flatnonzero(np.diff(h.row_ptr) > 0)
    assert live[0] == 0
    back = gsw.read_mtx(p, header=True, rinse=False)
    assert back.n == live[-1] + 1 < h.n
    assert np.array_equal(back.row_ptr, h.row_ptr[:back.n + 1])
    assert np.array_equal(back.col, h.col)
    # rinse: isolated vertices removed, the rest renumbered in order
    r = gsw.read_mtx(p, header=True)
    assert r.n == len(live)
    assert np.array_equal(np.diff(r.row_ptr), np.diff(h.row_ptr)[live])
    rename = np.full(h.n, -1, np.int64)
    rename[live] = np.arange(len(live))
    assert np.array_equal(r.col, rename[h.col])


def test_record_file_parsing(gsw, tmp_path):
    """Record files hold concatenated JSON objects (the reference's json_t
    dump layout); setting and clearing the file needs no GPU."""
    p = tmp_path / "r.json"
    p.write_text('{\n"0": 5,\n"22": 0.25\n}\n{\n"0": 6,\n"22": 1e-05\n}\n')
    recs = gsw.read_records(str(p))
    assert [r["0"] for r in recs] == [5, 6] and recs[1]["22"] == 1e-05
    gsw.record_features(str(tmp_path / "x.json"))
    gsw.record_features(None)


def test_csr_cache_round_trip(gsw, tmp_path):
    """The binary CSR cache gives back every array: undirected, weighted,
    and directed (with its reverse CSR)."""
    cases = [gsw.rmat(10, 8, seed=2), gsw.rmat(10, 8, seed=2, weighted=True)]
    p = tmp_path / "d.txt"
    p.write_text("0 1\n1 2\n2 0\n2 3\n")
    cases.append(gsw.read_mtx(str(p), directed=True))
    for i, h in enumerate(cases):
        f = str(tmp_path / ("c%d.bin" % i))
        h.save(f)
        b = gsw.load_graph(f)
        assert (b.n, b.m, b.directed, b.weighted) == (h.n, h.m, h.directed, h.weighted)
        assert np.array_equal(b.row_ptr, h.row_ptr) and np.array_equal(b.col, h.col)
        if h.weighted:
            assert np.array_equal(b.w, h.w)
    d = gsw.load_graph(str(tmp_path / "c2.bin"))
    assert d.directed and list(d.col) == [1, 2, 0, 3]


def test_csr_cache_rejects_bad_files(gsw, tmp_path):
    h = gsw.rmat(8, 4, seed=1)
    f = tmp_path / "c.bin"
    h.save(str(f))
    raw = f.read_bytes()
    (tmp_path / "short.bin").write_bytes(raw[:-5])
    (tmp_path / "tag.bin").write_bytes(b"XXXXXXXX" + raw[8:])
    bad = bytearray(raw)
    bad[8 + 16 + 4 + 4 + 8] = 0xFF  # row_ptr[1] becomes huge
    (tmp_path / "rows.bin").write_bytes(bytes(bad))
    for name in ["short.bin", "tag.bin", "rows.bin", "missing.bin"]:
        with pytest.raises(gsw.GswError):
            gsw.load_graph(str(tmp_path / name))
