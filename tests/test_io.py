"""Operator and stencil-graph file I/O against the reference's formats: its
own data files (tests/golden/*.points, *.graph, copied from data/) and its
Matrix Market writer's output (golden files written by the reference's
compiled rbf_io.h, and the live reference build when present)."""
import os

import numpy as np
import pytest

import rbf_amd as R
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _g(name):
    return os.path.join(GOLD, name)


def _parse_graph(path):
    with open(path) as f:
        n, nnz = map(int, f.readline().split())
        rows = [list(map(int, f.readline().split())) for _ in range(n)]
    ia = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    return n, nnz, ia, np.array([j for r in rows for j in r], np.int32)


@pytest.mark.parametrize("stem,k", [("cartesian_32_21", 21), ("poisson_32_21", 0)])
def test_reads_reference_data_files(stem, k):
    x, y = R.read_points(_g(stem + ".points"))
    xy = np.loadtxt(_g(stem + ".points"), skiprows=1)
    np.testing.assert_array_equal(x, xy[:, 0])
    np.testing.assert_array_equal(y, xy[:, 1])
    ia, ja = R.read_graph(_g(stem + ".graph"), k)
    n, nnz, ia_ref, ja_ref = _parse_graph(_g(stem + ".graph"))
    assert ia.size == n + 1 == x.size + 1 and ja.size == nnz
    np.testing.assert_array_equal(ia, ia_ref)
    np.testing.assert_array_equal(ja, ja_ref)
    if k:
        np.testing.assert_array_equal(ia, R.make_row_ptr(n, k))
    if O.ref_available():
        import ctypes
        rn, rnnz = ctypes.c_int32(), ctypes.c_int64()
        L = O.ref()
        assert L.ref_read_graph(os.fsencode(_g(stem + ".graph")), k, ctypes.byref(rn),
                                ctypes.byref(rnnz), None, None) == 0
        ia2 = np.empty(rn.value + 1, np.int32)
        ja2 = np.empty(rnnz.value, np.int32)
        L.ref_read_graph(os.fsencode(_g(stem + ".graph")), k, ctypes.byref(rn),
                         ctypes.byref(rnnz), O._p(ia2), O._p(ja2))
        np.testing.assert_array_equal(ia, ia2)
        np.testing.assert_array_equal(ja, ja2)


@pytest.mark.parametrize("text,k,what", [
    ("3\n", 0, "line 1 must hold two counts"),
    ("2 3 4\n0 1\n1\n", 0, "line 1 must hold two counts"),
    ("2 3\n0 1\n\n", 0, "stencil row 1 is empty"),
    ("2 3\n0 1\n2 1\n", 0, "node index 2 lies outside 0..1"),
    ("2 3\n0 1\n-1 1\n", 0, "node index -1 lies outside 0..1"),
    ("2 3\n0 0\n1\n", 0, "stencil row 0 names node 0 twice"),
    ("2 4\n0 1\n1\n", 2, "stencil row 1 holds 1 nodes; fixed-size stencils hold 2"),
    ("2 5\n0 1\n1\n", 2, "fixed size k = 2 need 4 entries"),
    ("2 3\n0 1\n1\n1 0\n", 0, "text follows the last of the 2 stencil rows"),
    ("2 5\n0 1\n1\n", 0, "the rows hold 3 entries; line 1 declares 5"),
    ("2 2\n0 1\n1\n", 0, "more than the 2 entries"),
    ("2 3\n0 1\n1 x\n", 0, "'x' is not an integer node index"),
    ("2 3\n0 1.5\n1 0\n", 0, "'1.5' is not an integer node index"),
    ("3 4\n0 1\n1\n", 0, "the file ends after 2 of 3 stencil rows"),
])
def test_graph_validation_follows_the_format(tmp_path, text, k, what):
    # the rules of the graph file (rbf_io.h:630-648, docs/file_formats.md)
    p = tmp_path / "bad.graph"
    p.write_text(text)
    with pytest.raises(R.RbfError) as e:
        R.read_graph(str(p), k)
    assert e.value.status == R.ERR_FORMAT
    assert what in R.lib().rbfa_last_error().decode(), R.lib().rbfa_last_error()


def test_graph_reader_accepts_crlf_and_trailing_blank_lines(tmp_path):
    p = tmp_path / "ok.graph"
    p.write_bytes(b"3 6\r\n0 1\r\n1 2 0\r\n2\r\n\r\n  \n")
    ia, ja = R.read_graph(str(p))
    np.testing.assert_array_equal(ia, [0, 2, 5, 6])
    np.testing.assert_array_equal(ja, [0, 1, 1, 2, 0, 2])


@pytest.mark.parametrize("text,what", [
    ("", "line 1 must hold the node count"),
    ("2 3\n0 0\n1 1\n", "line 1 must hold the node count"),
    ("x\n", "line 1 must hold the node count"),
    ("3\n0 0\n1 1\n", "the file ends after 2 of 3 coordinate pairs"),
    ("2\n0 0\n1\n", "line 3: a node takes exactly two coordinates"),
    ("2\n0 0 0\n1 1\n", "line 2: a node takes exactly two coordinates"),
    ("2\n0 0\n1 nan\n", "line 3: coordinates must be finite"),
    ("2\n0 inf\n1 1\n", "line 2: coordinates must be finite"),
])
def test_points_validation_follows_the_format(tmp_path, text, what):
    # the rules of the points file (rbf_io.h:198-221, docs/file_formats.md)
    p = tmp_path / "bad.points"
    p.write_text(text)
    with pytest.raises(R.RbfError) as e:
        R.read_points(str(p))
    assert e.value.status == R.ERR_FORMAT
    assert what in R.lib().rbfa_last_error().decode(), R.lib().rbfa_last_error()


def test_points_after_the_last_pair_are_ignored(tmp_path):
    p = tmp_path / "ok.points"
    p.write_text("2\n0.5 -1e-3\n2 3\nanything at all\n")
    x, y = R.read_points(str(p))
    np.testing.assert_array_equal(x, [0.5, 2.0])
    np.testing.assert_array_equal(y, [-1e-3, 3.0])


def test_missing_file_is_an_io_error(tmp_path):
    with pytest.raises(R.RbfError) as e:
        R.read_points(str(tmp_path / "none.points"))
    assert e.value.status == R.ERR_IO and "cannot open" in R.lib().rbfa_last_error().decode()


def test_matrix_market_matches_the_reference_writer(tmp_path):
    g = np.load(_g("mm_small.npz"))
    for name, a in (("mm_small.mtx", g["a"]), ("mm_small_pattern.mtx", None)):
        out = tmp_path / name
        R.write_matrix_market(str(out), g["ia"], g["ja"], a, cols=40)
        assert out.read_bytes() == open(_g(name), "rb").read(), name
    # 1-based CSR gives the same file
    out = tmp_path / "b1.mtx"
    R.write_matrix_market(str(out), g["ia"] + 1, g["ja"] + 1, g["a"], cols=40, csr_base=1)
    assert out.read_bytes() == open(_g("mm_small.mtx"), "rb").read()


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_matrix_market_matches_the_reference_writer_on_a_larger_operator(tmp_path):
    rng = np.random.default_rng(3)
    n, k = 500, 21
    ia = R.make_row_ptr(n, k)
    ja = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for _ in range(n)]).astype(np.int32)
    a = rng.standard_normal(ja.size) * np.exp(rng.uniform(-40, 40, ja.size))
    mine, theirs = tmp_path / "m.mtx", tmp_path / "r.mtx"
    R.write_matrix_market(str(mine), ia, ja, a)
    assert O.ref().ref_write_mm(os.fsencode(str(theirs)), n, n, O._p(ia), O._p(ja), O._p(a)) == 0
    assert mine.read_bytes() == theirs.read_bytes()


def test_matrix_market_round_trip_is_exact(tmp_path):
    import scipy.io
    import scipy.sparse as sp
    g = np.load(_g("mm_small.npz"))
    out = tmp_path / "m.mtx"
    R.write_matrix_market(str(out), g["ia"], g["ja"], g["a"], cols=40)
    A = scipy.io.mmread(str(out)).tocsr()
    B = sp.csr_matrix((g["a"], g["ja"], g["ia"]), shape=(12, 40))
    assert (A != B).nnz == 0
    np.testing.assert_array_equal(A.sorted_indices().data, B.sorted_indices().data)


def test_writer_rejects_bad_structure(tmp_path):
    with pytest.raises(R.RbfError):
        R.write_matrix_market(str(tmp_path / "x.mtx"), [0, 2, 1], [0, 1], [1.0, 2.0])
    with pytest.raises(R.RbfError):
        R.write_matrix_market(str(tmp_path / "x.mtx"), [0, 1], [5], [1.0], cols=3)
