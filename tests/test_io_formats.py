"""The reference's on-disk formats of SparseMatrixCSR: the binary fm_csr container
(Container::_serialise<double, uint64_t>, container.hpp:320-447) checked byte for byte against an
independent restatement of its layout here, and Matrix Market (sparse_matrix_csr.hpp:1125-1340):
round trips, comments, symmetric expansion, repeated entries (first kept), bad files rejected."""
import ctypes
import math
import struct

import numpy as np
import pytest

from feat3_amd import capi
from feat3_amd import problems as pb


def feat_offset(gi):
    # container.hpp:422/527: the running count passes through float arithmetic
    return int(math.ceil(float(np.float32(np.float32(gi) * np.float32(8.0)) / np.float32(8.0))))


def reference_bytes(A):
    """fm_csr bytes of A written by the reference, restated from container.hpp:320-447."""
    nnz, rows, cols = A.nnz, A.rows, A.cols
    hdr = [0, 4, 8 | 1 << 33 | 1 << 34, 8 | 1 << 32, 1, 2, 1, 2, 4, 1]
    sizes = [nnz, nnz, rows + 1, rows * cols, rows, cols, nnz]
    gsize = (10 + 1 + 2 + 4 + 1 + nnz + nnz + rows + 1) * 8 + 16
    words = bytearray(gsize)
    hdr[0] = gsize
    struct.pack_into(f"<{len(hdr) + len(sizes)}Q", words, 0, *hdr, *sizes)
    gi = feat_offset(17)
    struct.pack_into("<d", words, gi * 8, 0.0)
    gi += 1
    struct.pack_into(f"<{nnz}d", words, gi * 8, *A.val)
    gi = feat_offset(gi + nnz)
    struct.pack_into(f"<{nnz}Q", words, gi * 8, *[int(c) for c in A.col_ind])
    gi += nnz
    struct.pack_into(f"<{rows + 1}Q", words, gi * 8, *[int(p) for p in A.row_ptr])
    return bytes(words)


def same(A, B):
    return (A.rows, A.cols) == (B.rows, B.cols) and np.array_equal(A.row_ptr, B.row_ptr) and \
        np.array_equal(A.col_ind, B.col_ind) and np.array_equal(A.val, B.val)


def test_binary_layout_bytes_and_round_trip(tmp_path):
    A = pb.hierarchy(2, 1, 3, 3).levels[0].A
    f = str(tmp_path / "a.csr")
    pb.write_matrix(f, A)
    with open(f, "rb") as fh:
        assert fh.read() == reference_bytes(A)
    assert same(pb.read_matrix(f), A)
    # a file in the reference's layout written independently is read back
    g = str(tmp_path / "b.csr")
    P = pb.prolongation(3, 1, 3)
    with open(g, "wb") as fh:
        fh.write(reference_bytes(P))
    assert same(pb.read_matrix(g), P)


def test_mtx_round_trip_general_and_symmetric(tmp_path):
    A = pb.hierarchy(3, 1, 2, 2).levels[0].A
    for sym in (False, True):
        f = str(tmp_path / f"a{int(sym)}.mtx")
        pb.write_matrix(f, A, symmetric=sym)
        B = pb.read_matrix(f)
        assert same(B, A)   # %.17e keeps every double; the symmetric file is expanded again


def test_mtx_parsing_rules(tmp_path):
    f = str(tmp_path / "m.mtx")
    with open(f, "w") as fh:
        fh.write("%%MatrixMarket matrix coordinate real symmetric\n% a comment\n\n% another\n3 3 4\n"
                 "1 1 2.0\n2 1 -1.5\n3 3 4.0\n2 1 9.0\n")
    B = pb.read_matrix(f)
    D = B.to_dense()
    assert np.array_equal(D, np.array([[2.0, -1.5, 0.0], [-1.5, 0.0, 0.0], [0.0, 0.0, 4.0]]))   # first of the repeats
    assert B.nnz == 4 and B.row_ptr[2] == B.row_ptr[1] + 1    # row 2 holds only (2, 1)
    bad = str(tmp_path / "bad.mtx")
    with open(bad, "w") as fh:
        fh.write("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    with pytest.raises(capi.Feat3HipError):
        pb.read_matrix(bad)
    with pytest.raises(capi.Feat3HipError):
        pb.read_matrix(str(tmp_path / "missing.csr"))


# ---- DenseVector files (dense_vector.hpp:587-780) ---------------------------------------------
def reference_dv_bytes(v):
    """fm_dv bytes of a DenseVector<double> written by the reference, restated from
    container.hpp:320-447: 10 header words, elements_size {n}, scalar_index {n}, the values,
    16 bytes of padding."""
    n = v.size
    gsize = (4 + 6 + 1 + 1) * 8 + n * 8 + 16
    hdr = [gsize, 1, 8 | 1 << 33 | 1 << 34, 8 | 1 << 32, 1, 0, 1, 0, 1, 0, n, n]
    out = bytearray(gsize)
    struct.pack_into(f"<{len(hdr)}Q", out, 0, *hdr)
    struct.pack_into(f"<{n}d", out, feat_offset(12) * 8, *v)
    return bytes(out)


def test_vector_dv_layout_bytes_and_round_trip(tmp_path):
    v = np.random.default_rng(3).uniform(-1, 1, 37)
    v[5] = 0.0
    v[6] = -1e-300
    f = str(tmp_path / "v.dv")
    pb.write_vector(f, v, "dv")
    with open(f, "rb") as fh:
        assert fh.read() == reference_dv_bytes(v)
    assert np.array_equal(pb.read_vector(f, "dv"), v)
    g = str(tmp_path / "w.dv")
    w = np.arange(5, dtype=np.float64) - 2.5
    with open(g, "wb") as fh:
        fh.write(reference_dv_bytes(w))
    assert np.array_equal(pb.read_vector(g, "dv"), w)
    # empty vector
    pb.write_vector(g, np.zeros(0), "dv")
    assert pb.read_vector(g, "dv").size == 0


def test_vector_exp_and_mtx_text(tmp_path):
    v = np.array([1.0, -2.5e-7, 3.14159265358979, 0.0, 6.02214076e23])
    f = str(tmp_path / "v.exp")
    pb.write_vector(f, v, "exp")
    # std::scientific with the default precision 6, one value per line
    assert open(f).read() == "".join(f"{x:e}\n" for x in v)
    assert np.allclose(pb.read_vector(f, "exp"), v, rtol=5e-7, atol=0)
    g = str(tmp_path / "v.mtx")
    pb.write_vector(g, v, "mtx")
    assert open(g).read() == "%%MatrixMarket matrix array real general\n5 1\n" + "".join(f"{x:e}\n" for x in v)
    assert np.allclose(pb.read_vector(g, "mtx"), v, rtol=5e-7, atol=0)


def test_vector_exp_reader_rules(tmp_path):
    """The reference's fm_exp loop: a line holding '#' is skipped, leading blanks are ignored,
    an empty line reads as 0, and a last line without a newline is dropped."""
    f = str(tmp_path / "r.exp")
    with open(f, "w") as fh:
        fh.write("# header\n  1.5\n\n-2e3   # trailing comment skips the line\n4.25\n7.0")
    assert np.array_equal(pb.read_vector(f, "exp"), np.array([1.5, 0.0, 4.25]))


def test_vector_mtx_reader_rules(tmp_path):
    f = str(tmp_path / "r.mtx")
    with open(f, "w") as fh:
        fh.write("%%MatrixMarket matrix array real general\n% a comment\n3 1\n1.0\n  2.0 extra\n3.0\n")
    assert np.array_equal(pb.read_vector(f, "mtx"), np.array([1.0, 2.0, 3.0]))
    for bad in ("%%MatrixMarket matrix array real general\n3 2\n1\n2\n3\n",           # not n x 1
                "%%MatrixMarket matrix array real general\n3 1\n1\n2\n",              # too few values
                "%%MatrixMarket matrix coordinate real general\n3 1 1\n1 1 1.0\n"):   # not an array file
        with open(f, "w") as fh:
            fh.write(bad)
        with pytest.raises(Exception):
            pb.read_vector(f, "mtx")
    with pytest.raises(Exception):
        pb.read_vector(str(tmp_path / "missing.dv"), "dv")


@pytest.mark.parametrize("word,value", [
    (9, 2),                 # two scalar values: not a SparseMatrixCSR container
    (9, 1 << 62),           # an offset that would wrap
    (10, 1 << 62),          # nnz beyond any array
    (11, 5),                # index array 0 size differs from nnz
    (14, (1 << 63) + 3),    # rows near 2^64: rows + 1 wraps
    (16, 7),                # scalar nnz differs
])
def test_binary_reader_rejects_crafted_headers(tmp_path, word, value):
    """Header words come from the file: each is checked before it is used, so a crafted header is
    an error (code 1), never a read outside the file's words."""
    from feat3_amd.capi import Feat3HipError
    A = pb.hierarchy(2, 1, 2, 2).levels[0].A
    raw = bytearray(reference_bytes(A))
    struct.pack_into("<Q", raw, 8 * word, value)
    f = str(tmp_path / "bad.csr")
    with open(f, "wb") as fh:
        fh.write(bytes(raw))
    with pytest.raises(Feat3HipError):
        pb.read_matrix(f)


def test_binary_reader_rejects_truncated_and_changed_files(tmp_path):
    from feat3_amd import capi
    from feat3_amd.capi import Feat3HipError
    A = pb.hierarchy(2, 1, 2, 2).levels[0].A
    raw = reference_bytes(A)
    f = str(tmp_path / "t.csr")
    for cut in (8, 100, len(raw) - 8):
        with open(f, "wb") as fh:
            fh.write(raw[:cut])
        with pytest.raises(Feat3HipError):
            pb.read_matrix(f)
    # sizes read from one file, then the file replaced by a larger matrix before the read: the
    # read refuses instead of writing past the arrays sized for the first one
    with open(f, "wb") as fh:
        fh.write(raw)
    rows, cols, nnz = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    capi.call("feat3_io_csr_binary_sizes", f.encode(), ctypes.byref(rows), ctypes.byref(cols), ctypes.byref(nnz))
    B = pb.hierarchy(2, 1, 3, 3).levels[0].A
    pb.write_matrix(f, B)
    rp = np.empty(rows.value + 1, np.uint32)
    col = np.empty(nnz.value, np.uint32)
    val = np.empty(nnz.value)
    with pytest.raises(Feat3HipError):
        capi.call("feat3_io_read_csr_binary", f.encode(), rows.value, nnz.value, ctypes.c_void_p(rp.ctypes.data),
                  ctypes.c_void_p(col.ctypes.data), ctypes.c_void_p(val.ctypes.data))
    g = str(tmp_path / "t.mtx")
    pb.write_matrix(g, A)
    capi.call("feat3_io_mtx_sizes", g.encode(), ctypes.byref(rows), ctypes.byref(cols), ctypes.byref(nnz))
    pb.write_matrix(g, B)
    with pytest.raises(Feat3HipError):
        capi.call("feat3_io_read_mtx", g.encode(), rows.value, nnz.value, ctypes.c_void_p(rp.ctypes.data),
                  ctypes.c_void_p(col.ctypes.data), ctypes.c_void_p(val.ctypes.data))
