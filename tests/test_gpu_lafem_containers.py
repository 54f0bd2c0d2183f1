"""The Python device containers (feat3_amd.lafem) against the oracle's restatement of the
reference's container operations, through the lafem_hip_* C ABI.

Bars: FP64 within 1e-13 relative (device reductions sum in another order than the sequential CPU
loop), integer-valued reductions exact, filters and diagonals bit-exact.
"""
import numpy as np
import pytest

from feat3_amd import problems as pb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lafem(torch_cuda):
    from feat3_amd import lafem
    return lafem


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_dense_vector_blas1(lafem):
    rng = np.random.default_rng(2)
    n = 10007
    x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    X, Y, R = lafem.DenseVector(x), lafem.DenseVector(y), lafem.DenseVector(n)
    R.axpy(X, Y, 0.25)
    assert rel(R.to_numpy(), 0.25 * x + y) <= 1e-15
    R.axpy(X, Y, 1e-300)          # |alpha| < eps: copy of y (dense_vector.hpp:891-895)
    assert np.array_equal(R.to_numpy(), y)
    R.component_product(X, Y)
    assert np.array_equal(R.to_numpy(), x * y)
    R.scale(X, -2.5)
    assert np.array_equal(R.to_numpy(), -2.5 * x)
    R.component_invert(lafem.DenseVector(x + 3.0), 0.5)
    assert np.array_equal(R.to_numpy(), 0.5 / (x + 3.0))
    assert abs(X.dot(Y) - po.lib().orc_dot_f64(po.P(x), po.P(y), n)) <= 1e-13 * np.abs(x * y).sum()
    assert abs(X.norm2() - np.sqrt(np.dot(x, x))) <= 1e-13 * np.sqrt(np.dot(x, x))
    xi = rng.integers(-100, 100, n).astype(np.float64)
    assert lafem.DenseVector(xi).dot(lafem.DenseVector(xi)) == float(np.dot(xi, xi))
    F = lafem.DenseVector(n)
    F.format(3.0)
    assert np.all(F.to_numpy() == 3.0)
    import torch
    S = lafem.DenseVector(n, dtype=torch.float32)
    S.convert(X)
    assert np.array_equal(S.to_numpy(), x.astype(np.float32))


def test_sparse_matrix_csr_apply_and_filter(lafem):
    H = pb.hierarchy(2, 1, 5, 4)
    A = H.levels[0].A
    rng = np.random.default_rng(4)
    x, y = rng.uniform(-1, 1, A.cols), rng.uniform(-1, 1, A.rows)
    M = lafem.SparseMatrixCSR(A)
    R, X, Y = lafem.DenseVector(A.rows), lafem.DenseVector(x), lafem.DenseVector(y)
    M.apply(R, X)
    ref = np.zeros(A.rows)
    po.csr_apply(ref, 1.0, x, 0.0, ref, A.val, A.col_ind, A.row_ptr, A.rows, A.cols)
    assert rel(R.to_numpy(), ref) <= 1e-13
    M.apply(R, X, Y, -0.5)
    ref2 = y.copy()
    po.csr_apply(ref2, -0.5, x, 1.0, y, A.val, A.col_ind, A.row_ptr, A.rows, A.cols)
    assert rel(R.to_numpy(), ref2) <= 1e-13
    D = M.extract_diag()
    dref = np.array([A.val[A.row_ptr[i]:A.row_ptr[i + 1]][A.col_ind[A.row_ptr[i]:A.row_ptr[i + 1]] == i][0]
                     for i in range(A.rows)])
    assert np.array_equal(D.to_numpy(), dref)
    with pytest.raises(ValueError):
        M.apply(X, X)
    Fl = lafem.UnitFilter(H.levels[0].bnd, np.full(H.levels[0].bnd.size, 7.0))
    V = lafem.DenseVector(y)
    Fl.filter_rhs(V)
    want = y.copy()
    want[H.levels[0].bnd] = 7.0
    assert np.array_equal(V.to_numpy(), want)
    Fl.filter_def(V)
    want[H.levels[0].bnd] = 0.0
    assert np.array_equal(V.to_numpy(), want)
