"""Host-side CSR -> SELL-C-sigma conversion (the derived device format): every slot a kernel
reads addresses x inside [0, cols) -- for square system matrices and for the rectangular
transfer matrices P (fine x coarse) and R -- and the SELL product equals the CSR product."""
import ctypes

import numpy as np
import pytest

from feat3_amd import capi
from feat3_amd import problems as pb


def sell(m: pb.Csr, C: int, sigma: int, d16: bool = True):
    nch, vs = ctypes.c_uint64(), ctypes.c_uint64()
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    capi.call("feat3_sell_sizes", m.rows, P(m.row_ptr), C, sigma, ctypes.byref(nch), ctypes.byref(vs))
    cs = np.zeros(nch.value + 1, np.uint32)
    cl = np.zeros(max(nch.value, 1), np.uint32)
    perm = np.zeros(max(m.rows, 1), np.uint32)
    col = np.zeros(max(vs.value, 1), np.uint32)
    val = np.zeros(max(vs.value, 1))
    dd = np.zeros(max(vs.value, 1), np.int16)
    used = ctypes.c_int()
    capi.call("feat3_sell_build", m.rows, m.cols, P(m.row_ptr), P(m.col_ind), P(m.val), C, sigma, P(cs), P(cl),
              P(perm), P(col), P(val), P(dd) if d16 else ctypes.c_void_p(0), ctypes.byref(used))
    return cs, cl, perm[: m.rows], col, val, dd, bool(used.value)


def sell_apply(m, cs, cl, perm, col, val, dd, used, x, C):
    """The kernel's index arithmetic (one lane per position p, loop to the chunk width)."""
    y = np.zeros(m.rows)
    for p in range(m.rows):
        ch = p // C
        base = int(cs[ch]) + p % C
        r = int(perm[p])
        s = 0.0
        for k in range(int(cl[ch])):
            e = base + k * C
            c = r + int(dd[e]) if used else int(col[e])
            assert 0 <= c < m.cols, (p, k, c)
            s += val[e] * x[c]
        y[r] = s
    return y


@pytest.mark.parametrize("which", ["A", "P", "R", "Pcompact"])
@pytest.mark.parametrize("C,sigma", [(64, 1), (32, 64), (7, 3)])
def test_sell_indices_in_range_and_product(which, C, sigma):
    H = pb.hierarchy(3, 1, 3, 1)
    L = H.levels[0]
    m = {"A": L.A, "P": L.P, "R": L.R}.get(which)
    if which == "Pcompact":
        keep = L.P.val != 0.0
        rows = np.repeat(np.arange(L.P.rows), np.diff(L.P.row_ptr))
        rp = np.zeros(L.P.rows + 1, np.uint32)
        np.add.at(rp, rows[keep] + 1, 1)
        m = pb.Csr(L.P.rows, L.P.cols, np.cumsum(rp).astype(np.uint32), L.P.col_ind[keep], L.P.val[keep])
    x = np.random.default_rng(0).uniform(-1, 1, m.cols)
    ref = m.to_dense() @ x
    for d16 in (True, False):
        cs, cl, perm, col, val, dd, used = sell(m, C, sigma, d16)
        assert used == d16
        y = sell_apply(m, cs, cl, perm, col, val, dd, used, x, C)
        assert np.allclose(y, ref, rtol=0, atol=1e-13)
        if sigma == 1:
            assert np.array_equal(perm, np.arange(m.rows))


def test_sell_rejects_out_of_range_columns():
    m = pb.Csr(2, 2, np.array([0, 1, 2], np.uint32), np.array([0, 5], np.uint32), np.array([1.0, 2.0]))
    with pytest.raises(capi.Feat3HipError):
        sell(m, 64, 1)
