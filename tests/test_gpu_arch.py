"""Parity of the lafem_hip_* arch entry points with the CPU oracle (the reference's generic
kernels restated), called through the C ABI on device memory.

Bar: index arrays bit-exact; FP64 values within 1e-13 relative (the device reduces a row in a
different order than the sequential CPU loop); FP32 within 1e-5; binary16-stored values with
FP32 arithmetic against the oracle's identical storage rounding within 1e-5.
"""
import ctypes

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def vp(t):
    return ctypes.c_void_p(t.data_ptr())


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = max(np.linalg.norm(b), 1e-300)
    return np.linalg.norm(a - b) / den


def random_csr(rng, rows, cols, max_row, empty_rows=True, dtype=np.float64, itype=np.uint32):
    rp = [0]
    cols_l, vals = [], []
    for i in range(rows):
        k = int(rng.integers(0, max_row + 1)) if empty_rows else int(rng.integers(1, max_row + 1))
        k = min(k, cols)
        c = np.sort(rng.choice(cols, size=k, replace=False))
        cols_l.extend(c)
        vals.extend(rng.uniform(-1, 1, size=k))
        rp.append(rp[-1] + k)
    return (np.array(rp, itype), np.array(cols_l, itype), np.array(vals, dtype))


@pytest.fixture(scope="module")
def T(torch_cuda):
    from feat3_amd import capi
    return torch_cuda, capi


_alive = []


def dev(torch, a):
    """Device copy of a numpy array; kept alive until the end of the module so that a pointer
    taken from a temporary is never handed to the next allocation."""
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    _alive.append(t)
    return t


@pytest.fixture(autouse=True)
def _release():
    yield
    _alive.clear()


def host(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("itype", [np.uint32, np.uint64])
@pytest.mark.parametrize("ab", [(1.0, 0.0), (-1.0, 1.0), (4711.1, -0.5), (0.5, 1e-300)])
@pytest.mark.parametrize("shape", [(1, 1, 1), (257, 129, 7), (3000, 3000, 27), (1000, 50, 80)])
def test_csr_apply_f64(T, itype, ab, shape):
    torch, capi = T
    rng = np.random.default_rng(hash((shape, ab)) % 2**32)
    rows, cols, mr = shape
    rp, col, val = random_csr(rng, rows, cols, mr, itype=itype)
    x = rng.uniform(-1, 1, cols)
    y = rng.uniform(-1, 1, rows)
    a, b = ab
    ref = y.copy()
    po.csr_apply(ref, a, x, b, ref, val, col, rp, rows, cols)
    dr = dev(torch, y.copy())
    dx, dval, dcol, drp = dev(torch, x), dev(torch, val), dev(torch, col.astype(np.int64) if itype == np.uint64 else col.astype(np.int32)), None
    drp = dev(torch, rp.astype(np.int64) if itype == np.uint64 else rp.astype(np.int32))
    fn = "lafem_hip_apply_csr_f64_u64" if itype == np.uint64 else "lafem_hip_apply_csr_f64_u32"
    # r aliases y (the reference's in-place defect form)
    capi.call(fn, None, vp(dr), a, vp(dx), b, vp(dr), vp(dval), vp(dcol), vp(drp), rows, cols, int(rp[-1]), 0)
    torch.cuda.synchronize()
    err = rel(host(dr), ref)
    assert err <= 1e-13, err


@pytest.mark.parametrize("itype", [np.uint32, np.uint64])
def test_csr_apply_long_and_empty_rows(T, itype):
    """Rows of very different lengths (empty, short, and a few with thousands of entries): the
    streaming CSR kernel stages a block's products in LDS when its entries fit and sums rows
    with 8 lanes each when they do not; both paths (and blocks mixing them) match the oracle."""
    torch, capi = T
    rng = np.random.default_rng(11)
    rows, cols = 3000, 9000
    lens = rng.integers(0, 4, rows)
    lens[[5, 700, 701, 2999]] = [6000, 9000, 4500, 7000]
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(itype)
    col = np.concatenate([np.sort(rng.choice(cols, size=int(k), replace=False)) for k in lens]).astype(itype)
    val = rng.uniform(-1, 1, int(rp[-1]))
    x = rng.uniform(-1, 1, cols)
    y = rng.uniform(-1, 1, rows)
    it = np.int64 if itype == np.uint64 else np.int32
    suf = "u64" if itype == np.uint64 else "u32"
    for a, b in [(1.0, 0.0), (-0.5, 2.0)]:
        ref = y.copy()
        po.csr_apply(ref, a, x, b, ref, val, col, rp, rows, cols)
        dr = dev(torch, y.copy())
        capi.call(f"lafem_hip_apply_csr_f64_{suf}", None, vp(dr), a, vp(dev(torch, x)), b, vp(dr), vp(dev(torch, val)),
                  vp(dev(torch, col.astype(it))), vp(dev(torch, rp.astype(it))), rows, cols, int(rp[-1]), 0)
        torch.cuda.synchronize()
        assert rel(host(dr), ref) <= 1e-13
    if suf != "u32":
        return
    col32, rp32 = col.astype(np.uint32), rp.astype(np.uint32)
    # FP32 values and vectors
    x32, v32 = x.astype(np.float32), val.astype(np.float32)
    ref32 = y.astype(np.float32)
    po.lib().orc_csr_apply_f32_u32(po.P(ref32), 1.0, po.P(x32), 0.0, po.P(ref32), po.P(v32), po.P(col32), po.P(rp32),
                                   rows, cols, 0)
    dr = dev(torch, y.astype(np.float32))
    capi.call("lafem_hip_apply_csr_f32_u32", None, vp(dr), 1.0, vp(dev(torch, x32)), 0.0, vp(dr), vp(dev(torch, v32)),
              vp(dev(torch, col.astype(np.int32))), vp(dev(torch, rp.astype(np.int32))), rows, cols, int(rp[-1]), 0)
    torch.cuda.synchronize()
    assert rel(host(dr), ref32) <= 1e-5
    # csrsb with 3-vectors
    x3 = rng.uniform(-1, 1, 3 * cols)
    y3 = rng.uniform(-1, 1, 3 * rows)
    ref = y3.copy()
    po.lib().orc_csrsb_apply_f64_u32(po.P(ref), 1.5, po.P(x3), -1.0, po.P(ref), po.P(val), po.P(col32),
                                     po.P(rp32), rows, 3)
    dr = dev(torch, y3.copy())
    capi.call("lafem_hip_apply_csrsb_f64_u32", None, vp(dr), 1.5, vp(dev(torch, x3)), -1.0, vp(dr), vp(dev(torch, val)),
              vp(dev(torch, col.astype(np.int32))), vp(dev(torch, rp.astype(np.int32))), rows, cols, int(rp[-1]), 3)
    torch.cuda.synchronize()
    assert rel(host(dr), ref) <= 1e-13


def test_csr_apply_transposed_and_f32(T):
    torch, capi = T
    rng = np.random.default_rng(7)
    rows, cols = 500, 300
    rp, col, val = random_csr(rng, rows, cols, 12)
    x = rng.uniform(-1, 1, rows)
    y = rng.uniform(-1, 1, cols)
    ref = y.copy()
    po.csr_apply(ref, 2.0, x, -1.0, ref, val, col, rp, rows, cols, transposed=True)
    dr = dev(torch, y.copy())
    capi.call("lafem_hip_apply_csr_f64_u32", None, vp(dr), 2.0, vp(dev(torch, x)), -1.0, vp(dr), vp(dev(torch, val)),
              vp(dev(torch, col.astype(np.int32))), vp(dev(torch, rp.astype(np.int32))), rows, cols, int(rp[-1]), 1)
    torch.cuda.synchronize()
    err = rel(host(dr), ref)
    assert err <= 1e-12, err
    # FP32
    x32 = rng.uniform(-1, 1, cols).astype(np.float32)
    y32 = rng.uniform(-1, 1, rows).astype(np.float32)
    v32 = val.astype(np.float32)
    ref32 = y32.copy()
    po.csr_apply(ref32, np.float32(-1.0), x32, np.float32(1.0), ref32, v32, col, rp, rows, cols)
    dr = dev(torch, y32.copy())
    capi.call("lafem_hip_apply_csr_f32_u32", None, vp(dr), -1.0, vp(dev(torch, x32)), 1.0, vp(dr), vp(dev(torch, v32)),
              vp(dev(torch, col.astype(np.int32))), vp(dev(torch, rp.astype(np.int32))), rows, cols, int(rp[-1]), 0)
    torch.cuda.synchronize()
    err = rel(host(dr), ref32)
    assert err <= 1e-5, err


def test_csr_apply_h16(T):
    torch, capi = T
    rng = np.random.default_rng(11)
    rows = cols = 4000
    rp, col, val = random_csr(rng, rows, cols, 27, empty_rows=False)
    h = val.astype(np.float16)
    x = rng.uniform(-1, 1, cols).astype(np.float32)
    ref = np.zeros(rows, np.float32)
    po.lib().orc_csr_apply_h16_u32(po.P(ref), 1.0, po.P(x), 0.0, po.P(ref), po.P(h.view(np.uint16)), po.P(col), po.P(rp),
                                   rows, cols)
    dr = torch.zeros(rows, dtype=torch.float32, device="cuda")
    capi.call("lafem_hip_apply_csr_h16f32_u32", None, vp(dr), 1.0, vp(dev(torch, x)), 0.0, vp(dr),
              vp(dev(torch, h.view(np.int16))), vp(dev(torch, col.astype(np.int32))), vp(dev(torch, rp.astype(np.int32))),
              rows, cols, int(rp[-1]))
    torch.cuda.synchronize()
    err = rel(host(dr), ref)
    assert err <= 1e-5, err


@pytest.mark.parametrize("C", [4, 32, 64, 7])
def test_ell_and_sell_apply(T, C):
    torch, capi = T
    rng = np.random.default_rng(C)
    rows = cols = 1031
    rp, col, val = random_csr(rng, rows, cols, 20)
    L = po.lib()
    vs = L.orc_ell_val_size_u32(C, rows, po.P(rp))
    nch = (rows + C - 1) // C
    cs = np.zeros(nch + 1, np.uint32)
    cl = np.zeros(nch, np.uint32)
    rl = np.zeros(rows, np.uint32)
    ev = np.zeros(vs)
    ec = np.zeros(vs, np.uint32)
    L.orc_ell_from_csr_f64_u32(C, rows, po.P(rp), po.P(col), po.P(val), po.P(cs), po.P(cl), po.P(rl), po.P(ev), po.P(ec))
    x = rng.uniform(-1, 1, cols)
    y = rng.uniform(-1, 1, rows)
    ref = y.copy()
    L.orc_ell_apply_f64_u32(po.P(ref), -1.0, po.P(x), 1.0, po.P(ref), po.P(ev), po.P(ec), po.P(cs), po.P(cl), C, rows)
    # the ELL product equals the CSR product (format conversion is exact)
    ref_csr = y.copy()
    po.csr_apply(ref_csr, -1.0, x, 1.0, ref_csr, val, col, rp, rows, cols)
    assert rel(ref, ref_csr) <= 1e-14
    i32 = lambda a: dev(torch, a.astype(np.int32))
    dr = dev(torch, y.copy())
    capi.call("lafem_hip_apply_ell_f64_u32", None, vp(dr), -1.0, vp(dev(torch, x)), 1.0, vp(dr), vp(dev(torch, ev)),
              vp(i32(ec)), vp(i32(cs)), vp(i32(cl)), C, rows)
    torch.cuda.synchronize()
    err = rel(host(dr), ref)
    assert err <= 1e-13, err
    # sigma = 1 SELL with true row lengths and identity permutation
    dr = dev(torch, y.copy())
    perm = np.arange(rows, dtype=np.uint32)
    capi.call("lafem_hip_apply_sell_f64_u32", None, vp(dr), -1.0, vp(dev(torch, x)), 1.0, vp(dr), vp(dev(torch, ev)),
              vp(i32(ec)), vp(i32(cs)), vp(i32(cl)), vp(i32(rl)), vp(i32(perm)), C, rows)
    torch.cuda.synchronize()
    err = rel(host(dr), ref)
    assert err <= 1e-13, err
    # ELL diagonal
    dd = torch.zeros(rows, dtype=torch.float64, device="cuda")
    capi.call("lafem_hip_diagonal_ell_f64_u32", None, vp(dd), vp(dev(torch, ev)), vp(i32(ec)), vp(i32(cs)), vp(i32(cl)), C, rows)
    dref = np.zeros(rows)
    L.orc_diag_csr_f64_u32(po.P(dref), po.P(val), po.P(col), po.P(rp), rows)
    torch.cuda.synchronize()
    assert np.array_equal(host(dd), dref)


@pytest.mark.parametrize("shape", [(1031, 1031), (700, 1100), (900, 400)])
def test_banded_apply(T, shape):
    """Apply::banded (SparseMatrixBanded): FP64 u32/u64 offsets within 1e-13 and FP32 within 1e-5 of
    the oracle; diagonals leaving the matrix at both ends; r aliasing y."""
    torch, capi = T
    rows, cols = shape
    rng = np.random.default_rng(rows)
    offs = sorted(set(int(k) for k in rng.integers(-rows + 1, cols, size=9)) | {0, 1, -1})
    noo = len(offs)
    o = np.array([k + rows - 1 for k in offs], np.uint32)
    val = rng.uniform(-1, 1, noo * rows)
    x = rng.uniform(-1, 1, cols)
    y = rng.uniform(-1, 1, rows)
    L = po.lib()
    for a, b in [(1.0, 0.0), (-0.75, 1.5)]:
        ref = y.copy()
        L.orc_banded_apply_f64_u32(po.P(ref), a, po.P(x), b, po.P(ref), po.P(val), po.P(o), noo, rows, cols)
        for name, oo in [("lafem_hip_apply_banded_f64_u32", o), ("lafem_hip_apply_banded_f64_u64", o.astype(np.uint64))]:
            dr = dev(torch, y.copy())
            capi.call(name, None, vp(dr), a, vp(dev(torch, x)), b, vp(dr), vp(dev(torch, val)), vp(dev(torch, oo.view(np.int64) if oo.dtype == np.uint64 else oo.astype(np.int32))), noo, rows, cols)
            torch.cuda.synchronize()
            err = rel(host(dr), ref)
            assert err <= 1e-13, (name, err)
        ref32 = y.astype(np.float32)
        v32, x32 = val.astype(np.float32), x.astype(np.float32)
        L.orc_banded_apply_f32_u32(po.P(ref32), a, po.P(x32), b, po.P(ref32), po.P(v32), po.P(o), noo, rows, cols)
        dr = dev(torch, y.astype(np.float32))
        capi.call("lafem_hip_apply_banded_f32_u32", None, vp(dr), a, vp(dev(torch, x32)), b, vp(dr), vp(dev(torch, v32)),
                  vp(dev(torch, o.astype(np.int32))), noo, rows, cols)
        torch.cuda.synchronize()
        assert rel(host(dr), ref32) <= 1e-5


@pytest.mark.parametrize("bs", [1, 2, 3, 4])
def test_csrb_and_csrsb_apply(T, bs):
    torch, capi = T
    rng = np.random.default_rng(100 + bs)
    rows = cols = 700
    rp, col, _ = random_csr(rng, rows, cols, 15)
    nnzb = int(rp[-1])
    val = rng.uniform(-1, 1, nnzb * bs * bs)
    x = rng.uniform(-1, 1, cols * bs)
    y = rng.uniform(-1, 1, rows * bs)
    ref = y.copy()
    po.lib().orc_csrb_apply_f64_u32(po.P(ref), 0.75, po.P(x), -2.0, po.P(ref), po.P(val), po.P(col), po.P(rp), rows, bs, bs)
    i32 = lambda a: dev(torch, a.astype(np.int32))
    dr = dev(torch, y.copy())
    capi.call("lafem_hip_apply_csrb_f64_u32", None, vp(dr), 0.75, vp(dev(torch, x)), -2.0, vp(dr), vp(dev(torch, val)),
              vp(i32(col)), vp(i32(rp)), rows, cols, nnzb, bs, bs)
    torch.cuda.synchronize()
    err = rel(host(dr), ref)
    assert err <= 1e-13, err
    # diagonal of the diagonal blocks
    dd = torch.zeros(rows * bs, dtype=torch.float64, device="cuda")
    capi.call("lafem_hip_diagonal_csrb_f64_u32", None, vp(dd), vp(dev(torch, val)), vp(i32(col)), vp(i32(rp)), rows, bs, bs)
    dref = np.zeros(rows * bs)
    po.lib().orc_diag_csrb_f64_u32(po.P(dref), po.P(val), po.P(col), po.P(rp), rows, bs, bs)
    torch.cuda.synchronize()
    assert np.array_equal(host(dd), dref)
    if bs <= 3:
        sval = rng.uniform(-1, 1, nnzb)
        ref = y.copy()
        po.lib().orc_csrsb_apply_f64_u32(po.P(ref), 1.0, po.P(x), 1.0, po.P(ref), po.P(sval), po.P(col), po.P(rp), rows, bs)
        dr = dev(torch, y.copy())
        capi.call("lafem_hip_apply_csrsb_f64_u32", None, vp(dr), 1.0, vp(dev(torch, x)), 1.0, vp(dr), vp(dev(torch, sval)),
                  vp(i32(col)), vp(i32(rp)), rows, cols, nnzb, bs)
        torch.cuda.synchronize()
        err = rel(host(dr), ref)
    assert err <= 1e-13, err


@pytest.mark.parametrize("n", [0, 1, 63, 1000, 1 << 20, 3_000_001])
def test_blas1_f64(T, n):
    torch, capi = T
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, n)
    y = rng.uniform(-1, 1, n)
    dx, dy = dev(torch, x), dev(torch, y)
    L = po.lib()
    # axpy r = a x + y (and the aliased forms)
    r = np.zeros(n)
    L.orc_axpy_f64(po.P(r), 0.3, po.P(x), po.P(y), n)
    dr = torch.zeros(n, dtype=torch.float64, device="cuda")
    capi.call("lafem_hip_axpy_f64", None, vp(dr), 0.3, vp(dx), vp(dy), n)
    if n:
        dz = dev(torch, y.copy())
        capi.call("lafem_hip_axpy_f64", None, vp(dz), 0.3, vp(dx), vp(dz), n)  # r == y
        torch.cuda.synchronize()
        err = rel(host(dz), r)
        assert err <= 1e-15, err
    torch.cuda.synchronize()
    err = rel(host(dr), r)
    assert err <= 1e-15, err
    # dot and norm (device scalar and host value)
    dres = torch.zeros(1, dtype=torch.float64, device="cuda")
    capi.call("lafem_hip_dot_f64", None, vp(dx), vp(dy), n, vp(dres))
    hv = ctypes.c_double()
    capi.call("lafem_hip_norm2_value_f64", None, vp(dx), n, ctypes.byref(hv))
    torch.cuda.synchronize()
    dref = L.orc_dot_f64(po.P(x), po.P(y), n)
    nref = L.orc_norm2_f64(po.P(x), n)
    assert abs(float(dres.item()) - dref) <= 1e-12 * max(1.0, np.abs(x * y).sum())
    assert abs(hv.value - nref) <= 1e-13 * max(nref, 1e-300) or n == 0
    if n:
        # integer-valued inputs: the sums are exact in any order
        xi = rng.integers(-1000, 1000, n).astype(np.float64)
        yi = rng.integers(-1000, 1000, n).astype(np.float64)
        hv2 = ctypes.c_double()
        capi.call("lafem_hip_dot_value_f64", None, vp(dev(torch, xi)), vp(dev(torch, yi)), n, ctypes.byref(hv2))
        assert hv2.value == float(np.dot(xi.astype(np.int64), yi.astype(np.int64)))
        # component product / invert / scale
        cp = np.zeros(n)
        L.orc_component_product_f64(po.P(cp), po.P(x), po.P(y), n)
        dcp = torch.zeros(n, dtype=torch.float64, device="cuda")
        capi.call("lafem_hip_component_product_f64", None, vp(dcp), vp(dx), vp(dy), n)
        ci = np.zeros(n)
        xs = x + 2.0
        L.orc_component_invert_f64(po.P(ci), po.P(xs), 0.7, n)
        dci = torch.zeros(n, dtype=torch.float64, device="cuda")
        capi.call("lafem_hip_component_invert_f64", None, vp(dci), vp(dev(torch, xs)), 0.7, n)
        sc = np.zeros(n)
        L.orc_scale_f64(po.P(sc), po.P(x), -3.5, n)
        dsc = torch.zeros(n, dtype=torch.float64, device="cuda")
        capi.call("lafem_hip_scale_f64", None, vp(dsc), vp(dx), -3.5, n)
        torch.cuda.synchronize()
        assert np.array_equal(host(dcp), cp)
        assert np.array_equal(host(dci), ci)
        assert np.array_equal(host(dsc), sc)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_blas1_unaligned_pointers(T, dtype):
    """axpy and dot use 16-byte accesses when every pointer is 16-byte aligned and element
    accesses otherwise: views starting one element into a buffer take the second path, with
    the same results as the aligned call on the same data."""
    torch, capi = T
    rng = np.random.default_rng(3)
    n = 100_003
    suf = "f64" if dtype == np.float64 else "f32"
    tol = 1e-15 if dtype == np.float64 else 1e-6
    x = rng.uniform(-1, 1, n + 1).astype(dtype)
    y = rng.uniform(-1, 1, n + 1).astype(dtype)
    dx, dy = dev(torch, x), dev(torch, y)
    es = dx.element_size()
    px, py = ctypes.c_void_p(dx.data_ptr() + es), ctypes.c_void_p(dy.data_ptr() + es)
    out = torch.zeros(n + 1, dtype=dx.dtype, device="cuda")
    capi.call(f"lafem_hip_axpy_{suf}", None, ctypes.c_void_p(out.data_ptr() + es), 0.25, px, py, n)
    res = torch.zeros(2, dtype=dx.dtype, device="cuda")
    capi.call(f"lafem_hip_dot_{suf}", None, px, py, n, ctypes.c_void_p(res.data_ptr()))
    xa, ya = dev(torch, x[1:].copy()), dev(torch, y[1:].copy())
    outa = torch.zeros(n, dtype=dx.dtype, device="cuda")
    capi.call(f"lafem_hip_axpy_{suf}", None, vp(outa), 0.25, vp(xa), vp(ya), n)
    capi.call(f"lafem_hip_dot_{suf}", None, vp(xa), vp(ya), n, ctypes.c_void_p(res.data_ptr() + es))
    torch.cuda.synchronize()
    ref = (0.25 * x[1:].astype(np.float64)) + y[1:]
    assert rel(host(out)[1:], ref) <= tol and rel(host(outa), ref) <= tol
    d = float(np.dot(x[1:].astype(np.float64), y[1:].astype(np.float64)))
    r = host(res).astype(np.float64)
    assert abs(r[0] - d) <= (1e-12 if dtype == np.float64 else 1e-3) * np.abs(x * y).sum()
    assert abs(r[1] - d) <= (1e-12 if dtype == np.float64 else 1e-3) * np.abs(x * y).sum()


def test_blas1_f32_and_convert(T):
    torch, capi = T
    rng = np.random.default_rng(5)
    n = 777777
    x = rng.uniform(-1, 1, n)
    dx = dev(torch, x)
    d32 = torch.zeros(n, dtype=torch.float32, device="cuda")
    capi.call("lafem_hip_convert_f32_f64", None, vp(d32), vp(dx), n)
    torch.cuda.synchronize()
    assert np.array_equal(host(d32), x.astype(np.float32))
    d64 = torch.zeros(n, dtype=torch.float64, device="cuda")
    capi.call("lafem_hip_convert_f64_f32", None, vp(d64), vp(d32), n)
    d16 = torch.zeros(n, dtype=torch.int16, device="cuda")
    capi.call("lafem_hip_convert_h16_f32", None, vp(d16), vp(d32), n)
    torch.cuda.synchronize()
    assert np.array_equal(host(d64), x.astype(np.float32).astype(np.float64))
    assert np.array_equal(host(d16).view(np.float16), x.astype(np.float32).astype(np.float16))
    hv = ctypes.c_float()
    capi.call("lafem_hip_norm2_value_f32", None, vp(d32), n, ctypes.byref(hv))
    # the oracle's sequential FP32 sum of 777777 squares carries an O(n eps) rounding error
    # (about 5e-2 relative here); the device tree reduction is checked against the exact FP64
    # norm of the same FP32 values instead, and the oracle's error is bounded by n*eps
    x32 = x.astype(np.float32)
    exact = float(np.sqrt(np.sum(x32.astype(np.float64) ** 2)))
    assert abs(hv.value - exact) <= 1e-5 * exact
    seq = po.lib().orc_norm2_f32(po.P(x32), n)
    assert abs(seq - exact) <= n * np.finfo(np.float32).eps * exact
    # set_memory
    capi.call("lafem_hip_set_memory_f64", None, vp(d64), 2.5, n)
    torch.cuda.synchronize()
    assert np.all(host(d64) == 2.5)


def test_unit_filter_and_diag(T):
    torch, capi = T
    rng = np.random.default_rng(3)
    n = 5000
    idx = np.sort(rng.choice(n, 700, replace=False)).astype(np.uint32)
    sv = rng.uniform(-1, 1, idx.size)
    v = rng.uniform(-1, 1, n)
    ref = v.copy()
    po.lib().orc_unit_filter_rhs_f64_u32(po.P(ref), po.P(sv), po.P(idx), idx.size)
    dv = dev(torch, v.copy())
    capi.call("lafem_hip_unit_filter_rhs_f64_u32", None, vp(dv), vp(dev(torch, sv)), vp(dev(torch, idx.astype(np.int32))),
              idx.size, 1)
    torch.cuda.synchronize()
    assert np.array_equal(host(dv), ref)
    ref2 = v.copy()
    po.lib().orc_unit_filter_def_f64_u32(po.P(ref2), po.P(idx), idx.size)
    dv = dev(torch, v.copy())
    capi.call("lafem_hip_unit_filter_def_f64_u32", None, vp(dv), vp(dev(torch, idx.astype(np.int32))), idx.size, 1)
    torch.cuda.synchronize()
    assert np.array_equal(host(dv), ref2)
    # blocked filter (3 components per node)
    vb = rng.uniform(-1, 1, 3 * n)
    dv = dev(torch, vb.copy())
    capi.call("lafem_hip_unit_filter_def_f64_u32", None, vp(dv), vp(dev(torch, idx.astype(np.int32))), idx.size, 3)
    torch.cuda.synchronize()
    exp = vb.reshape(n, 3).copy()
    exp[idx] = 0.0
    assert np.array_equal(host(dv), exp.ravel())
    # diagonal (missing diagonal entries give 0)
    rp, col, val = random_csr(rng, 900, 900, 10)
    dref = np.zeros(900)
    po.lib().orc_diag_csr_f64_u32(po.P(dref), po.P(val), po.P(col), po.P(rp), 900)
    dd = torch.zeros(900, dtype=torch.float64, device="cuda")
    capi.call("lafem_hip_diagonal_csr_f64_u32", None, vp(dd), vp(dev(torch, val)), vp(dev(torch, col.astype(np.int32))),
              vp(dev(torch, rp.astype(np.int32))), 900)
    torch.cuda.synchronize()
    assert np.array_equal(host(dd), dref)


def test_memory_pool_refcount(T):
    torch, capi = T
    p = ctypes.c_void_p()
    capi.call("feat3_hip_malloc", ctypes.byref(p), 4096)
    capi.call("feat3_hip_increase", p)
    capi.call("feat3_hip_free", p)   # refcount 2 -> 1, still allocated
    host_buf = np.arange(512, dtype=np.float64)
    out = np.zeros(512)
    capi.call("feat3_hip_upload", None, p, ctypes.c_void_p(host_buf.ctypes.data), 4096)
    capi.call("feat3_hip_download", None, ctypes.c_void_p(out.ctypes.data), p, 4096)
    capi.call("feat3_hip_synchronise", None)
    assert np.array_equal(out, host_buf)
    capi.call("feat3_hip_free", p)
    assert capi.lib.feat3_hip_free(p) == 1  # released: unknown pointer


def test_reductions_on_two_streams(T):
    """Dots and norms issued back to back on two streams share the device's reduction workspace:
    each waits for the previous one, so both results are those of the sequential oracle sums
    (integer-valued inputs: exact), also when the first one is still running when the second is
    issued (a large first reduction)."""
    torch, capi = T
    rng = np.random.default_rng(11)
    n1, n2 = 1 << 24, 1 << 20
    x1 = rng.integers(-8, 8, n1).astype(np.float64)
    y1 = rng.integers(-8, 8, n1).astype(np.float64)
    x2 = rng.integers(-8, 8, n2).astype(np.float64)
    d1x, d1y, d2x = (torch.from_numpy(v).cuda() for v in (x1, y1, x2))
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    for rep in range(3):
        r1 = torch.zeros(1, dtype=torch.float64, device="cuda")
        r2 = torch.zeros(1, dtype=torch.float64, device="cuda")
        capi.call("lafem_hip_dot_f64", ctypes.c_void_p(s1.cuda_stream), vp(d1x), vp(d1y), n1, vp(r1))
        capi.call("lafem_hip_norm2_f64", ctypes.c_void_p(s2.cuda_stream), vp(d2x), n2, vp(r2))
        hv = ctypes.c_double()
        capi.call("lafem_hip_dot_value_f64", ctypes.c_void_p(s1.cuda_stream), vp(d2x), vp(d2x), n2, ctypes.byref(hv))
        torch.cuda.synchronize()
        assert r1.item() == float(np.dot(x1.astype(np.int64), y1.astype(np.int64))), rep
        assert r2.item() == float(np.sqrt(float(np.dot(x2.astype(np.int64), x2.astype(np.int64))))), rep
        assert hv.value == float(np.dot(x2.astype(np.int64), x2.astype(np.int64))), rep


@pytest.mark.parametrize("bh,bw", [(2, 3), (3, 1), (5, 7), (9, 9), (9, 2), (1, 9), (6, 6)])
@pytest.mark.parametrize("suf", ["f64_u32", "f64_u64", "f32_u32"])
def test_csrb_any_block_shape(T, bh, bw, suf):
    """Apply::csrb for every block shape up to 9 x 9 (apply.hpp:277-281) against the oracle."""
    torch, capi = T
    rng = np.random.default_rng(bh * 100 + bw)
    rows, cols = 500, 310
    itype = np.uint64 if suf.endswith("u64") else np.uint32
    rp, col, _ = random_csr(rng, rows, cols, 12, itype=itype)
    nnzb = int(rp[-1])
    val = rng.uniform(-1, 1, nnzb * bh * bw)
    x = rng.uniform(-1, 1, cols * bw)
    y = rng.uniform(-1, 1, rows * bh)
    ref = y.copy()
    # the 32-bit copies are held in names: a pointer taken from a temporary would dangle
    col32, rp32 = col.astype(np.uint32), rp.astype(np.uint32)
    po.lib().orc_csrb_apply_f64_u32(po.P(ref), 0.75, po.P(x), -2.0, po.P(ref), po.P(val), po.P(col32), po.P(rp32),
                                    rows, bh, bw)
    dt = np.float32 if suf.startswith("f32") else np.float64
    idt = np.int64 if itype == np.uint64 else np.int32
    dr = dev(torch, y.astype(dt))
    capi.call(f"lafem_hip_apply_csrb_{suf}", None, vp(dr), 0.75, vp(dev(torch, x.astype(dt))), -2.0, vp(dr),
              vp(dev(torch, val.astype(dt))), vp(dev(torch, col.astype(idt))), vp(dev(torch, rp.astype(idt))), rows, cols,
              nnzb, bh, bw)
    torch.cuda.synchronize()
    err = rel(host(dr), ref)
    assert err <= (1e-13 if dt == np.float64 else 1e-5), err


def test_csrb_rejects_shapes_beyond_9(T):
    torch, capi = T
    d = torch.zeros(16, dtype=torch.float64, device="cuda")
    i = torch.zeros(2, dtype=torch.int32, device="cuda")
    for bh, bw in ((10, 2), (2, 10), (0, 3)):
        with pytest.raises(capi.Feat3HipError):
            capi.call("lafem_hip_apply_csrb_f64_u32", None, vp(d), 1.0, vp(d), 0.0, vp(d), vp(d), vp(i), vp(i), 1, 1, 0,
                      bh, bw)
