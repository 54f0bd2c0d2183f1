"""Pins the CPU oracle and the host setup against what the reference's own tests and the
mathematics fix (CPU only).

Reference-held known answers used here:
  * kernel/solver/basic_solver-test.cpp:100-160,278-390 -- iteration counts on
    PointstarFactoryFD(17, 2) with a Q2-bubble solution: CG 28, PCG-JAC 28, Richardson-JAC
    1175 (max_iter 1500), each +-2, and ||x - x_ref||^2 <= sqrt(eps)
  * kernel/lafem/pointstar_factory-test.cpp:43-83 + pointstar_factory.hpp:388-411 -- the
    smallest eigenpair of the FD pointstar matrix in closed form
  * kernel/assembly/grid_transfer-test.cpp:65-140 -- the 1-cell -> 4-cell Q1 prolongation:
    dense 9x4 pattern (row_ptr[i] = 4i, col = j) and values {1, 1/2, 1/4, 0}
  * applications/CMakeLists.txt:141-150 + poisson_dirichlet.cpp:266-276 -- PCG-MG with
    Richardson(Jacobi 0.7) V(4,4) on Q1 Poisson, level 5 .. 0, ExpBubble: 4 +- 1 iterations
Closed forms / invariants: the textbook Q1 stencils (2D: 8/3 and -1/3; 3D: 8h/3, 0, -h/6,
-h/12), zero row sums, symmetry, exactness of the prolongation on (bi/tri)linear and
biquadratic functions, R = P^T, exact two-grid identities and dense brute force.
"""
import numpy as np
import pytest

from feat3_amd import problems as pb
from oracle import pyoracle as po

EPS = np.finfo(np.float64).eps


def dense(rp, col, val, n, m=None):
    m = n if m is None else m
    d = np.zeros((n, m))
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            d[i, col[k]] += val[k]
    return d


# ---------------------------------------------------------------------------------------------
# arch kernels: brute force against dense products
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("ab", [(1.0, 0.0), (-1.0, 1.0), (4711.1, -0.5), (2.0, 1e-300)])
def test_csr_apply_brute_force(ab):
    rng = np.random.default_rng(1)
    n, m = 37, 23
    d = rng.uniform(-1, 1, (n, m)) * (rng.uniform(size=(n, m)) < 0.3)
    rp = np.concatenate([[0], np.cumsum((d != 0).sum(1))]).astype(np.uint32)
    col = np.nonzero(d)[1].astype(np.uint32)
    val = d[d != 0]
    x = rng.uniform(-1, 1, m)
    y = rng.uniform(-1, 1, n)
    a, b = ab
    r = y.copy()
    po.csr_apply(r, a, x, b, r, val, col, rp, n, m)
    exp = a * (d @ x) + (b * y if abs(b) >= EPS else 0.0)
    assert np.max(np.abs(r - exp)) <= 1e-14 * max(1.0, np.max(np.abs(exp))) * m
    # transposed
    xt = rng.uniform(-1, 1, n)
    yt = rng.uniform(-1, 1, m)
    rt = yt.copy()
    po.csr_apply(rt, a, xt, b, rt, val, col, rp, n, m, transposed=True)
    expt = a * (d.T @ xt) + (b * yt if abs(b) >= EPS else 0.0)
    assert np.max(np.abs(rt - expt)) <= 1e-14 * max(1.0, np.max(np.abs(expt))) * n


def test_bcsr_csrsb_ell_brute_force():
    rng = np.random.default_rng(2)
    nb, bs = 11, 3
    pat = rng.uniform(size=(nb, nb)) < 0.4
    np.fill_diagonal(pat, True)
    rp = np.concatenate([[0], np.cumsum(pat.sum(1))]).astype(np.uint32)
    col = np.nonzero(pat)[1].astype(np.uint32)
    blocks = rng.uniform(-1, 1, (col.size, bs, bs))
    D = np.zeros((nb * bs, nb * bs))
    for i in range(nb):
        for k in range(rp[i], rp[i + 1]):
            D[i * bs:(i + 1) * bs, col[k] * bs:(col[k] + 1) * bs] = blocks[k]
    x = rng.uniform(-1, 1, nb * bs)
    r = np.zeros(nb * bs)
    L = po.lib()
    bflat = blocks.ravel().copy()     # kept alive across the call (a raw pointer does not hold it)
    L.orc_csrb_apply_f64_u32(po.P(r), 1.0, po.P(x), 0.0, po.P(r), po.P(bflat), po.P(col), po.P(rp), nb, bs, bs)
    assert np.allclose(r, D @ x, rtol=1e-14, atol=1e-14)
    # diagonal of the diagonal blocks
    dg = np.zeros(nb * bs)
    L.orc_diag_csrb_f64_u32(po.P(dg), po.P(bflat), po.P(col), po.P(rp), nb, bs, bs)
    assert np.array_equal(dg, np.diag(D))
    # csrsb: scalar matrix times blocked vector == kron(A, I)
    sval = rng.uniform(-1, 1, col.size)
    A = dense(rp, col, sval, nb)
    r2 = np.zeros(nb * bs)
    L.orc_csrsb_apply_f64_u32(po.P(r2), 1.0, po.P(x), 0.0, po.P(r2), po.P(sval), po.P(col), po.P(rp), nb, bs)
    assert np.allclose(r2, np.kron(A, np.eye(bs)) @ x, rtol=1e-14, atol=1e-14)
    # ELL (SELL-C, sigma = 1) round trip for several chunk heights
    for C in (1, 2, 4, 5, 32):
        vs = L.orc_ell_val_size_u32(C, nb, po.P(rp))
        nch = (nb + C - 1) // C
        cs, cl, rl = np.zeros(nch + 1, np.uint32), np.zeros(nch, np.uint32), np.zeros(nb, np.uint32)
        ev, ec = np.zeros(vs), np.zeros(vs, np.uint32)
        L.orc_ell_from_csr_f64_u32(C, nb, po.P(rp), po.P(col), po.P(sval), po.P(cs), po.P(cl), po.P(rl), po.P(ev), po.P(ec))
        xs = x[:nb].copy()
        r3 = np.zeros(nb)
        L.orc_ell_apply_f64_u32(po.P(r3), 1.0, po.P(xs), 0.0, po.P(r3), po.P(ev), po.P(ec), po.P(cs), po.P(cl), C, nb)
        assert np.allclose(r3, A @ xs, rtol=1e-14, atol=1e-14)
        assert np.array_equal(rl, np.diff(rp))


def test_blas1_exact_on_integers_and_alias_semantics():
    rng = np.random.default_rng(3)
    n = 10007
    x = rng.integers(-100, 100, n).astype(np.float64)
    y = rng.integers(-100, 100, n).astype(np.float64)
    L = po.lib()
    assert L.orc_dot_f64(po.P(x), po.P(y), n) == float(np.dot(x.astype(np.int64), y.astype(np.int64)))
    assert L.orc_norm2_f64(po.P(x), n) == np.sqrt(float(np.dot(x.astype(np.int64), x.astype(np.int64))))
    r = y.copy()  # r == y: r += a x
    L.orc_axpy_f64(po.P(r), 2.0, po.P(x), po.P(r), n)
    assert np.array_equal(r, y + 2.0 * x)
    r = x.copy()  # r == x: r = r*a + y
    L.orc_axpy_f64(po.P(r), 3.0, po.P(r), po.P(y), n)
    assert np.array_equal(r, 3.0 * x + y)
    z = np.abs(x) + 1.0
    ci = np.zeros(n)
    L.orc_component_invert_f64(po.P(ci), po.P(z), 0.5, n)
    assert np.array_equal(ci, 0.5 / z)
    # scale (dense_vector.hpp:955-963, scale_generic.hpp: r_i = s * x_i) and component_product
    # (component_product_generic.hpp:21-51: r_i = x_i * y_i): one rounding each, so numpy's
    # elementwise products must agree bit for bit on arbitrary values (a dropped or doubled
    # factor, or a wrong index, would not)
    u = rng.uniform(-1e3, 1e3, n)
    w = rng.uniform(-1e3, 1e3, n)
    sc = np.zeros(n)
    L.orc_scale_f64(po.P(sc), po.P(u), -3.7, n)
    assert np.array_equal(sc, u * -3.7)
    u0 = u.copy()
    L.orc_scale_f64(po.P(u), po.P(u), 0.5, n)   # in place (r == x)
    assert np.array_equal(u, u0 * 0.5)
    cp = np.zeros(n)
    L.orc_component_product_f64(po.P(cp), po.P(u), po.P(w), n)
    assert np.array_equal(cp, u * w)


def test_unit_filter_and_transpose():
    rng = np.random.default_rng(4)
    v = rng.uniform(size=50)
    idx = np.array([0, 7, 49], np.uint32)
    sv = np.array([1.0, 2.0, 3.0])
    w = v.copy()
    po.lib().orc_unit_filter_rhs_f64_u32(po.P(w), po.P(sv), po.P(idx), 3)
    assert w[0] == 1.0 and w[7] == 2.0 and w[49] == 3.0 and np.array_equal(np.delete(w, idx), np.delete(v, idx))
    po.lib().orc_unit_filter_def_f64_u32(po.P(w), po.P(idx), 3)
    assert not np.any(w[idx])
    P = pb.prolongation(2, 1, 3)
    trp, tcol, tval = np.zeros(P.cols + 1, np.uint32), np.zeros(P.nnz, np.uint32), np.zeros(P.nnz)
    po.lib().orc_transpose_csr_f64_u32(P.rows, P.cols, po.P(P.row_ptr), po.P(P.col_ind), po.P(P.val), po.P(trp), po.P(tcol), po.P(tval))
    R = pb.transpose(P)
    assert np.array_equal(trp, R.row_ptr) and np.array_equal(tcol, R.col_ind) and np.array_equal(tval, R.val)
    assert np.array_equal(dense(trp, tcol, tval, P.cols, P.rows), P.to_dense().T)


def test_half_conversion_exhaustive():
    L = po.lib()
    bits = np.arange(65536, dtype=np.uint16)
    ref = bits.view(np.float16).astype(np.float32)
    got = np.array([L.orc_half_to_float(int(b)) for b in bits], np.float32)
    fin = np.isfinite(ref)
    assert np.array_equal(got[fin], ref[fin])
    assert np.all(np.isnan(got[~fin & np.isnan(ref)]))
    # float -> half round to nearest even against numpy on a dense sample
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(-7e4, 7e4, 20000), rng.uniform(-1e-4, 1e-4, 20000),
                         (bits[:31744].view(np.float16).astype(np.float32) * np.float32(1.0004))]).astype(np.float32)
    h = np.array([L.orc_float_to_half(float(v)) for v in xs], np.uint16)
    assert np.array_equal(h, xs.astype(np.float16).view(np.uint16))


# ---------------------------------------------------------------------------------------------
# pointstar factory and solver KATs (basic_solver-test.cpp)
# ---------------------------------------------------------------------------------------------
def test_pointstar_fd_structure_and_eigenpair():
    for m, d in [(17, 2), (9, 3), (5, 1)]:
        rp, col, val = po.pointstar_fd(m, d)
        n = m ** d
        A = dense(rp, col, val, n)
        assert rp[-1] == po.lib().orc_pointstar_fd_nnz(m, d)
        assert np.array_equal(A, A.T)
        assert np.all(np.diff(col.astype(np.int64))[np.diff(np.repeat(np.arange(n), np.diff(rp))) == 0] > 0)
        # closed-form smallest eigenpair (pointstar_factory.hpp:388-411)
        lmin = 2 * d * (1 - np.cos(np.pi / (m + 1)))
        i = np.arange(n)
        v = np.ones(n)
        q = 1
        while q < n:
            v *= np.sin(np.pi / (m + 1) * ((i // q) % m + 1))
            q *= m
        assert np.linalg.norm(A @ v - lmin * v) <= 1e-12 * np.linalg.norm(v)


@pytest.mark.parametrize("name,kind,ref_iters,max_iter", [
    ("CG", po.CG, 28, 100), ("PCG-JAC", po.PCG_JACOBI, 28, 100), ("Richardson-JAC", po.RICHARDSON_JACOBI, 1175, 1500)])
def test_basic_solver_kat(name, kind, ref_iters, max_iter):
    m, d = 17, 2
    rp, col, val = po.pointstar_fd(m, d)
    n = m ** d
    ref = po.pointstar_q2_bubble(m, d)
    b = np.zeros(n)
    po.csr_apply(b, 1.0, ref, 0.0, b, val, col, rp, n, n)
    s = po.System()
    s.add_level(n, rp, col, val, np.zeros(0, np.uint32))
    x = np.zeros(n)
    st, it, hist = s.solve(x, b, po.params(kind=kind, tol_rel=np.sqrt(EPS), max_iter=max_iter, omega_jac=1.0, use_def=0))
    assert st in (2, 5, 6)
    assert abs(it - ref_iters) <= 2, (name, it)
    assert np.sum((x - ref) ** 2) <= np.sqrt(EPS)


def test_solver_status_semantics():
    """iterative.hpp:744-832: zero rhs -> success at 0 iterations; NaN -> aborted; max_iter."""
    H = pb.hierarchy(2, 1, 4, 1)
    S = po.System.from_hierarchy(H)
    x = np.zeros(H.n)
    st, it, _ = S.solve(x, np.zeros(H.n), po.params())
    assert (st, it) == (2, 0)
    b = pb.filtered_rhs(H, 0)
    bn = b.copy()
    bn[H.n // 2] = np.nan
    st, it, _ = S.solve(np.zeros(H.n), bn, po.params())
    assert st == 3
    st, it, _ = S.solve(np.zeros(H.n), b, po.params(max_iter=1))
    assert (st, it) == (5, 1)
    # an unstable Richardson damping grows the defect past div_rel * def_init = def_init / eps
    st, it, _ = S.solve(np.zeros(H.n), b, po.params(max_iter=100, omega_rich=25.0, kind=po.RICHARDSON_JACOBI))
    assert st == 4 and it < 100  # diverged


@pytest.mark.parametrize("level", [1, 2])
def test_nan_in_a_coarser_level_aborts_after_one_iteration(level):
    """A NaN coupling in a coarser level matrix leaves the initial defect finite but makes the
    V-cycle's correction NaN, so gamma and alpha are NaN.  DenseVector::axpy copies only when
    |alpha| < eps (dense_vector.hpp:891); a NaN alpha fails that test, updates x and r, and the
    first new defect norm is NaN: aborted after one iteration (iterative.hpp:794-795).  Skipping
    the update for a NaN alpha (an `|alpha| >= eps` test) would instead run to max_iter."""
    H = pb.hierarchy(2, 1, 5, 1)
    A = H.levels[level].A
    m = 2 ** (5 - level) + 1
    r = (m // 2) * m + m // 2
    ks = [k for k in range(A.row_ptr[r], A.row_ptr[r + 1]) if A.col_ind[k] != r]
    A.val[ks[0]] = np.nan
    S = po.System.from_hierarchy(H)
    st, it, hist = S.solve(np.zeros(H.n), pb.filtered_rhs(H, 0), po.params(max_iter=50))
    assert (st, it) == (3, 1)
    assert np.isfinite(hist[0]) and np.isnan(hist[1])


# ---------------------------------------------------------------------------------------------
# host setup against closed forms and the reference's golden prolongation
# ---------------------------------------------------------------------------------------------
def test_q1_stencils_closed_form():
    A, bnd = pb.laplace(2, 1, 3)
    m = 9
    D = A.to_dense()
    c = 4 * m + 4  # an interior node
    row = D[c].reshape(m, m)
    assert np.isclose(row[4, 4], 8 / 3, rtol=1e-14)
    nb = [row[4 + dy, 4 + dx] for dy in (-1, 0, 1) for dx in (-1, 0, 1) if dx or dy]
    assert np.allclose(nb, -1 / 3, rtol=1e-14)
    assert np.allclose(D.sum(1), 0, atol=1e-13) and np.allclose(D, D.T, atol=1e-15)
    # 3D: 8h/3 centre, face 0, edge -h/6, corner -h/12 (element-based sparsity keeps the zeros)
    A3, _ = pb.laplace(3, 1, 2)
    h = 0.25
    m = 5
    D3 = A3.to_dense()
    c = 2 + 2 * m + 2 * m * m
    row = D3[c].reshape(m, m, m)
    assert np.isclose(row[2, 2, 2], 8 * h / 3, rtol=1e-14)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                k = abs(dx) + abs(dy) + abs(dz)
                expect = {0: 8 * h / 3, 1: 0.0, 2: -h / 6, 3: -h / 12}[k]
                assert abs(row[2 + dz, 2 + dy, 2 + dx] - expect) <= 1e-15
    assert (A3.row_ptr[c + 1] - A3.row_ptr[c]) == 27


@pytest.mark.parametrize("cfg,n,nnz,nbnd", [(1, 4225, 37249, 256), (2, 1050625, 16785409, 4096),
                                            (3, 2146689, 57066625, 98306), (4, 1050625, 9443329, 4096),
                                            (5, 274625, 7189057, 24578)])
def test_config_sizes(cfg, n, nnz, nbnd):
    """the per-config sizes (n, nnz(A), boundary DOFs = n - (m-2)^d)"""
    import ctypes
    c = pb.CONFIGS[cfg]
    a, b, d = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    from feat3_amd import capi
    capi.call("feat3_setup_laplace_sizes", c["dim"], c["degree"], c["level"], ctypes.byref(a), ctypes.byref(b), ctypes.byref(d))
    assert (a.value, b.value, d.value) == (n, nnz, nbnd)


def test_grid_transfer_golden_1_to_4_cells():
    """grid_transfer-test.cpp:65-140: the unit square refined once, P from level 0 to level 1 in
    the reference's numbering (StandardRefinery: corners 0..3, then the midpoints of edges
    (0,1), (2,3), (0,2), (1,3), then the centre) -- pattern and values compared as stored."""
    H = pb.hierarchy(2, 1, 1, 0, numbering="feat")
    P = H.levels[0].P
    assert (P.rows, P.cols, P.nnz) == (9, 4, 36)
    assert np.array_equal(P.row_ptr, 4 * np.arange(10))
    assert np.array_equal(P.col_ind, np.tile(np.arange(4), 9))
    d0, d1, d2, d4 = 0.0, 1.0, 0.5, 0.25
    data_ref = np.array([d1, d0, d0, d0, d0, d1, d0, d0, d0, d0, d1, d0, d0, d0, d0, d1,
                         d2, d2, d0, d0, d0, d0, d2, d2, d2, d0, d2, d0, d0, d2, d0, d2, d4, d4, d4, d4])
    eps = EPS ** 0.8
    assert np.max(np.abs(P.val - data_ref)) <= eps
    # R = P^T in the same numbering (LAFEM transpose)
    assert np.array_equal(H.levels[0].R.to_dense(), P.to_dense().T)


@pytest.mark.parametrize("dim,deg,lvl", [(2, 1, 4), (3, 1, 3), (2, 2, 3)])
def test_prolongation_interpolates_exactly(dim, deg, lvl):
    P = pb.prolongation(dim, deg, lvl)
    R = pb.transpose(P)
    assert np.array_equal(R.to_dense(), P.to_dense().T)
    mf = deg * 2 ** lvl + 1
    mc = deg * 2 ** (lvl - 1) + 1

    def grid(m):
        t = np.linspace(0, 1, m)
        return np.meshgrid(*([t] * dim), indexing="ij")

    def f(c):  # exact in the coarse space: (bi/tri)linear for Q1, biquadratic for Q2
        if deg == 1:
            v = 1.0 + sum((i + 2) * ci for i, ci in enumerate(c))
            return v + c[0] * c[1] * (c[2] if dim == 3 else 1.0)
        return 1.0 + c[0] ** 2 * c[1] - 3 * c[1] ** 2 + c[0] * c[1] ** 2

    fc = np.moveaxis(f(grid(mc)[::-1]), 0, 0)
    # lexicographic order: x fastest -> flatten the reversed-axis grid in C order
    cc = [g for g in grid(mc)][::-1]
    fcoarse = f([cc[dim - 1 - k] for k in range(dim)]).ravel()
    ff = [g for g in grid(mf)][::-1]
    ffine = f([ff[dim - 1 - k] for k in range(dim)]).ravel()
    assert np.allclose(P.to_dense() @ fcoarse, ffine, atol=1e-13)
    del fc
    assert np.allclose(P.to_dense().sum(1), 1.0, atol=1e-14)


def test_two_grid_exact_coarse_identity():
    """With a single coarse level solved exactly and no smoothing the V-cycle is the exact
    coarse-grid correction: MG(d) = P A_c^{-1} R d on the filtered system."""
    H = pb.hierarchy(2, 1, 3, 2)
    S = po.System.from_hierarchy(H)
    rng = np.random.default_rng(6)
    d = rng.uniform(-1, 1, H.n)
    d[H.levels[0].bnd] = 0.0
    cor = np.zeros(H.n)
    # pre = post = 1 smoothing step with omega_rich = 0 leaves only the coarse correction
    S.vcycle(cor, d, po.params(pre=1, post=1, omega_rich=0.0))
    Lc = H.levels[1]
    Ac = Lc.A.to_dense()
    for i in Lc.bnd:
        Ac[i, :] = 0.0
        Ac[i, i] = 1.0
    rc = H.levels[0].R.to_dense() @ d
    rc[Lc.bnd] = 0.0
    xc = np.linalg.solve(Ac, rc)
    xc[Lc.bnd] = 0.0
    exp = H.levels[0].P.to_dense() @ xc
    exp[H.levels[0].bnd] = 0.0
    assert np.allclose(cor, exp, atol=1e-12)


def test_poisson_dirichlet_kat():
    """applications/CMakeLists.txt:141-150: level 5..0, Jacobi 0.7, V(4,4), smoother as coarse
    solver, ExpBubble: 4 +- 1 PCG iterations."""
    H = pb.hierarchy(2, 1, 5, 0)
    b = pb.filtered_rhs(H, 2)
    S = po.System.from_hierarchy(H)
    x = np.zeros(H.n)
    st, it, hist = S.solve(x, b, po.params(omega_jac=0.7, coarse=po.COARSE_SMOOTHER))
    assert st == 2 and abs(it - 4) <= 1
    assert np.all(np.diff(hist) < 0)


@pytest.mark.parametrize("inner", [0, 1, 2])
def test_mixed_precision_oracle_converges_like_fp64(inner):
    H = pb.hierarchy(2, 1, 6, 1)
    b = pb.filtered_rhs(H, 0)
    S = po.System.from_hierarchy(H)
    x64 = np.zeros(H.n)
    _, i64, _ = S.solve(x64, b, po.params(inner=0))
    x = np.zeros(H.n)
    st, it, hist = S.solve(x, b, po.params(inner=inner))
    assert st == 2 and abs(it - i64) <= 1
    assert np.linalg.norm(x - x64) <= 1e-8 * np.linalg.norm(x64)


@pytest.mark.parametrize("inner", [0, 1])
@pytest.mark.parametrize("comp", [0, 2])
def test_blocked_vector_laplacian_reduces_to_scalar(inner, comp):
    """The 3x3-block Laplacian has blocks a_ij * I: with a load in one component only, the
    blocked solve (BCSR apply, blocked Jacobi diagonal, csrsb transfers, UnitFilterBlocked)
    must reproduce the scalar solve in that component and leave the others zero."""
    H = pb.hierarchy(3, 1, 4, 1)
    vals = [pb.block_identity_values(L.A, 3) for L in H.levels]
    b = pb.filtered_rhs(H, 0)
    x = np.zeros(H.n)
    st, it, hist = po.System.from_hierarchy(H).solve(x, b, po.params(inner=inner))
    B = np.zeros((H.n, 3))
    B[:, comp] = b
    X = np.zeros(3 * H.n)
    stb, itb, histb = po.System.from_hierarchy(H, block=3, values=vals).solve(X, B.reshape(-1).copy(),
                                                                              po.params(inner=inner))
    X = X.reshape(H.n, 3)
    assert (st, it) == (stb, itb) == (2, it)
    assert np.max(np.abs(X[:, comp] - x)) <= 1e-13 * np.max(np.abs(x))
    assert np.all(np.delete(X, comp, axis=1) == 0.0)
    assert np.allclose(hist, histb, rtol=1e-12)


def test_banded_apply_brute_force_and_csr_identity():
    """Banded apply (apply_generic.hpp:519-598): dense brute force on random banded matrices
    (square and rectangular, diagonals leaving the matrix at both ends), beta/alias semantics,
    and bit-identity with the CSR apply on the Q1 matrix (same entries, same summation order,
    the fill adds exact zeros)."""
    L = po.lib()
    rng = np.random.default_rng(7)
    for rows, cols, offs in [(17, 17, [-5, -1, 0, 2, 9]), (12, 20, [-3, 0, 4, 15]), (20, 9, [-19, -4, 0, 8])]:
        noo = len(offs)
        o = np.array([k + rows - 1 for k in offs], np.uint32)
        val = rng.standard_normal(noo * rows)
        x = rng.standard_normal(cols)
        y = rng.standard_normal(rows)
        dense = np.zeros((rows, cols))
        for d, k in enumerate(offs):
            for l in range(rows):
                if 0 <= l + k < cols:
                    dense[l, l + k] = val[d * rows + l]
        for a, b in [(1.0, 0.0), (-0.5, 2.0)]:
            r = np.empty(rows)
            L.orc_banded_apply_f64_u32(po.P(r), a, po.P(x), b, po.P(y), po.P(val), po.P(o), noo, rows, cols)
            assert np.allclose(r, a * dense @ x + b * y, rtol=1e-14, atol=1e-14)
        r = y.copy()                       # r aliases y
        L.orc_banded_apply_f64_u32(po.P(r), 1.5, po.P(x), -1.0, po.P(r), po.P(val), po.P(o), noo, rows, cols)
        assert np.allclose(r, 1.5 * dense @ x - y, rtol=1e-14, atol=1e-14)
    A = pb.hierarchy(3, 1, 3, 3).levels[0].A
    o, val = pb.to_banded(A)
    assert o.size == 27
    x = rng.standard_normal(A.rows)
    rb = np.empty(A.rows)
    rc = np.empty(A.rows)
    L.orc_banded_apply_f64_u32(po.P(rb), 1.0, po.P(x), 0.0, po.P(x), po.P(val), po.P(o), o.size, A.rows, A.rows)
    po.csr_apply(rc, 1.0, x, 0.0, x, A.val, A.col_ind, A.row_ptr, A.rows, A.rows)
    assert np.array_equal(rb, rc)


@pytest.mark.parametrize("bh,bw", [(2, 3), (3, 1), (5, 7), (9, 9), (9, 2), (1, 9)])
def test_bcsr_non_square_brute_force(bh, bw):
    """apply_generic.hpp:121-154 for any block shape: the oracle's csrb against the dense product
    of the assembled block matrix (nb x mb blocks of bh x bw), with alpha/beta."""
    rng = np.random.default_rng(bh * 10 + bw)
    nb, mb = 13, 9
    pat = rng.uniform(size=(nb, mb)) < 0.35
    rp = np.concatenate([[0], np.cumsum(pat.sum(1))]).astype(np.uint32)
    col = np.nonzero(pat)[1].astype(np.uint32)
    blocks = rng.uniform(-1, 1, (col.size, bh, bw))
    D = np.zeros((nb * bh, mb * bw))
    for i in range(nb):
        for k in range(rp[i], rp[i + 1]):
            D[i * bh:(i + 1) * bh, col[k] * bw:(col[k] + 1) * bw] = blocks[k]
    x = rng.uniform(-1, 1, mb * bw)
    y = rng.uniform(-1, 1, nb * bh)
    r = y.copy()
    bflat = blocks.ravel().copy()
    po.lib().orc_csrb_apply_f64_u32(po.P(r), 0.5, po.P(x), -1.5, po.P(r), po.P(bflat), po.P(col), po.P(rp), nb, bh, bw)
    assert np.allclose(r, 0.5 * (D @ x) - 1.5 * y, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("dim,deg,lvl", [(2, 1, 5), (3, 1, 4), (2, 2, 4), (3, 2, 3)])
def test_transfers_are_kronecker_products(dim, deg, lvl):
    """In lattice order the prolongation of a tensor-product element on the unit square or cube
    is Pz (x) Py (x) Px with the 1D prolongation on every axis, and R = P^T (the structure the
    solver's tensor-product transfer evaluation checks for at init).  The 1D factor is the
    restriction of P to the first line of fine nodes along x; both P and R equal the Kronecker
    products of it entry for entry (the factors' values are dyadic, so the products are exact)."""
    import scipy.sparse as sp
    H = pb.hierarchy(dim, deg, lvl, 1)
    for L in H.levels[:-1]:
        P = sp.csr_matrix((L.P.val, L.P.col_ind, L.P.row_ptr), shape=(L.P.rows, L.P.cols))
        R = sp.csr_matrix((L.R.val, L.R.col_ind, L.R.row_ptr), shape=(L.R.rows, L.R.cols))
        P.eliminate_zeros()
        R.eliminate_zeros()
        mf, mc = round(P.shape[0] ** (1 / dim)), round(P.shape[1] ** (1 / dim))
        assert mf ** dim == P.shape[0] and mc ** dim == P.shape[1]
        P1 = P[:mf, :mc]
        assert P1.nnz == P[:mf, :].nnz          # the first fine x-line only reaches the first coarse x-line
        K = P1
        for _ in range(dim - 1):
            K = sp.kron(P1, K, format="csr")
        assert (abs(K - P)).max() == 0.0 and K.nnz == P.nnz
        assert (abs(K.T.tocsr() - R)).max() == 0.0
        # 1D weights: 1 at coinciding nodes; Q1 halves, Q2 the quadratic interpolation weights
        w = np.unique(P1.data)
        assert set(w) <= ({0.5, 1.0} if deg == 1 else {-0.125, 0.375, 0.75, 1.0})
