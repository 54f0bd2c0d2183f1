"""Parity of the batched device MG-PCG (K systems in lockstep) with K independent oracle solves.

Each system of a batch must behave exactly as its own solve: FP64 inner path: identical status
and iteration count, relative solution difference <= 1e-10 and defect history within 1e-9
relative; FP32 inner path: iterations within +-1 of the oracle's identical mixed algorithm and
relative difference to the FP64 oracle solution <= 1e-8.  Cases: K right-hand sides of one
scalar matrix, of the 3x3-block vector Laplacian (BCSR, UnitFilterBlocked), and K coefficient
fields on one mesh; K from 1 to 40 (idle lanes, all three lane-group widths), systems that
converge at different iterations, a zero right-hand side (success at iteration 0).
"""
import numpy as np
import pytest

from feat3_amd import problems as pb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def torch(torch_cuda):
    return torch_cuda


def gpu_batch(torch, H, B, nsys, values=None, block=1, shared=True, **kw):
    from feat3_amd.batch import BatchMGPCG
    s = BatchMGPCG(H, nsys, values=values, block=block, shared=shared, **kw).init()
    x = torch.zeros(s.shape, dtype=torch.float64, device="cuda")
    st, it = s.correct(x, torch.from_numpy(np.ascontiguousarray(B)).cuda())
    out = x.cpu().numpy()
    hists = [s.history(k) for k in range(nsys)]
    s.done()
    return st, it, out, hists


def orc_solve(H, b, block=1, values=None, **kw):
    S = po.System.from_hierarchy(H, block=block, values=values)
    x = np.zeros(b.size)
    st, it, hist = S.solve(x, np.ascontiguousarray(b), po.params(**kw))
    return st, it, x, hist


def check_system(k, g, o, inner):
    (sg, ig, xg, hg), (so, io, xo, ho) = g, o
    print(f"  system {k}: gpu status {sg} it {ig}; oracle status {so} it {io}; rel {rel(xg, xo):.2e}", flush=True)
    assert sg == so
    if inner == "f64":
        assert ig == io
        assert rel(xg, xo) <= 1e-10
        m = min(len(hg), len(ho))
        assert np.allclose(hg[:m], ho[:m], rtol=1e-9, atol=1e-300)
    else:
        assert abs(int(ig) - int(io)) <= 1
        assert rel(xg, xo) <= 1e-8


# K covers every lane layout: one system per lane (1, 3, 5, 40), S = 4 floats per lane with
# K/4 lanes per row (f32: 4, 16, 64), S = 2 doubles per lane (f64: 16, 32, 64)
@pytest.mark.parametrize("inner,nsys", [("f64", 1), ("f64", 3), ("f32", 4), ("f32", 5), ("f32", 16), ("f64", 16),
                                        ("f64", 32), ("f32", 40), ("f32", 64), ("f64", 64)])
def test_shared_scalar_rhs(torch, inner, nsys):
    print(f"case shared scalar 2D Q1 L5 inner {inner} K {nsys}", flush=True)
    H = pb.hierarchy(2, 1, 5, 1)
    B = pb.random_rhs(H.n, 1, nsys, H.levels[0].bnd, seed=1234 + nsys)
    B[:, :, 0] = pb.filtered_rhs(H, 0)[:, None]   # the sine-bubble load as system 0
    sg, ig, X, hg = gpu_batch(torch, H, B, nsys, inner=inner)
    for k in range(nsys):
        b = B[:, 0, k].copy()
        o64 = orc_solve(H, b)
        ref = o64 if inner == "f64" else orc_solve(H, b, inner=1)
        g = (sg[k], ig[k], X[:, 0, k], hg[k])
        if inner == "f64":
            check_system(k, g, ref, inner)
        else:
            check_system(k, (sg[k], ig[k], X[:, 0, k], hg[k]), (ref[0], ref[1], o64[2], ref[3]), inner)


@pytest.mark.parametrize("inner,nsys", [("f64", 8), ("f32", 8), ("f32", 32), ("f64", 3)])
def test_vector_laplacian_bcsr(torch, inner, nsys):
    print(f"case 3x3 BCSR vector Laplacian 3D Q1 L4 inner {inner} K {nsys}", flush=True)
    H = pb.hierarchy(3, 1, 4, 1)
    vals = [pb.block_identity_values(L.A, 3) for L in H.levels]
    B = pb.random_rhs(H.n, 3, nsys, H.levels[0].bnd, seed=77)
    sg, ig, X, hg = gpu_batch(torch, H, B, nsys, values=vals, block=3, inner=inner)
    for k in range(nsys):
        b = B[:, :, k].reshape(-1).copy()
        o64 = orc_solve(H, b, block=3, values=vals)
        ref = o64 if inner == "f64" else orc_solve(H, b, block=3, values=vals, inner=1)
        check_system(k, (sg[k], ig[k], X[:, :, k].reshape(-1), hg[k]), (ref[0], ref[1], o64[2], ref[3]), inner)


@pytest.mark.parametrize("inner,nsys", [("f64", 6), ("f32", 6), ("f64", 16), ("f32", 16)])
def test_coefficient_fields(torch, inner, nsys):
    lvl = 5
    print(f"case {nsys} coefficient fields 2D Q1 L{lvl} inner {inner}", flush=True)
    K = np.exp(np.stack([pb.gaussian_random_field(2, lvl, seed=100 + k) for k in range(nsys)]))
    Hs = pb.field_hierarchies(2, lvl, K)
    vals = [pb.stacked_values([Hk.levels[i].A for Hk in Hs]) for i in range(len(Hs[0].levels))]
    b = pb.filtered_rhs(Hs[0], 1)       # f = 1
    B = np.repeat(b[:, None, None], nsys, axis=2)
    sg, ig, X, hg = gpu_batch(torch, Hs[0], B, nsys, values=vals, shared=False, inner=inner)
    for k in range(nsys):
        o64 = orc_solve(Hs[k], b)
        ref = o64 if inner == "f64" else orc_solve(Hs[k], b, inner=1)
        check_system(k, (sg[k], ig[k], X[:, 0, k], hg[k]), (ref[0], ref[1], o64[2], ref[3]), inner)


def test_mixed_convergence_and_zero_rhs(torch):
    """Systems stop at different iterations; a zero right-hand side ends at iteration 0; the
    frozen systems are untouched while the others iterate."""
    print("case mixed convergence 2D Q1 L5 K 4", flush=True)
    H = pb.hierarchy(2, 1, 5, 1)
    nsys = 4
    B = np.zeros((H.n, 1, nsys))
    B[:, 0, 1] = pb.filtered_rhs(H, 0)
    B[:, 0, 2] = pb.filtered_rhs(H, 2)
    B[:, 0, 3] = pb.random_rhs(H.n, 1, 1, H.levels[0].bnd, seed=5)[:, 0, 0]
    sg, ig, X, hg = gpu_batch(torch, H, B, nsys, inner="f64")
    assert sg[0] == 2 and ig[0] == 0 and np.all(X[:, 0, 0] == 0.0)
    for k in range(1, nsys):
        check_system(k, (sg[k], ig[k], X[:, 0, k], hg[k]), orc_solve(H, B[:, 0, k].copy()), "f64")
    assert len(set(int(i) for i in ig[1:])) >= 1


def test_graph_and_stream_replay_identical(torch):
    print("case graph vs stream 2D Q1 L5 K 3", flush=True)
    H = pb.hierarchy(2, 1, 5, 1)
    B = pb.random_rhs(H.n, 1, 3, H.levels[0].bnd, seed=9)
    a = gpu_batch(torch, H, B, 3, inner="f32", graph=True)
    b = gpu_batch(torch, H, B, 3, inner="f32", graph=False)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])


def test_max_iter_status(torch):
    print("case max-iter 2D Q1 L5 K 2", flush=True)
    H = pb.hierarchy(2, 1, 5, 1)
    B = pb.random_rhs(H.n, 1, 2, H.levels[0].bnd, seed=11)
    sg, ig, X, hg = gpu_batch(torch, H, B, 2, inner="f64", max_iter=2, tol_rel=1e-14)
    for k in range(2):
        check_system(k, (sg[k], ig[k], X[:, 0, k], hg[k]),
                     orc_solve(H, B[:, 0, k].copy(), max_iter=2, tol_rel=1e-14), "f64")
        assert sg[k] == 5 and ig[k] == 2


def test_tile_order_does_not_change_results(torch):
    """The locality order changes which rows run together, not a row's arithmetic: iteration
    counts agree and solutions differ only by the reductions' summation order."""
    print("case tile order 3D Q1 L4 vector K 8", flush=True)
    H = pb.hierarchy(3, 1, 4, 1)
    vals = [pb.block_identity_values(L.A, 3) for L in H.levels]
    B = pb.random_rhs(H.n, 3, 8, H.levels[0].bnd, seed=21)
    a = gpu_batch(torch, H, B, 8, values=vals, block=3, inner="f64", options={"tile_rows": 0})
    b = gpu_batch(torch, H, B, 8, values=vals, block=3, inner="f64", options={"tile_rows": 64})
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert rel(b[2], a[2]) <= 1e-13


def test_inputs_from_the_default_stream_are_awaited(torch):
    """Regression: the solver runs on a stream of its own and must wait for the caller's stream
    also when that is the default (null) stream, whose handle is 0.  x is refilled with zeros
    behind a long kernel on the default stream; the solve has to start from zero, i.e. repeat
    the reference iterations and solution exactly."""
    from feat3_amd.batch import BatchMGPCG
    H = pb.hierarchy(2, 1, 7, 1)
    nsys = 4
    b = pb.filtered_rhs(H, 0)
    B = torch.from_numpy(np.ascontiguousarray(np.repeat(b[:, None, None], nsys, axis=2))).cuda()
    s = BatchMGPCG(H, nsys, inner="f64").init()
    x = torch.zeros(s.shape, dtype=torch.float64, device="cuda")
    st0, it0 = s.correct(x, B)
    ref = x.clone()
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(3):
        # x holds the converged solution; keep the default stream busy, then refill x with zeros
        big = torch.randn(4096, 4096, device="cuda")
        for _ in range(4):
            big = big @ big * 1e-3
        x.zero_()
        st, it = s.correct(x, B, stream=stream)
        assert list(it) == list(it0) and np.array_equal(x.cpu().numpy(), ref.cpu().numpy())
    s.done()


def test_thinned_graph_refuses_new_values(torch):
    """With shared FP32 values the inner graph of the 3D vector Laplacian leaves out the face
    blocks (roundoff residues below eps_f32^2 of the diagonal); the solve keeps the parity bar
    (test_vector_laplacian_bcsr) and a later set_values on such a level is refused, since the
    thinning was chosen for the values given at init."""
    import ctypes
    from feat3_amd import capi
    from feat3_amd.batch import BatchMGPCG
    H = pb.hierarchy(3, 1, 3, 1)
    vals = [pb.block_identity_values(L.A, 3) for L in H.levels]
    s = BatchMGPCG(H, 4, values=vals, block=3, inner="f32").init()
    dv = torch.from_numpy(np.ascontiguousarray(vals[0], np.float64)).cuda()
    with pytest.raises(Exception):
        capi.call("feat3_batch_set_values", s._h, 0, ctypes.c_void_p(dv.data_ptr()))
    s.done()
    # FP64 keeps the full graph, so new values are taken -- with full blocks; the block-diagonal
    # storage of these diagonal blocks (the default) was chosen for the values given at init too
    s = BatchMGPCG(H, 4, values=vals, block=3, inner="f64", options={"diag_blocks": 0}).init()
    capi.call("feat3_batch_set_values", s._h, 0, ctypes.c_void_p(dv.data_ptr()))
    s.done()
    s = BatchMGPCG(H, 4, values=vals, block=3, inner="f64").init()
    with pytest.raises(Exception):
        capi.call("feat3_batch_set_values", s._h, 0, ctypes.c_void_p(dv.data_ptr()))
    s.done()


@pytest.mark.parametrize("inner", ["f64", "f32", "f16"])
@pytest.mark.parametrize("nsys", [4, 32])
def test_diagonal_block_storage_is_identical(torch, inner, nsys):
    """The vector Laplacian's 3x3 blocks are a_ij * I: stored as block diagonals (option
    diag_blocks, default) the solve is bit for bit the full-block solve -- the left-out terms are
    products with exact zeros, which change no partial sum."""
    print(f"case diag blocks {inner} K={nsys}", flush=True)
    H = pb.hierarchy(3, 1, 4, 1)
    vals = [pb.block_identity_values(L.A, 3) for L in H.levels]
    B = pb.random_rhs(H.n, 3, nsys, H.levels[0].bnd, seed=3)
    s1, i1, x1, _ = gpu_batch(torch, H, B, nsys, values=vals, block=3, inner=inner)
    s0, i0, x0, _ = gpu_batch(torch, H, B, nsys, values=vals, block=3, inner=inner, options={"diag_blocks": 0})
    assert list(s1) == list(s0) and list(i1) == list(i0)
    assert np.array_equal(x1, x0)


@pytest.mark.parametrize("inner", ["f64", "f32", "f16"])
@pytest.mark.parametrize("dim,level,block,nsys,tail", [(3, 5, 3, 32, 1000), (3, 4, 3, 5, 1000), (3, 3, 3, 4, 1000),
                                                        (2, 6, 1, 8, 1000), (2, 5, 1, 3, 2000)])
def test_coarse_tail_is_identical(torch, inner, dim, level, block, nsys, tail):
    """Shared values: the levels of at most tail_max_n nodes (3D L5: the levels of 729, 125 and
    27 nodes; 3D L3 and 2D L5 with 2000: the whole hierarchy, the finest level included)
    run in one LDS-resident workgroup per system; the solve is bit for bit that of the per-level
    launches (tail_max_n 0)."""
    print(f"case coarse tail {inner} {dim}D L{level} B={block} K={nsys} tail {tail}", flush=True)
    H = pb.hierarchy(dim, 1, level, 1)
    vals = [pb.block_identity_values(L.A, 3) for L in H.levels] if block == 3 else None
    B = pb.random_rhs(H.n, block, nsys, H.levels[0].bnd, seed=17)
    s1, i1, x1, h1 = gpu_batch(torch, H, B, nsys, values=vals, block=block, inner=inner, options={"tail_max_n": tail})
    s0, i0, x0, h0 = gpu_batch(torch, H, B, nsys, values=vals, block=block, inner=inner, options={"tail_max_n": 0})
    assert list(s1) == list(s0) and list(i1) == list(i0)
    assert np.array_equal(x1, x0)
    assert all(np.array_equal(a, b) for a, b in zip(h1, h0))


@pytest.mark.parametrize("inner", ["f64", "f32"])
def test_coupled_blocks_keep_full_storage(torch, inner):
    """Blocks a_ij * M with a coupling M (SPD, tridiagonal) are not diagonal: the full 3x3 blocks
    are kept, and every system meets its oracle solve (apply_generic.hpp:121-154 block products)."""
    print(f"case coupled blocks {inner}", flush=True)
    H = pb.hierarchy(3, 1, 3, 1)
    M = np.array([[1.0, 0.25, 0.0], [0.25, 1.0, 0.25], [0.0, 0.25, 1.0]])
    vals = [np.kron(L.A.val, M.ravel()) for L in H.levels]
    nsys = 4
    B = pb.random_rhs(H.n, 3, nsys, H.levels[0].bnd, seed=9)
    g = gpu_batch(torch, H, B, nsys, values=vals, block=3, inner=inner)
    for k in range(nsys):
        o = orc_solve(H, B[:, :, k].reshape(-1).copy(), block=3, values=vals)
        check_system(k, (g[0][k], g[1][k], g[2][:, :, k].reshape(-1), g[3][k]), o, inner)
