"""Parity of the device MG-PCG solver with the oracle's restatement of FEAT's PCG / MultiGrid /
Richardson / Jacobi / ConvertPrecond on the same hierarchies.

Bars (stated per test): FP64 path: identical iteration count and status, relative solution
difference <= 1e-10, defect history within 1e-9 relative; mixed FP32 / FP16-stored paths:
iteration count within +-1 of the oracle's identical mixed algorithm and relative solution
difference to the FP64 oracle solution <= 1e-8 (both solves stop at relative residual 1e-8).
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


def solve_gpu(torch, H, b, x0=None, **kw):
    from feat3_amd.solver import MGPCG
    s = MGPCG(H, **kw).init()
    x = torch.zeros(H.n, dtype=torch.float64, device="cuda") if x0 is None else torch.from_numpy(x0.copy()).cuda()
    db = torch.from_numpy(b).cuda()
    st, it = s.correct(x, db)
    out = x.cpu().numpy()
    hist = s.history()
    s.done()
    return st, it, out, hist


def solve_orc(H, b, **kw):
    S = po.System.from_hierarchy(H)
    x = np.zeros(H.n)
    st, it, hist = S.solve(x, b, po.params(**kw))
    return st, it, x, hist


CASES = [
    ("2D Q1 L6 (config 1)", 2, 1, 6),
    ("2D Q2 L5", 2, 2, 5),
    ("3D Q1 L5", 3, 1, 5),
]


@pytest.mark.parametrize("name,dim,deg,lvl", CASES)
def test_fp64_solve_matches_oracle(torch, name, dim, deg, lvl):
    print(f"case {name}", flush=True)
    H = pb.hierarchy(dim, deg, lvl, 1)
    b = pb.filtered_rhs(H, 0)
    so, io, xo, ho = solve_orc(H, b)
    sg, ig, xg, hg = solve_gpu(torch, H, b)
    assert so == 2 and sg == 2
    assert ig == io
    assert rel(xg, xo) <= 1e-10
    assert np.allclose(hg[: io + 1], ho[: io + 1], rtol=1e-9, atol=0)


@pytest.mark.parametrize("inner,code", [("f32", 1), ("f16", 2)])
@pytest.mark.parametrize("name,dim,deg,lvl", CASES)
def test_mixed_solve_matches_oracle(torch, inner, code, name, dim, deg, lvl):
    print(f"case {name} inner {inner}", flush=True)
    H = pb.hierarchy(dim, deg, lvl, 1)
    b = pb.filtered_rhs(H, 0)
    so, io, xo, _ = solve_orc(H, b, inner=code)
    s64, i64, x64, _ = solve_orc(H, b)
    sg, ig, xg, hg = solve_gpu(torch, H, b, inner=inner)
    assert sg == 2
    assert abs(ig - io) <= 1
    assert hg[ig] <= 1e-8 * hg[0]
    assert rel(xg, x64) <= 1e-8


def test_vcycle_matches_oracle(torch):
    from feat3_amd.solver import MGPCG
    H = pb.hierarchy(2, 1, 6, 1)
    rng = np.random.default_rng(1)
    d = rng.uniform(-1, 1, H.n)
    d[H.levels[0].bnd] = 0.0
    for inner, code, tol in [("f64", 0, 1e-12), ("f32", 1, 1e-5), ("f16", 2, 1e-5)]:
        S = po.System.from_hierarchy(H)
        cor = np.zeros(H.n)
        S.vcycle(cor, d, po.params(inner=code))
        s = MGPCG(H, inner=inner).init()
        dd = torch.from_numpy(d).cuda()
        dc = torch.zeros_like(dd)
        s.vcycle(dc, dd)
        assert rel(dc.cpu().numpy(), cor) <= tol, inner
        s.done()


def test_sell_graph_and_lanes_variants_agree(torch):
    H = pb.hierarchy(3, 1, 4, 1)
    b = pb.filtered_rhs(H, 0)
    ref = solve_gpu(torch, H, b)
    for kw in [dict(fmt="sell"), dict(graph=False), dict(lanes=16), dict(fmt="sell", sell_C=32, sell_sigma=1),
               dict(fmt="sell", sell_sigma=64), dict(options=dict(tail_max_n=0)), dict(options=dict(tail_max_n=10 ** 6)),
               dict(fmt="sell", options=dict(index16=0, compact_transfer=0)),
               dict(fmt="sell", options=dict(transfer_csr=0, transfer_sigma=64)),
               dict(fmt="banded", options=dict(transfer_csr=0, transfer_sigma=256)), dict(fmt="sell", options=dict(transfer_csr=0)),
               dict(fmt="banded", options=dict(rows_per_block=128)),
               dict(options=dict(transfer_lanes=4, transfer_rows_per_block=0)),
               dict(fmt="banded", options=dict(transfer_rows_per_block=512))]:
        st, it, x, _ = solve_gpu(torch, H, b, **kw)
        assert st == 2 and it == ref[1], kw
        assert rel(x, ref[2]) <= 1e-12, kw
    # graph replay is bitwise identical to kernel-by-kernel launch (deterministic reductions)
    st, it, x, _ = solve_gpu(torch, H, b, graph=False)
    assert np.array_equal(x, ref[2])


def test_cg_pointstar_kat(torch):
    """basic_solver-test.cpp:136-140: plain CG on PointstarFactoryFD(17, 2), 28 +- 2 iterations."""
    m = 17
    rp, col, val = po.pointstar_fd(m, 2)
    n = m * m
    ref = po.pointstar_q2_bubble(m, 2)
    b = np.zeros(n)
    po.csr_apply(b, 1.0, ref, 0.0, b, val, col, rp, n, n)
    H = pb.Hierarchy(2, 1, [pb.Level(pb.Csr(n, n, rp, col, val), np.zeros(0, np.uint32))], 0, 0)
    st, it, x, _ = solve_gpu(torch, H, b, precond="none", tol_rel=np.sqrt(np.finfo(float).eps), max_iter=100)
    assert st == 2
    assert abs(it - 28) <= 2
    assert np.sum((x - ref) ** 2) <= np.sqrt(np.finfo(float).eps)


def test_status_semantics(torch):
    """iterative.hpp:744-832: zero rhs -> success after 0 iterations; NaN -> aborted;
    max_iter reached -> max-iter."""
    H = pb.hierarchy(2, 1, 4, 1)
    z = np.zeros(H.n)
    st, it, x, _ = solve_gpu(torch, H, z)
    assert st == 2 and it == 0 and not np.any(x)
    bn = pb.filtered_rhs(H, 0)
    bn[H.n // 2] = np.nan
    st, it, _, _ = solve_gpu(torch, H, bn)
    assert st == 3
    b = pb.filtered_rhs(H, 0)
    st, it, _, _ = solve_gpu(torch, H, b, max_iter=1)
    assert st == 5 and it == 1
    so, io, _, _ = solve_orc(H, b, max_iter=1)
    assert so == 5 and io == 1


def test_apply_and_restart(torch):
    """apply(cor, def) starts from zero (pcg.hpp:141-156); correct(x, b) starts from the given
    x and measures the relative tolerance against its own initial defect."""
    from feat3_amd.solver import MGPCG
    H = pb.hierarchy(2, 2, 4, 1)
    b = pb.filtered_rhs(H, 0)
    s = MGPCG(H).init()
    db = torch.from_numpy(b).cuda()
    cor = torch.full_like(db, 7.0)
    st, it = s.apply(cor, db)
    assert st == 2
    so, io, xo, _ = solve_orc(H, b, use_def=0)
    assert it == io and rel(cor.cpu().numpy(), xo) <= 1e-10
    x = cor.clone()
    s.set_tolerances(1e-8)
    st, it2 = s.correct(x, db)
    assert st == 2 and it2 >= 1
    assert rel(x.cpu().numpy(), xo) <= 1e-9
    s.done()


@pytest.mark.parametrize("dim,deg,lvl", [(3, 1, 5), (2, 2, 5), (2, 1, 6)])
@pytest.mark.parametrize("inner", ["f64", "f32", "f16"])
def test_banded_format_matches_oracle(torch, dim, deg, lvl, inner):
    """The banded level-matrix format keeps the parity bars: FP64 identical iterations and
    <= 1e-10; mixed paths iterations +-1 of the oracle's mixed algorithm and <= 1e-8 against the
    FP64 oracle solution.  In lexicographic numbering Q1 matrices are banded (27 / 9 diagonals,
    little fill); the FP32 / binary16 band of the 3D matrix leaves out the 6 face diagonals, whose
    values are roundoff residues of about 2e-17 of the diagonal (below eps_f32^2), so it has 21.
    The 2D Q2 matrix has 25 diagonals but rows of 9..25 entries (56% fill), so a band of values
    falls back to SELL-C-sigma there; the value-coded band (default) takes it."""
    print(f"case banded {dim}D Q{deg} L{lvl} {inner}", flush=True)
    from feat3_amd.solver import MGPCG
    H = pb.hierarchy(dim, deg, lvl, 1)
    b = pb.filtered_rhs(H, 0)
    so, io, xo, _ = solve_orc(H, b)
    s = MGPCG(H, inner=inner, fmt="banded").init()
    info = s.info()
    s.done()
    s = MGPCG(H, inner=inner, fmt="banded", options={"band_dict": 0}).init()
    info0 = s.info()
    s.done()
    if deg == 1:
        # every diagonal holds at most 16 distinct values: value-coded band (3); without it the
        # assembled Laplacians, symmetric bit for bit, keep the upper half of the FP64 band (2)
        assert info["banded"] == 3
        assert info0["banded"] == (2 if inner == "f64" else 1)
        assert info["ndiag"] == info0["ndiag"] == ((27 if inner == "f64" else 21) if dim == 3 else 9)
    else:
        # 25 diagonals at 57% fill: too much fill for a band of values (SELL-C-sigma), within the
        # 4 slots per entry a value-coded band may carry
        assert info["banded"] == 3 and info["ndiag"] == 25
        assert info0["banded"] == 0 and info0["sell_chunks"] > 0
    sg, ig, xg, _ = solve_gpu(torch, H, b, inner=inner, fmt="banded")
    assert so == 2 and sg == 2
    if inner == "f64":
        assert ig == io and rel(xg, xo) <= 1e-10
    else:
        _, im, _, _ = solve_orc(H, b, inner={"f32": 1, "f16": 2}[inner])
        assert abs(ig - im) <= 1 and rel(xg, xo) <= 1e-8


@pytest.mark.parametrize("dim,deg,lvl,tail_levels", [(3, 1, 5, 3), (2, 2, 5, 4), (2, 1, 6, 5), (2, 1, 4, 4)])
@pytest.mark.parametrize("inner", ["f64", "f32", "f16"])
def test_coarse_tail_matches_per_level_launches(torch, dim, deg, lvl, tail_levels, inner):
    """The single-workgroup coarse tail (levels of at most 1100 rows, vectors in LDS, ELL
    matrices) takes the same V-cycle steps as the per-level launches: FP64 identical iteration
    count and solutions within 1e-12 (only the order of a row's terms differs); FP32 / FP16-stored
    V-cycles identical iteration count and within 1e-8 of the FP64 oracle solution.  On 2D Q1 L4
    every level fits, so the tail holds the whole hierarchy."""
    print(f"case tail {dim}D Q{deg} L{lvl} {inner}", flush=True)
    from feat3_amd.solver import MGPCG
    H = pb.hierarchy(dim, deg, lvl, 1)
    b = pb.filtered_rhs(H, 0)
    s = MGPCG(H, inner=inner).init()
    info = s.info()
    s.done()
    assert info["tail_levels"] == tail_levels, info
    assert info["tail_lds"] == 5 * {"f64": 8, "f32": 4, "f16": 4}[inner] * sum(L.A.rows for L in H.levels[-tail_levels:])
    st0, it0, x0, _ = solve_gpu(torch, H, b, inner=inner, options=dict(tail_max_n=0))
    st1, it1, x1, _ = solve_gpu(torch, H, b, inner=inner)
    assert st0 == 2 and st1 == 2 and it1 == it0
    if inner == "f64":
        assert rel(x1, x0) <= 1e-12
    else:
        _, _, xo, _ = solve_orc(H, b)
        assert rel(x1, xo) <= 1e-8


def test_time_kernel_entries(torch):
    """feat3_mgpcg_time_kernel: every hot-path kernel it times runs and reports its algorithmic
    bytes; timing leaves a following solve unchanged (bitwise)."""
    from feat3_amd.solver import MGPCG
    H = pb.hierarchy(3, 1, 4, 1)
    b = pb.filtered_rhs(H, 0)
    ref = solve_gpu(torch, H, b, inner="f32", fmt="banded")
    s = MGPCG(H, inner="f32", fmt="banded", options={"band_dict": 0}).init()
    A = H.levels[0].A
    n, nnz = H.n, A.nnz
    # the FP32 band keeps the entries above eps_f32^2 of their row's diagonal (21 of 27 per row)
    rows = np.repeat(np.arange(A.rows), np.diff(A.row_ptr.astype(np.int64)))
    diag = np.zeros(A.rows)
    diag[rows[A.col_ind == rows]] = A.val[A.col_ind == rows]
    keep = np.abs(A.val) > np.finfo(np.float32).eps ** 2 * np.abs(diag[rows])
    kept = int(np.count_nonzero(keep))
    assert kept < nnz
    # the FP64 band of the outer product has symmetric storage: the entries on and above the diagonal
    up64 = int(np.count_nonzero(A.col_ind >= rows))
    for which, min_bytes in ((0, kept * 4 + 4 * n * 4), (1, kept * 4 + 3 * n * 4), (2, up64 * 8 + 2 * n * 8), (3, 0)):
        ms, nb = s.time_kernel(which, 3)
        assert ms > 0.0, which
        assert nb >= min_bytes, (which, nb)
    x = torch.zeros(H.n, dtype=torch.float64, device="cuda")
    st, it = s.correct(x, torch.from_numpy(b).cuda())
    s.done()
    assert st == 2 and it == ref[1]
    assert np.array_equal(x.cpu().numpy(), ref[2])
    # value-coded bands: 4-bit codes, 8 diagonals per word (FP32: 21 diagonals, 3 words; FP64:
    # 27, 4 words), the offsets and per word 8 tables of 16 values, then the vectors
    s = MGPCG(H, inner="f32", fmt="banded").init()
    assert s.info()["banded"] == 3
    mask = (n + 31) // 32 * 4
    band32 = 3 * n * 4 + 21 * 4 + 3 * 8 * 16 * 4
    band64 = 4 * n * 4 + 27 * 4 + 4 * 8 * 16 * 8
    for which, nbytes in ((0, band32 + 4 * n * 4), (1, band32 + 3 * n * 4 + mask), (2, band64 + 2 * n * 8 + mask)):
        ms, nb = s.time_kernel(which, 3)
        assert ms > 0.0 and nb == nbytes, (which, nb, nbytes)
    x.zero_()
    st, it = s.correct(x, torch.from_numpy(b).cuda())
    s.done()
    assert st == 2 and it == ref[1]
    assert np.array_equal(x.cpu().numpy(), ref[2])
    with pytest.raises(Exception):
        MGPCG(H).init().time_kernel(7, 1)


@pytest.mark.parametrize("dim,lvl", [(3, 4), (2, 6)])
def test_symmetric_band_is_bitwise_the_full_band(torch, dim, lvl):
    """The FP64 band of a matrix equal to its transpose bit for bit stores only the center and
    upper diagonals; the row sums take the same values in the same order, so the FP64 solve is
    bitwise that of the full band (option symmetric_band = 0), for the host- and the
    device-built solver."""
    print(f"case symmetric band {dim}D Q1 L{lvl}", flush=True)
    from feat3_amd.solver import MGPCG
    H = pb.hierarchy(dim, 1, lvl, 1)
    b = pb.filtered_rhs(H, 0)
    s = MGPCG(H, fmt="banded", options=dict(band_dict=0)).init()
    assert s.info()["banded"] == 2
    s.done()
    ref = solve_gpu(torch, H, b, fmt="banded", options=dict(symmetric_band=0, band_dict=0))
    sym = solve_gpu(torch, H, b, fmt="banded", options=dict(band_dict=0))
    assert ref[1] == sym[1] and np.array_equal(ref[2], sym[2])
    levels = pb.hierarchy_device(dim, 1, lvl, 1)
    sd = MGPCG(levels, fmt="banded", options=dict(band_dict=0)).init()
    assert sd.info()["banded"] == 2
    x = torch.zeros(H.n, dtype=torch.float64, device="cuda")
    st, it = sd.correct(x, torch.from_numpy(b).cuda())
    sd.done()
    assert it == ref[1] and np.array_equal(x.cpu().numpy(), ref[2])


def test_nonsymmetric_band_keeps_the_full_band(torch):
    """A banded matrix that is not symmetric keeps the full band and the FP64 parity bars."""
    print("case non-symmetric band 2D Q1 L5", flush=True)
    from feat3_amd.solver import MGPCG
    H = pb.hierarchy(2, 1, 5, 1)
    A = H.levels[0].A
    rows = np.repeat(np.arange(A.rows), np.diff(A.row_ptr.astype(np.int64)))
    A.val = np.where(A.col_ind > rows, A.val * (1.0 + 1e-3), A.val)
    b = pb.filtered_rhs(H, 0)
    s = MGPCG(H, fmt="banded", options=dict(band_dict=0)).init()
    assert s.info()["banded"] == 1
    s.done()
    so, io, xo, _ = solve_orc(H, b)
    for opts in (dict(band_dict=0), dict(band_dict=1)):
        sg, ig, xg, _ = solve_gpu(torch, H, b, fmt="banded", options=opts)
        assert so == sg == 2 and ig == io
        assert rel(xg, xo) <= 1e-10


@pytest.mark.parametrize("fmt", ["csr", "sell"])
def test_fp32_level_copies_drop_unresolvable_couplings(torch, fmt):
    """The FP32 level matrices of every format keep only the entries above eps_f32^2 of their
    row's diagonal (3D Q1: the 6 face couplings are roundoff residues); the finest sweep streams
    exactly that many entries (CSR: kept * (4 + 4) + 4 (n + 1) + 4 n * 4 bytes), the FP64 copy keeps
    every nonzero, the host- and device-built solvers agree bitwise and the solve keeps the
    oracle's iteration count."""
    print(f"case FP32 thinned {fmt} 3D Q1 L4", flush=True)
    from feat3_amd.solver import MGPCG
    H = pb.hierarchy(3, 1, 4, 1)
    A = H.levels[0].A
    n, nnz = H.n, A.nnz
    rows = np.repeat(np.arange(A.rows), np.diff(A.row_ptr.astype(np.int64)))
    diag = np.zeros(A.rows)
    diag[rows[A.col_ind == rows]] = A.val[A.col_ind == rows]
    kept = int(np.count_nonzero(np.abs(A.val) > np.finfo(np.float32).eps ** 2 * np.abs(diag[rows])))
    assert kept < nnz
    if fmt == "csr":
        s = MGPCG(H, inner="f32", fmt=fmt).init()
        assert s.time_kernel(0, 1)[1] == kept * 8 + 4 * (n + 1) + 4 * n * 4
        s.done()
        s = MGPCG(H, inner="f64", fmt=fmt).init()
        assert s.time_kernel(0, 1)[1] == nnz * 12 + 4 * (n + 1) + 8 * n * 4
        s.done()
    b = pb.filtered_rhs(H, 0)
    so, io, xo, _ = solve_orc(H, b)
    sg, ig, xg, _ = solve_gpu(torch, H, b, inner="f32", fmt=fmt)
    assert so == sg == 2 and ig == io
    assert rel(xg, xo) <= 1e-8
    sd = MGPCG(pb.hierarchy_device(3, 1, 4, 1), inner="f32", fmt=fmt).init()
    x = torch.zeros(H.n, dtype=torch.float64, device="cuda")
    st, it = sd.correct(x, torch.from_numpy(b).cuda())
    sd.done()
    assert st == 2 and it == ig and np.array_equal(x.cpu().numpy(), xg)


@pytest.mark.parametrize("fmt", ["csr", "sell", "banded"])
@pytest.mark.parametrize("inner", ["f64", "f32", "f16"])
@pytest.mark.parametrize("level", [0, 1])
def test_nan_matrix_aborts_in_every_format(torch, fmt, inner, level):
    """A NaN coupling in a level matrix (the finest: the outer operator and the V-cycle; the next
    coarser: the V-cycle only) makes the solve report 'aborted' (iterative.hpp:757-758,794-795;
    pcg.hpp:220-224,266-269), as the oracle does, in every level format and precision: no copy of
    a level matrix leaves the non-finite value out."""
    print(f"case fmt {fmt} inner {inner} level {level}", flush=True)
    H = pb.hierarchy(2, 1, 5, 1)
    A = H.levels[level].A
    m = (2 ** (5 - level) + 1)
    r = (m // 2) * m + m // 2                      # an interior row
    ks = [k for k in range(A.row_ptr[r], A.row_ptr[r + 1]) if A.col_ind[k] != r]
    A.val[ks[0]] = np.nan
    b = pb.filtered_rhs(H, 0)
    so, io, _, _ = solve_orc(H, b)
    sg, ig, xg, _ = solve_gpu(torch, H, b, inner=inner, fmt=fmt)
    assert so == 3
    assert sg == 3, f"status {sg} after {ig} iterations"
    assert ig == io


@pytest.mark.parametrize("dim,deg,lvl", [(3, 1, 5), (2, 2, 5), (2, 1, 6)])
@pytest.mark.parametrize("inner", ["f64", "f32", "f16"])
def test_band_x_staged_in_lds_is_bitwise_identical(torch, dim, deg, lvl, inner):
    """Option band_stage_x: the banded row kernels read x from LDS windows staged per round of
    256 rows (1: the FP64 outer matrix, 2: every banded matrix).  The row sums take the same terms
    in the same order, so a V-cycle (no reductions) is bit for bit the unstaged one; a solve, whose
    dot products sum block partials of a grid of 512-row blocks instead, keeps the iteration count
    and agrees to 1e-12."""
    from feat3_amd.solver import MGPCG
    print(f"case band stage {dim}D Q{deg} L{lvl} {inner}", flush=True)
    H = pb.hierarchy(dim, deg, lvl, 1)
    b = pb.filtered_rhs(H, 0)
    db = torch.from_numpy(b).cuda()
    ref = None
    for stage in (0, 1, 2):
        s = MGPCG(H, inner=inner, fmt="banded", options={"band_stage_x": stage}).init()
        c = torch.zeros_like(db)
        s.vcycle(c, db)
        torch.cuda.synchronize()
        s.done()
        st, it, x, _ = solve_gpu(torch, H, b, inner=inner, fmt="banded", options={"band_stage_x": stage})
        assert st == 2
        if ref is None:
            ref = (c.cpu().numpy(), it, x)
        else:
            assert np.array_equal(c.cpu().numpy(), ref[0]), stage
            assert it == ref[1] and rel(x, ref[2]) <= 1e-12, stage
