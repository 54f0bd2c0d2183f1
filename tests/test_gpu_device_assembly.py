"""The device assembly of a level (graph, values, Dirichlet list, prolongation and restriction)
equals the host setup bit for bit, with and without a coefficient field; the whole device
hierarchy equals hierarchy().  Bar: every index array and every value bitwise equal."""
import numpy as np
import pytest

from feat3_amd import problems as pb

pytestmark = pytest.mark.gpu


def same_csr(d, h, what):
    assert d.rows == h.rows and d.cols == h.cols, what
    assert np.array_equal(d.row_ptr, h.row_ptr), what
    assert np.array_equal(d.col_ind, h.col_ind), what
    assert np.array_equal(d.val.view(np.uint64), h.val.view(np.uint64)), what


@pytest.mark.parametrize("dim,deg,lvl", [(2, 1, 0), (2, 1, 5), (2, 2, 1), (2, 2, 4), (3, 1, 1), (3, 1, 4)])
def test_level_arrays_bitwise(torch_cuda, dim, deg, lvl):
    torch = torch_cuda
    print(f"case {dim}D Q{deg} L{lvl}", flush=True)
    A, bnd = pb.laplace(dim, deg, lvl)
    Ad, bd = pb.laplace_device(dim, deg, lvl)
    same_csr(Ad.to_host(), A, "A")
    assert np.array_equal(bd.cpu().numpy().view(np.uint32), bnd)
    # coefficient-weighted values on the same graph
    rng = np.random.default_rng(lvl)
    k = np.exp(rng.normal(size=(1 << lvl) ** dim))
    Ak, _ = pb.laplace(dim, deg, lvl, k)
    Akd, _ = pb.laplace_device(dim, deg, lvl, torch.from_numpy(k).cuda())
    same_csr(Akd.to_host(), Ak, "A(k)")
    if lvl >= 1:
        P = pb.prolongation(dim, deg, lvl)
        same_csr(pb.transfer_device(dim, deg, lvl).to_host(), P, "P")
        same_csr(pb.transfer_device(dim, deg, lvl, transpose=True).to_host(), pb.transpose(P), "R")


@pytest.mark.parametrize("dim,deg,lvl", [(3, 1, 6), (2, 2, 7)])
def test_hierarchy_bitwise(torch_cuda, dim, deg, lvl):
    torch = torch_cuda
    print(f"case hierarchy {dim}D Q{deg} L{lvl}", flush=True)
    rng = np.random.default_rng(5)
    k = np.exp(rng.normal(size=(1 << lvl) ** dim))
    H = pb.hierarchy(dim, deg, lvl, 1, kcell=k)
    Hd = pb.device_levels_to_host(dim, deg, lvl, 1, pb.hierarchy_device(dim, deg, lvl, 1, torch.from_numpy(k).cuda()))
    assert len(Hd.levels) == len(H.levels)
    for i, (a, b) in enumerate(zip(Hd.levels, H.levels)):
        same_csr(a.A, b.A, f"A level {i}")
        assert np.array_equal(a.bnd, b.bnd)
        if b.P is not None:
            same_csr(a.P, b.P, f"P level {i}")
            same_csr(a.R, b.R, f"R level {i}")


@pytest.mark.parametrize("dim,deg,lvl,fmt,inner", [(3, 1, 6, "banded", "f32"), (3, 1, 6, "csr", "f64"),
                                                   (3, 1, 6, "sell", "f16"), (2, 2, 7, "banded", "f32"),
                                                   (2, 1, 9, "banded", "f64")])
def test_solver_from_device_levels(torch_cuda, dim, deg, lvl, fmt, inner):
    """A solver initialised from the device-assembled levels (formats built on the device for
    levels above 65536 rows) solves bitwise like the one initialised from the host hierarchy."""
    torch = torch_cuda
    print(f"case device init {dim}D Q{deg} L{lvl} {fmt} {inner}", flush=True)
    from feat3_amd.solver import MGPCG
    H = pb.hierarchy(dim, deg, lvl, 1)
    levels = pb.hierarchy_device(dim, deg, lvl, 1)
    db = torch.from_numpy(pb.filtered_rhs(H, 0)).cuda()
    out = []
    for src in (H, levels):
        s = MGPCG(src, inner=inner, fmt=fmt).init()
        info = s.info()
        x = torch.zeros(H.n, dtype=torch.float64, device="cuda")
        st, it = s.correct(x, db)
        out.append((st, it, x.cpu().numpy(), info))
        s.done()
    assert out[0][0] == out[1][0] == 2 and out[0][1] == out[1][1]
    assert out[0][3] == out[1][3]
    assert np.array_equal(out[0][2], out[1][2])


def test_device_levels_are_validated(torch_cuda):
    """A device level whose columns leave the matrix is rejected by init (validated on the
    device before any kernel uses it)."""
    torch = torch_cuda
    from feat3_amd import capi
    from feat3_amd.solver import MGPCG
    levels = pb.hierarchy_device(3, 1, 6, 1)
    levels[0].A.col_ind[7] = levels[0].A.rows + 5
    with pytest.raises(capi.Feat3HipError):
        MGPCG(levels, inner="f32", fmt="banded").init()
