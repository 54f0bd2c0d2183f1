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
