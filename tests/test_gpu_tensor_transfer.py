"""Tensor-product evaluation of the grid transfers (option "tensor_transfer", default on).

The prolongation of a Q1/Q2 hierarchy on the unit square or cube in lattice order is
Pz (x) Py (x) Px (transfer.hpp:249-285 applies it as a stored CSR matrix); the solver detects that
at init, checks every stored entry of P and R against the product of the 1D factors, and then
evaluates them from the factors.  Bar: the solution of every solve is bit for bit that of the
same solve with the stored matrices (the terms of each row sum are the same, in the same order);
the FEAT-numbered hierarchy keeps the stored matrices unless the solver works in the lattice
order; the solve meets the oracle as every solve does.
"""
import numpy as np
import pytest

from feat3_amd import problems as pb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch(torch_cuda):
    return torch_cuda


def solve(torch, H, b, **kw):
    from feat3_amd.solver import MGPCG
    s = MGPCG(H, **kw).init()
    x = torch.zeros(H.n, dtype=torch.float64, device="cuda")
    st, it = s.correct(x, torch.from_numpy(b).cuda())
    info = s.info()
    s.done()
    return st, it, x.cpu().numpy(), info


@pytest.mark.parametrize("dim,deg,lvl", [(2, 1, 6), (2, 2, 5), (3, 1, 5)])
@pytest.mark.parametrize("inner", ["f64", "f32", "f16"])
@pytest.mark.parametrize("fmt", ["banded", "sell"])
def test_tensor_transfers_give_identical_solves(torch, dim, deg, lvl, inner, fmt):
    print(f"case tensor transfer {dim}D Q{deg} L{lvl} {inner} {fmt}", flush=True)
    H = pb.hierarchy(dim, deg, lvl, 1)
    b = pb.filtered_rhs(H, 0)
    st1, it1, x1, i1 = solve(torch, H, b, inner=inner, fmt=fmt)
    st0, it0, x0, i0 = solve(torch, H, b, inner=inner, fmt=fmt, options={"tensor_transfer": 0})
    assert i1["tensor_transfers"] == len(H.levels) - 1
    assert i0["tensor_transfers"] == 0
    assert (st1, it1) == (st0, it0) and st1 == 2
    assert np.array_equal(x1, x0)


def test_tensor_transfers_follow_the_level_order(torch):
    """FEAT-numbered levels: the stored matrices without a level order (no lattice), the tensor
    evaluation with the lattice order; both solves meet the oracle (FP64: same iterations,
    1e-10)."""
    H = pb.hierarchy(3, 1, 5, 1, numbering="feat")
    b = pb.filtered_rhs(H, 0)
    S = po.System.from_hierarchy(H)
    xo = np.zeros(H.n)
    so, io, _ = S.solve(xo, b, po.params())
    for orders, expect in (("auto", len(H.levels) - 1), (None, 0)):
        st, it, x, info = solve(torch, H, b, inner="f64", fmt="banded", level_orders=orders)
        assert info["tensor_transfers"] == expect
        assert st == so == 2 and it == io
        assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)


def test_vcycle_with_tensor_transfers_matches_oracle(torch):
    """One FP64 V-cycle on 3D Q1 L5 against the oracle's (multigrid.hpp:1711-2000), 1e-12."""
    from feat3_amd.solver import MGPCG
    H = pb.hierarchy(3, 1, 5, 1)
    d = pb.filtered_rhs(H, 0)
    S = po.System.from_hierarchy(H)
    co = np.zeros(H.n)
    S.vcycle(co, d.copy(), po.params())
    s = MGPCG(H, inner="f64", fmt="banded").init()
    assert s.info()["tensor_transfers"] == len(H.levels) - 1
    cg = torch.zeros(H.n, dtype=torch.float64, device="cuda")
    s.vcycle(cg, torch.from_numpy(d).cuda())
    torch.cuda.synchronize()
    s.done()
    assert np.linalg.norm(cg.cpu().numpy() - co) <= 1e-12 * np.linalg.norm(co)
