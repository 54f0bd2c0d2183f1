"""A solver built by the SolverFactory from an INI configuration is the solver built directly
with the same parameters: identical iteration count and bitwise identical solution; the FP64
configuration meets the oracle bar (identical iterations, <= 1e-10)."""
import numpy as np
import pytest

from feat3_amd import factory as fa
from feat3_amd import problems as pb
from feat3_amd.solver import MGPCG
from oracle import pyoracle as po

from test_factory import POISSON

pytestmark = pytest.mark.gpu


def _solve(torch, s, H, b):
    s.init()
    x = torch.zeros(H.n, dtype=torch.float64, device="cuda")
    st, it = s.correct(x, torch.from_numpy(b).cuda())
    out = x.cpu().numpy()
    s.done()
    return st, it, out


def test_factory_solver_equals_direct_construction(torch_cuda):
    torch = torch_cuda
    H = pb.hierarchy(3, 1, 5, 1)
    b = pb.filtered_rhs(H, 0)
    sf, itf, xf = _solve(torch, fa.create_scalar_solver(H, POISSON), H, b)
    sd, itd, xd = _solve(torch, MGPCG(H, inner="f32", tol_rel=1e-8, max_iter=1000, omega_jacobi=0.5,
                                      omega_richardson=1.0, pre=4, post=4), H, b)
    assert sf == sd == 2 and itf == itd
    assert np.array_equal(xf, xd)


@pytest.mark.parametrize("cycle,code", [("v", po.CYCLE_V), ("f", po.CYCLE_F), ("w", po.CYCLE_W)])
def test_factory_fp64_meets_the_oracle(torch_cuda, cycle, code):
    torch = torch_cuda
    H = pb.hierarchy(2, 2, 5, 1)
    b = pb.filtered_rhs(H, 0)
    cfg = POISSON.replace("datatype = float", "datatype = double").replace("cycle = v", f"cycle = {cycle}")
    st, it, x = _solve(torch, fa.create_scalar_solver(H, cfg), H, b)
    S = po.System.from_hierarchy(H)
    xo = np.zeros(H.n)
    so, io, _ = S.solve(xo, b, po.params(cycle=code))
    assert st == so == 2 and it == io
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-10
