"""Parity of the device multigrid F- and W-cycles (kernel/solver/multigrid.hpp:1366-1519) with the
oracle's restatement on the same hierarchies.

Bars: one cycle in FP64 within 1e-12 relative of the oracle's (1e-5 in FP32 and with FP16-stored
values, compared with the oracle's identical mixed algorithm); PCG preconditioned by an FP64
F- or W-cycle: identical iteration count and status, relative solution difference <= 1e-10,
defect history within 1e-9 relative (floor 1e-14 of the initial defect).  Smoothing V(2,3) makes the peak smoothing odd (pre + post
= 5 sweeps), so the top level's solution ends in the other buffer of its pair and is copied back;
V(4,4) keeps it in place.  Each case runs with the small levels in the single-workgroup tail
kernel (segments of the cycle's steps) and with every level launched on its own.
"""
import numpy as np
import pytest

from feat3_amd import problems as pb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

CYC = {"f": po.CYCLE_F, "w": po.CYCLE_W}


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def torch(torch_cuda):
    return torch_cuda


def _defect(H, seed):
    d = np.random.default_rng(seed).uniform(-1, 1, H.n)
    d[H.levels[0].bnd] = 0.0
    return d


@pytest.mark.parametrize("tail", [True, False])
@pytest.mark.parametrize("cycle", ["f", "w"])
@pytest.mark.parametrize("pre,post", [(2, 3), (4, 4)])
@pytest.mark.parametrize("fmt", ["csr", "sell", "banded"])
def test_one_cycle_matches_oracle(torch, cycle, pre, post, fmt, tail):
    from feat3_amd.solver import MGPCG
    print(f"cycle {cycle} V({pre},{post}) {fmt} tail {tail}", flush=True)
    H = pb.hierarchy(2, 1, 6, 1)
    d = _defect(H, 4)
    for inner, code, tol in [("f64", 0, 1e-12), ("f32", 1, 1e-5), ("f16", 2, 1e-5)]:
        S = po.System.from_hierarchy(H)
        cor = np.zeros(H.n)
        S.vcycle(cor, d, po.params(inner=code, pre=pre, post=post, cycle=CYC[cycle]))
        s = MGPCG(H, inner=inner, fmt=fmt, pre=pre, post=post, cycle=cycle,
                  options=None if tail else dict(tail_max_n=0)).init()
        assert (s.info()["tail_levels"] > 0) == tail
        dd = torch.from_numpy(d).cuda()
        dc = torch.zeros_like(dd)
        s.vcycle(dc, dd)
        assert rel(dc.cpu().numpy(), cor) <= tol, inner
        s.done()


@pytest.mark.parametrize("tail", [True, False])
@pytest.mark.parametrize("cycle", ["f", "w"])
@pytest.mark.parametrize("name,dim,deg,lvl", [("2D Q1 L6", 2, 1, 6), ("2D Q2 L5", 2, 2, 5), ("3D Q1 L4", 3, 1, 4),
                                                ("2D Q1 L4, every level in the tail", 2, 1, 4)])
def test_pcg_with_cycle_matches_oracle(torch, cycle, name, dim, deg, lvl, tail):
    from feat3_amd.solver import MGPCG
    print(f"case {name} cycle {cycle} tail {tail}", flush=True)
    H = pb.hierarchy(dim, deg, lvl, 1)
    b = pb.filtered_rhs(H, 0)
    S = po.System.from_hierarchy(H)
    xo = np.zeros(H.n)
    so, io, ho = S.solve(xo, b.copy(), po.params(cycle=CYC[cycle]))
    s = MGPCG(H, cycle=cycle, options=None if tail else dict(tail_max_n=0)).init()
    x = torch.zeros(H.n, dtype=torch.float64, device="cuda")
    st, it = s.correct(x, torch.from_numpy(b).cuda())
    hg = s.history()
    info = s.info()
    s.done()
    assert so == 2 and st == 2
    assert it == io
    assert rel(x.cpu().numpy(), xo) <= 1e-10
    # the W-cycle reaches 1e-11 of the initial defect in two iterations; that deep the history
    # entries are rounding-level (observed 3e-19 apart at 1.5e-12), hence the floor 1e-14 * |r_0|
    assert np.allclose(hg[: io + 1], ho[: io + 1], rtol=1e-9, atol=1e-14 * ho[0])
    assert (info["tail_levels"] > 0) == tail



@pytest.mark.parametrize("tail", [True, False])
@pytest.mark.parametrize("cycle", ["f", "w"])
@pytest.mark.parametrize("fine,coarse", [(2, 2), (3, 2), (4, 2)])
def test_short_hierarchies(torch, cycle, fine, coarse, tail):
    """One level (the cycle is the exact coarse solve), two levels (F = V; the W-cycle's one peak
    is the top level itself) and three levels, FP64, against the oracle."""
    from feat3_amd.solver import MGPCG
    print(f"cycle {cycle} levels {fine}..{coarse} tail {tail}", flush=True)
    H = pb.hierarchy(2, 1, fine, coarse)
    d = _defect(H, fine)
    S = po.System.from_hierarchy(H)
    cor = np.zeros(H.n)
    S.vcycle(cor, d, po.params(pre=2, post=3, cycle=CYC[cycle]))
    s = MGPCG(H, pre=2, post=3, cycle=cycle, options=None if tail else dict(tail_max_n=0)).init()
    dd = torch.from_numpy(d).cuda()
    dc = torch.zeros_like(dd)
    s.vcycle(dc, dd)
    s.done()
    assert rel(dc.cpu().numpy(), cor) <= 1e-12
