"""The device solver on hierarchies in the reference's own DOF numbering (StandardRefinery +
Lagrange DOF mapping), with the lexicographic lattice order as the solver's internal level order
(feat3_mgpcg_set_level_order), so that the level matrices are banded.

Bars: FP64 path against the oracle on the same FEAT-numbered hierarchy: identical iterations,
relative solution difference <= 1e-10; FP32 / FP16-stored V-cycle: iterations +-1, solution
within 1e-8 of the FP64 oracle.  Renumbering moves values without arithmetic, so the FEAT-numbered
solve must equal the lexicographic solve bit for bit after the permutation, for correct, apply
and the V-cycle alone, from host and from device-resident levels.
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


def gpu_solve(torch, H, b, **kw):
    from feat3_amd.solver import MGPCG
    s = MGPCG(H, **kw).init()
    x = torch.zeros(H.n if hasattr(H, "n") else H[0].A.rows, dtype=torch.float64, device="cuda")
    st, it = s.correct(x, torch.from_numpy(b).cuda())
    info = s.info()
    s.done()
    return st, it, x.cpu().numpy(), info


CASES = [("2D Q1 L6", 2, 1, 6), ("2D Q2 L5", 2, 2, 5), ("3D Q1 L5", 3, 1, 5)]


@pytest.mark.parametrize("name,dim,deg,lvl", CASES)
def test_feat_numbered_solve_matches_oracle(torch, name, dim, deg, lvl):
    print(f"case {name}", flush=True)
    H = pb.hierarchy(dim, deg, lvl, 1, numbering="feat")
    b = pb.filtered_rhs(H, 0)
    S = po.System.from_hierarchy(H)
    xo = np.zeros(H.n)
    so, io, _ = S.solve(xo, b, po.params())
    assert so == 2
    for inner in ("f64", "f32", "f16"):
        st, it, xg, info = gpu_solve(torch, H, b, inner=inner, fmt="banded")
        assert st == 2, inner
        if deg == 1:
            assert info["banded"] > 0, "the lattice order must give the banded level format"
        if inner == "f64":
            assert it == io
            assert rel(xg, xo) <= 1e-10
        else:
            assert abs(it - io) <= 1
            assert rel(xg, xo) <= 1e-8


@pytest.mark.parametrize("name,dim,deg,lvl", CASES)
def test_feat_numbered_solve_is_permuted_lex_solve(torch, name, dim, deg, lvl):
    from feat3_amd.solver import MGPCG
    print(f"case {name}", flush=True)
    G = pb.hierarchy(dim, deg, lvl, 1, numbering="feat")
    H = pb.hierarchy(dim, deg, lvl, 1)
    o = G.orders[0]
    bg, bh = pb.filtered_rhs(G, 0), pb.filtered_rhs(H, 0)
    assert np.array_equal(bg, bh[o])
    for inner in ("f64", "f32"):
        _, ig, xg, _ = gpu_solve(torch, G, bg, inner=inner, fmt="banded")
        _, ih, xh, _ = gpu_solve(torch, H, bh, inner=inner, fmt="banded")
        assert ig == ih and np.array_equal(xg, xh[o]), inner
    # apply (zero start) and the V-cycle alone, same comparison
    rng = np.random.default_rng(3)
    d = rng.uniform(-1, 1, H.n)
    d[H.levels[0].bnd] = 0.0
    sg = MGPCG(G, inner="f32", fmt="banded").init()
    sh = MGPCG(H, inner="f32", fmt="banded").init()
    dg, dh = torch.from_numpy(d[o].copy()).cuda(), torch.from_numpy(d).cuda()
    cg, ch = torch.zeros_like(dg), torch.zeros_like(dh)
    sg.apply(cg, dg)
    sh.apply(ch, dh)
    assert np.array_equal(cg.cpu().numpy(), ch.cpu().numpy()[o])
    sg.vcycle(cg, dg)
    sh.vcycle(ch, dh)
    assert np.array_equal(cg.cpu().numpy(), ch.cpu().numpy()[o])
    sg.done()
    sh.done()


def test_nonzero_start_in_feat_numbering(torch):
    """correct() from a non-zero x: x is renumbered at entry as b is."""
    G = pb.hierarchy(3, 1, 4, 1, numbering="feat")
    H = pb.hierarchy(3, 1, 4, 1)
    o = G.orders[0]
    rng = np.random.default_rng(5)
    x0 = rng.uniform(-1, 1, H.n)
    from feat3_amd.solver import MGPCG
    out = []
    for hh, x in ((G, x0[o]), (H, x0)):
        s = MGPCG(hh, inner="f64", fmt="banded").init()
        dx = torch.from_numpy(x.copy()).cuda()
        st, it = s.correct(dx, torch.from_numpy(pb.filtered_rhs(hh, 0)).cuda())
        out.append((st, it, dx.cpu().numpy()))
        s.done()
    assert out[0][:2] == out[1][:2]
    assert np.array_equal(out[0][2], out[1][2][o])


@pytest.mark.parametrize("dim,deg,lvl", [(3, 1, 4), (2, 2, 4)])
def test_device_renumbering_equals_host(torch, dim, deg, lvl):
    """feat3_device_permute_csr == feat3_setup_permute_csr bit for bit, and the device-assembled
    FEAT-numbered hierarchy equals the host one."""
    G = pb.hierarchy(dim, deg, lvl, 1, numbering="feat")
    D = pb.hierarchy_device(dim, deg, lvl, 1, numbering="feat")
    Dh = pb.device_levels_to_host(dim, deg, lvl, 1, D)
    for L, M in zip(G.levels, Dh.levels):
        for a, b in ((L.A, M.A), (L.P, M.P), (L.R, M.R)):
            if a is None:
                assert b is None
                continue
            assert np.array_equal(a.row_ptr, b.row_ptr)
            assert np.array_equal(a.col_ind, b.col_ind)
            assert np.array_equal(a.val, b.val)
        assert np.array_equal(L.bnd, M.bnd)


def test_device_levels_with_orders_equal_host_levels(torch):
    """A solver initialised from device-resident FEAT-numbered levels (renumbered on the device at
    init) solves bit for bit as the one from host levels."""
    dim, deg, lvl = 3, 1, 6      # the finest levels exceed the host-copy threshold of push_level_device
    G = pb.hierarchy(dim, deg, lvl, 1, numbering="feat")
    D = pb.hierarchy_device(dim, deg, lvl, 1, numbering="feat")
    b = pb.filtered_rhs(G, 0)
    _, ih, xh, infh = gpu_solve(torch, G, b, inner="f32", fmt="banded")
    _, idv, xd, infd = gpu_solve(torch, D, b, inner="f32", fmt="banded")
    assert infd["banded"] > 0 and infh["banded"] > 0
    assert ih == idv and np.array_equal(xh, xd)


def test_set_level_order_rejects_bad_orders(torch):
    import ctypes
    from feat3_amd import capi
    from feat3_amd.solver import MGPCG
    H = pb.hierarchy(2, 1, 3, 1)
    s = MGPCG(H, level_orders=None)
    n = H.n
    bad = np.arange(n, dtype=np.uint32)
    bad[1] = 0          # repeated index
    with pytest.raises(capi.Feat3HipError):
        capi.call("feat3_mgpcg_set_level_order", s._h, 0, ctypes.c_void_p(bad.ctypes.data))
    with pytest.raises(capi.Feat3HipError):
        capi.call("feat3_mgpcg_set_level_order", s._h, 7, ctypes.c_void_p(bad.ctypes.data))
    ok = np.arange(n, dtype=np.uint32)[::-1].copy()
    capi.call("feat3_mgpcg_set_level_order", s._h, 0, ctypes.c_void_p(ok.ctypes.data))
    with pytest.raises(capi.Feat3HipError):   # once per level
        capi.call("feat3_mgpcg_set_level_order", s._h, 0, ctypes.c_void_p(ok.ctypes.data))
    s.init()
    s.done()
