"""Value-coded bands (option band_dict, mg::DictBandRows): a banded level matrix whose every
diagonal holds at most 16 distinct values is stored as 4-bit codes into per-diagonal tables.
The row sums read exactly the band's values and take the same terms in the same order, so a
V-cycle and a whole solve are bit for bit those of the value-stored band (band_dict 0, which
keeps the full or, for FP64, the symmetric band), in every inner precision.  A matrix with more
distinct values on some diagonal keeps the value-stored band and still matches the oracle.
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


def run(torch, H, b, inner, opts):
    from feat3_amd.solver import MGPCG
    db = torch.from_numpy(b).cuda()
    s = MGPCG(H, inner=inner, fmt="banded", options=opts).init()
    info = s.info()
    c = torch.zeros_like(db)
    s.vcycle(c, db)
    x = torch.zeros_like(db)
    st, it = s.correct(x, db)
    out = (info, c.cpu().numpy(), int(st), int(it), x.cpu().numpy())
    s.done()
    return out


@pytest.mark.parametrize("dim,deg,lvl", [(3, 1, 5), (2, 1, 6), (3, 1, 6)])
@pytest.mark.parametrize("inner", ["f64", "f32", "f16"])
def test_value_coded_band_is_bitwise_the_value_band(torch, dim, deg, lvl, inner):
    print(f"case band_dict {dim}D Q{deg} L{lvl} {inner}", flush=True)
    H = pb.hierarchy(dim, deg, lvl, 1)
    b = pb.filtered_rhs(H, 0)
    i1, c1, st1, it1, x1 = run(torch, H, b, inner, {"band_dict": 1})
    i0, c0, st0, it0, x0 = run(torch, H, b, inner, {"band_dict": 0})
    assert i1["banded"] == 3, i1
    assert i0["banded"] in (1, 2), i0
    assert i1["ndiag"] == i0["ndiag"]
    assert np.array_equal(c1, c0)
    assert st1 == st0 == 2 and it1 == it0
    assert np.array_equal(x1, x0)


def test_many_distinct_values_keep_the_value_band(torch):
    """A finest matrix with a distinct diagonal entry per row (A + diag(t), t > 0 random): the
    main diagonal has more than 16 values, so the finest level stays value-stored; the solve
    matches the FP64 oracle (identical iterations, 1e-10)."""
    print("case band_dict fallback 3D Q1 L5", flush=True)
    H = pb.hierarchy(3, 1, 5, 1)
    A = H.levels[0].A
    t = np.random.default_rng(3).uniform(0.0, 0.05, A.rows)
    rows = np.repeat(np.arange(A.rows), np.diff(A.row_ptr))
    on_diag = A.col_ind == rows
    A.val[on_diag] += t[rows[on_diag]] * A.val[on_diag]
    b = pb.filtered_rhs(H, 0)
    info, _, st, it, x = run(torch, H, b, "f64", {"band_dict": 1})
    assert info["banded"] in (1, 2), info
    S = po.System.from_hierarchy(H)
    xo = np.zeros(H.n)
    so, io, _ = S.solve(xo, b, po.params())
    assert st == so == 2 and it == io
    assert rel(x, xo) <= 1e-10


@pytest.mark.parametrize("lvl", [5, 6])
@pytest.mark.parametrize("inner", ["f64", "f32"])
def test_lane_shifted_runs_are_bitwise_the_loaded_ones(torch, lvl, inner):
    """FP64 value-coded rows of the 3D Q1 stencil (27 diagonals in nine runs of three) take a
    run's outer x from the neighbouring lanes (option lane_shift, default 1): the same x values,
    so the V-cycle and the solve (the FP64 outer product, and with FP64 inner arithmetic every
    level) are bit for bit those of three loads per run (lane_shift 0)."""
    print(f"case lane shift 3D Q1 L{lvl} {inner}", flush=True)
    H = pb.hierarchy(3, 1, lvl, 1)
    b = pb.filtered_rhs(H, 0)
    i1, c1, st1, it1, x1 = run(torch, H, b, inner, {"lane_shift": 1})
    i0, c0, st0, it0, x0 = run(torch, H, b, inner, {"lane_shift": 0})
    assert i1["banded"] == 3 and i1 == i0
    assert np.array_equal(c1, c0)
    assert st1 == st0 == 2 and it1 == it0
    assert np.array_equal(x1, x0)
