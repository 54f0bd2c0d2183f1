"""The launch-shape autotuner on the GPU: it times real configurations, returns the fastest of
its trials, and the tuned solver (and every rows-per-block setting) keeps the FP64 parity with the
oracle: same iteration count, solution within 1e-10."""
import json

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
    s.done()
    return st, it, x.cpu().numpy()


def test_tuner_picks_fastest_trial_and_tuned_solver_matches_oracle(torch, tmp_path):
    from feat3_amd import tune
    print("case tune 2D Q1 L7 f64", flush=True)
    H = pb.hierarchy(2, 1, 7, 1)
    cache = str(tmp_path / "tune.json")
    res = tune.tune_mgpcg(H, inner="f64", reps=3, cache=cache,
                          formats=[dict(fmt="csr", lanes=8), dict(fmt="sell", sell_C=64, sell_sigma=1),
                                   dict(fmt="sell", sell_C=128, sell_sigma=64)],
                          rows_per_block=(0, 256, 1024), transfer_rows_per_block=(0, 512))
    ms = [t[1] for t in res["trials"]]
    assert len(ms) == 7 and all(m > 0 for m in ms)
    assert res["ms"] == min(ms)
    with open(cache) as f:
        assert len(json.load(f)) == 1
    again = tune.tune_mgpcg(H, inner="f64", cache=cache)      # served from the cache
    assert again["config"] == res["config"]
    b = pb.filtered_rhs(H, 0)
    xo = np.zeros(H.n)
    so, io, _ = po.System.from_hierarchy(H).solve(xo, b, po.params())
    st, it, x = solve(torch, H, b, inner="f64", **tune.solver_kwargs(res))
    assert st == so == 2 and it == io
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)


@pytest.mark.parametrize("fmt", ["csr", "sell"])
@pytest.mark.parametrize("rpb", [64, 192, 4096, 10 ** 7])
def test_rows_per_block_keeps_parity(torch, fmt, rpb):
    print(f"case rows_per_block {rpb} {fmt} 3D Q1 L4", flush=True)
    H = pb.hierarchy(3, 1, 4, 1)
    b = pb.filtered_rhs(H, 0)
    xo = np.zeros(H.n)
    so, io, _ = po.System.from_hierarchy(H).solve(xo, b, po.params())
    st, it, x = solve(torch, H, b, inner="f64", fmt=fmt, options={"rows_per_block": rpb})
    assert st == so == 2 and it == io
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)
