"""lafem_hip_dot_noalloc_f64: the dot product without any allocation (block partials in a static
device array) against lafem_hip_dot_f64, which uses the initialised workspace.

Bar: the same kernels, grid and summation order, so the two results are equal bit for bit, on
two calls in a row (the second with a smaller input, which reads fewer partials than the first
left behind), for aligned and unaligned vectors and for n = 0; integer-valued inputs are also
checked against the exact sum.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def vp(t, offset=0):
    return ctypes.c_void_p(t.data_ptr() + 8 * offset)


@pytest.fixture(scope="module")
def T(torch_cuda):
    from feat3_amd import capi
    return torch_cuda, capi


def dots(torch, capi, x, y, n, offset=0):
    r0 = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    r1 = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    capi.call("lafem_hip_dot_f64", None, vp(x, offset), vp(y, offset), n, vp(r0))
    capi.call("lafem_hip_dot_noalloc_f64", None, vp(x, offset), vp(y, offset), n, vp(r1))
    torch.cuda.synchronize()
    return r0.item(), r1.item()


def test_noalloc_dot_equals_workspace_dot(T):
    torch, capi = T
    rng = np.random.default_rng(17)
    for n in (3_000_001, 777):          # a grid of 2048 blocks, then one of 4 blocks
        print(f"case noalloc dot n={n}", flush=True)
        xh, yh = rng.uniform(-1, 1, n + 1), rng.uniform(-1, 1, n + 1)
        x, y = torch.from_numpy(xh).cuda(), torch.from_numpy(yh).cuda()
        for off in (0, 1):               # 16-byte aligned (vector loads) and 8-byte aligned
            d0, d1 = dots(torch, capi, x, y, n, off)
            assert d1 == d0, (n, off, d0, d1)
            ref = float(np.dot(xh[off:off + n], yh[off:off + n]))
            assert abs(d1 - ref) <= 1e-13 * np.sqrt(n) * max(1.0, abs(ref))


def test_noalloc_dot_exact_on_integers_and_empty(T):
    torch, capi = T
    rng = np.random.default_rng(5)
    for n in (1 << 22, 1000, 1):
        xi, yi = rng.integers(-8, 8, n), rng.integers(-8, 8, n)
        x = torch.from_numpy(xi.astype(np.float64)).cuda()
        y = torch.from_numpy(yi.astype(np.float64)).cuda()
        d0, d1 = dots(torch, capi, x, y, n)
        assert d1 == d0 == float(np.dot(xi, yi))
    e = torch.zeros(1, dtype=torch.float64, device="cuda")
    d0, d1 = dots(torch, capi, e, e, 0)
    assert d1 == d0 == 0.0
