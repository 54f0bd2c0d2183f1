"""Every benchmark configuration at its stated size, against the oracle (the reference's FP64 CPU
path restated) on the same system.

  config 1  2D Q1 L6 (4,225 DOF)           -- test_gpu_mgpcg / smoke()
  config 2  2D Q2 L9 (1,050,625 DOF)       FEAT numbering, FP64 / FP32 / FP16-stored V-cycle
  config 3  3D Q1 L7 (2,146,689 DOF)       FEAT numbering + lattice order (banded), FP64 / FP32 / FP16
  config 4  2D Q1 L10, 16 coefficient fields: adjoint gradient vs the FP64 CPU gradient (2 fields),
            central finite differences (one random direction), the identity for all 16 fields
  config 5  3D Q1 L6 vector Laplacian (3x3 BCSR), 32 right-hand sides: 4 systems against oracle
            solves, all 32 by their final filtered relative defect

Bars: FP64 path identical iterations and relative solution difference <= 1e-10; FP32 / FP16-stored
V-cycle iterations +-1 of the oracle's FP64 count and <= 1e-8 from the FP64 oracle solution;
gradients: 1e-8 vs the CPU gradient, 1e-6 for the identity, and vs central finite differences
1e-5 at h = 1e-6 (the solves' noise floor at 1M DOF, measured 2.7e-6) and 1e-6 at h = 1e-5.
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


def solve_gpu(torch, H, b, **kw):
    from feat3_amd.solver import MGPCG
    s = MGPCG(H, **kw).init()
    x = torch.zeros(H.n, dtype=torch.float64, device="cuda")
    st, it = s.correct(x, torch.from_numpy(b).cuda())
    info = s.info()
    s.done()
    return st, it, x.cpu().numpy(), info


@pytest.mark.parametrize("cfg", [2, 3])
def test_scalar_config_full_size(torch, cfg):
    c = pb.CONFIGS[cfg]
    print(f"case config {cfg}: {c['name']}, FEAT numbering", flush=True)
    H = pb.config_hierarchy(cfg, numbering="feat")
    b = pb.filtered_rhs(H, 0)
    S = po.System.from_hierarchy(H)
    xo = np.zeros(H.n)
    so, io, _ = S.solve(xo, b, po.params())
    assert so == 2
    print(f"  oracle: {io} iterations", flush=True)
    for inner in ("f64", "f32", "f16"):
        st, it, xg, info = solve_gpu(torch, H, b, inner=inner, fmt="banded")
        err = rel(xg, xo)
        print(f"  {inner}: status {st}, {it} iterations, rel. diff {err:.2e}, banded {info['banded']}", flush=True)
        assert st == 2
        if c["dim"] == 3:
            assert info["banded"] > 0
        if inner == "f64":
            assert it == io and err <= 1e-10
        else:
            assert abs(it - io) <= 1 and err <= 1e-8


def test_config5_vector_laplacian_32_rhs(torch):
    import scipy.sparse as sp
    from feat3_amd.batch import BatchMGPCG
    print("case config 5: 3D Q1 L6 vector Laplacian, 32 right-hand sides", flush=True)
    K = 32
    H = pb.hierarchy(3, 1, 6, 1)
    vals = [pb.block_identity_values(L.A, 3) for L in H.levels]
    B = pb.random_rhs(H.n, 3, K, H.levels[0].bnd, seed=5)
    A = H.levels[0].A
    As = sp.csr_matrix((A.val, A.col_ind, A.row_ptr), shape=(A.rows, A.cols))
    bnd = H.levels[0].bnd
    picks = (0, 7, 19, 31)
    oracle = {}
    for k in picks:
        b = B[:, :, k].reshape(-1).copy()
        xo = np.zeros(b.size)
        so, io, _ = po.System.from_hierarchy(H, block=3, values=vals).solve(xo, b, po.params())
        assert so == 2
        oracle[k] = (io, xo)
    for inner in ("f64", "f32"):
        s = BatchMGPCG(H, K, values=vals, block=3, inner=inner).init()
        X = torch.zeros(s.shape, dtype=torch.float64, device="cuda")
        st, it = s.correct(X, torch.from_numpy(B).cuda())
        x = X.cpu().numpy()
        s.done()
        assert all(int(v) == 2 for v in st), st
        # every system: the filtered defect of the block system (a_ij I blocks: A per component)
        for k in range(K):
            r = B[:, :, k] - As @ x[:, :, k]
            r[bnd] = 0.0
            assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(B[:, :, k]) * 1.0001, (inner, k)
        # four systems against independent oracle solves of the BCSR system
        for k in picks:
            io, xo = oracle[k]
            xg = x[:, :, k].reshape(-1)
            print(f"  {inner} system {k}: gpu {int(it[k])} it, oracle {io} it, rel {rel(xg, xo):.2e}", flush=True)
            if inner == "f64":
                assert int(it[k]) == io and rel(xg, xo) <= 1e-10
            else:
                assert abs(int(it[k]) - io) <= 1 and rel(xg, xo) <= 1e-8


def test_config4_gradient_16_fields(torch):
    from feat3_amd.autograd import DiffusionProblem, diffusion_solve
    print("case config 4: 2D Q1 L10, 16 coefficient fields, adjoint gradient", flush=True)
    lvl = 10
    prob = DiffusionProblem(dim=2, level=lvl, inner="f64", tol_rel=1e-13, max_iter=400)
    K = np.exp(np.stack([pb.gaussian_random_field(2, lvl, seed=s) for s in range(16)]))
    kobs = np.exp(pb.gaussian_random_field(2, lvl, seed=97))
    Hobs = pb.hierarchy(2, 1, lvl, 1, kcell=kobs)
    uobs = np.zeros(Hobs.n)
    st, _, _ = po.System.from_hierarchy(Hobs).solve(uobs, prob.b.copy(), po.params(tol_rel=1e-13, max_iter=400))
    assert st == 2
    kt = torch.from_numpy(K).cuda().requires_grad_(True)
    u = diffusion_solve(kt, prob)
    uo = torch.from_numpy(uobs).cuda()[:, None]
    L = 0.5 * ((u - uo) ** 2).sum()
    L.backward()
    g = kt.grad.cpu().numpy()
    un = u.detach().cpu().numpy()
    # identity sum_e k_e dL/dk_e = -(u - u_obs)^T u for every field (A linear in k)
    for j in range(16):
        lhs, rhs = K[j] @ g[j], -np.dot(un[:, j] - uobs, un[:, j])
        assert abs(lhs - rhs) <= 1e-6 * abs(rhs), (j, lhs, rhs)
    # the FP64 CPU gradient (oracle solves + element restatement) of two fields
    for j in (0, 11):
        H = pb.hierarchy(2, 1, lvl, 1, kcell=K[j])
        S = po.System.from_hierarchy(H)
        prm = po.params(tol_rel=1e-13, max_iter=400)
        uc = np.zeros(H.n)
        S.solve(uc, prob.b.copy(), prm)
        r = uc - uobs
        r[H.levels[0].bnd] = 0.0
        lam = np.zeros(H.n)
        S.solve(lam, r, prm)
        gc = po.cell_gradient(2, 1, lvl, prob.elem, uc, lam)
        err = np.linalg.norm(g[j] - gc) / np.linalg.norm(gc)
        print(f"  field {j}: adjoint vs CPU gradient rel {err:.2e}", flush=True)
        assert err <= 1e-8
    # central finite differences along one random direction of all 16 fields
    h = 1e-6
    d = np.random.default_rng(4).standard_normal(K.shape)

    def loss(Kp):
        with torch.no_grad():
            up = diffusion_solve(torch.from_numpy(Kp).cuda(), prob)
            return float(0.5 * ((up - uo) ** 2).sum())
    fd = (loss(K + h * d) - loss(K - h * d)) / (2 * h)
    ad = float(np.sum(g * d))
    # at 1M DOF the quotient's noise floor is the solves' own error (relative residual 1e-13 on a
    # matrix of condition ~1e6) divided by 2h; the h = 1e-5 quotient has a tenth of that noise and
    # still a negligible O(h^2) truncation term, so it must come closer to the adjoint value
    fd5 = (loss(K + 1e-5 * d) - loss(K - 1e-5 * d)) / 2e-5
    print(f"  FD(h=1e-6) {fd:.12e} FD(h=1e-5) {fd5:.12e} adjoint {ad:.12e}", flush=True)
    assert abs(fd - ad) <= 1e-5 * abs(fd)
    assert abs(fd5 - ad) <= 1e-6 * abs(fd5)
