"""The differentiable batched diffusion solve (torch.autograd.Function over the batched device
MG-PCG and the HIP adjoint kernel) against
  * central finite differences (h = 1e-6) of the loss L = 1/2 sum_j |u_j - u_obs|^2 through
    forward solves, FP64 inner path at relative tolerance 1e-13 (bar: 1e-6 relative);
  * the FP64 CPU gradient of the oracle (oracle solves + the oracle's element restatement):
    FP64 inner path at tolerance 1e-12 (bar: 1e-8 relative), FP32-inner path at tolerance 1e-8
    (bar: 1e-6 relative, the solves' own accuracy);
  * the identity sum_e k_e dL/dk_e = -(u - u_obs)^T u (A linear in k, symmetric) for 16 fields;
  * the HIP element kernel alone against the oracle's element restatement (2D Q1, 2D Q2, 3D Q1).
"""
import numpy as np
import pytest

from feat3_amd import problems as pb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch(torch_cuda):
    return torch_cuda


def fields(dim, lvl, seeds):
    return np.exp(np.stack([pb.gaussian_random_field(dim, lvl, seed=s) for s in seeds]))


def u_obs_for(prob, seed):
    k = fields(prob.dim, prob.level, [seed])[0]
    H = pb.hierarchy(prob.dim, prob.degree, prob.level, prob.level_coarse, kcell=k)
    x = np.zeros(H.n)
    st, _, _ = po.System.from_hierarchy(H).solve(x, prob.b.copy(), po.params(tol_rel=1e-14, max_iter=400))
    assert st in (2, 5, 6)
    return x


def run(torch, prob, K, uobs):
    from feat3_amd.autograd import diffusion_solve
    kt = torch.from_numpy(K).cuda().requires_grad_(True)
    u = diffusion_solve(kt, prob)
    uo = torch.from_numpy(uobs).cuda()[:, None]
    L = 0.5 * ((u - uo) ** 2).sum()
    L.backward()
    return float(L.detach()), kt.grad.cpu().numpy(), u.detach().cpu().numpy()


def loss_only(torch, prob, K, uobs):
    from feat3_amd.autograd import diffusion_solve
    with torch.no_grad():
        u = diffusion_solve(torch.from_numpy(K).cuda(), prob)
        return float(0.5 * ((u - torch.from_numpy(uobs).cuda()[:, None]) ** 2).sum())


def test_gradient_vs_central_fd(torch):
    from feat3_amd.autograd import DiffusionProblem
    print("case autograd FD 2D Q1 L4 K 2 fp64", flush=True)
    prob = DiffusionProblem(dim=2, level=4, inner="f64", tol_rel=1e-13, max_iter=400)
    K = fields(2, 4, [11, 12])
    uobs = u_obs_for(prob, 99)
    L0, g, _ = run(torch, prob, K, uobs)
    h = 1e-6
    d = np.random.default_rng(2).standard_normal(K.shape)
    fd = (loss_only(torch, prob, K + h * d, uobs) - loss_only(torch, prob, K - h * d, uobs)) / (2 * h)
    print(f"  FD {fd:.12e} adjoint {np.sum(g * d):.12e}", flush=True)
    assert abs(fd - np.sum(g * d)) <= 1e-6 * abs(fd)
    for (j, c) in ((0, 0), (1, 37), (1, K.shape[1] - 1)):
        e = np.zeros(K.shape)
        e[j, c] = 1.0
        fdc = (loss_only(torch, prob, K + h * e, uobs) - loss_only(torch, prob, K - h * e, uobs)) / (2 * h)
        assert abs(fdc - g[j, c]) <= 1e-6 * max(abs(fdc), np.max(np.abs(g)))


def cpu_gradient(prob, k, uobs):
    H = pb.hierarchy(prob.dim, prob.degree, prob.level, prob.level_coarse, kcell=k)
    S = po.System.from_hierarchy(H)
    prm = po.params(tol_rel=1e-14, max_iter=400)
    u = np.zeros(H.n)
    S.solve(u, prob.b.copy(), prm)
    r = u - uobs
    r[H.levels[0].bnd] = 0.0
    lam = np.zeros(H.n)
    S.solve(lam, r, prm)
    return po.cell_gradient(prob.dim, prob.degree, prob.level, prob.elem, u, lam)


@pytest.mark.parametrize("inner,tol,bar", [("f64", 1e-12, 1e-8), ("f32", 1e-8, 1e-6)])
def test_gradient_vs_fp64_cpu_gradient(torch, inner, tol, bar):
    from feat3_amd.autograd import DiffusionProblem
    print(f"case autograd vs CPU gradient 2D Q1 L5 K 3 inner {inner}", flush=True)
    prob = DiffusionProblem(dim=2, level=5, inner=inner, tol_rel=tol)
    K = fields(2, 5, [21, 22, 23])
    uobs = u_obs_for(prob, 98)
    _, g, _ = run(torch, prob, K, uobs)
    for j in range(K.shape[0]):
        gc = cpu_gradient(prob, K[j], uobs)
        err = np.linalg.norm(g[j] - gc) / np.linalg.norm(gc)
        print(f"  field {j}: rel {err:.2e}", flush=True)
        assert err <= bar


def test_homogeneity_identity_16_fields(torch):
    from feat3_amd.autograd import DiffusionProblem
    print("case autograd identity 2D Q1 L7 K 16 inner f32", flush=True)
    prob = DiffusionProblem(dim=2, level=7, inner="f32", tol_rel=1e-8)
    K = fields(2, 7, range(16))
    uobs = u_obs_for(prob, 97)
    _, g, u = run(torch, prob, K, uobs)
    for j in range(16):
        lhs = K[j] @ g[j]
        rhs = -np.dot(u[:, j] - uobs, u[:, j])
        assert abs(lhs - rhs) <= 1e-6 * abs(rhs), (j, lhs, rhs)


@pytest.mark.parametrize("dim,deg,lvl", [(2, 1, 5), (2, 2, 4), (3, 1, 3)])
def test_element_kernel_matches_oracle(torch, dim, deg, lvl):
    from feat3_amd.autograd import DiffusionProblem, cell_gradient
    print(f"case element kernel dim {dim} degree {deg} level {lvl}", flush=True)
    prob = DiffusionProblem(dim=dim, degree=deg, level=lvl)
    rng = np.random.default_rng(dim * 10 + deg)
    K = 5
    U = rng.standard_normal((prob.nodes, K))
    Lm = rng.standard_normal((prob.nodes, K))
    g = cell_gradient(prob, torch.from_numpy(U).cuda(), torch.from_numpy(Lm).cuda()).cpu().numpy()
    for j in range(K):
        gc = po.cell_gradient(dim, deg, lvl, prob.elem, U[:, j].copy(), Lm[:, j].copy())
        assert np.allclose(g[:, j], gc, rtol=1e-13, atol=1e-13 * np.max(np.abs(gc)))


@pytest.mark.parametrize("dim,deg,lvl", [(2, 1, 5), (3, 1, 3), (2, 2, 4)])
def test_device_assembly_is_bitwise_host_assembly(torch, dim, deg, lvl):
    """feat3_assemble_laplace_fields / feat3_coarsen_cellfields reproduce the host setup's values
    bit for bit on every level (same cell order, no contraction of the products)."""
    from feat3_amd.autograd import DiffusionProblem
    print(f"case device assembly {dim}D Q{deg} L{lvl}", flush=True)
    prob = DiffusionProblem(dim=dim, degree=deg, level=lvl)
    K = fields(dim, lvl, [31, 32, 33])
    dev = prob.assemble(torch.from_numpy(K).cuda())
    Hs = pb.field_hierarchies(dim, lvl, K, 1, deg)
    for i in range(len(Hs[0].levels)):
        host = pb.stacked_values([H.levels[i].A for H in Hs]).reshape(-1)
        assert np.array_equal(dev[i].cpu().numpy(), host), i


def test_device_assembly_path_matches_host_path(torch):
    from feat3_amd.autograd import DiffusionProblem
    print("case device vs host assembly path 2D Q1 L6 K 4", flush=True)
    K = fields(2, 6, [41, 42, 43, 44])
    res = []
    for dev in (True, False):
        prob = DiffusionProblem(dim=2, level=6, inner="f32", device_assembly=dev)
        uobs = u_obs_for(prob, 96)
        res.append(run(torch, prob, K, uobs))
    (L1, g1, u1), (L2, g2, u2) = res
    assert np.array_equal(u1, u2) and np.array_equal(g1, g2) and L1 == L2
