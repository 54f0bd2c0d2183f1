"""CPU pinning of the oracle's adjoint gradient dL/dk_e = -lambda_e^T E u_e (no reference
counterpart) for L(k) = 1/2 |u(k) - u_obs|^2, A(k) u = b, by
  * central finite differences (h = 1e-6) of L through two FP64 oracle solves at tolerance 1e-14,
    along a random direction and along single cells,
  * the identity sum_e k_e dL/dk_e = -lambda^T A u = -(u - u_obs)^T u (A is linear in k and
    symmetric), which a dropped term, a sign error or a wrong local node order breaks.
"""
import numpy as np
import pytest

from feat3_amd import problems as pb
from oracle import pyoracle as po

TOL = dict(tol_rel=1e-14, max_iter=400)


def solve(H, b):
    x = np.zeros(H.n)
    st, it, _ = po.System.from_hierarchy(H).solve(x, b, po.params(**TOL))
    assert st in (2, 5, 6)
    return x


def loss_and_grad(dim, deg, lvl, k, uobs, b):
    H = pb.hierarchy(dim, deg, lvl, 1, kcell=k)
    u = solve(H, b)
    r = u - uobs
    r[H.levels[0].bnd] = 0.0
    lam = solve(H, r)
    E = pb.ref_stiffness(dim, deg) * (1.0 / (1 << lvl)) ** (dim - 2)
    g = po.cell_gradient(dim, deg, lvl, E, u, lam)
    return 0.5 * np.dot(u - uobs, u - uobs), g, u, lam


def loss(dim, deg, lvl, k, uobs, b):
    u = solve(pb.hierarchy(dim, deg, lvl, 1, kcell=k), b)
    return 0.5 * np.dot(u - uobs, u - uobs)


@pytest.mark.parametrize("dim,deg,lvl", [(2, 1, 4), (2, 2, 3), (3, 1, 3)])
def test_adjoint_gradient_vs_central_fd(dim, deg, lvl):
    k = np.exp(pb.gaussian_random_field(dim, lvl, seed=3))
    kobs = np.exp(pb.gaussian_random_field(dim, lvl, seed=4))
    H = pb.hierarchy(dim, deg, lvl, 1, kcell=k)
    b = pb.filtered_rhs(H, 1)
    uobs = solve(pb.hierarchy(dim, deg, lvl, 1, kcell=kobs), b)
    L0, g, u, lam = loss_and_grad(dim, deg, lvl, k, uobs, b)
    h = 1e-6
    d = np.random.default_rng(0).standard_normal(k.size)
    fd = (loss(dim, deg, lvl, k + h * d, uobs, b) - loss(dim, deg, lvl, k - h * d, uobs, b)) / (2 * h)
    assert abs(fd - g @ d) <= 1e-6 * abs(fd)
    for c in (0, k.size // 2 + 3, k.size - 1):
        e = np.zeros(k.size)
        e[c] = 1.0
        fdc = (loss(dim, deg, lvl, k + h * e, uobs, b) - loss(dim, deg, lvl, k - h * e, uobs, b)) / (2 * h)
        assert abs(fdc - g[c]) <= 1e-6 * max(abs(fdc), np.max(np.abs(g)))
    # homogeneity of A in k
    lhs = k @ g
    rhs = -np.dot(u - uobs, u)
    assert abs(lhs - rhs) <= 1e-10 * abs(rhs)


def test_planted_sign_error_is_caught():
    """A gradient with the wrong sign (planted in a copy of the oracle's result) fails the
    finite-difference comparison used above."""
    dim, deg, lvl = 2, 1, 3
    k = np.exp(pb.gaussian_random_field(dim, lvl, seed=5))
    H = pb.hierarchy(dim, deg, lvl, 1, kcell=k)
    b = pb.filtered_rhs(H, 1)
    uobs = 0.5 * solve(H, b)
    _, g, _, _ = loss_and_grad(dim, deg, lvl, k, uobs, b)
    d = np.random.default_rng(1).standard_normal(k.size)
    h = 1e-6
    fd = (loss(dim, deg, lvl, k + h * d, uobs, b) - loss(dim, deg, lvl, k - h * d, uobs, b)) / (2 * h)
    assert abs(fd - g @ d) <= 1e-6 * abs(fd)
    assert not abs(fd - (-g) @ d) <= 1e-6 * abs(fd)
