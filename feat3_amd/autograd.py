"""Differentiable batched diffusion solve: -div(k grad u) = f on the unit square/cube with
u = 0 on the boundary, one solution per coefficient field, as a torch.autograd.Function.

forward:  k (K, cells) FP64 cuda tensor of per-cell coefficients (cells x fastest) ->
          u (nodes, K): K systems A(k_j) u_j = b solved together by the batched device MG-PCG
          (one multigrid hierarchy per field, coarse coefficients = mean of the children).  The
          level matrices are assembled on the GPU from k (feat3_assemble_laplace_fields) and
          loaded into one persistent solver object (feat3_batch_set_values); the host assembly
          path stays available (device_assembly=False).
backward: dL/du (nodes, K) -> the adjoint solves A(k_j) lambda_j = filter(dL/du_j) on the same
          solver object (A is symmetric), then the HIP kernel dL/dk_e = -lambda_e^T E u_e
          (E = h^(d-2) K_ref, the element stiffness of one cell).

The reference has no coefficient operator and no adjoint; its Poisson solver tree and
element assembly (assembly/common_operators.hpp:36-97) are what A(k) generalises.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import capi
from . import problems as pb
from .batch import BatchMGPCG
from .solver import status_success


@dataclass
class DiffusionProblem:
    dim: int = 2
    level: int = 10
    degree: int = 1
    level_coarse: int = 1
    inner: str = "f32"
    tol_rel: float = 1e-8
    max_iter: int = 1000
    load: int = 1             # 1: f = 1, 0: sine bubble
    device_assembly: bool = True   # matrices of k assembled on the GPU (False: on the host)

    def __post_init__(self):
        H = pb.hierarchy(self.dim, self.degree, self.level, self.level_coarse)
        self.H = H
        self.nodes = H.n
        self.bnd = H.levels[0].bnd
        self.b = pb.filtered_rhs(H, self.load)
        self.cells = (1 << self.level) ** self.dim
        h = 1.0 / (1 << self.level)
        self.ref = np.ascontiguousarray(pb.ref_stiffness(self.dim, self.degree))
        self.elem = np.ascontiguousarray(self.ref * h ** (self.dim - 2))
        self.last_status = None
        self.last_iters = None
        self._solver = None
        self._graphs = None

    def make_solver(self, kfields: np.ndarray) -> BatchMGPCG:
        """A solver for the fields assembled on the host (the reference's assembly step)."""
        Hs = pb.field_hierarchies(self.dim, self.level, kfields, self.level_coarse, self.degree)
        vals = [pb.stacked_values([H.levels[i].A for H in Hs]) for i in range(len(Hs[0].levels))]
        return BatchMGPCG(Hs[0], len(Hs), values=vals, shared=False, inner=self.inner,
                          tol_rel=self.tol_rel, max_iter=self.max_iter).init()

    def assemble(self, k):
        """Level values [entry][field] of A(k) on the device for every level (finest first):
        the fields are coarsened level by level (mean of the children) and each level assembled by
        feat3_assemble_laplace_fields, bit-identical to the host assembly."""
        import torch
        K = k.shape[0]
        if self._graphs is None:
            self._graphs = [(torch.from_numpy(L.A.row_ptr.view(np.int32)).cuda(),
                             torch.from_numpy(L.A.col_ind.view(np.int32)).cuda()) for L in self.H.levels]
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        kl = k.detach().contiguous()
        out = []
        for i, L in enumerate(self.H.levels):
            lvl = self.level - i
            if i > 0:
                kc = torch.empty((K, (1 << lvl) ** self.dim), dtype=torch.float64, device=k.device)
                capi.call("feat3_coarsen_cellfields", stream, self.dim, lvl + 1, K, ctypes.c_void_p(kl.data_ptr()),
                          ctypes.c_void_p(kc.data_ptr()))
                kl = kc
            val = torch.empty(L.A.nnz * K, dtype=torch.float64, device=k.device)
            rp, ci = self._graphs[i]
            capi.call("feat3_assemble_laplace_fields", stream, self.dim, self.degree, lvl, K,
                      ctypes.c_void_p(self.ref.ctypes.data), ctypes.c_void_p(kl.data_ptr()),
                      ctypes.c_void_p(rp.data_ptr()), ctypes.c_void_p(ci.data_ptr()), L.A.rows,
                      ctypes.c_void_p(val.data_ptr()))
            out.append(val)
        return out

    def rhs_for(self, K: int, device):
        """The filtered load for K systems, [node][system], kept on the device."""
        import torch
        key = (K, str(device))
        if getattr(self, "_rhs_key", None) != key:
            self._rhs = torch.from_numpy(np.ascontiguousarray(np.repeat(self.b[:, None], K, axis=1))).to(device)
            self._rhs_key = key
        return self._rhs

    def bnd_index(self, device):
        import torch
        if getattr(self, "_bnd_dev", None) is None or self._bnd_dev.device != torch.device(device):
            self._bnd_dev = torch.from_numpy(self.bnd.astype(np.int64)).to(device)
        return self._bnd_dev

    def solver_for(self, k) -> BatchMGPCG:
        """The batched solver holding A(k) for the (K, cells) cuda tensor k."""
        if not self.device_assembly:
            return self.make_solver(k.detach().cpu().numpy())
        import torch
        K = k.shape[0]
        if self._solver is None or self._solver.nsys != K:
            vals = [np.ascontiguousarray(np.repeat(L.A.val[:, None], K, axis=1)) for L in self.H.levels]
            self._solver = BatchMGPCG(self.H, K, values=vals, shared=False, inner=self.inner,
                                      tol_rel=self.tol_rel, max_iter=self.max_iter).init()
        vals = self.assemble(k)
        torch.cuda.current_stream().synchronize()
        for i, v in enumerate(vals):
            capi.call("feat3_batch_set_values", self._solver._h, i, ctypes.c_void_p(v.data_ptr()))
        return self._solver


def _solve(torch, prob: DiffusionProblem, solver: BatchMGPCG, rhs):
    x = torch.zeros_like(rhs)
    st, it = solver.correct(x, rhs, stream=torch.cuda.current_stream().cuda_stream)
    prob.last_status, prob.last_iters = st, it
    if not all(status_success(int(s)) for s in st):
        raise RuntimeError(f"batched MG-PCG did not converge: statuses {st.tolist()}")
    return x


def _make_function():
    import torch

    class DiffusionSolve(torch.autograd.Function):
        @staticmethod
        def forward(ctx, k, prob: DiffusionProblem):
            if k.dim() != 2 or k.shape[1] != prob.cells or k.dtype != torch.float64 or not k.is_cuda:
                raise ValueError("k must be a (K, cells) float64 cuda tensor")
            solver = prob.solver_for(k)
            rhs = prob.rhs_for(k.shape[0], k.device)
            u = _solve(torch, prob, solver, rhs)
            ctx.prob, ctx.solver = prob, solver
            ctx.save_for_backward(u, k.detach())
            return u

        @staticmethod
        def backward(ctx, gu):
            u, k = ctx.saved_tensors
            prob, solver = ctx.prob, ctx.solver
            if prob.device_assembly:
                solver = prob.solver_for(k)   # the shared solver may hold another k by now
            rhs = gu.detach().to(torch.float64).contiguous().clone()
            rhs[prob.bnd_index(rhs.device)] = 0.0
            lam = _solve(torch, prob, solver, rhs)
            return cell_gradient(prob, u, lam).t().contiguous(), None

    return DiffusionSolve


_FN = None


def diffusion_solve(k, prob: DiffusionProblem):
    """u = A(k)^-1 b for every row of k; differentiable in k."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(k, prob)


def cell_gradient(prob: DiffusionProblem, u, lam, stream: Optional[int] = None):
    """dL/dk_e = -lambda_e^T E u_e for u, lambda (nodes, K) FP64 cuda tensors -> (cells, K)."""
    import torch
    if u.shape != lam.shape or u.shape[0] != prob.nodes or u.dtype != torch.float64 or lam.dtype != torch.float64:
        raise ValueError("u and lam must be (nodes, K) float64 tensors")
    uc, lc = u.contiguous(), lam.contiguous()
    K = u.shape[1]
    g = torch.empty((prob.cells, K), dtype=torch.float64, device=u.device)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream if stream is None else stream)
    capi.call("feat3_adjoint_cell_gradient", s, prob.dim, prob.degree, prob.level, K,
              ctypes.c_void_p(prob.elem.ctypes.data), ctypes.c_void_p(uc.data_ptr()),
              ctypes.c_void_p(lc.data_ptr()), ctypes.c_void_p(g.data_ptr()))
    return g
