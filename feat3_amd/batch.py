"""Batched device MG-PCG (feat3_batch_* C ABI): K systems on one sparsity graph in lockstep.

Two uses:
  * K right-hand sides of one matrix (``values`` shared), e.g. 32 right-hand sides of the 3x3
    block vector Laplacian (the reference's SparseMatrixBCSR<3,3>, applied once per right-hand
    side there, sparse_matrix_bcsr.hpp:1295-1600);
  * K coefficient fields on one mesh (``values`` per system, scalar blocks).
Every system runs the solver of ``solver.MGPCG`` (PCG(MG-V(Richardson(Jacobi))), pcg.hpp:188-288)
with its own scalars, status and defect history.  Vectors are FP64 device arrays laid out
``[node][component][system]``.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional

import numpy as np

from . import capi
from .problems import Hierarchy
from .solver import INNER, _ptr


class BatchMGPCG:
    """values: one array per level of the hierarchy (finest first): shape (nnz, B, B) for
    shared values (B = block) or (nnz, K) for per-system scalar values.  ``None`` uses the
    hierarchy's scalar values as shared values."""

    def __init__(self, hierarchy: Hierarchy, nsys: int, values: Optional[List[np.ndarray]] = None,
                 block: int = 1, shared: bool = True, inner: str = "f64",
                 omega_jacobi: float = 0.5, omega_richardson: float = 1.0, pre: int = 4, post: int = 4,
                 tol_rel: float = 1e-8, tol_abs: float = 0.0, tol_abs_low: float = 0.0,
                 min_iter: int = 0, max_iter: int = 1000, graph: bool = True, options: Optional[dict] = None):
        self.hierarchy = hierarchy
        self.nsys, self.block, self.shared = nsys, block, shared
        self.max_iter = max_iter
        h = ctypes.c_void_p()
        capi.call("feat3_batch_create", ctypes.byref(h), INNER[inner], block, nsys, int(shared))
        self._h = h
        for i, L in enumerate(hierarchy.levels):
            A = L.A
            v = A.val if values is None else values[i]
            v = np.ascontiguousarray(v, np.float64)
            per_entry = block * block * (1 if shared else nsys)
            if v.size != A.nnz * per_entry:
                raise ValueError(f"level {i}: {v.size} values, expected {A.nnz * per_entry}")
            if L.P is not None:
                P, R = L.P, L.R
                nc = P.cols
                pa = (_ptr(P.row_ptr), _ptr(P.col_ind), _ptr(P.val), _ptr(R.row_ptr), _ptr(R.col_ind), _ptr(R.val))
            else:
                nc = 0
                pa = (ctypes.c_void_p(0),) * 6
            capi.call("feat3_batch_push_level", h, A.rows, _ptr(A.row_ptr), _ptr(A.col_ind), _ptr(v),
                      L.bnd.size, _ptr(L.bnd), nc, *pa)
        capi.call("feat3_batch_set_smoother", h, omega_jacobi, omega_richardson, pre, post)
        capi.call("feat3_batch_set_tolerances", h, tol_rel, tol_abs, tol_abs_low, min_iter, max_iter)
        capi.call("feat3_batch_set_graph", h, int(graph))
        for k, v in (options or {}).items():
            capi.call("feat3_batch_set_option", h, k.encode(), float(v))
        self.status = np.zeros(nsys, np.int32)
        self.iters = np.zeros(nsys, np.uint64)

    @property
    def shape(self):
        """shape of a solution / right-hand side array"""
        return (self.hierarchy.n, self.block, self.nsys)

    def init(self) -> "BatchMGPCG":
        capi.call("feat3_batch_init", self._h)
        return self

    def set_tolerances(self, tol_rel: float, tol_abs: float = 0.0, tol_abs_low: float = 0.0,
                       min_iter: int = 0, max_iter: Optional[int] = None) -> None:
        mi = self.max_iter if max_iter is None else max_iter
        capi.call("feat3_batch_set_tolerances", self._h, tol_rel, tol_abs, tol_abs_low, min_iter, mi)

    def correct(self, x, b, stream=None):
        """x <- solutions of the K systems (FP64 device arrays [node][component][system])."""
        capi.call("feat3_batch_correct", self._h, _ptr(stream), _ptr(x), _ptr(b),
                  _ptr(self.status), _ptr(self.iters))
        return self.status.copy(), self.iters.copy()

    def history(self, system: int) -> np.ndarray:
        out = np.zeros(self.max_iter + 2)
        n = ctypes.c_uint64()
        capi.call("feat3_batch_history", self._h, system, _ptr(out), out.size, ctypes.byref(n))
        return out[: min(n.value, out.size)]

    def time_kernel(self, which: int, reps: int):
        ms, nb = ctypes.c_double(), ctypes.c_uint64()
        capi.call("feat3_batch_time_kernel", self._h, which, reps, ctypes.byref(ms), ctypes.byref(nb))
        return ms.value, nb.value

    def done(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            capi.lib.feat3_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.done()
        except Exception:
            pass
