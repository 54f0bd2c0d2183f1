"""The device MG-PCG solver object (feat3_mgpcg_* C ABI) with the reference's Solver lifecycle.

Mirrors, on one GPU, the solver tree of applications/poisson_dirichlet.cpp:152-190:
``PCG( MultiGrid-V( Richardson(omega_R, Jacobi(omega_J)) x pre/post ), coarse = exact )``, with
FEAT's method names: ``init()`` (init_symbolic + init_numeric), ``correct(x, b)``
(Solver::solve for iterative solvers, iterative.hpp:926-935), ``apply(cor, def)``
(SolverBase::apply), and the statuses of solver/base.hpp:48-64.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import capi
from .problems import Hierarchy

STATUS = {0: "undefined", 1: "progress", 2: "success", 3: "aborted", 4: "diverged", 5: "max-iter", 6: "stagnated"}
INNER = {"f64": 0, "f32": 1, "f16": 2}
FORMAT = {"csr": 0, "sell": 1, "banded": 2}
CYCLE = {"v": 0, "f": 1, "w": 2}          # MultiGridCycle, kernel/solver/multigrid.hpp:26-34


def status_success(status: int) -> bool:
    """solver/base.hpp:110-123"""
    return STATUS.get(status) in ("success", "max-iter", "stagnated")


def _ptr(t) -> ctypes.c_void_p:
    if t is None:
        return ctypes.c_void_p(0)
    if isinstance(t, int):
        return ctypes.c_void_p(t)
    if isinstance(t, np.ndarray):
        return ctypes.c_void_p(t.ctypes.data)
    return ctypes.c_void_p(t.data_ptr())


class MGPCG:
    """PCG preconditioned by a device multigrid cycle (V by default, cycle='f' or 'w' for the
    reference's F- and W-cycles; plain CG with precond='none')."""

    def __init__(self, hierarchy: Hierarchy, inner: str = "f64", fmt: str = "csr",
                 omega_jacobi: float = 0.5, omega_richardson: float = 1.0, pre: int = 4, post: int = 4,
                 tol_rel: float = 1e-8, tol_abs: float = 0.0, tol_abs_low: float = 0.0,
                 min_iter: int = 0, max_iter: int = 1000, precond: str = "mg",
                 graph: bool = True, lanes: int = 0, sell_C: int = 64, sell_sigma: int = 1,
                 options: Optional[dict] = None, cycle: str = "v", level_orders="auto"):
        """level_orders: the solver's internal DOF order per level (feat3_mgpcg_set_level_order):
        "auto" takes the hierarchy's own (problems.Hierarchy.orders of a FEAT-numbered hierarchy:
        the lexicographic lattice order, which makes the level matrices banded), None keeps the
        given numbering, or a list of arrays (finest first; None entries keep a level's)."""
        self.hierarchy = hierarchy
        h = ctypes.c_void_p()
        capi.call("feat3_mgpcg_create", ctypes.byref(h), INNER[inner], FORMAT[fmt])
        self._h = h
        self.inner = inner
        self.max_iter = max_iter
        # a Hierarchy (host arrays) or the device-assembled levels of problems.hierarchy_device
        levels = hierarchy.levels if hasattr(hierarchy, "levels") else hierarchy
        for L in levels:
            A = L.A
            on_device = not isinstance(A.row_ptr, np.ndarray)
            if L.P is not None:
                P, R = L.P, L.R
                nc = P.cols
                pa = (_ptr(P.row_ptr), _ptr(P.col_ind), _ptr(P.val), _ptr(R.row_ptr), _ptr(R.col_ind), _ptr(R.val))
            else:
                nc = 0
                pa = (ctypes.c_void_p(0),) * 6
            nb = L.bnd.numel() if on_device else L.bnd.size
            capi.call("feat3_mgpcg_push_level_device" if on_device else "feat3_mgpcg_push_level", h, A.rows,
                      _ptr(A.row_ptr), _ptr(A.col_ind), _ptr(A.val), nb, _ptr(L.bnd), nc, *pa)
        if isinstance(level_orders, str):
            if level_orders != "auto":
                raise ValueError("level_orders must be 'auto', None or a list of arrays")
            level_orders = getattr(hierarchy, "orders", None)
        for i, o in enumerate(level_orders or []):
            if o is not None:
                oa = np.ascontiguousarray(o, np.uint32)
                capi.call("feat3_mgpcg_set_level_order", h, i, _ptr(oa))
        capi.call("feat3_mgpcg_set_smoother", h, omega_jacobi, omega_richardson, pre, post)
        capi.call("feat3_mgpcg_set_tolerances", h, tol_rel, tol_abs, tol_abs_low, min_iter, max_iter)
        capi.call("feat3_mgpcg_set_precond", h, {"mg": 0, "none": 1}[precond])
        capi.call("feat3_mgpcg_set_launch", h, int(graph), lanes, sell_C, sell_sigma)
        if cycle != "v":
            capi.call("feat3_mgpcg_set_option", h, b"cycle", float(CYCLE[cycle]))
        for k, v in (options or {}).items():
            capi.call("feat3_mgpcg_set_option", h, k.encode(), float(v))
        self.status = 0
        self.iters = 0

    # ---- lifecycle ------------------------------------------------------------------------
    def init(self) -> "MGPCG":
        """init_symbolic + init_numeric; device-assembled levels are converted on the device here
        (they must stay alive until init returns)."""
        capi.call("feat3_mgpcg_init", self._h)
        return self

    def set_tolerances(self, tol_rel: float, tol_abs: float = 0.0, tol_abs_low: float = 0.0,
                       min_iter: int = 0, max_iter: Optional[int] = None) -> None:
        mi = self.max_iter if max_iter is None else max_iter
        capi.call("feat3_mgpcg_set_tolerances", self._h, tol_rel, tol_abs, tol_abs_low, min_iter, mi)

    def done(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            capi.lib.feat3_mgpcg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.done()
        except Exception:
            pass

    # ---- solving --------------------------------------------------------------------------
    def correct(self, x, b, stream=None):
        """x <- solution of A x = b starting from x (FP64 device vectors)."""
        st, it = ctypes.c_int(), ctypes.c_uint64()
        capi.call("feat3_mgpcg_correct", self._h, _ptr(stream), _ptr(x), _ptr(b), ctypes.byref(st), ctypes.byref(it))
        self.status, self.iters = st.value, it.value
        return self.status, self.iters

    def apply(self, cor, defect, stream=None):
        st, it = ctypes.c_int(), ctypes.c_uint64()
        capi.call("feat3_mgpcg_apply", self._h, _ptr(stream), _ptr(cor), _ptr(defect), ctypes.byref(st), ctypes.byref(it))
        self.status, self.iters = st.value, it.value
        return self.status, self.iters

    def vcycle(self, cor, defect, stream=None) -> None:
        capi.call("feat3_mgpcg_vcycle", self._h, _ptr(stream), _ptr(cor), _ptr(defect))

    def history(self) -> np.ndarray:
        out = np.zeros(self.max_iter + 2)
        n = ctypes.c_uint64()
        capi.call("feat3_mgpcg_history", self._h, _ptr(out), out.size, ctypes.byref(n))
        return out[: min(n.value, out.size)]

    def time_kernel(self, which: int, reps: int):
        ms, nb = ctypes.c_double(), ctypes.c_uint64()
        capi.call("feat3_mgpcg_time_kernel", self._h, which, reps, ctypes.byref(ms), ctypes.byref(nb))
        return ms.value, nb.value

    def statistics(self, reps: int = 10) -> dict:
        """Per-level cycle statistics (feat3_mgpcg_statistics): times in ms per cycle and counts,
        bytes, flops per cycle, lists indexed by level (finest first)."""
        nl = self.info()["levels"]
        out = np.zeros(3 + 10 * nl)
        capi.call("feat3_mgpcg_statistics", self._h, reps, _ptr(out), out.size)
        per = out[3:].reshape(nl, 10)
        keys = ["time_smooth", "time_defect", "time_transfer", "time_coarse", "num_smooth", "num_defect",
                "num_transfer", "num_coarse", "bytes", "flops"]
        d = {k: per[:, i].copy() for i, k in enumerate(keys)}
        d.update(levels=nl, ms_per_cycle=float(out[1]), operations=int(out[2]))
        return d

    def info(self) -> dict:
        out = np.zeros(17, np.uint64)
        capi.call("feat3_mgpcg_info", self._h, _ptr(out), 17)
        keys = ["levels", "n", "nnz", "lanes", "sell_size", "kernels_per_vcycle", "vcycle_bytes",
                "outer_bytes", "sizeof_inner", "sizeof_value", "index16", "sell_chunks", "banded", "ndiag",
                "tail_levels", "tail_lds", "tensor_transfers"]
        return {k: int(v) for k, v in zip(keys, out)}
