"""ctypes binding of the CPU oracle (liborc.so / liborc_omp.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

c_vp = ctypes.c_void_p
c_u64 = ctypes.c_uint64
c_d = ctypes.c_double
c_f = ctypes.c_float
c_i = ctypes.c_int

_PROTOS = {
    "orc_csr_apply_f64_u32": (None, [c_vp, c_d, c_vp, c_d, c_vp, c_vp, c_vp, c_vp, c_u64, c_u64, c_i]),
    "orc_csr_apply_f64_u64": (None, [c_vp, c_d, c_vp, c_d, c_vp, c_vp, c_vp, c_vp, c_u64, c_u64, c_i]),
    "orc_csr_apply_f32_u32": (None, [c_vp, c_f, c_vp, c_f, c_vp, c_vp, c_vp, c_vp, c_u64, c_u64, c_i]),
    "orc_csr_apply_h16_u32": (None, [c_vp, c_f, c_vp, c_f, c_vp, c_vp, c_vp, c_vp, c_u64, c_u64]),
    "orc_csrb_apply_f64_u32": (None, [c_vp, c_d, c_vp, c_d, c_vp, c_vp, c_vp, c_vp, c_u64, c_i, c_i]),
    "orc_csrb_apply_f32_u32": (None, [c_vp, c_f, c_vp, c_f, c_vp, c_vp, c_vp, c_vp, c_u64, c_i, c_i]),
    "orc_csrsb_apply_f64_u32": (None, [c_vp, c_d, c_vp, c_d, c_vp, c_vp, c_vp, c_vp, c_u64, c_i]),
    "orc_banded_apply_f64_u32": (None, [c_vp, c_d, c_vp, c_d, c_vp, c_vp, c_vp, c_u64, c_u64, c_u64]),
    "orc_banded_apply_f32_u32": (None, [c_vp, c_f, c_vp, c_f, c_vp, c_vp, c_vp, c_u64, c_u64, c_u64]),
    "orc_ell_val_size_u32": (c_u64, [c_u64, c_u64, c_vp]),
    "orc_ell_from_csr_f64_u32": (None, [c_u64, c_u64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "orc_ell_apply_f64_u32": (None, [c_vp, c_d, c_vp, c_d, c_vp, c_vp, c_vp, c_vp, c_vp, c_u64, c_u64]),
    "orc_axpy_f64": (None, [c_vp, c_d, c_vp, c_vp, c_u64]),
    "orc_axpy_f32": (None, [c_vp, c_f, c_vp, c_vp, c_u64]),
    "orc_dot_f64": (c_d, [c_vp, c_vp, c_u64]),
    "orc_dot_f32": (c_f, [c_vp, c_vp, c_u64]),
    "orc_norm2_f64": (c_d, [c_vp, c_u64]),
    "orc_norm2_f32": (c_f, [c_vp, c_u64]),
    "orc_component_product_f64": (None, [c_vp, c_vp, c_vp, c_u64]),
    "orc_component_invert_f64": (None, [c_vp, c_vp, c_d, c_u64]),
    "orc_scale_f64": (None, [c_vp, c_vp, c_d, c_u64]),
    "orc_diag_csr_f64_u32": (None, [c_vp, c_vp, c_vp, c_vp, c_u64]),
    "orc_diag_csrb_f64_u32": (None, [c_vp, c_vp, c_vp, c_vp, c_u64, c_i, c_i]),
    "orc_unit_filter_rhs_f64_u32": (None, [c_vp, c_vp, c_vp, c_u64]),
    "orc_unit_filter_def_f64_u32": (None, [c_vp, c_vp, c_u64]),
    "orc_transpose_csr_f64_u32": (None, [c_u64, c_u64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "orc_float_to_half": (ctypes.c_uint16, [c_f]),
    "orc_half_to_float": (c_f, [ctypes.c_uint16]),
    "orc_pointstar_fd_nnz": (c_u64, [c_u64, c_u64]),
    "orc_pointstar_fd_f64_u32": (None, [c_u64, c_u64, c_vp, c_vp, c_vp]),
    "orc_pointstar_q2_bubble_f64": (None, [c_u64, c_u64, c_vp]),
    "orc_system_create": (c_vp, []),
    "orc_system_destroy": (None, [c_vp]),
    "orc_system_add_level": (c_i, [c_vp, c_u64, c_i, c_vp, c_vp, c_vp, c_u64, c_vp, c_u64, c_vp, c_vp, c_vp,
                                   c_vp, c_vp, c_vp]),
    "orc_system_solve": (c_u64, [c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(c_i), c_vp, c_u64, ctypes.POINTER(c_u64)]),
    "orc_system_vcycle": (c_i, [c_vp, c_vp, c_vp, c_vp]),
    "orc_last_solve_seconds": (c_d, []),
    "orc_cell_gradient": (c_i, [c_i, c_i, c_i, c_vp, c_vp, c_vp, c_vp]),
}


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load(name: str):
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    for fn, (res, args) in _PROTOS.items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
    return lib


_libs = {}


def lib(omp: bool = False):
    key = "liborc_omp.so" if omp else "liborc.so"
    if key not in _libs:
        _libs[key] = _load(key)
    return _libs[key]


def P(a):
    """ctypes pointer of a numpy array (None -> NULL)."""
    return c_vp(0) if a is None else c_vp(a.ctypes.data)


# ---- convenience wrappers ----------------------------------------------------------------
def csr_apply(r, a, x, b, y, val, col, rp, rows, cols, transposed=False, omp=False):
    L = lib(omp)
    if val.dtype == np.float64:
        fn = L.orc_csr_apply_f64_u64 if col.dtype == np.uint64 else L.orc_csr_apply_f64_u32
    else:
        fn = L.orc_csr_apply_f32_u32
    fn(P(r), a, P(x), b, P(y), P(val), P(col), P(rp), rows, cols, int(transposed))


def pointstar_fd(m: int, d: int):
    L = lib()
    neq = m ** d
    nnz = L.orc_pointstar_fd_nnz(m, d)
    rp = np.empty(neq + 1, np.uint32)
    col = np.empty(nnz, np.uint32)
    val = np.empty(nnz, np.float64)
    L.orc_pointstar_fd_f64_u32(m, d, P(rp), P(col), P(val))
    return rp, col, val


def pointstar_q2_bubble(m: int, d: int):
    v = np.empty(m ** d, np.float64)
    lib().orc_pointstar_q2_bubble_f64(m, d, P(v))
    return v


# solver kinds and parameter vector of orc_system_solve
PCG_MG, PCG_JACOBI, CG, RICHARDSON_JACOBI = 0, 1, 2, 3
COARSE_EXACT, COARSE_SMOOTHER, COARSE_NONE = 0, 1, 2


CYCLE_V, CYCLE_F, CYCLE_W = 0, 1, 2


def params(kind=PCG_MG, tol_rel=1e-8, max_iter=1000, omega_jac=0.5, omega_rich=1.0, pre=4, post=4,
           coarse=COARSE_EXACT, inner=0, tol_abs=0.0, min_iter=0, use_def=1, cycle=CYCLE_V):
    return np.array([kind, tol_rel, max_iter, omega_jac, omega_rich, pre, post, coarse, inner, tol_abs,
                     min_iter, use_def, cycle], np.float64)


class System:
    """A level hierarchy for the oracle's solvers (finest first)."""

    def __init__(self, omp: bool = False):
        self.L = lib(omp)
        self.h = self.L.orc_system_create()
        self.nlev = 0

    def add_level(self, n, rp, col, val, bnd, block=1, P_=None, R_=None):
        if P_ is not None:
            nc = P_.cols
            args = (P(P_.row_ptr), P(P_.col_ind), P(P_.val), P(R_.row_ptr), P(R_.col_ind), P(R_.val))
        else:
            nc = 0
            args = (c_vp(0),) * 6
        rc = self.L.orc_system_add_level(self.h, n, block, P(rp), P(col), P(val), bnd.size, P(bnd), nc, *args)
        assert rc == 0
        self.nlev += 1

    @classmethod
    def from_hierarchy(cls, H, omp: bool = False, block: int = 1, values=None):
        """values: optional per-level value arrays replacing L.A.val (block x block blocks,
        row-major, for block > 1)."""
        s = cls(omp)
        for i, L in enumerate(H.levels):
            v = L.A.val if values is None else np.ascontiguousarray(values[i], np.float64)
            s.add_level(L.A.rows, L.A.row_ptr, L.A.col_ind, v, L.bnd, block, L.P, L.R)
        return s

    def solve(self, x, b, prm):
        st = c_i()
        hist = np.zeros(int(prm[2]) + 2)
        hl = c_u64()
        it = self.L.orc_system_solve(self.h, P(prm), P(x), P(b), ctypes.byref(st), P(hist), hist.size, ctypes.byref(hl))
        return st.value, int(it), hist[: min(hl.value, hist.size)]

    def vcycle(self, cor, d, prm):
        return self.L.orc_system_vcycle(self.h, P(prm), P(cor), P(d))

    def __del__(self):
        try:
            self.L.orc_system_destroy(self.h)
        except Exception:
            pass


def cell_gradient(dim, degree, level, E, u, lam):
    """dL/dk_e = -lambda_e^T E u_e for every cell (one system)."""
    m = 1 << level
    g = np.empty(m ** dim)
    E = np.ascontiguousarray(E, np.float64)
    u = np.ascontiguousarray(u, np.float64)      # held while the call reads them
    lam = np.ascontiguousarray(lam, np.float64)
    rc = lib().orc_cell_gradient(dim, degree, level, P(E), P(u), P(lam), P(g))
    assert rc == 0
    return g
