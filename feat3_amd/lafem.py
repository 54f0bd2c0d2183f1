"""Device containers with the reference's LAFEM method names, over the lafem_hip_* C ABI.

The reference's containers (kernel/lafem/dense_vector.hpp, sparse_matrix_csr.hpp,
sparse_matrix_ell.hpp, sparse_matrix_bcsr.hpp, unit_filter.hpp) own their arrays and forward
every operation to Arch::<Op><Mem_>; these classes hold torch device tensors and forward to the
same operations of the HIP library (the Mem::HIP specialisations).  Semantics follow the
reference's member functions:

  DenseVector.axpy(x, y, alpha)         this = alpha x + y; |alpha| < eps copies y (dense_vector.hpp:883-906)
  DenseVector.component_product(x, y)   this_i = x_i y_i (:913-922)
  DenseVector.component_invert(x, a)    this_i = a / x_i (:936-944)
  DenseVector.scale(x, a)               this = a x (:955-963)
  DenseVector.dot(x), norm2()           host values (:994-1026)
  SparseMatrixCSR.apply(r, x)           r = A x (sparse_matrix_csr.hpp:1899-1928)
  SparseMatrixCSR.apply(r, x, y, a)     r = y + a A x (:1973-2010); transposed likewise
  SparseMatrixCSR.extract_diag()        (:2321-2327)
  UnitFilter.filter_rhs/sol/def/cor     (unit_filter.hpp:399-441)

All work is enqueued on torch's current stream of the vector's device; the host-returning
reductions synchronise it.  No CPU path: without the HIP library the import fails.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import capi

_SUF = {"float64": "f64", "float32": "f32"}


def _torch():
    import torch
    return torch


def _vp(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _suf(t) -> str:
    return _SUF[str(t.dtype).replace("torch.", "")]


class DenseVector:
    """DenseVector<Mem::HIP, DT, u32>: one FP64 or FP32 device array."""

    def __init__(self, data, dtype=None):
        torch = _torch()
        if isinstance(data, int):
            self.t = torch.zeros(data, dtype=dtype or torch.float64, device="cuda")
        elif isinstance(data, np.ndarray):
            self.t = torch.from_numpy(np.ascontiguousarray(data)).to("cuda", dtype=dtype or torch.float64)
        else:
            self.t = data.to("cuda", dtype=dtype or data.dtype).contiguous()
        if _suf(self.t) not in ("f64", "f32"):
            raise TypeError("DenseVector holds float64 or float32 values")
        capi.call("feat3_hip_initialise", self.t.device.index or 0)

    # ---- container ------------------------------------------------------------------------
    def size(self) -> int:
        return self.t.numel()

    def elements(self):
        return self.t

    def clone(self) -> "DenseVector":
        return DenseVector(self.t.clone())

    def to_numpy(self) -> np.ndarray:
        return self.t.cpu().numpy()

    def format(self, value: float = 0.0) -> None:
        capi.call(f"lafem_hip_set_memory_{_suf(self.t)}", _stream(), _vp(self.t), value, self.size())

    def copy(self, x: "DenseVector") -> None:
        self._same(x)
        self.t.copy_(x.t)

    def convert(self, x: "DenseVector") -> None:
        """Container::assign across value types (container.hpp:185-310): FP64 <-> FP32."""
        if x.size() != self.size():
            raise ValueError("Vector size does not match!")
        a, b = _suf(self.t), _suf(x.t)
        if a == b:
            self.t.copy_(x.t)
        else:
            capi.call(f"lafem_hip_convert_{a}_{b}", _stream(), _vp(self.t), _vp(x.t), self.size())

    # ---- BLAS-1 ---------------------------------------------------------------------------
    def axpy(self, x: "DenseVector", y: "DenseVector", alpha: float = 1.0) -> None:
        self._same(x, y)
        capi.call(f"lafem_hip_axpy_{_suf(self.t)}", _stream(), _vp(self.t), alpha, _vp(x.t), _vp(y.t), self.size())

    def component_product(self, x: "DenseVector", y: "DenseVector") -> None:
        self._same(x, y)
        capi.call(f"lafem_hip_component_product_{_suf(self.t)}", _stream(), _vp(self.t), _vp(x.t), _vp(y.t), self.size())

    def component_invert(self, x: "DenseVector", alpha: float = 1.0) -> None:
        self._same(x)
        capi.call(f"lafem_hip_component_invert_{_suf(self.t)}", _stream(), _vp(self.t), _vp(x.t), alpha, self.size())

    def scale(self, x: "DenseVector", alpha: float) -> None:
        self._same(x)
        capi.call(f"lafem_hip_scale_{_suf(self.t)}", _stream(), _vp(self.t), _vp(x.t), alpha, self.size())

    def dot(self, x: "DenseVector") -> float:
        self._same(x)
        out = ctypes.c_double() if _suf(self.t) == "f64" else ctypes.c_float()
        capi.call(f"lafem_hip_dot_value_{_suf(self.t)}", _stream(), _vp(self.t), _vp(x.t), self.size(), ctypes.byref(out))
        return out.value

    def norm2(self) -> float:
        out = ctypes.c_double() if _suf(self.t) == "f64" else ctypes.c_float()
        capi.call(f"lafem_hip_norm2_value_{_suf(self.t)}", _stream(), _vp(self.t), self.size(), ctypes.byref(out))
        return out.value

    def _same(self, *vs: "DenseVector") -> None:
        for v in vs:
            if v.size() != self.size():
                raise ValueError("Vector size does not match!")
            if v.t.dtype != self.t.dtype:
                raise TypeError("Vector value types do not match!")


class SparseMatrixCSR:
    """SparseMatrixCSR<Mem::HIP, DT, u32> from a host problems.Csr (or row_ptr/col_ind/val
    arrays); values converted to dtype (float64 or float32)."""

    def __init__(self, csr, dtype=None):
        torch = _torch()
        dtype = dtype or torch.float64
        self._rows, self._cols = int(csr.rows), int(csr.cols)
        dev = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype=dt)  # noqa: E731
        self.row_ptr = dev(np.asarray(csr.row_ptr, np.uint32).view(np.int32), torch.int32)
        self.col_ind = dev(np.asarray(csr.col_ind, np.uint32).view(np.int32), torch.int32)
        self.val = dev(np.asarray(csr.val, np.float64), dtype)

    def rows(self) -> int:
        return self._rows

    def columns(self) -> int:
        return self._cols

    def used_elements(self) -> int:
        return self.val.numel()

    def apply(self, r: DenseVector, x: DenseVector, y: Optional[DenseVector] = None, alpha: float = 1.0,
              transposed: bool = False) -> None:
        """r = A x, or r = y + alpha A x (transposed: with A^T)."""
        n_out, n_in = (self._cols, self._rows) if transposed else (self._rows, self._cols)
        if r.size() != n_out or x.size() != n_in or (y is not None and y.size() != n_out):
            raise ValueError("Vector size does not match!")
        if r.t.data_ptr() == x.t.data_ptr():
            raise ValueError("Vector x and r must not be the same!")
        suf = _suf(self.val)
        a, b, yy = (1.0, 0.0, r) if y is None else (alpha, 1.0, y)
        capi.call(f"lafem_hip_apply_csr_{suf}_u32", _stream(), _vp(r.t), a, _vp(x.t), b, _vp(yy.t), _vp(self.val),
                  _vp(self.col_ind), _vp(self.row_ptr), self._rows, self._cols, self.used_elements(), int(transposed))

    def extract_diag(self) -> DenseVector:
        d = DenseVector(self._rows, self.val.dtype)
        capi.call(f"lafem_hip_diagonal_csr_{_suf(self.val)}_u32", _stream(), _vp(d.t), _vp(self.val), _vp(self.col_ind),
                  _vp(self.row_ptr), self._rows)
        return d


class UnitFilter:
    """UnitFilter<Mem::HIP, DT, u32>: Dirichlet indices (ascending) with prescribed values."""

    def __init__(self, indices, values=None, dtype=None):
        torch = _torch()
        idx = np.ascontiguousarray(indices, np.uint32)
        self.idx = torch.from_numpy(idx.view(np.int32)).cuda()
        vals = np.zeros(idx.size) if values is None else np.asarray(values, np.float64)
        self.val = torch.from_numpy(np.ascontiguousarray(vals)).to("cuda", dtype=dtype or torch.float64)

    def used_elements(self) -> int:
        return self.idx.numel()

    def filter_rhs(self, v: DenseVector) -> None:
        capi.call(f"lafem_hip_unit_filter_rhs_{_suf(v.t)}_u32", _stream(), _vp(v.t), _vp(self.val.to(v.t.dtype)),
                  _vp(self.idx), self.used_elements(), 1)

    filter_sol = filter_rhs

    def filter_def(self, v: DenseVector) -> None:
        capi.call(f"lafem_hip_unit_filter_def_{_suf(v.t)}_u32", _stream(), _vp(v.t), _vp(self.idx),
                  self.used_elements(), 1)

    filter_cor = filter_def
