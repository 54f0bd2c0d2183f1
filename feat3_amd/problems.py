"""Host-side problem setup: the systems of the benchmark configurations as CSR hierarchies.

This wraps the C setup routines (``feat3_setup_*``): Q1/Q2 Lagrange spaces on the refined unit
square/cube, Laplace matrices, two-level prolongations with the reference's structural pattern,
R = P^T, Dirichlet lists and load vectors.  Everything here runs once before the solve (the
reference's assembly step, kernel/assembly + kernel/geometry) and returns numpy arrays.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import capi

U32 = np.uint32
F64 = np.float64


def _p(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


@dataclass
class Csr:
    rows: int
    cols: int
    row_ptr: np.ndarray
    col_ind: np.ndarray
    val: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[self.rows])

    def to_dense(self) -> np.ndarray:
        d = np.zeros((self.rows, self.cols))
        for i in range(self.rows):
            s, e = self.row_ptr[i], self.row_ptr[i + 1]
            np.add.at(d[i], self.col_ind[s:e], self.val[s:e])
        return d


@dataclass
class Level:
    A: Csr
    bnd: np.ndarray
    P: Optional[Csr] = None   # to the next coarser level (fine x coarse)
    R: Optional[Csr] = None   # coarse x fine
    kcell: Optional[np.ndarray] = None


@dataclass
class Hierarchy:
    dim: int
    degree: int
    levels: List[Level] = field(default_factory=list)   # finest first
    level_fine: int = 0
    level_coarse: int = 0
    # "lex": DOFs numbered lexicographically on the tensor lattice (x fastest); "feat": the
    # reference's numbering (StandardRefinery + DofMappingUniform), with orders[l][i] = the
    # lexicographic lattice index of DOF i of level l (finest first)
    numbering: str = "lex"
    orders: Optional[List[np.ndarray]] = None

    @property
    def n(self) -> int:
        return self.levels[0].A.rows


def laplace(dim: int, degree: int, level: int, kcell: Optional[np.ndarray] = None):
    n, nnz, nb = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    capi.call("feat3_setup_laplace_sizes", dim, degree, level, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(nb))
    rp = np.empty(n.value + 1, U32)
    col = np.empty(nnz.value, U32)
    val = np.empty(nnz.value, F64)
    bnd = np.empty(nb.value, U32)
    kc = None if kcell is None else np.ascontiguousarray(kcell, F64)
    capi.call("feat3_setup_laplace", dim, degree, level, _p(kc), _p(rp), _p(col), _p(val), _p(bnd))
    return Csr(n.value, n.value, rp, col, val), bnd


def prolongation(dim: int, degree: int, level_fine: int) -> Csr:
    nf, nc, nnz = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    capi.call("feat3_setup_prol_sizes", dim, degree, level_fine, ctypes.byref(nf), ctypes.byref(nc), ctypes.byref(nnz))
    rp = np.empty(nf.value + 1, U32)
    col = np.empty(nnz.value, U32)
    val = np.empty(nnz.value, F64)
    capi.call("feat3_setup_prol", dim, degree, level_fine, _p(rp), _p(col), _p(val))
    return Csr(nf.value, nc.value, rp, col, val)


def transpose(m: Csr) -> Csr:
    rp = np.empty(m.cols + 1, U32)
    col = np.empty(m.nnz, U32)
    val = np.empty(m.nnz, F64)
    capi.call("feat3_setup_transpose", m.rows, m.cols, _p(m.row_ptr), _p(m.col_ind), _p(m.val), _p(rp), _p(col), _p(val))
    return Csr(m.cols, m.rows, rp, col, val)


def load_vector(dim: int, degree: int, level: int, func: int = 0) -> np.ndarray:
    """func: 0 sine bubble (f = dim*pi^2*prod sin(pi x_i)), 1 constant one, 2 exp bubble."""
    n, nnz, nb = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    capi.call("feat3_setup_laplace_sizes", dim, degree, level, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(nb))
    b = np.empty(n.value, F64)
    capi.call("feat3_setup_rhs", dim, degree, level, func, _p(b))
    return b


def coarsen_cellfield(dim: int, level_fine: int, kf: np.ndarray) -> np.ndarray:
    kc = np.empty(kf.size >> dim, F64)
    kfc = np.ascontiguousarray(kf, F64)   # held while the call reads it
    capi.call("feat3_setup_coarsen_cellfield", dim, level_fine, _p(kfc), _p(kc))
    return kc


def ref_stiffness(dim: int, degree: int) -> np.ndarray:
    nloc = (degree + 1) ** dim
    K = np.empty(nloc * nloc, F64)
    capi.call("feat3_setup_ref_stiffness", dim, degree, _p(K))
    return K.reshape(nloc, nloc)


def hierarchy(dim: int, degree: int, level_fine: int, level_coarse: int = 1,
              kcell: Optional[np.ndarray] = None, numbering: str = "lex") -> Hierarchy:
    """Levels level_fine (finest, index 0) down to level_coarse.  kcell: per-cell coefficient
    on the finest level (cells x fastest); coarser levels use the mean of the children.
    numbering "feat" renumbers every level's DOFs as the reference does (feat_order)."""
    if numbering not in ("lex", "feat"):
        raise ValueError(f"numbering must be 'lex' or 'feat', not {numbering!r}")
    H = Hierarchy(dim, degree, [], level_fine, level_coarse)
    k = kcell
    for lvl in range(level_fine, level_coarse - 1, -1):
        A, bnd = laplace(dim, degree, lvl, k)
        L = Level(A, bnd, kcell=k)
        if lvl > level_coarse:
            L.P = prolongation(dim, degree, lvl)
            L.R = transpose(L.P)
            if k is not None:
                k = coarsen_cellfield(dim, lvl, k)
        H.levels.append(L)
    return to_feat_numbering(H) if numbering == "feat" else H


# -------------------------------------------------------------------------------------------
# the reference's numbering: StandardRefinery + Lagrange DOF mapping
# -------------------------------------------------------------------------------------------
def feat_order(dim: int, degree: int, level: int) -> np.ndarray:
    """lex_of_feat[i]: the lexicographic lattice index of DOF i in the reference's numbering of
    the Lagrange-`degree` space on the unit square/cube refined `level` times
    (feat3_setup_feat_order)."""
    m = (degree << level) + 1
    out = np.empty(m ** dim, U32)
    capi.call("feat3_setup_feat_order", dim, degree, level, _p(out))
    return out


def inverse_order(order: np.ndarray) -> np.ndarray:
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size, dtype=order.dtype)
    return inv


def permute_csr(M: Csr, row_order: Optional[np.ndarray], col_order: Optional[np.ndarray]) -> Csr:
    """Old row i -> row row_order[i], old column j -> column col_order[j] (None: unchanged),
    columns ascending (feat3_setup_permute_csr)."""
    rp = np.empty(M.rows + 1, U32)
    col = np.empty(M.nnz, U32)
    val = np.empty(M.nnz, F64)
    ro = None if row_order is None else np.ascontiguousarray(row_order, U32)
    co = None if col_order is None else np.ascontiguousarray(col_order, U32)
    capi.call("feat3_setup_permute_csr", M.rows, M.cols, _p(M.row_ptr), _p(M.col_ind), _p(M.val), _p(ro), _p(co),
              _p(rp), _p(col), _p(val))
    return Csr(M.rows, M.cols, rp, col, val)


def to_feat_numbering(H: Hierarchy) -> Hierarchy:
    """The same hierarchy with every level in the reference's DOF numbering: matrices, transfers
    and Dirichlet lists permuted (values moved unchanged), orders[l] = lex_of_feat of level l."""
    if H.numbering == "feat":
        return H
    orders = [feat_order(H.dim, H.degree, lvl) for lvl in range(H.level_fine, H.level_coarse - 1, -1)]
    inv = [inverse_order(o) for o in orders]   # feat index of each lexicographic index
    G = Hierarchy(H.dim, H.degree, [], H.level_fine, H.level_coarse, "feat", orders)
    for i, L in enumerate(H.levels):
        A = permute_csr(L.A, inv[i], inv[i])
        bnd = np.sort(inv[i][L.bnd]).astype(U32)
        P = R = None
        if L.P is not None:
            P = permute_csr(L.P, inv[i], inv[i + 1])
            R = permute_csr(L.R, inv[i + 1], inv[i])
        G.levels.append(Level(A, bnd, P, R, kcell=L.kcell))
    return G


def to_feat_vector(H: Hierarchy, v_lex: np.ndarray, level: int = 0) -> np.ndarray:
    """A vector of the lexicographic numbering in H's numbering (identity for "lex")."""
    return v_lex if H.numbering == "lex" else np.ascontiguousarray(v_lex[H.orders[level]])


def lattice_order(coords: np.ndarray) -> np.ndarray:
    """order[i] = rank of node i by its coordinates (rows of coords, last coordinate slowest):
    the lexicographic lattice index of a tensor lattice's nodes (feat3_setup_lattice_order)."""
    c = np.ascontiguousarray(coords, F64)
    n, dim = c.shape
    out = np.empty(n, U32)
    capi.call("feat3_setup_lattice_order", dim, n, _p(c), _p(out))
    return out


def refine_mesh(dim: int, mesh: dict) -> dict:
    """One StandardRefinery step of a conformal quad/hexa mesh given as the reference's index
    sets: keys vtx, v_e, v_q, e_q (+ v_c, e_c, q_c for dim 3), flat arrays."""
    nv = len(mesh["vtx"]) // dim
    counts = np.array([nv, len(mesh["v_e"]) // 2, len(mesh["v_q"]) // 4,
                       len(mesh["v_c"]) // 8 if dim == 3 else 0], np.uint64)
    out = np.empty(4, np.uint64)
    capi.call("feat3_mesh_refine_sizes", dim, _p(counts), _p(out))
    fv, fe, fq, fh = (int(x) for x in out)
    a = {k: np.ascontiguousarray(mesh[k], F64 if k == "vtx" else U32) for k in mesh}
    f = {"vtx": np.empty(fv * dim, F64), "v_e": np.empty(2 * fe, U32), "v_q": np.empty(4 * fq, U32),
         "e_q": np.empty(4 * fq, U32)}
    if dim == 3:
        f.update(v_c=np.empty(8 * fh, U32), e_c=np.empty(12 * fh, U32), q_c=np.empty(6 * fh, U32))
    g = lambda d, k: _p(d.get(k))  # noqa: E731
    capi.call("feat3_mesh_refine", dim, _p(counts), g(a, "vtx"), g(a, "v_e"), g(a, "v_q"), g(a, "e_q"), g(a, "v_c"),
              g(a, "e_c"), g(a, "q_c"), g(f, "vtx"), g(f, "v_e"), g(f, "v_q"), g(f, "e_q"), g(f, "v_c"), g(f, "e_c"),
              g(f, "q_c"))
    return f


# -------------------------------------------------------------------------------------------
# the same hierarchy assembled on the device (bit-identical arrays)
# -------------------------------------------------------------------------------------------
@dataclass
class DeviceCsr:
    """CSR arrays in device memory (torch tensors; the u32 index arrays held as int32)."""
    rows: int
    cols: int
    row_ptr: object
    col_ind: object
    val: object

    def to_host(self) -> Csr:
        return Csr(self.rows, self.cols, self.row_ptr.cpu().numpy().view(U32), self.col_ind.cpu().numpy().view(U32),
                   self.val.cpu().numpy())


def _dptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def laplace_device(dim: int, degree: int, level: int, kcell=None, stream=None):
    """(DeviceCsr A, Dirichlet list) assembled on the device: feat3_device_laplace, the same
    arrays as laplace(); kcell: device coefficient per cell or None for k = 1."""
    import torch
    n, nnz, nb = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    capi.call("feat3_device_laplace_sizes", dim, degree, level, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(nb))
    dev = torch.device("cuda")
    rp = torch.empty(n.value + 1, dtype=torch.int32, device=dev)
    col = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=dev)
    val = torch.empty(max(nnz.value, 1), dtype=torch.float64, device=dev)
    bnd = torch.empty(max(nb.value, 1), dtype=torch.int32, device=dev)
    K = ref_stiffness(dim, degree)
    capi.call("feat3_device_laplace", _dptr(stream) if stream is not None else None, dim, degree, level, _p(K), _dptr(kcell),
              _dptr(rp), _dptr(col), _dptr(val), _dptr(bnd))
    return DeviceCsr(n.value, n.value, rp, col[: nnz.value], val[: nnz.value]), bnd[: nb.value]


def transfer_device(dim: int, degree: int, level_fine: int, transpose: bool = False) -> DeviceCsr:
    """P (fine x coarse) or R = P^T assembled on the device (feat3_device_transfer), the same
    arrays as prolongation() / transpose(prolongation())."""
    import torch
    nf, nc, nnz = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    capi.call("feat3_device_prol_sizes", dim, degree, level_fine, ctypes.byref(nf), ctypes.byref(nc), ctypes.byref(nnz))
    rows, cols = (nc.value, nf.value) if transpose else (nf.value, nc.value)
    dev = torch.device("cuda")
    rp = torch.empty(rows + 1, dtype=torch.int32, device=dev)
    col = torch.empty(nnz.value, dtype=torch.int32, device=dev)
    val = torch.empty(nnz.value, dtype=torch.float64, device=dev)
    capi.call("feat3_device_transfer", None, dim, degree, level_fine, int(transpose), _dptr(rp), _dptr(col), _dptr(val))
    return DeviceCsr(rows, cols, rp, col, val)


@dataclass
class DeviceLevel:
    A: DeviceCsr
    bnd: object
    P: Optional[DeviceCsr] = None
    R: Optional[DeviceCsr] = None


class DeviceLevels(list):
    """Device levels (finest first) of a FEAT-numbered hierarchy with its level orders."""

    def __init__(self, orders=None):
        super().__init__()
        self.orders = orders
        self.numbering = "feat" if orders is not None else "lex"


def permute_device_csr(M: DeviceCsr, row_order, col_order) -> DeviceCsr:
    """feat3_device_permute_csr: old row i -> row_order[i], old column j -> col_order[j]
    (device u32 tensors held as int32, or None), columns ascending."""
    import torch
    rp = torch.empty_like(M.row_ptr)
    col = torch.empty_like(M.col_ind)
    val = torch.empty_like(M.val)
    capi.call("feat3_device_permute_csr", None, M.rows, M.cols, _dptr(M.row_ptr), _dptr(M.col_ind), _dptr(M.val),
              _dptr(row_order), _dptr(col_order), _dptr(rp), _dptr(col), _dptr(val))
    return DeviceCsr(M.rows, M.cols, rp, col, val)


def hierarchy_device(dim: int, degree: int, level_fine: int, level_coarse: int = 1, kcell=None,
                     numbering: str = "lex", orders=None):
    """The levels of hierarchy() assembled on the device (finest first); kcell: device
    coefficient per finest cell or None.  numbering "feat": every level renumbered on the device
    into the reference's numbering (the host computes the orders, feat_order); the returned
    list then carries .orders like a FEAT-numbered Hierarchy (orders: those lists if already
    computed).  device_levels_to_host gives the Hierarchy of the same arrays."""
    if numbering == "feat":
        import torch
        lev = hierarchy_device(dim, degree, level_fine, level_coarse, kcell)
        if orders is None:
            orders = [feat_order(dim, degree, lvl) for lvl in range(level_fine, level_coarse - 1, -1)]
        inv = [torch.from_numpy(inverse_order(o).view(np.int32)).cuda() for o in orders]
        out = DeviceLevels(orders=orders)
        for i, L in enumerate(lev):
            A = permute_device_csr(L.A, inv[i], inv[i])
            bnd = torch.sort(inv[i][L.bnd.long()])[0]
            P = R = None
            if L.P is not None:
                P = permute_device_csr(L.P, inv[i], inv[i + 1])
                R = permute_device_csr(L.R, inv[i + 1], inv[i])
            out.append(DeviceLevel(A, bnd, P, R))
        torch.cuda.synchronize()
        return out
    if numbering != "lex":
        raise ValueError(f"numbering must be 'lex' or 'feat', not {numbering!r}")
    import torch
    levels = []
    k = kcell
    for lvl in range(level_fine, level_coarse - 1, -1):
        A, bnd = laplace_device(dim, degree, lvl, k)
        L = DeviceLevel(A, bnd)
        if lvl > level_coarse:
            L.P = transfer_device(dim, degree, lvl, False)
            L.R = transfer_device(dim, degree, lvl, True)
            if k is not None:
                kc = torch.empty(k.numel() >> dim, dtype=torch.float64, device=k.device)
                capi.call("feat3_coarsen_cellfields", None, dim, lvl, 1, _dptr(k), _dptr(kc))
                k = kc
        levels.append(L)
    return levels


def device_levels_to_host(dim: int, degree: int, level_fine: int, level_coarse: int, levels) -> Hierarchy:
    H = Hierarchy(dim, degree, [], level_fine, level_coarse, getattr(levels, "numbering", "lex"),
                  getattr(levels, "orders", None))
    for L in levels:
        H.levels.append(Level(L.A.to_host(), L.bnd.cpu().numpy().view(U32),
                              None if L.P is None else L.P.to_host(), None if L.R is None else L.R.to_host()))
    return H


def filtered_rhs(H: Hierarchy, func: int = 0) -> np.ndarray:
    """Load vector on the finest level with homogeneous Dirichlet values (filter_rhs), in H's
    numbering."""
    b = to_feat_vector(H, load_vector(H.dim, H.degree, H.level_fine, func))
    b[H.levels[0].bnd] = 0.0
    return b


# -------------------------------------------------------------------------------------------
# the benchmark configurations
# -------------------------------------------------------------------------------------------
CONFIGS = {
    1: dict(dim=2, degree=1, level=6, name="2D Poisson Q1 L6 (4225 DOF)"),
    2: dict(dim=2, degree=2, level=9, name="2D Poisson Q2 L9 (1,050,625 DOF)"),
    3: dict(dim=3, degree=1, level=7, name="3D Poisson Q1 L7 (2,146,689 DOF)"),
    4: dict(dim=2, degree=1, level=10, name="2D variable-coefficient Q1 L10 (1,050,625 DOF)"),
    5: dict(dim=3, degree=1, level=6, name="3D vector Laplacian Q1 L6 (274,625 nodes x 3)"),
}


def config_hierarchy(cfg: int, level: Optional[int] = None, numbering: str = "lex") -> Hierarchy:
    c = CONFIGS[cfg]
    lvl = c["level"] if level is None else level
    return hierarchy(c["dim"], c["degree"], lvl, 1, numbering=numbering)


# -------------------------------------------------------------------------------------------
# batched systems: vector Laplacian blocks, coefficient fields, seeded inputs
# -------------------------------------------------------------------------------------------
def block_identity_values(A: Csr, block: int = 3) -> np.ndarray:
    """Values [entry][h][w] of the blocked Laplacian: every block is a_ij * I (the reference's
    LaplaceOperatorBlocked, assembly/common_operators.hpp:172-181), stored as full row-major
    blocks as in SparseMatrixBCSR<block,block>."""
    v = np.zeros((A.nnz, block, block), F64)
    for h in range(block):
        v[:, h, h] = A.val
    return v


def stacked_values(mats: List[Csr]) -> np.ndarray:
    """Values [entry][system] of K matrices that share one graph."""
    A0 = mats[0]
    for M in mats[1:]:
        if not (np.array_equal(M.row_ptr, A0.row_ptr) and np.array_equal(M.col_ind, A0.col_ind)):
            raise ValueError("matrices of a batch must share the sparsity graph")
    return np.ascontiguousarray(np.stack([M.val for M in mats], axis=1))


class XorShift64Star:
    """The reference's random number generator (util/random.hpp:62-110): xorshift64* with
    multiplier 2685821657736338717; doubles are (next() & 0xFFFFFFFF) / 0xFFFFFFFF and ranged
    doubles a + x (b - a) (random.hpp:259-280)."""

    MULT = np.uint64(2685821657736338717)
    DEF_SEED = 28054777172512

    def __init__(self, seed: int = DEF_SEED):
        self.x = np.uint64(seed if seed != 0 else self.DEF_SEED)

    def next_block(self, n: int) -> np.ndarray:
        out = np.empty(n, np.uint64)
        x = self.x
        with np.errstate(over="ignore"):
            for i in range(n):
                x ^= x >> np.uint64(12)
                x ^= x << np.uint64(25)
                x ^= x >> np.uint64(27)
                out[i] = x * self.MULT
        self.x = x
        return out

    def uniform(self, n: int, a: float, b: float) -> np.ndarray:
        u = (self.next_block(n) & np.uint64(0xFFFFFFFF)).astype(F64) * (1.0 / float(0xFFFFFFFF))
        return a + u * (b - a) if a + np.finfo(F64).eps < b else np.full(n, a)


def random_rhs(nodes: int, block: int, nsys: int, bnd: np.ndarray, seed: int) -> np.ndarray:
    """nsys right-hand sides uniform in [-1, 1] (system fastest: [node][component][system]),
    zero on Dirichlet nodes (filter_rhs)."""
    b = np.empty(nodes * block * nsys, F64)
    capi.call("feat3_setup_random_uniform", seed, b.size, -1.0, 1.0, _p(b))
    b = b.reshape(nodes, block, nsys)
    b[bnd] = 0.0
    return np.ascontiguousarray(b)


_GRF_SPECTRA = {}


def _grf_spectrum(dim: int, level: int, sigma: float, length: float) -> np.ndarray:
    key = (dim, level, sigma, length)
    if key not in _GRF_SPECTRA:
        m = 1 << level
        h = 1.0 / m
        M = 2 * m
        d = np.minimum(np.arange(M), M - np.arange(M)) * h
        r2 = np.zeros((M,) * dim)
        for ax in range(dim):
            shape = [1] * dim
            shape[ax] = M
            r2 = r2 + (d ** 2).reshape(shape)
        lam = np.fft.fftn(sigma ** 2 * np.exp(-r2 / (2.0 * length ** 2))).real
        _GRF_SPECTRA[key] = np.sqrt(np.maximum(lam, 0.0) / lam.size)
    return _GRF_SPECTRA[key]


def gaussian_random_field(dim: int, level: int, seed: int, sigma: float = 1.0, length: float = 0.1) -> np.ndarray:
    """A stationary Gaussian field with covariance sigma^2 exp(-|x-y|^2 / (2 length^2)) on the cell
    centres of the level's mesh (x fastest), sampled by circulant embedding: the covariance of
    the periodic grid of twice the extent is diagonalised by the FFT; the few negative
    eigenvalues of the embedding (of order 1e-16 for length 0.1) are set to zero."""
    m = 1 << level
    M = 2 * m
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((M,) * dim) + 1j * rng.standard_normal((M,) * dim)
    f = np.fft.fftn(_grf_spectrum(dim, level, sigma, length) * z).real
    g = f[tuple(slice(0, m) for _ in range(dim))]
    # array axes are (x, y[, z]); cells are numbered x fastest
    return np.ascontiguousarray(g.transpose(tuple(range(dim - 1, -1, -1))).reshape(-1))


def field_hierarchies(dim: int, level_fine: int, kfields: np.ndarray, level_coarse: int = 1,
                      degree: int = 1) -> List[Hierarchy]:
    """One hierarchy per coefficient field (rows of kfields: per-cell values on the finest level).
    The graph, Dirichlet lists and transfers do not depend on the field and are shared."""
    kf = np.atleast_2d(kfields)
    base = hierarchy(dim, degree, level_fine, level_coarse, kcell=kf[0])
    out = [base]
    for k in kf[1:]:
        H = Hierarchy(dim, degree, [], level_fine, level_coarse)
        kl = k
        for i, lvl in enumerate(range(level_fine, level_coarse - 1, -1)):
            A, _ = laplace(dim, degree, lvl, kl)
            B = base.levels[i]
            H.levels.append(Level(A, B.bnd, B.P, B.R, kcell=kl))
            if lvl > level_coarse:
                kl = coarsen_cellfield(dim, lvl, kl)
        out.append(H)
    return out


def to_banded(A: Csr, max_diag: int = 64):
    """SparseMatrixBanded form of a CSR matrix (sparse_matrix_banded.hpp): sorted offsets stored as
    offset + rows - 1 (u32) and val[d * rows + row] with zeros where the matrix has no entry.
    Returns None when the matrix has more than max_diag distinct diagonals."""
    rows = A.rows
    r = np.repeat(np.arange(rows, dtype=np.int64), np.diff(A.row_ptr.astype(np.int64)))
    off = A.col_ind.astype(np.int64) - r
    offs = np.unique(off)
    if offs.size > max_diag:
        return None
    d = np.searchsorted(offs, off)
    val = np.zeros(offs.size * rows)
    val[d * rows + r] = A.val
    return (offs + rows - 1).astype(np.uint32), val


def to_sell(A: Csr, C: int = 32, sigma: int = 1):
    """SELL-C-sigma arrays of a CSR matrix (feat3_sell_build; sigma = 1 is the reference's
    SparseMatrixELL layout, sparse_matrix_ell.hpp:865-960): chunk starts cs, chunk widths cl,
    row at each position perm, true row length at each position rl, columns, values."""
    nch, size = ctypes.c_uint64(), ctypes.c_uint64()
    capi.call("feat3_sell_sizes", A.rows, _p(A.row_ptr), C, sigma, ctypes.byref(nch), ctypes.byref(size))
    cs = np.empty(nch.value + 1, U32)
    cl = np.empty(max(nch.value, 1), U32)
    perm = np.empty(max(A.rows, 1), U32)
    col = np.empty(max(size.value, 1), U32)
    val = np.empty(max(size.value, 1), F64)
    used = ctypes.c_int()
    capi.call("feat3_sell_build", A.rows, A.cols, _p(A.row_ptr), _p(A.col_ind), _p(A.val), C, sigma, _p(cs), _p(cl),
              _p(perm), _p(col), _p(val), None, ctypes.byref(used))
    rl = np.diff(A.row_ptr.astype(np.int64))[perm[: A.rows]].astype(U32)
    return cs, cl, perm, col, val, rl


# -------------------------------------------------------------------------------------------
# the reference's on-disk matrix formats (Matrix Market, binary fm_csr)
# -------------------------------------------------------------------------------------------
def read_matrix(path: str) -> Csr:
    """A CSR matrix from a Matrix Market file (.mtx) or the reference's binary fm_csr file."""
    mtx = path.endswith(".mtx")
    rows, cols, nnz = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    capi.call("feat3_io_mtx_sizes" if mtx else "feat3_io_csr_binary_sizes", path.encode(),
              ctypes.byref(rows), ctypes.byref(cols), ctypes.byref(nnz))
    rp = np.empty(rows.value + 1, U32)
    col = np.empty(nnz.value, U32)
    val = np.empty(nnz.value, F64)
    capi.call("feat3_io_read_mtx" if mtx else "feat3_io_read_csr_binary", path.encode(), rows.value, nnz.value, _p(rp),
              _p(col), _p(val))
    return Csr(rows.value, cols.value, rp, col, val)


VECTOR_MODES = {"exp": 0, "dv": 1, "mtx": 2}   # FileMode fm_exp, fm_dv, fm_mtx


def read_vector(path: str, mode: str) -> np.ndarray:
    """A DenseVector file of the reference (mode "exp", "dv" or "mtx") as FP64 values."""
    n = ctypes.c_uint64()
    capi.call("feat3_io_vector_size", path.encode(), VECTOR_MODES[mode], ctypes.byref(n))
    out = np.empty(n.value, F64)
    capi.call("feat3_io_read_vector", path.encode(), VECTOR_MODES[mode], _p(out), n.value)
    return out


def write_vector(path: str, v, mode: str) -> None:
    """Writes v as the reference's DenseVector file (text modes keep 6 significant digits)."""
    v = np.ascontiguousarray(v, F64)
    capi.call("feat3_io_write_vector", path.encode(), VECTOR_MODES[mode], v.size, _p(v))


def write_matrix(path: str, A: Csr, symmetric: bool = False) -> None:
    """Writes A as Matrix Market (.mtx; symmetric: lower triangle only) or binary fm_csr."""
    if path.endswith(".mtx"):
        capi.call("feat3_io_write_mtx", path.encode(), A.rows, A.cols, _p(A.row_ptr), _p(A.col_ind), _p(A.val),
                  int(symmetric))
    else:
        capi.call("feat3_io_write_csr_binary", path.encode(), A.rows, A.cols, _p(A.row_ptr), _p(A.col_ind), _p(A.val))
