"""The entry counts behind bench.py's algorithmic sweep bytes (CPU only).

band_nnz counts the entries the solver's level copy keeps: |a_ij| > eps^2 |a_ii|.  For the
trilinear (Q1) Laplacian the couplings of axis neighbours (nodes one step apart along one
axis) are zero analytically (27-point stencil 8/3, 0 x 6, -1/6 x 12, -1/12 x 8), so in an m^3
grid the FP32 copy keeps nnz - 6 m^2 (m - 1) entries; the bilinear 2D stencil (8/3, -1/3 x 8)
has no zero coupling.  FP64 keeps every stored entry, and a symmetric pattern keeps
(kept + n) / 2 entries on and above the diagonal.
"""
import pytest

import bench
from feat3_amd import problems as pb


@pytest.mark.parametrize("dim,lvl", [(2, 4), (3, 3), (3, 4)])
def test_kept_entries_closed_form(dim, lvl):
    A = pb.hierarchy(dim, 1, lvl, lvl).levels[0].A
    m = 2 ** lvl + 1
    assert A.rows == m ** dim
    kept32 = A.nnz - (6 * m * m * (m - 1) if dim == 3 else 0)
    assert bench.band_nnz(A, "f32") == kept32
    assert bench.band_nnz(A, "f16") == kept32
    assert bench.band_nnz(A, "f64") == A.nnz
    assert bench.band_nnz(A, "f32", upper=True) == (kept32 + A.rows) // 2
    assert bench.band_nnz(A, "f64", upper=True) == (A.nnz + A.rows) // 2


def test_sweep_bytes_per_format():
    n, nnz = 1000, 9000
    csr = dict(sizeof_value=4, sizeof_inner=4, banded=0, sell_chunks=0)
    assert bench.sweep_algorithmic_bytes(csr, n, nnz, 0) == nnz * 8 + 4 * (n + 1) + 16 * n
    sell = dict(csr, sell_chunks=16, index16=1)
    assert bench.sweep_algorithmic_bytes(sell, n, nnz, 16) == nnz * 6 + 8 * 16 + 16 * n
    band = dict(csr, banded=1, ndiag=9)
    assert bench.sweep_algorithmic_bytes(band, n, nnz, 0) == nnz * 4 + 4 * 9 + 16 * n
