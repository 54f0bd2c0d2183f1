"""The C++ interface (Mem::HIP containers, op-by-op solver classes, ConvertPrecond and the fused
solver behind FEAT's SolverBase lifecycle) used like an application
(applications/poisson_dirichlet.cpp:131-200), against the oracle.

Problem: 2D Q1 Poisson levels 5..0, ExpBubble load, Jacobi 0.7, V(4,4), the smoother as coarse
solver -- the configuration of the reference's integration test (applications/CMakeLists.txt:
141-150, 4 +- 1 iterations).  The coarsest level (one cell) is all-Dirichlet, so the exact
coarse solve of the fused solver and the smoother give the same (zero) correction there.
Bars: FP64 paths: iterations equal to the oracle's, relative solution difference <= 1e-10;
FP32 multigrid inside FP64 PCG: iterations +-1, relative difference to the FP64 oracle <= 1e-8.
"""
import os
import subprocess

import numpy as np
import pytest

from feat3_amd import problems as pb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "lafem_hip_driver")


def test_cpp_application_level_solvers(tmp_path, torch_cuda):
    assert os.path.exists(DRIVER), "build tests/cpp/lafem_hip_driver first (__graft_entry__.build())"
    out = subprocess.run([DRIVER, "5", "0", "0.7", str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0, out.stderr
    res = {l.split()[0]: [int(v) for v in l.split()[1:]] for l in out.stdout.strip().splitlines()}
    H = pb.hierarchy(2, 1, 5, 0)
    b = pb.filtered_rhs(H, 2)
    S = po.System.from_hierarchy(H)
    xo = np.zeros(H.n)
    so, io, _ = S.solve(xo, b, po.params(omega_jac=0.7, coarse=po.COARSE_SMOOTHER))
    assert so == 2 and abs(io - 4) <= 1

    def load(name):
        return np.fromfile(str(tmp_path / f"{name}.bin"), dtype=np.float64)

    def rel(a):
        return np.linalg.norm(a - xo) / np.linalg.norm(xo)

    assert res["generic64"] == [2, io]
    assert rel(load("generic64")) <= 1e-10
    assert res["fused64"] == [2, io]
    assert rel(load("fused64")) <= 1e-10
    assert res["fusedbanded64"] == [2, io]
    assert rel(load("fusedbanded64")) <= 1e-10
    assert res["banded_vs_csr"][0] <= 1000 and res["banded_vs_csr"][1] == 9   # within 1e-13 of max |entry|
    assert res["mixed32"][0] == 2 and abs(res["mixed32"][1] - io) <= 1
    assert rel(load("mixed32")) <= 1e-8
    assert res["file_roundtrip"][0] == 1   # fm_csr and symmetric fm_mtx round trips, bitwise products
    assert res["vector_files"] == [1, H.n]  # DenseVector fm_dv bitwise, fm_exp / fm_mtx to 6 digits
    # the SolverFactory-built solver is the directly built one (same INI parameters)
    assert res["factory64"] == res["fused64"]
    assert np.array_equal(load("factory64"), load("fused64"))
    assert res["factory_rejects_unknown_cycle"] == [1]
    # cycle = w: the oracle's W-cycle PCG
    xw = np.zeros(H.n)
    sw, iw, _ = S.solve(xw, b, po.params(omega_jac=0.7, coarse=po.COARSE_SMOOTHER, cycle=po.CYCLE_W))
    assert sw == 2 and iw <= io
    assert res["factoryw64"] == [2, iw]
    assert np.linalg.norm(load("factoryw64") - xw) <= 1e-10 * np.linalg.norm(xw)
    assert res["leaked"] == [0]
