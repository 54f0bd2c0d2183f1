"""MI355X-native LAFEM + Solver hot path of FEAT3.

``capi``      ctypes binding of libfeat3_hip.so (hand-written HIP kernels for gfx950)
``problems``  host-side setup of the benchmark systems (CSR hierarchies, FEAT's DOF numbering)
``lafem``     device containers (DenseVector, SparseMatrixCSR, UnitFilter) over the lafem_hip_* entry points
``solver``    the device MG-PCG solver object with the reference's Solver lifecycle
``batch``     the batched solver: many right-hand sides or coefficient fields in one launch sequence
``autograd``  the differentiable variable-coefficient solve (adjoint gradient)
``factory``   the reference's INI solver configuration (SolverFactory) for the device path
``tune``      the launch-shape autotuner
"""
from . import capi  # noqa: F401  (fails loudly when the HIP library is missing)

__all__ = ["capi", "problems", "lafem", "solver", "batch", "autograd", "factory", "tune"]
