"""MI355X-native LAFEM + Solver hot path of FEAT3.

``capi``      ctypes binding of libfeat3_hip.so (hand-written HIP kernels for gfx950)
``problems``  host-side setup of the benchmark systems (CSR hierarchies)
``solver``    the device MG-PCG solver object with the reference's Solver lifecycle
``lafem``     thin device containers over the lafem_hip_* arch entry points
"""
from . import capi  # noqa: F401  (fails loudly when the HIP library is missing)

__all__ = ["capi", "problems", "solver", "lafem"]
