"""MI355X-native CG solves and semi-Lagrangian advection for the P2-P1
finite-element Navier-Stokes solver of vasil-pashov/Navier_Stokes_FEM.

The compute path is libnsfem_hip.so (hand-written HIP for gfx950) behind the
C ABI in include/nsfem_gpu.h; this package is its Python front-end:

* :mod:`.mesh`        cavity / channel meshes and the reference JSON format
* :mod:`.device`      ``GPUSimulationDevice`` mirror (uploads, SpMV, BLAS-1,
                      CG, advection), host KD-tree build and FEM assembly
* :mod:`.simulation`  device-resident time stepping
"""
from . import mesh  # noqa: F401  (pure numpy, always importable)

__all__ = ["mesh", "device", "simulation"]


def __getattr__(name):
    if name in ("device", "simulation", "_native"):
        import importlib

        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
