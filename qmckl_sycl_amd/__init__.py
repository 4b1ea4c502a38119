"""MI355X-native per-electron basis evaluation (en_distance -> ao_vgl ->
mo_vgl) behind the qmckl GPU context/provider C ABI (``include/qmckl_gpu.h``).

The compute lives in ``libqmckl_gpu_amd.so`` (HIP kernels for gfx950 and the
C++ provider layer); this package holds its ctypes binding (``_lib``), a thin
Pythonic mirror of the API (``context``), the system generators
(``systems``) and the build script (``build``).
"""
from . import systems  # noqa: F401

__all__ = ["systems", "context", "build"]
