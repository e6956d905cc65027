"""DG-BGK explicit time march on MI355X (gfx950): Python host side.

The time march runs in libdgbgk.so (HIP kernels behind the C ABI declared
in include/dgbgk.h).  This package holds the host mirror of the
reference's driver (``solver``), the run-deck / mesh / velocity-file
formats (``io``), synthetic meshes (``meshgen``) and workloads (``cases``).
"""
from . import cases, meshgen  # noqa: F401

__all__ = ["cases", "meshgen"]
