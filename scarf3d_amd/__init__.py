"""MI355X-native SCARF3D SRM: Python host mirror of the library's C ABI.

The library itself is C++/HIP (scarf3d_amd/csrc, built into
scarf3d_amd/lib/libscarf3df.so and libscarf3d.so). This module binds the same
entry points with ctypes and offers the Fortran module's generic interface
(src/scarflib.f90:48-62: scarf_initialize / scarf_execute / scarf_finalize,
optional arguments by keyword) on top of it.
"""
from .api import (  # noqa: F401
    KERNELS,
    Library,
    LibraryMissing,
    load,
)
