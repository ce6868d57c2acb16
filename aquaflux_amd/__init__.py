"""aquaflux_amd -- MI355X-native (gfx950, HIP) finite-volume residual R(state, params), its tangent
and adjoint, and the block-CSR Newton-Krylov solve, behind the aquaflux residual/operator API.

The compute path is the HIP library ``libaquaflux_amd.so`` (C ABI in ``include/aquaflux_amd.h``),
reached through ctypes; torch supplies device memory and streams.  There is no CPU fallback.
"""

from .mesh import (  # noqa: F401
    CellZones,
    FacePatches,
    Mesh,
    perturbed_grid_2d,
    perturbed_grid_3d,
    structured_grid_2d,
    structured_grid_3d,
    tet_grid_3d,
)

__version__ = "0.1.0"
