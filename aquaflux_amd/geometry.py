"""Face and cell geometry on the GPU (``aquaflux/mesh/geometry.py``, ``mesh/mesh.py:341-387``).

Storage is structure-of-arrays (``(dim, n)`` tensors) because that is what the kernels read
coalesced; the reference's ``(n, dim)`` layout is exposed as transposed views.
"""

from __future__ import annotations

import ctypes
import dataclasses

import torch

from . import _lib


@dataclasses.dataclass
class FaceGeometry:
    area: torch.Tensor  # (F,)
    centroid_soa: torch.Tensor  # (d, F)
    normal_soa: torch.Tensor  # (d, F)

    @property
    def centroid(self) -> torch.Tensor:
        return self.centroid_soa.T

    @property
    def normal(self) -> torch.Tensor:
        return self.normal_soa.T


@dataclasses.dataclass
class CellGeometry:
    volume: torch.Tensor  # (C,)
    centroid_soa: torch.Tensor  # (d, C)

    @property
    def centroid(self) -> torch.Tensor:
        return self.centroid_soa.T


@dataclasses.dataclass
class MeshGeometry:
    """Bundled face and cell metrics (``aquaflux/mesh/geometry.py``)."""

    face: FaceGeometry
    cell: CellGeometry
    node_coords_soa: torch.Tensor | None = None  # (d, N), the coordinates this was derived from

    def c_struct(self) -> _lib.AfxGeometry:
        return _lib.AfxGeometry(
            self.face.area.data_ptr(),
            self.face.normal_soa.data_ptr(),
            self.face.centroid_soa.data_ptr(),
            self.cell.volume.data_ptr(),
            self.cell.centroid_soa.data_ptr(),
        )


def _nodes_soa(mesh, node_coords, device):
    if node_coords is None:
        x = torch.as_tensor(mesh.node_coords, dtype=torch.float64)
        return x.T.contiguous().to(device)
    if not isinstance(node_coords, torch.Tensor):
        node_coords = torch.as_tensor(node_coords, dtype=torch.float64)
    if node_coords.shape != (mesh.n_nodes, mesh.dim):
        raise ValueError(f"node_coords must have shape {(mesh.n_nodes, mesh.dim)}, got {tuple(node_coords.shape)}")
    return node_coords.detach().to(device=device, dtype=torch.float64).T.contiguous()


def compute_geometry(mesh, node_coords=None, device=None) -> MeshGeometry:
    """Derive face/cell geometry from node coordinates with the HIP kernels."""
    dm = mesh.device_mesh(device)
    dev = dm.device
    X = _nodes_soa(mesh, node_coords, dev)
    d, F, C = mesh.dim, mesh.n_faces, mesh.n_cells
    face = FaceGeometry(
        area=torch.empty(F, dtype=torch.float64, device=dev),
        centroid_soa=torch.empty((d, F), dtype=torch.float64, device=dev),
        normal_soa=torch.empty((d, F), dtype=torch.float64, device=dev),
    )
    cell = CellGeometry(
        volume=torch.empty(C, dtype=torch.float64, device=dev),
        centroid_soa=torch.empty((d, C), dtype=torch.float64, device=dev),
    )
    geom = MeshGeometry(face=face, cell=cell, node_coords_soa=X)
    gs = geom.c_struct()
    _lib.check(
        _lib.lib().afx_geometry_compute(dm.handle, X.data_ptr(), ctypes.byref(gs), _lib.stream_of(dev)),
        "afx_geometry_compute",
    )
    return geom
