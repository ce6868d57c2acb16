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
    device_mesh: object = dataclasses.field(default=None, repr=False)
    _pack: tuple | None = dataclasses.field(default=None, repr=False)

    def face_pack(self):
        """Face area, normal and centroid in the residual kernels' internal face order (None when
        the mesh needs no permutation).  Rebuilt when a face tensor was modified in place."""
        dm = self.device_mesh
        if dm is None or not dm.face_permuted:
            return None
        f = self.face
        key = (f.area.data_ptr(), f.normal_soa.data_ptr(), f.centroid_soa.data_ptr(), f.area._version,
               f.normal_soa._version, f.centroid_soa._version)
        if self._pack is None or self._pack[1] != key:
            pack = torch.empty((2 * dm.dim + 1) * dm.n_faces, dtype=torch.float64, device=f.area.device)
            gs = _lib.AfxGeometry(f.area.data_ptr(), f.normal_soa.data_ptr(), f.centroid_soa.data_ptr(), None, None,
                                  None)
            _lib.check(_lib.lib().afx_geometry_pack(dm.handle, ctypes.byref(gs), pack.data_ptr(),
                                                    _lib.stream_of(f.area.device)), "afx_geometry_pack")
            self._pack = (pack, key)
        return self._pack[0]

    def c_struct(self) -> _lib.AfxGeometry:
        return _lib.AfxGeometry(
            self.face.area.data_ptr(),
            self.face.normal_soa.data_ptr(),
            self.face.centroid_soa.data_ptr(),
            self.cell.volume.data_ptr(),
            self.cell.centroid_soa.data_ptr(),
            _lib.ptr(self.face_pack()),
        )


def _nodes_soa(mesh, node_coords, device):
    """Node coordinates as a (dim, n_nodes) float64 device tensor; a tensor input keeps its graph."""
    if node_coords is None:
        x = torch.as_tensor(mesh.node_coords, dtype=torch.float64)
        return x.T.contiguous().to(device)
    if not isinstance(node_coords, torch.Tensor):
        node_coords = torch.as_tensor(node_coords, dtype=torch.float64)
    if node_coords.shape != (mesh.n_nodes, mesh.dim):
        raise ValueError(f"node_coords must have shape {(mesh.n_nodes, mesh.dim)}, got {tuple(node_coords.shape)}")
    return node_coords.to(device=device, dtype=torch.float64).T.contiguous()


def _empty_geometry(d, F, C, dev):
    e = lambda *shape: torch.empty(shape, dtype=torch.float64, device=dev)  # noqa: E731
    return e(F), e(d, F), e(d, F), e(C), e(d, C)


class _GeometryFn(torch.autograd.Function):
    """Node coordinates (dim, N) -> (area, normal, face centroid, volume, cell centroid), with the
    hand-written HIP adjoint (afx_geometry_vjp) as the backward and the dual-number HIP tangent
    (afx_geometry_jvp) as the forward derivative."""

    @staticmethod
    def forward(ctx, dm, X):
        d, F, C = dm.dim, dm.n_faces, dm.n_cells
        area, nrm, fc, vol, cc = _empty_geometry(d, F, C, X.device)
        gs = _lib.AfxGeometry(area.data_ptr(), nrm.data_ptr(), fc.data_ptr(), vol.data_ptr(), cc.data_ptr())
        ctx.record = None
        if ctx.needs_input_grad[1]:  # keep the approximate centroids for the reverse pass
            ctx.record = torch.empty((d, C), dtype=torch.float64, device=X.device)
            _lib.check(
                _lib.lib().afx_geometry_compute_record(dm.handle, X.data_ptr(), ctypes.byref(gs),
                                                       ctx.record.data_ptr(), _lib.stream_of(X.device)),
                "afx_geometry_compute_record",
            )
        else:
            _lib.check(
                _lib.lib().afx_geometry_compute(dm.handle, X.data_ptr(), ctypes.byref(gs), _lib.stream_of(X.device)),
                "afx_geometry_compute",
            )
        ctx.dm = dm
        ctx.save_for_backward(X, area, nrm, fc, vol, cc)
        ctx.save_for_forward(X)
        # an output nobody differentiates arrives as None (the adjoint takes a null cotangent)
        # rather than as a materialised zero array
        ctx.set_materialize_grads(False)
        return area, nrm, fc, vol, cc

    @staticmethod
    def backward(ctx, gA, gN, gF, gV, gC):
        X, area, nrm, fc, vol, cc = ctx.saved_tensors
        dm = ctx.dm
        bars = [None if g is None else g.contiguous() for g in (gA, gN, gF, gV, gC)]
        gs = _lib.AfxGeometry(area.data_ptr(), nrm.data_ptr(), fc.data_ptr(), vol.data_ptr(), cc.data_ptr())
        gb = _lib.AfxGeometryCotangents(*[_lib.ptr(b) for b in bars])
        nbytes = ctypes.c_int64()
        L = _lib.lib()
        _lib.check(L.afx_geometry_vjp_workspace_bytes(dm.handle, ctypes.byref(nbytes)), "afx_geometry_vjp_workspace_bytes")
        work = dm.workspace("geom_vjp", int(nbytes.value))
        out = torch.empty_like(X)
        if ctx.record is None:
            rc = L.afx_geometry_vjp(dm.handle, X.data_ptr(), ctypes.byref(gs), ctypes.byref(gb), out.data_ptr(),
                                    work.data_ptr(), work.numel(), _lib.stream_of(X.device))
        else:
            rc = L.afx_geometry_vjp_record(dm.handle, X.data_ptr(), ctypes.byref(gs), ctx.record.data_ptr(),
                                           ctypes.byref(gb), out.data_ptr(), work.data_ptr(), work.numel(),
                                           _lib.stream_of(X.device))
        _lib.check(rc, "afx_geometry_vjp" if ctx.record is None else "afx_geometry_vjp_record")
        return None, out

    @staticmethod
    def jvp(ctx, _dm_t, dX):
        (X,) = ctx.saved_tensors
        dm = ctx.dm
        d, F, C = dm.dim, dm.n_faces, dm.n_cells
        vals = _empty_geometry(d, F, C, X.device)  # recomputed values, discarded
        tans = _empty_geometry(d, F, C, X.device)
        gs = _lib.AfxGeometry(*[t.data_ptr() for t in vals], None)
        gt = _lib.AfxGeometry(*[t.data_ptr() for t in tans], None)
        nbytes = ctypes.c_int64()
        L = _lib.lib()
        _lib.check(L.afx_geometry_jvp_workspace_bytes(dm.handle, ctypes.byref(nbytes)), "afx_geometry_jvp_workspace_bytes")
        work = dm.workspace("geom_jvp", int(nbytes.value))
        _lib.check(
            L.afx_geometry_jvp(dm.handle, X.data_ptr(), dX.contiguous().data_ptr(), ctypes.byref(gs), ctypes.byref(gt),
                               work.data_ptr(), work.numel(), _lib.stream_of(X.device)),
            "afx_geometry_jvp",
        )
        return tans


def compute_geometry(mesh, node_coords=None, device=None) -> MeshGeometry:
    """Derive face/cell geometry from node coordinates with the HIP kernels.  With ``node_coords``
    a tensor that requires grad, the geometry tensors carry the hand-written adjoint."""
    dm = mesh.device_mesh(device)
    dev = dm.device
    X = _nodes_soa(mesh, node_coords, dev)
    area, nrm, fc, vol, cc = _GeometryFn.apply(dm, X)
    if dm.block_order is None and dm.auto_block_order:
        # first geometry of this handle: order the cell kernels' blocks by these centroids
        from .device import slab_tile_block_order

        dm.set_block_order(slab_tile_block_order(cc.detach().T.cpu().numpy()))
        dm.auto_block_order = False
    face = FaceGeometry(area=area, centroid_soa=fc, normal_soa=nrm)
    cell = CellGeometry(volume=vol, centroid_soa=cc)
    return MeshGeometry(face=face, cell=cell, node_coords_soa=X, device_mesh=dm)
