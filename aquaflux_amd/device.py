"""Device-resident mesh topology: the C ABI mesh handle plus torch views of the index arrays."""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib


class DeviceMesh:
    """Owns one ``afx_mesh`` handle on one GPU for a :class:`~aquaflux_amd.mesh.Mesh`."""

    @staticmethod
    def device_key(device) -> int:
        if not torch.cuda.is_available():
            raise RuntimeError("aquaflux_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        if device is None:
            return torch.cuda.current_device()
        return torch.device(device).index if torch.device(device).index is not None else torch.cuda.current_device()

    def __init__(self, mesh, device=None):
        index = self.device_key(device)
        self.device = torch.device("cuda", index)
        self.dim = mesh.dim
        self.n_cells = mesh.n_cells
        self.n_faces = mesh.n_faces
        self.n_nodes = mesh.n_nodes
        L = _lib.lib()
        handle = ctypes.c_void_p()
        owner = np.ascontiguousarray(mesh.owner, dtype=np.int32)
        nb = np.ascontiguousarray(mesh.neighbour, dtype=np.int32)
        off = np.ascontiguousarray(mesh.face_node_offsets, dtype=np.int64)
        idx = np.ascontiguousarray(mesh.face_node_indices, dtype=np.int32)
        with torch.cuda.device(self.device):
            _lib.check(
                L.afx_mesh_create(
                    mesh.dim, mesh.n_cells, mesh.n_faces, mesh.n_nodes,
                    owner.ctypes.data, nb.ctypes.data, off.ctypes.data, idx.ctypes.data,
                    ctypes.byref(handle),
                ),
                "afx_mesh_create",
            )
        self.handle = handle
        self.owner = torch.from_numpy(owner.astype(np.int64)).to(self.device)
        self.neighbour = torch.from_numpy(nb.astype(np.int64)).to(self.device)
        info = (ctypes.c_int64 * 7)()
        _lib.check(L.afx_mesh_info(handle, ctypes.addressof(info)), "afx_mesh_info")
        self.n_cell_face = int(info[4])
        self.max_faces = int(info[5])
        permuted = ctypes.c_int32()
        _lib.check(L.afx_mesh_face_order(handle, ctypes.byref(permuted), None, None), "afx_mesh_face_order")
        # the residual kernels visit faces sorted by owner; geometry then carries a packed copy
        self.face_permuted = bool(permuted.value)
        self._work = {}

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib._LIB is not None:
            try:
                _lib.lib().afx_mesh_destroy(h)
            except Exception:  # interpreter teardown
                pass
            self.handle = None

    def workspace(self, key, nbytes: int) -> torch.Tensor:
        """A cached device byte buffer of at least ``nbytes`` (torch caching allocator)."""
        buf = self._work.get(key)
        if buf is None or buf.numel() < nbytes:
            buf = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
            self._work[key] = buf
        return buf

    def face_order(self):
        """Caller face index of each internal face (int32 device tensor), or None when the mesh's
        own face order is the internal one."""
        if not self.face_permuted:
            return None
        fo = torch.empty(self.n_faces, dtype=torch.int32, device=self.device)
        flag = ctypes.c_int32()
        _lib.check(_lib.lib().afx_mesh_face_order(self.handle, ctypes.byref(flag), fo.data_ptr(),
                                                  _lib.stream_of(self.device)), "afx_mesh_face_order")
        return fo

    def cell_faces(self):
        """(cell_ptr, cell_faces) int32 device tensors of the cell -> face list."""
        cptr = torch.empty(self.n_cells + 1, dtype=torch.int32, device=self.device)
        cf = torch.empty(self.n_cell_face, dtype=torch.int32, device=self.device)
        _lib.check(
            _lib.lib().afx_mesh_cell_faces(self.handle, cptr.data_ptr(), cf.data_ptr(), _lib.stream_of(self.device)),
            "afx_mesh_cell_faces",
        )
        return cptr, cf
