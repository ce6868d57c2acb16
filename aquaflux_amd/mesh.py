"""Unstructured cell-centred meshes: host-side integer connectivity plus device geometry.

A :class:`Mesh` holds the topology the residual kernels consume -- owner / neighbour per face
(``neighbour == -1`` marks a boundary face), the face -> node ring in compressed-sparse-row form,
named face patches and cell zones -- as host ``numpy`` arrays, exactly the arrays the reference
mesh carries (``aquaflux/mesh/mesh.py:76-206``, ``aquaflux/mesh/connectivity.py:114-136``).
Geometry is derived on the GPU by :meth:`Mesh.geometry` from the node coordinates, which stay a
differentiable input (``aquaflux/mesh/mesh.py:341-387``).

The structured generators reproduce the reference's face numbering bit for bit: faces come in
families normal to x, then y, then z; inside a family the lattice index along the family axis
varies slowest; the owner is the lower cell and the neighbour the higher one (``-1`` on the domain
boundary); patches ``left/right/bottom/top/back/front`` name the six sides
(``aquaflux/mesh/structured.py:63-375``).  That numbering is a data format shared with the
reference, which is why it is kept identical here.
"""

from __future__ import annotations

import dataclasses

import numpy as np

__all__ = [
    "CellZones",
    "FacePatches",
    "Mesh",
    "structured_grid_2d",
    "structured_grid_3d",
    "perturbed_grid_2d",
    "perturbed_grid_3d",
    "tet_grid_3d",
    "index_dtype",
]


def index_dtype(bound: int):
    """``int32`` when every index below ``bound`` fits, else ``int64``
    (``aquaflux/mesh/connectivity.py:32-52``)."""
    return np.int32 if bound < np.iinfo(np.int32).max else np.int64


class _Groups:
    """A partition of elements into named groups: one integer label per element."""

    def __init__(self, label: np.ndarray, names: tuple[str, ...]):
        self.label = np.asarray(label, dtype=np.int64)
        self.names = tuple(names)

    @property
    def n_groups(self) -> int:
        return len(self.names)

    def id_of(self, name: str) -> int:
        try:
            return self.names.index(name)
        except ValueError:
            raise ValueError(f"no group named '{name}'; groups are {self.names}") from None

    def mask(self, name: str) -> np.ndarray:
        return self.label == self.id_of(name)

    def indices(self, name: str) -> np.ndarray:
        """Element indices of the named group, ascending."""
        return np.nonzero(self.mask(name))[0]

    def size(self, name: str) -> int:
        return int(np.count_nonzero(self.mask(name)))


class CellZones(_Groups):
    """Named cell zones; unlisted cells fall in ``"default"`` (``aquaflux/mesh/groups.py:86-92``)."""

    @classmethod
    def default(cls, n_cells: int) -> "CellZones":
        return cls(np.zeros(n_cells, dtype=np.int64), ("default",))

    @classmethod
    def from_dict(cls, n_cells: int, groups: dict) -> "CellZones":
        names = ["default"] + [n for n in groups if n != "default"]
        ids = {n: i for i, n in enumerate(names)}
        label = np.zeros(n_cells, dtype=np.int64)
        taken = np.zeros(n_cells, dtype=bool)
        for name, idx in groups.items():
            idx = np.asarray(list(idx), dtype=np.int64)
            if idx.size == 0:
                continue
            if idx.min() < 0 or idx.max() >= n_cells:
                raise ValueError(f"group '{name}' has an index outside [0, {n_cells})")
            if np.any(taken[idx]):
                raise ValueError(f"group '{name}' overlaps another group (partition violated)")
            taken[idx] = True
            label[idx] = ids[name]
        return cls(label, tuple(names))


class FacePatches(_Groups):
    """Named face patches over the default ``interior``/``boundary`` split
    (``aquaflux/mesh/groups.py:128-170``)."""

    @classmethod
    def from_dict(cls, neighbour: np.ndarray, groups: dict) -> "FacePatches":
        nb = np.asarray(neighbour)
        n_faces = nb.shape[0]
        reserved = [n for n in groups if n in ("interior", "boundary")]
        if reserved:
            raise ValueError(
                f"patch name(s) {reserved} are reserved (interior/boundary are assigned "
                "automatically); use a different name"
            )
        names = ["interior", "boundary"] + list(groups)
        label = np.where(nb < 0, 1, 0).astype(np.int64)
        taken = np.zeros(n_faces, dtype=bool)
        for gid, (name, idx) in enumerate(groups.items(), start=2):
            idx = np.asarray(list(idx) if not isinstance(idx, np.ndarray) else idx, dtype=np.int64)
            if idx.size == 0:
                continue
            if idx.min() < 0 or idx.max() >= n_faces:
                raise ValueError(f"patch '{name}' has an index outside [0, {n_faces})")
            if np.any(taken[idx]):
                raise ValueError(f"patch '{name}' overlaps another named patch")
            taken[idx] = True
            label[idx] = gid
        return cls(label, tuple(names))

    def is_boundary_patch(self, name: str, neighbour) -> bool:
        """True when the named patch is non-empty and all its faces are boundary faces
        (``aquaflux/mesh/groups.py:182-197``); ``neighbour`` is the per-face neighbour array or a
        :class:`Mesh`."""
        nb = np.asarray(getattr(neighbour, "neighbour", neighbour))
        m = self.mask(name)
        return bool(m.any()) and bool(np.all(nb[m] < 0))


@dataclasses.dataclass(eq=False)
class Mesh:
    """Cell-centred unstructured mesh (topology on the host, geometry on the device).

    Attributes
    ----------
    node_coords : np.ndarray, shape ``(n_nodes, dim)``, float64
    owner, neighbour : np.ndarray of int32, shape ``(n_faces,)``; ``neighbour < 0`` on boundary faces
    face_node_offsets : np.ndarray of int64, shape ``(n_faces + 1,)``
    face_node_indices : np.ndarray of int32, shape ``(n_incidences,)``
    n_cells : int
    face_patches : FacePatches
    cell_zones : CellZones
    """

    node_coords: np.ndarray
    owner: np.ndarray
    neighbour: np.ndarray
    face_node_offsets: np.ndarray
    face_node_indices: np.ndarray
    n_cells: int
    face_patches: FacePatches
    cell_zones: CellZones
    _device_cache: dict = dataclasses.field(default_factory=dict, repr=False)

    # --- construction -------------------------------------------------------------------
    @classmethod
    def from_csr(
        cls,
        node_coords,
        face_node_offsets,
        face_node_indices,
        owner,
        neighbour,
        n_cells: int,
        *,
        cell_zones: dict | None = None,
        face_patches: dict | None = None,
        neighbour_offset=None,
    ) -> "Mesh":
        """Build from CSR face connectivity (``aquaflux/mesh/mesh.py:137-206``)."""
        if neighbour_offset is not None and np.any(np.asarray(neighbour_offset) != 0):
            raise ValueError("periodic meshes (a non-zero neighbour_offset) are not supported")
        owner = np.asarray(owner)
        neighbour = np.asarray(neighbour)
        for name, arr in (("owner", owner), ("neighbour", neighbour)):
            if not np.issubdtype(arr.dtype, np.integer):
                raise ValueError(f"{name} must have an integer dtype; got {arr.dtype}")
        offsets = np.asarray(face_node_offsets)
        indices = np.asarray(face_node_indices)
        if offsets.ndim != 1 or offsets.shape[0] < 1 or int(offsets[0]) != 0:
            raise ValueError("face_node_offsets must be a 1-D CSR pointer array starting at 0")
        if np.any(np.diff(offsets) < 0):
            raise ValueError("face_node_offsets must be non-decreasing (CSR row pointers)")
        if int(offsets[-1]) != indices.shape[0]:
            raise ValueError(
                f"face_node_offsets[-1] ({int(offsets[-1])}) must equal "
                f"len(face_node_indices) ({indices.shape[0]})"
            )
        zones = CellZones.from_dict(n_cells, cell_zones) if cell_zones else CellZones.default(n_cells)
        patches = FacePatches.from_dict(neighbour, face_patches or {})
        mesh = cls(
            node_coords=np.ascontiguousarray(np.asarray(node_coords, dtype=np.float64)),
            owner=np.ascontiguousarray(owner.astype(np.int32)),
            neighbour=np.ascontiguousarray(neighbour.astype(np.int32)),
            face_node_offsets=np.ascontiguousarray(offsets.astype(np.int64)),
            face_node_indices=np.ascontiguousarray(indices.astype(np.int32)),
            n_cells=int(n_cells),
            face_patches=patches,
            cell_zones=zones,
        )
        mesh.validate()
        return mesh

    @classmethod
    def from_faces(cls, node_coords, faces, owner, neighbour, n_cells: int, **kw) -> "Mesh":
        """Build from a ragged list of per-face node lists (``aquaflux/mesh/mesh.py:76-135``)."""
        offsets = [0]
        flat: list[int] = []
        for ring in faces:
            flat.extend(int(i) for i in ring)
            offsets.append(len(flat))
        return cls.from_csr(node_coords, offsets, flat, owner, neighbour, n_cells, **kw)

    def validate(self) -> None:
        """Structural checks that raise ``ValueError`` (``aquaflux/mesh/mesh.py:208-319``)."""
        nc = self.node_coords
        if nc.ndim != 2 or nc.shape[1] not in (2, 3):
            raise ValueError(f"node_coords must be (n_nodes, 2 or 3); got shape {nc.shape}")
        if not np.all(np.isfinite(nc)):
            raise ValueError("node_coords contains a non-finite value (NaN or inf)")
        n_nodes = nc.shape[0]
        if self.face_node_indices.size and (
            self.face_node_indices.min() < 0 or self.face_node_indices.max() >= n_nodes
        ):
            raise ValueError(f"face_node_indices out of range [0, {n_nodes})")
        counts = np.diff(self.face_node_offsets)
        if self.dim == 2 and np.any(counts != 2):
            bad = int(np.argmax(counts != 2))
            raise ValueError(f"face {bad} has {int(counts[bad])} node(s); a 2D face needs exactly 2")
        if self.dim == 3 and np.any(counts < 3):
            bad = int(np.argmax(counts < 3))
            raise ValueError(f"face {bad} has {int(counts[bad])} node(s); a 3D face needs >= 3")
        n_faces = self.owner.shape[0]
        if self.neighbour.shape[0] != n_faces or counts.shape[0] != n_faces:
            raise ValueError("owner, neighbour and face_node_offsets disagree on n_faces")
        if self.n_cells <= 0:
            raise ValueError(f"n_cells must be positive; got {self.n_cells}")
        if n_faces and (self.owner.min() < 0 or self.owner.max() >= self.n_cells):
            raise ValueError(f"owner index out of range [0, {self.n_cells})")
        if n_faces and (self.neighbour.min() < -1 or self.neighbour.max() >= self.n_cells):
            raise ValueError(f"neighbour index out of range [-1, {self.n_cells}) (only -1 = boundary)")
        interior = self.neighbour >= 0
        if np.any(self.owner[interior] == self.neighbour[interior]):
            raise ValueError("an interior face has owner == neighbour")
        referenced = np.zeros(self.n_cells, dtype=bool)
        referenced[self.owner] = True
        referenced[self.neighbour[interior]] = True
        if not np.all(referenced):
            raise ValueError(f"cell {int(np.argmin(referenced))} is not referenced by any face")

    # --- sizes --------------------------------------------------------------------------
    @property
    def dim(self) -> int:
        return int(self.node_coords.shape[1])

    @property
    def n_faces(self) -> int:
        return int(self.owner.shape[0])

    @property
    def n_nodes(self) -> int:
        return int(self.node_coords.shape[0])

    @property
    def interior(self) -> np.ndarray:
        return self.neighbour >= 0

    @property
    def safe_neighbour(self) -> np.ndarray:
        """Neighbour with boundary faces replaced by their owner (``connectivity.py:126-136``)."""
        return np.where(self.neighbour >= 0, self.neighbour, self.owner).astype(np.int32)

    def with_node_coords(self, node_coords) -> "Mesh":
        """Same topology, new node coordinates (the mesh-motion / sensitivity seam)."""
        return dataclasses.replace(
            self,
            node_coords=np.ascontiguousarray(np.asarray(node_coords, dtype=np.float64)),
            _device_cache={},
        )

    # --- device -------------------------------------------------------------------------
    def device_mesh(self, device=None):
        """The device-resident topology handle (built once per device and cached)."""
        from .device import DeviceMesh

        key = DeviceMesh.device_key(device)
        dm = self._device_cache.get(key)
        if dm is None:
            dm = DeviceMesh(self, device)
            self._device_cache[key] = dm
        return dm

    def geometry(self, node_coords=None, device=None):
        """Face and cell geometry computed on the GPU (``aquaflux/mesh/mesh.py:341-387``).

        ``node_coords`` may be a CUDA float64 tensor ``(n_nodes, dim)`` requiring grad, in which
        case the returned geometry is differentiable with respect to it.
        """
        from .geometry import compute_geometry

        return compute_geometry(self, node_coords=node_coords, device=device)


# --- structured generators ----------------------------------------------------------------


def _assemble(families, coords, n_cells, named_boundaries):
    """Concatenate face families ``(rings, lo, hi, lo_ok, hi_ok, lo_name, hi_name)``."""
    rings, owners, neighbours, sides = [], [], [], {}
    base = 0
    for ring, lo, hi, lo_ok, hi_ok, lo_name, hi_name in families:
        owners.append(np.where(lo_ok, lo, hi))
        neighbours.append(np.where(lo_ok & hi_ok, hi, -1))
        rings.append(ring)
        gidx = base + np.arange(lo_ok.size)
        sides[lo_name] = gidx[~lo_ok]
        sides[hi_name] = gidx[~hi_ok]
        base += lo_ok.size
    ring = np.concatenate(rings, axis=0)
    k = ring.shape[1]
    offsets = np.arange(ring.shape[0] + 1, dtype=np.int64) * k
    return Mesh.from_csr(
        coords,
        offsets,
        ring.reshape(-1),
        np.concatenate(owners),
        np.concatenate(neighbours),
        n_cells,
        face_patches=sides if named_boundaries else None,
    )


def _lattice_axis(nodes, n, length):
    if nodes is None:
        return np.linspace(0.0, length, n + 1)
    nodes = np.asarray(nodes, dtype=float)
    if nodes.shape != (n + 1,):
        raise ValueError(f"node coordinates must have shape ({n + 1},), got {nodes.shape}")
    if np.any(np.diff(nodes) <= 0.0):
        raise ValueError("node coordinates must be strictly increasing")
    return nodes


def structured_grid_2d(
    nx: int,
    ny: int,
    lx: float = 1.0,
    ly: float = 1.0,
    named_boundaries: bool = False,
    *,
    x_nodes=None,
    y_nodes=None,
    periodic: tuple = (),
) -> Mesh:
    """Quad grid on ``[0,lx] x [0,ly]``, face numbering as ``aquaflux/mesh/structured.py:155-282``."""
    if periodic:
        raise ValueError("periodic structured grids are not supported")
    x = _lattice_axis(x_nodes, nx, lx)
    y = _lattice_axis(y_nodes, ny, ly)
    # Nodes run x fastest: node (i, j) -> j*(nx+1) + i.
    coords = np.empty(((nx + 1) * (ny + 1), 2))
    coords[:, 0] = np.tile(x, ny + 1)
    coords[:, 1] = np.repeat(y, nx + 1)

    def node(i, j):
        return j * (nx + 1) + i

    def cell(i, j):
        return j * nx + i

    # x-normal family: face (i, j), i in [0, nx] slowest, j fastest.
    i = np.repeat(np.arange(nx + 1), ny)
    j = np.tile(np.arange(ny), nx + 1)
    fam_x = (
        np.stack([node(i, j), node(i, j + 1)], axis=1),
        cell(np.maximum(i - 1, 0), j),
        cell(np.minimum(i, nx - 1), j),
        i > 0,
        i < nx,
        "left",
        "right",
    )
    # y-normal family: face (i, j), i in [0, nx) slowest, j in [0, ny] fastest.
    i = np.repeat(np.arange(nx), ny + 1)
    j = np.tile(np.arange(ny + 1), nx)
    fam_y = (
        np.stack([node(i, j), node(i + 1, j)], axis=1),
        cell(i, np.maximum(j - 1, 0)),
        cell(i, np.minimum(j, ny - 1)),
        j > 0,
        j < ny,
        "bottom",
        "top",
    )
    return _assemble([fam_x, fam_y], coords, nx * ny, named_boundaries)


def structured_grid_3d(
    nx: int,
    ny: int,
    nz: int,
    lx: float = 1.0,
    ly: float = 1.0,
    lz: float = 1.0,
    named_boundaries: bool = False,
) -> Mesh:
    """Hex grid on a box, face numbering as ``aquaflux/mesh/structured.py:285-375``."""
    hx, hy, hz = lx / nx, ly / ny, lz / nz
    n_nodes = (nx + 1) * (ny + 1) * (nz + 1)
    lin = np.arange(n_nodes)
    ii = lin % (nx + 1)
    jj = (lin // (nx + 1)) % (ny + 1)
    kk = lin // ((nx + 1) * (ny + 1))
    coords = np.stack([ii * hx, jj * hy, kk * hz], axis=1)

    def node(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i

    def cell(i, j, k):
        return (k * ny + j) * nx + i

    def grid(a, b, c):
        # index triples with the first axis slowest and the last fastest
        A, B, C = np.meshgrid(np.arange(a), np.arange(b), np.arange(c), indexing="ij")
        return A.ravel(), B.ravel(), C.ravel()

    i, j, k = grid(nx + 1, ny, nz)
    fam_x = (
        np.stack([node(i, j, k), node(i, j + 1, k), node(i, j + 1, k + 1), node(i, j, k + 1)], 1),
        cell(np.maximum(i - 1, 0), j, k),
        cell(np.minimum(i, nx - 1), j, k),
        i > 0,
        i < nx,
        "left",
        "right",
    )
    i, j, k = grid(nx, ny + 1, nz)
    fam_y = (
        np.stack([node(i, j, k), node(i + 1, j, k), node(i + 1, j, k + 1), node(i, j, k + 1)], 1),
        cell(i, np.maximum(j - 1, 0), k),
        cell(i, np.minimum(j, ny - 1), k),
        j > 0,
        j < ny,
        "bottom",
        "top",
    )
    i, j, k = grid(nx, ny, nz + 1)
    fam_z = (
        np.stack([node(i, j, k), node(i + 1, j, k), node(i + 1, j + 1, k), node(i, j + 1, k)], 1),
        cell(i, j, np.maximum(k - 1, 0)),
        cell(i, j, np.minimum(k, nz - 1)),
        k > 0,
        k < nz,
        "back",
        "front",
    )
    return _assemble([fam_x, fam_y, fam_z], coords, nx * ny * nz, named_boundaries)


def _lattice_jitter(counts, extents, perturb, seed):
    """Seeded U(+-perturb*h) offsets of the strictly interior lattice nodes, zero on the boundary.

    Draw order and masking follow ``tests/support/meshes.py:24-68``: one uniform draw per axis
    (x, y[, z]) over the whole lattice in node order, multiplied by the interior mask.
    """
    steps = [length / n for length, n in zip(extents, counts)]
    shape = tuple(n + 1 for n in reversed(counts))
    grids = np.meshgrid(*[np.arange(c) for c in shape], indexing="ij")
    interior = np.ones(shape, dtype=bool)
    for g, n in zip(grids, reversed(counts)):
        interior &= (g > 0) & (g < n)
    rng = np.random.default_rng(seed)
    cols = [(interior * rng.uniform(-perturb * h, perturb * h, shape)).ravel() for h in steps]
    return np.stack(cols, axis=1)


def perturbed_grid_2d(nx, ny, lx=1.0, ly=1.0, perturb=0.2, seed=0, named_boundaries=False) -> Mesh:
    """Structured quad grid with seeded interior-node jitter (``tests/support/meshes.py:71-100``)."""
    mesh = structured_grid_2d(nx, ny, lx, ly, named_boundaries=named_boundaries)
    if not perturb:
        return mesh
    return mesh.with_node_coords(mesh.node_coords + _lattice_jitter((nx, ny), (lx, ly), perturb, seed))


def perturbed_grid_3d(
    nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, perturb=0.2, seed=0, named_boundaries=False
) -> Mesh:
    """Structured hex grid with seeded interior-node jitter (``tests/support/meshes.py:103-134``)."""
    mesh = structured_grid_3d(nx, ny, nz, lx, ly, lz, named_boundaries=named_boundaries)
    if not perturb:
        return mesh
    return mesh.with_node_coords(
        mesh.node_coords + _lattice_jitter((nx, ny, nz), (lx, ly, lz), perturb, seed)
    )


# Kuhn (Freudenthal) split of a unit cube into 6 tetrahedra sharing the v000 -> v111 diagonal:
# one tet per ordering of the three axes, walking v000 -> +e_a -> +e_a+e_b -> v111.
_KUHN_ORDERS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def tet_grid_3d(
    n: int,
    ny: int | None = None,
    nz: int | None = None,
    lx: float = 1.0,
    ly: float = 1.0,
    lz: float = 1.0,
    perturb: float = 0.0,
    seed: int = 0,
    named_boundaries: bool = True,
) -> Mesh:
    """Conforming tetrahedral mesh: every hex of an ``n x ny x nz`` box split into 6 Kuhn tets.

    Cells are numbered hex-major (``6 * hex + tet``); faces are triangles sorted by owner cell
    (owner = lower cell index), so each cell's faces sit close together in memory.  Interior
    lattice nodes are jittered by ``U(+-perturb * h)`` with the structured-grid recipe.  The
    reference has no tetrahedral generator; the mesh is a valid input to its ``Mesh.from_csr``.
    """
    nx = n
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    hx, hy, hz = lx / nx, ly / ny, lz / nz
    n_nodes = (nx + 1) * (ny + 1) * (nz + 1)
    lin = np.arange(n_nodes)
    ii = lin % (nx + 1)
    jj = (lin // (nx + 1)) % (ny + 1)
    kk = lin // ((nx + 1) * (ny + 1))
    coords = np.stack([ii * hx, jj * hy, kk * hz], axis=1)
    if perturb:
        coords = coords + _lattice_jitter((nx, ny, nz), (lx, ly, lz), perturb, seed)

    def node(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i

    n_hex = nx * ny * nz
    hexes = np.arange(n_hex, dtype=np.int64)
    hi = hexes % nx
    hj = (hexes // nx) % ny
    hk = hexes // (nx * ny)
    tets = np.empty((n_hex, 6, 4), dtype=np.int64)
    for t, order in enumerate(_KUHN_ORDERS):
        off = np.zeros(3, dtype=np.int64)
        verts = [node(hi, hj, hk)]
        for axis in order:
            off[axis] += 1
            verts.append(node(hi + off[0], hj + off[1], hk + off[2]))
        tets[:, t, :] = np.stack(verts, axis=1)
    tets = tets.reshape(-1, 4)
    n_cells = tets.shape[0]
    # The four triangles of each tet (opposite each vertex), as sorted node triples.
    tri = np.stack([tets[:, [1, 2, 3]], tets[:, [0, 2, 3]], tets[:, [0, 1, 3]], tets[:, [0, 1, 2]]], 1)
    tri = np.sort(tri.reshape(-1, 3), axis=1)
    cell_of = np.repeat(np.arange(n_cells, dtype=np.int64), 4)
    key = (tri[:, 0] << 42) | (tri[:, 1] << 21) | tri[:, 2]
    order = np.argsort(key, kind="stable")
    key_sorted = key[order]
    first = np.ones(key_sorted.size, dtype=bool)
    first[1:] = key_sorted[1:] != key_sorted[:-1]
    starts = np.nonzero(first)[0]
    counts = np.diff(np.append(starts, key_sorted.size))
    if np.any(counts > 2):
        raise RuntimeError("non-manifold tetrahedral split")
    c0 = cell_of[order[starts]]
    c1 = np.where(counts == 2, cell_of[order[np.minimum(starts + 1, key_sorted.size - 1)]], -1)
    own = np.where(c1 >= 0, np.minimum(c0, c1), c0)
    nbr = np.where(c1 >= 0, np.maximum(c0, c1), -1)
    rings = tri[order[starts]]
    # Owner-sorted face order (ties: interior before boundary, then by neighbour).
    face_order = np.lexsort((np.where(nbr >= 0, nbr, np.iinfo(np.int64).max), own))
    own, nbr, rings = own[face_order], nbr[face_order], rings[face_order]
    patches = None
    if named_boundaries:
        bnd = np.nonzero(nbr < 0)[0]
        lat = np.stack([ii, jj, kk], axis=1)[rings[bnd]]  # (nb, 3 nodes, 3 axes)
        patches = {}
        for axis, (lo, hi_name, top) in enumerate(
            (("left", "right", nx), ("bottom", "top", ny), ("back", "front", nz))
        ):
            patches[lo] = bnd[np.all(lat[:, :, axis] == 0, axis=1)]
            patches[hi_name] = bnd[np.all(lat[:, :, axis] == top, axis=1)]
    offsets = np.arange(own.size + 1, dtype=np.int64) * 3
    return Mesh.from_csr(
        coords, offsets, rings.reshape(-1), own, nbr, n_cells, face_patches=patches
    )
