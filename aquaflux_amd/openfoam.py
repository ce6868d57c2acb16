"""OpenFOAM ASCII polyMesh import: a case or polyMesh directory to a :class:`~aquaflux_amd.mesh.Mesh`.

The reader mirrors ``aquaflux.io.openfoam`` (``io/openfoam/reader.py:50-146``,
``assembler.py:24-115``, ``foamfile.py:22-120``, ``grammar.py:27-136``, ``records.py:15-84``) and
``aquaflux.mesh.collapse_extruded_direction`` (``mesh/collapse.py:37-260``): same entry points,
records, argument meaning and errors.  The parsed face lists go straight into the mesh handle the
residual kernels use, so general polyhedra (any number of faces per cell, any number of nodes per
face) reach the GPU geometry and residual kernels without a conversion step.

Parsing here is numpy-first: a list body's element count is checked from bracket counts, its
numbers are converted in one ``numpy`` call, and the closing bracket of a list is found from a
vectorised nesting-depth scan, so multi-million-face files parse in seconds.
"""

from __future__ import annotations

import re
from pathlib import Path
from typing import NamedTuple

import numpy as np

from .mesh import Mesh

__all__ = [
    "FoamFile",
    "FoamPatch",
    "CellZone",
    "PolyMeshData",
    "strip_comments",
    "parse_dict_entries",
    "parse_foamfile",
    "is_binary",
    "read_foam_body",
    "list_envelope",
    "parse_vector_list",
    "parse_scalar_list",
    "parse_face_list",
    "parse_boundary",
    "parse_cell_zones",
    "assemble",
    "collapse_extruded_direction",
    "OpenFOAMReader",
    "read_openfoam",
    "parse_scalar_field",
    "read_volume_scalar_field",
    "read_surface_scalar_field",
]

# patch names the mesh assigns by itself from the boundary mask
RESERVED_PATCH_NAMES = ("interior", "boundary")
# cap faces are flat along the extruded axis to this fraction of the mesh extent, and in-plane node
# positions are merged at the same resolution (``mesh/collapse.py:33-35``)
COLLAPSE_TOLERANCE = 1e-8


# --- records (io/openfoam/records.py:15-84) ----------------------------------------------------
class FoamPatch(NamedTuple):
    """A boundary patch: faces ``[start_face, start_face + n_faces)``; ``type_`` is carried through
    (only ``empty`` is acted on: it marks the caps of a two-dimensional case)."""

    name: str
    type_: str
    start_face: int
    n_faces: int


class CellZone(NamedTuple):
    name: str
    cell_labels: np.ndarray


class PolyMeshData(NamedTuple):
    """The parsed arrays of a polyMesh directory before assembly.  ``neighbour_internal`` has one
    entry per interior face (the raw ``neighbour`` file)."""

    points: np.ndarray
    face_node_offsets: np.ndarray
    face_node_indices: np.ndarray
    owner: np.ndarray
    neighbour_internal: np.ndarray
    patches: tuple
    cell_zones: tuple


class FoamFile(NamedTuple):
    header: dict
    body: str


# --- file envelope (io/openfoam/foamfile.py:22-120) ----------------------------------------------
_COMMENT = re.compile(r"/\*.*?\*/|//[^\n]*", re.DOTALL)
_HEADER = re.compile(r"\bFoamFile\s*\{([^{}]*)\}", re.DOTALL)
_ENTRY = re.compile(r"(\w+)\s+([^;{}]+);")


def strip_comments(text: str) -> str:
    """``/* ... */`` and ``// ...`` comments replaced by one space each."""
    return _COMMENT.sub(" ", text)


def parse_dict_entries(block: str) -> dict:
    """``key value;`` pairs of a flat dictionary body, values stripped of blanks and quotes."""
    return {k: v.strip().strip('"') for k, v in _ENTRY.findall(block)}


def parse_foamfile(text: str) -> FoamFile:
    """Split a file into its ``FoamFile { ... }`` header and the payload after it.

    Raises ``ValueError`` when there is no header."""
    clean = strip_comments(text)
    m = _HEADER.search(clean)
    if m is None:
        raise ValueError("not an OpenFOAM file: no 'FoamFile { ... }' header found")
    return FoamFile(parse_dict_entries(m.group(1)), clean[m.end():])


def is_binary(foam: FoamFile) -> bool:
    return foam.header.get("format", "ascii").strip().lower() == "binary"


def read_foam_body(path) -> str:
    """The payload of an ASCII OpenFOAM file.  ``FileNotFoundError`` when absent,
    ``NotImplementedError`` for ``format binary;`` (detected, never misread)."""
    path = Path(path)
    if not path.exists():
        raise FileNotFoundError(f"OpenFOAM file not found: {path}")
    foam = parse_foamfile(path.read_text())
    if is_binary(foam):
        raise NotImplementedError(f"{path} is in binary format; only ASCII files are supported")
    return foam.body


# --- list grammars (io/openfoam/grammar.py:27-136) -----------------------------------------------
_COUNT_OPEN = re.compile(r"(\d+)\s*\(")
_BLOCK = re.compile(r"(\w+)\s*\{(.*?)\}", re.DOTALL)
_CELL_LABELS = re.compile(r"cellLabels\s+(?:List<label>\s*)?(\d+)\s*\((.*?)\)", re.DOTALL)


def list_envelope(body: str) -> tuple[int, str]:
    """The declared count ``N`` and the text inside the brackets of the leading ``N ( ... )`` list,
    brackets matched by nesting depth.  ``ValueError`` when absent or unbalanced."""
    m = _COUNT_OPEN.search(body)
    if m is None:
        raise ValueError("expected a 'N ( ... )' list body")
    start = m.end() - 1
    raw = np.frombuffer(body[start:].encode("latin-1", "replace"), dtype=np.uint8)
    step = (raw == ord("(")).astype(np.int32) - (raw == ord(")")).astype(np.int32)
    depth = np.cumsum(step)
    closed = np.flatnonzero(depth == 0)
    if closed.size == 0:
        raise ValueError("unbalanced parentheses in list body")
    return int(m.group(1)), body[start + 1: start + int(closed[0])]


def _count_check(kind: str, declared: int, found: int) -> None:
    if declared != found:
        raise ValueError(f"{kind} declares {declared} entries but lists {found}")


def _numbers(text: str, dtype) -> np.ndarray:
    tokens = text.replace("(", " ").replace(")", " ").split()
    return np.array(tokens, dtype=dtype) if tokens else np.zeros(0, dtype=dtype)


def parse_vector_list(body: str) -> np.ndarray:
    """A ``vectorField`` body (``points``) as ``(n, 3)`` float64."""
    count, inner = list_envelope(body)
    _count_check("vector list", count, inner.count("("))
    if count == 0:
        return np.zeros((0, 3), dtype=np.float64)
    values = _numbers(inner, np.float64)
    if values.size != 3 * count:
        raise ValueError("each point must have three coordinates")
    return values.reshape(count, 3)


def parse_scalar_list(body: str) -> np.ndarray:
    """A ``labelList`` body (``owner`` / ``neighbour``) as int64."""
    count, inner = list_envelope(body)
    values = _numbers(inner, np.int64)
    _count_check("label list", count, values.size)
    return values


def parse_face_list(body: str) -> tuple[np.ndarray, np.ndarray]:
    """A ``faceList`` body (``faces``: ``n(i0 ... i_{n-1})`` per face, nodes in perimeter order) as
    CSR ``(offsets, indices)``."""
    count, inner = list_envelope(body)
    groups = re.findall(r"\(([^()]*)\)", inner)
    _count_check("face list", count, len(groups))
    listed = np.array([len(g.split()) for g in groups], dtype=np.int64)
    offsets = np.zeros(count + 1, dtype=np.int64)
    np.cumsum(listed, out=offsets[1:])
    # after bracket removal the tokens run n, i0 .. i_{n-1}, n, ...: the size tokens sit at
    # offsets[f] + f
    flat = _numbers(inner, np.int64)
    if flat.size != int(offsets[-1]) + count:
        raise ValueError("every face must be written as n(i0 ... i_{n-1})")
    heads = offsets[:-1] + np.arange(count, dtype=np.int64)
    bad = np.flatnonzero(flat[heads] != listed)
    if bad.size:
        f = int(bad[0])
        raise ValueError(f"a face declares {int(flat[heads[f]])} nodes but lists {int(listed[f])}")
    keep = np.ones(flat.size, dtype=bool)
    keep[heads] = False
    return offsets, flat[keep]


def parse_boundary(body: str) -> tuple:
    """A ``polyBoundaryMesh`` body (``boundary``) as :class:`FoamPatch` records in file order."""
    count, inner = list_envelope(body)
    patches = []
    for name, block in _BLOCK.findall(inner):
        e = {k: v.strip() for k, v in _ENTRY.findall(block)}
        missing = [k for k in ("startFace", "nFaces") if k not in e]
        if missing:
            raise ValueError(f"boundary patch '{name}' is missing '{missing[0]}'")
        patches.append(FoamPatch(name, e.get("type", ""), int(e["startFace"]), int(e["nFaces"])))
    _count_check("boundary", count, len(patches))
    return tuple(patches)


def parse_cell_zones(body: str) -> tuple:
    """A ``cellZones`` body as :class:`CellZone` records (name and cell labels)."""
    count, inner = list_envelope(body)
    zones = []
    for name, block in _BLOCK.findall(inner):
        m = _CELL_LABELS.search(block)
        if m is None:
            raise ValueError(f"cellZone '{name}' has no cellLabels list")
        labels = _numbers(m.group(2), np.int64)
        _count_check(f"cellZone '{name}'", int(m.group(1)), labels.size)
        zones.append(CellZone(name, labels))
    _count_check("cellZones", count, len(zones))
    return tuple(zones)


# --- assembly (io/openfoam/assembler.py:24-115) ------------------------------------------------
def assemble(data: PolyMeshData) -> Mesh:
    """The 3D mesh of parsed polyMesh arrays: the interior-only neighbour list padded with ``-1``
    (OpenFOAM numbers interior faces first), ``n_cells`` one past the largest cell label, patches
    as contiguous face ranges, zones by label.  ``ValueError`` for an empty face list, a neighbour
    list longer than the face list, or a patch named ``interior`` / ``boundary``; topology errors
    come from :meth:`Mesh.from_csr`."""
    owner = np.asarray(data.owner, dtype=np.int64)
    nbi = np.asarray(data.neighbour_internal, dtype=np.int64)
    F = owner.shape[0]
    if F == 0:
        raise ValueError("polyMesh has no faces")
    if nbi.shape[0] > F:
        raise ValueError(f"neighbour list ({nbi.shape[0]}) is longer than the face count ({F})")
    neighbour = np.full(F, -1, dtype=np.int64)
    neighbour[: nbi.shape[0]] = nbi
    n_cells = int(max(owner.max(), nbi.max() if nbi.size else -1)) + 1
    clash = [p.name for p in data.patches if p.name in RESERVED_PATCH_NAMES]
    if clash:
        raise ValueError(
            f"boundary patch name(s) {clash} collide with the reserved patch names {list(RESERVED_PATCH_NAMES)}; "
            "rename them in the polyMesh boundary file"
        )
    patches = {p.name: np.arange(p.start_face, p.start_face + p.n_faces) for p in data.patches} or None
    zones = {z.name: np.asarray(z.cell_labels) for z in data.cell_zones} or None
    return Mesh.from_csr(data.points, data.face_node_offsets, data.face_node_indices, owner, neighbour,
                         n_cells=n_cells, cell_zones=zones, face_patches=patches)


# --- two-dimensional cases (mesh/collapse.py:37-260) ---------------------------------------------
def collapse_extruded_direction(mesh: Mesh, removed_patch_names) -> Mesh:
    """The 2D mesh of a one-cell-thick extrusion capped by the named patches (two caps or one
    ``frontAndBack`` patch): the caps and the extruded axis are dropped, coincident front / back
    nodes merge, every side quad becomes the edge between its two in-plane positions.  Cells, their
    order and zones are kept; surviving patches are renumbered onto the kept faces.

    ``ValueError`` if the mesh is not 3D, no caps are named or they hold no faces, a cap is not
    flat along exactly one shared axis, the caps do not lie on exactly two planes, or a side face
    does not reduce to two in-plane nodes."""
    removed = list(removed_patch_names)
    if not removed:
        raise ValueError("expected at least one capping patch to remove; got none")
    X = np.asarray(mesh.node_coords)
    if X.shape[1] != 3:
        raise ValueError(f"collapse_extruded_direction expects a 3D mesh; got dim {mesh.dim}")
    extent = float(np.max(X.max(axis=0) - X.min(axis=0)))
    if extent <= 0.0:
        raise ValueError("mesh has zero spatial extent")
    tol = COLLAPSE_TOLERANCE * extent
    off = np.asarray(mesh.face_node_offsets)
    idx = np.asarray(mesh.face_node_indices)
    is_cap = np.zeros(mesh.n_faces, dtype=bool)
    for name in removed:
        is_cap |= mesh.face_patches.mask(name)
    caps = np.flatnonzero(is_cap)
    if caps.size == 0:
        raise ValueError(f"capping patches {removed} contain no faces")

    # the axis every cap is flat along, and the cap planes' positions on it
    axis = None
    levels = np.empty(caps.size)
    for n, f in enumerate(caps):
        pts = X[idx[off[f]: off[f + 1]]]
        flat_axes = np.flatnonzero(pts.max(axis=0) - pts.min(axis=0) <= tol)
        if flat_axes.size != 1:
            raise ValueError(
                f"capping face {int(f)} is not planar and normal to a single axis (constant along "
                f"{flat_axes.size} axes); patches {removed} are not extrusion caps"
            )
        if axis is None:
            axis = int(flat_axes[0])
        elif int(flat_axes[0]) != axis:
            raise ValueError(
                f"capping faces are normal to different axes ({axis} and {int(flat_axes[0])}); "
                f"patches {removed} must be the two ends of one extrusion"
            )
        levels[n] = pts[:, axis].mean()
    lv = np.sort(levels)
    n_planes = 1 + int(np.count_nonzero(np.diff(lv) > tol))
    if n_planes != 2:
        raise ValueError(
            f"capping patches {removed} must lie on exactly two parallel planes normal to one axis "
            f"(the front and back of the extrusion); found {n_planes}"
        )

    # merge nodes that share an in-plane position (quantised at the tolerance)
    keep_axes = [a for a in range(3) if a != axis]
    plane = X[:, keep_axes]
    key = np.round(plane / tol).astype(np.int64)
    _, node_of = np.unique(key, axis=0, return_inverse=True)
    node_of = node_of.reshape(-1)
    coords2 = np.zeros((int(node_of.max()) + 1, 2), dtype=np.float64)
    coords2[node_of] = plane

    # each side face -> the edge between its first node and the first node at another position
    kept = np.flatnonzero(~is_cap)
    edges = np.empty((kept.size, 2), dtype=np.int64)
    for r, f in enumerate(kept):
        ring = node_of[idx[off[f]: off[f + 1]]]
        first = ring[0]
        others = ring[ring != first]
        distinct = 1 + np.unique(others).size
        if distinct != 2:
            raise ValueError(
                f"face {int(f)} reduces to {distinct} distinct in-plane node(s), not 2; the mesh is "
                "not a one-cell-thick extrusion of the removed patches"
            )
        edges[r] = (first, others[0])

    renumber = np.full(mesh.n_faces, -1, dtype=np.int64)
    renumber[kept] = np.arange(kept.size)
    skip = set(RESERVED_PATCH_NAMES) | set(removed)
    patches = {name: renumber[mesh.face_patches.indices(name)] for name in mesh.face_patches.names
               if name not in skip} or None
    zones = {name: mesh.cell_zones.indices(name) for name in mesh.cell_zones.names if name != "default"} or None
    return Mesh.from_csr(coords2, np.arange(kept.size + 1, dtype=np.int64) * 2, edges.reshape(-1),
                         mesh.owner[kept], mesh.neighbour[kept], n_cells=mesh.n_cells, cell_zones=zones,
                         face_patches=patches)


# --- scalar fields on an imported mesh (io/openfoam/fields.py:1-229) ------------------------------
_BOUNDARY_FIELD = re.compile(r"\bboundaryField\s*\{")
_INTERNAL_FIELD = re.compile(r"\binternalField\b")
_UNIFORM = re.compile(r"\buniform\s+([-+0-9.eE]+)")


def _closing_brace(text: str, at: int) -> int:
    """Index of the ``}`` closing the ``{`` at ``at`` (vectorised nesting depth)."""
    raw = np.frombuffer(text[at:].encode("latin-1", "replace"), dtype=np.uint8)
    depth = np.cumsum((raw == ord("{")).astype(np.int32) - (raw == ord("}")).astype(np.int32))
    closed = np.flatnonzero(depth == 0)
    if closed.size == 0:
        raise ValueError("unbalanced braces")
    return at + int(closed[0])


def _entry_values(text: str, count: int, what: str) -> np.ndarray:
    """``count`` values of a ``uniform X`` or ``nonuniform List<scalar> N ( ... )`` entry (a file
    mixes both: walls are usually ``uniform 0``)."""
    if "nonuniform" in text:
        declared, inner = list_envelope(text)
        vals = _numbers(inner, np.float64)
        if declared != vals.size:
            raise ValueError(f"{what} declares {declared} values but lists {vals.size}")
        if declared != count:
            raise ValueError(f"{what} has {declared} values but the patch has {count} faces")
        return vals
    m = _UNIFORM.search(text)
    if m is None:
        raise ValueError(f"{what} is neither a 'uniform' nor a 'nonuniform' entry")
    return np.full(count, float(m.group(1)), dtype=np.float64)


def parse_scalar_field(body: str, n_internal: int, patch_sizes: dict) -> np.ndarray:
    """``[internal, patch_0, patch_1, ...]`` of a scalar-field body, patches in ``patch_sizes``
    order.  ``ValueError`` when the internal block or a listed patch is missing or of the wrong
    length."""
    mi = _INTERNAL_FIELD.search(body)
    if mi is None:
        raise ValueError("field has no internalField entry")
    mb = _BOUNDARY_FIELD.search(body)
    parts = [_entry_values(body[mi.end(): mb.start() if mb else len(body)], n_internal, "internalField")]
    if patch_sizes and mb is None:
        raise ValueError("field has no boundaryField entry")
    blocks = {}
    if mb is not None:
        lo = mb.end() - 1
        inner = body[lo + 1: _closing_brace(body, lo)]
        pos = 0
        while True:
            brace = inner.find("{", pos)
            if brace < 0:
                break
            words = inner[pos:brace].split()
            end = _closing_brace(inner, brace)
            blocks[words[-1] if words else ""] = inner[brace + 1: end]
            pos = end + 1
    for name, size in patch_sizes.items():
        if name not in blocks:
            raise ValueError(f"field has no boundaryField entry for patch '{name}'")
        parts.append(_entry_values(blocks[name], size, f"patch '{name}'"))
    return np.concatenate(parts)


def read_volume_scalar_field(path, mesh: Mesh) -> np.ndarray:
    """The ``internalField`` of a ``volScalarField`` on the mesh's cells (``(n_cells,)``).  Cell
    labels are OpenFOAM's own (the assembler keeps them), so no ordering check is needed."""
    return parse_scalar_field(read_foam_body(path), mesh.n_cells, {})


def read_surface_scalar_field(path, mesh: Mesh) -> np.ndarray:
    """A ``surfaceScalarField`` (the face flux ``phi``) on the mesh's faces, ``(n_faces,)``,
    owner-outward.  Value ``i`` lands on face ``i`` only when the mesh keeps OpenFOAM's face order
    (interior faces first, every patch a contiguous block); that order is checked, and a mesh
    without it (a generated grid, or a collapsed 2D case, which is renumbered) is refused."""
    interior = mesh.interior
    n_int = int(interior.sum())
    if not interior[:n_int].all() or interior[n_int:].any():
        raise ValueError(
            "mesh face ordering is not OpenFOAM's (interior faces are not the leading block), so field values "
            "cannot be placed by index; a 2D case collapsed from an empty-capped polyMesh is renumbered and "
            "cannot be used here"
        )
    sizes = {}
    for name in mesh.face_patches.names:
        if name == "interior":
            continue
        idx = mesh.face_patches.indices(name)
        if idx.size == 0:
            continue
        if name == "boundary":
            raise ValueError(f"{idx.size} boundary faces are not in a named patch, so the field has no values "
                             "for them; the polyMesh must tile its boundary with named patches")
        if idx[0] < n_int:
            raise ValueError(f"patch '{name}' includes an interior face; ordering is not OpenFOAM's")
        if idx[-1] - idx[0] + 1 != idx.size:
            raise ValueError(f"patch '{name}' is not a contiguous block of faces")
        sizes[name] = (int(idx[0]), int(idx.size))
    ordered = {name: n for name, (_, n) in sorted(sizes.items(), key=lambda kv: kv[1][0])}
    values = parse_scalar_field(read_foam_body(path), n_int, ordered)
    if values.shape[0] != mesh.n_faces:
        raise ValueError(f"field has {values.shape[0]} values but the mesh has {mesh.n_faces} faces")
    return values


# --- the reader (io/openfoam/reader.py:34-146) ---------------------------------------------------
def _polymesh_dir(path) -> Path:
    path = Path(path)
    for cand in (path, path / "constant" / "polyMesh"):
        if (cand / "points").exists():
            return cand
    raise FileNotFoundError(f"no polyMesh found at {path} (looked for a 'points' file there and in constant/polyMesh)")


class OpenFOAMReader:
    """Reads a polyMesh directory, or a case directory holding ``constant/polyMesh``."""

    def __init__(self, directory):
        self.directory = str(_polymesh_dir(directory))

    def _read(self, name, parser, required=True):
        path = Path(self.directory) / name
        if not path.exists():
            if required:
                raise FileNotFoundError(f"polyMesh file not found: {path}")
            return None
        return parser(read_foam_body(path))

    def read_polymesh(self) -> PolyMeshData:
        """Parsed points, faces, owner, interior neighbours, patches and zones (``neighbour`` and
        ``cellZones`` are optional)."""
        offsets, indices = self._read("faces", parse_face_list)
        nb = self._read("neighbour", parse_scalar_list, required=False)
        zones = self._read("cellZones", parse_cell_zones, required=False)
        return PolyMeshData(
            points=self._read("points", parse_vector_list),
            face_node_offsets=offsets,
            face_node_indices=indices,
            owner=self._read("owner", parse_scalar_list),
            neighbour_internal=np.zeros(0, dtype=np.int64) if nb is None else nb,
            patches=self._read("boundary", parse_boundary),
            cell_zones=() if zones is None else zones,
        )

    def read(self) -> Mesh:
        """The mesh, collapsed to 2D when ``empty`` patches cap a one-cell-thick case."""
        data = self.read_polymesh()
        mesh = assemble(data)
        empty = [p.name for p in data.patches if p.type_ == "empty"]
        return collapse_extruded_direction(mesh, empty) if empty else mesh


def read_openfoam(directory) -> Mesh:
    return OpenFOAMReader(directory).read()
