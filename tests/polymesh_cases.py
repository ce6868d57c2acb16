"""General polyhedral test meshes written as OpenFOAM polyMesh directories.

The "brick wall" mesh: rows of lx/nx x ly/ny bricks, every other row shifted by half a brick (half
bricks close the shifted rows), with a node at every half-brick position on each row line.  A full
brick is a hexagon in the plane (four corners and the two mid-points it shares with the shifted
rows above and below) and has six neighbours; extruded over nz layers it is an eight-faced prism
whose caps are hexagons.  Interior nodes are jittered in the plane (seeded).  The directory is
written in OpenFOAM order: interior faces first, sorted by (owner, neighbour) with owner <
neighbour, then the boundary patches left, right, bottom, top, back, front, each face's node ring
oriented so its right-hand normal leaves the owner.  With ``nz == 1`` and ``empty`` caps it is a
two-dimensional case that the reader collapses."""

from __future__ import annotations

import os
import tempfile

import numpy as np

PATCHES = ("left", "right", "bottom", "top", "back", "front")


def brick_polymesh(nx, ny, nz, lx=2.0, ly=1.0, lz=0.5, perturb=0.15, seed=0, empty_caps=False):
    """``points (N, 3)``, ``faces`` (list of node rings), ``owner``, ``neighbour_internal`` and
    ``patches`` [(name, type, start, count)] of the brick-wall mesh, in polyMesh order."""
    rng = np.random.default_rng(seed)
    K = 2 * nx + 1                      # half-brick positions per row line
    hx, hy = lx / nx, ly / ny
    node2 = lambda k, j: j * K + k      # noqa: E731
    xy = np.zeros((K * (ny + 1), 2))
    for j in range(ny + 1):
        for k in range(K):
            x, y = 0.5 * k * hx, j * hy
            if 0 < k < K - 1 and 0 < j < ny:
                x += rng.uniform(-perturb, perturb) * 0.5 * hx
                y += rng.uniform(-perturb, perturb) * hy
            xy[node2(k, j)] = (x, y)
    # 2D cells as (k_first, k_last, row); rings counter-clockwise
    cells2 = []
    for j in range(ny):
        if j % 2 == 0:
            spans = [(2 * i, 2 * i + 2) for i in range(nx)]
        else:
            spans = [(0, 1)] + [(2 * i - 1, 2 * i + 1) for i in range(1, nx)] + [(K - 2, K - 1)]
        cells2 += [(a, b, j) for a, b in spans]
    rings2 = [[node2(k, j) for k in range(a, b + 1)] + [node2(k, j + 1) for k in range(b, a - 1, -1)]
              for a, b, j in cells2]
    C2, N2 = len(rings2), xy.shape[0]
    # 2D edges -> cells
    edge_cells: dict = {}
    for c, ring in enumerate(rings2):
        for i in range(len(ring)):
            a, b = ring[i], ring[(i + 1) % len(ring)]
            edge_cells.setdefault((min(a, b), max(a, b)), []).append(c)

    def side_patch(a, b):
        (ka, ja), (kb, jb) = divmod(a, K)[::-1], divmod(b, K)[::-1]
        if ka == kb == 0:
            return "left"
        if ka == kb == K - 1:
            return "right"
        if ja == jb == 0:
            return "bottom"
        if ja == jb == ny:
            return "top"
        raise AssertionError("boundary edge off the domain sides")

    z = np.linspace(0.0, lz, nz + 1)
    pts = np.concatenate([np.column_stack([xy, np.full(N2, z[l])]) for l in range(nz + 1)])
    interior, boundary = [], {p: [] for p in PATCHES}
    for (a, b), cs in edge_cells.items():
        for l in range(nz):
            ring = [a + l * N2, b + l * N2, b + (l + 1) * N2, a + (l + 1) * N2]
            cells = [c + l * C2 for c in cs]
            if len(cells) == 2:
                interior.append((min(cells), max(cells), ring))
            else:
                boundary[side_patch(a, b)].append((cells[0], ring))
    for c, ring in enumerate(rings2):
        for l in range(nz + 1):
            lifted = [n + l * N2 for n in ring]
            if l == 0:
                boundary["back"].append((c, lifted))
            elif l == nz:
                boundary["front"].append((c + (l - 1) * C2, lifted))
            else:
                interior.append((c + (l - 1) * C2, c + l * C2, lifted))
    interior.sort(key=lambda t: (t[0], t[1]))
    faces, owner, neighbour = [], [], []
    for o, n, ring in interior:
        faces.append(ring)
        owner.append(o)
        neighbour.append(n)
    patches = []
    for p in PATCHES:
        rows = sorted(boundary[p], key=lambda t: t[0])
        kind = "empty" if empty_caps and p in ("back", "front") else ("wall" if p in ("bottom", "top") else "patch")
        patches.append((p, kind, len(faces), len(rows)))
        for o, ring in rows:
            faces.append(ring)
            owner.append(o)
    # orient every ring so its right-hand normal leaves the owner
    C = C2 * nz
    cell_sum = np.zeros((C, 3))
    cell_cnt = np.zeros(C)
    for f, ring in enumerate(faces):
        cell_sum[owner[f]] += pts[ring].mean(axis=0)
        cell_cnt[owner[f]] += 1
        if f < len(neighbour):
            cell_sum[neighbour[f]] += pts[ring].mean(axis=0)
            cell_cnt[neighbour[f]] += 1
    centre = cell_sum / cell_cnt[:, None]
    for f, ring in enumerate(faces):
        p = pts[ring]
        c = p.mean(axis=0)
        area = sum(np.cross(p[i] - c, p[(i + 1) % len(ring)] - c) for i in range(len(ring)))
        if np.dot(area, c - centre[owner[f]]) < 0.0:
            faces[f] = ring[::-1]
    return dict(points=pts, faces=faces, owner=np.asarray(owner), neighbour_internal=np.asarray(neighbour),
                patches=patches, n_cells=C)


def _header(cls, obj):
    return ("/* brick-wall polyhedral test mesh */\nFoamFile\n{\n    version     2.0;\n    format      ascii;\n"
            f"    class       {cls};\n    location    \"constant/polyMesh\";\n    object      {obj};\n}}\n\n")


def write_polymesh(directory, data):
    """Writes ``points``, ``faces``, ``owner``, ``neighbour`` and ``boundary`` under ``directory``."""
    os.makedirs(directory, exist_ok=True)
    pts = data["points"]

    def put(name, cls, rows):
        with open(os.path.join(directory, name), "w") as fh:
            fh.write(_header(cls, name))
            fh.write(f"{len(rows)}\n(\n")
            fh.write("\n".join(rows))
            fh.write("\n)\n")

    put("points", "vectorField", ["(" + " ".join(repr(float(v)) for v in p) + ")" for p in pts])
    put("faces", "faceList", [f"{len(r)}(" + " ".join(str(int(i)) for i in r) + ")" for r in data["faces"]])
    put("owner", "labelList", [str(int(o)) for o in data["owner"]])
    put("neighbour", "labelList", [str(int(n)) for n in data["neighbour_internal"]])
    with open(os.path.join(directory, "boundary"), "w") as fh:
        fh.write(_header("polyBoundaryMesh", "boundary"))
        fh.write(f"{len(data['patches'])}\n(\n")
        for name, kind, start, count in data["patches"]:
            fh.write(f"    {name}\n    {{\n        type            {kind};\n        nFaces          {count};\n"
                     f"        startFace       {start};\n    }}\n")
        fh.write(")\n")


def brick_mesh(dim, nx=6, ny=4, nz=3, seed=0, perturb=0.15):
    """The brick-wall mesh read back through the OpenFOAM reader: 3D (nz layers) or 2D (one layer
    with ``empty`` caps, collapsed by the reader)."""
    from aquaflux_amd.openfoam import read_openfoam

    data = brick_polymesh(nx, ny, nz if dim == 3 else 1, perturb=perturb, seed=seed, empty_caps=dim == 2)
    with tempfile.TemporaryDirectory() as tmp:
        write_polymesh(tmp, data)
        return read_openfoam(tmp)
