"""Oracle: loop-form restatement of the structured-grid connectivity (test infrastructure only).

Restates ``aquaflux/mesh/structured.py:155-375`` face by face with plain Python loops (small
grids only), independent of the vectorised generator it checks: families x, then y, then z; the
family-axis index slowest; owner = lower cell, neighbour = higher cell or -1; patches
left/right/bottom/top/back/front.  Pinned by the reference's structured-grid facts (face and
patch counts) in the tests.
"""

from __future__ import annotations


def structured_2d(nx, ny):
    node = lambda i, j: j * (nx + 1) + i  # noqa: E731
    cell = lambda i, j: j * nx + i  # noqa: E731
    rings, owner, nb = [], [], []
    patches = {"left": [], "right": [], "bottom": [], "top": []}
    for i in range(nx + 1):
        for j in range(ny):
            f = len(rings)
            rings.append([node(i, j), node(i, j + 1)])
            lo, hi = i - 1, i
            if lo >= 0 and hi < nx:
                owner.append(cell(lo, j)), nb.append(cell(hi, j))
            elif lo < 0:
                owner.append(cell(hi, j)), nb.append(-1), patches["left"].append(f)
            else:
                owner.append(cell(lo, j)), nb.append(-1), patches["right"].append(f)
    for i in range(nx):
        for j in range(ny + 1):
            f = len(rings)
            rings.append([node(i, j), node(i + 1, j)])
            lo, hi = j - 1, j
            if lo >= 0 and hi < ny:
                owner.append(cell(i, lo)), nb.append(cell(i, hi))
            elif lo < 0:
                owner.append(cell(i, hi)), nb.append(-1), patches["bottom"].append(f)
            else:
                owner.append(cell(i, lo)), nb.append(-1), patches["top"].append(f)
    return rings, owner, nb, patches


def structured_3d(nx, ny, nz):
    node = lambda i, j, k: (k * (ny + 1) + j) * (nx + 1) + i  # noqa: E731
    cell = lambda i, j, k: (k * ny + j) * nx + i  # noqa: E731
    rings, owner, nb = [], [], []
    patches = {n: [] for n in ("left", "right", "bottom", "top", "back", "front")}

    def add(ring, lo_cell, hi_cell, lo_ok, hi_ok, lo_name, hi_name):
        f = len(rings)
        rings.append(ring)
        if lo_ok and hi_ok:
            owner.append(lo_cell), nb.append(hi_cell)
        elif not lo_ok:
            owner.append(hi_cell), nb.append(-1), patches[lo_name].append(f)
        else:
            owner.append(lo_cell), nb.append(-1), patches[hi_name].append(f)

    for i in range(nx + 1):
        for j in range(ny):
            for k in range(nz):
                add([node(i, j, k), node(i, j + 1, k), node(i, j + 1, k + 1), node(i, j, k + 1)],
                    cell(i - 1, j, k), cell(i, j, k), i > 0, i < nx, "left", "right")
    for i in range(nx):
        for j in range(ny + 1):
            for k in range(nz):
                add([node(i, j, k), node(i + 1, j, k), node(i + 1, j, k + 1), node(i, j, k + 1)],
                    cell(i, j - 1, k), cell(i, j, k), j > 0, j < ny, "bottom", "top")
    for i in range(nx):
        for j in range(ny):
            for k in range(nz + 1):
                add([node(i, j, k), node(i + 1, j, k), node(i + 1, j + 1, k), node(i, j + 1, k)],
                    cell(i, j, k - 1), cell(i, j, k), k > 0, k < nz, "back", "front")
    return rings, owner, nb, patches
