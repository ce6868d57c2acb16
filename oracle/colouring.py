"""CPU restatement of the cell-block sparsity pattern and the saturation-degree colouring used for
compressed Jacobian assembly (test infrastructure only).

Follows ``aquaflux/solve/sparse_jacobian.py``: the pattern at ``reach`` is the ``reach``-th power
of the cell adjacency with self loops (``block_stencil_colouring``, ``:338-406``), the conflict
graph is the pattern squared, and the colouring picks, each round, the uncoloured vertex whose
neighbours show the most distinct colours, ties by degree (``_saturation_colouring``,
``:409-534``).  The selection here is a full scan with the smallest-index tie-break, like the
reference test's ``_naive_saturation_colouring`` (``tests/unit/test_sparse_jacobian.py:120-156``):
it pins validity and colour count, not a vertex-for-vertex order, which the reference itself
leaves to its bucket queue.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def adjacency(owner, nb, n):
    """Cell adjacency with self loops, as a 0/1 CSR matrix."""
    owner = np.asarray(owner, dtype=np.int64)
    nb = np.asarray(nb, dtype=np.int64)
    keep = owner != nb
    owner, nb = owner[keep], nb[keep]
    rows = np.concatenate([owner, nb, np.arange(n)])
    cols = np.concatenate([nb, owner, np.arange(n)])
    a = sp.coo_matrix((np.ones(rows.size), (rows, cols)), shape=(n, n)).tocsr()
    a.data[:] = 1.0
    return a


def pattern(owner, nb, n, reach):
    """The cell pairs within ``reach`` steps (``block_stencil_colouring``'s pattern)."""
    a = adjacency(owner, nb, n)
    p = a
    for _ in range(reach - 1):
        p = (p @ a).tocsr()
        p.data[:] = 1.0
    return p


def conflict(owner, nb, n, reach):
    """Pattern^T pattern: pairs that some row couples to both."""
    p = pattern(owner, nb, n, reach)
    c = (p.T @ p).tocsr()
    c.data[:] = 1.0
    return c


def saturation_colouring(conf, n):
    """Full-scan saturation-degree greedy (smallest index on ties)."""
    indptr, indices = conf.indptr, conf.indices
    degree = np.diff(indptr).astype(np.int64)
    colour = np.full(n, -1, dtype=np.int64)
    seen = [set() for _ in range(n)]
    for _ in range(n):
        best, key = -1, None
        for v in range(n):
            if colour[v] >= 0:
                continue
            k = (len(seen[v]), degree[v])
            if key is None or k > key:
                best, key = v, k
        c = 0
        while c in seen[best]:
            c += 1
        colour[best] = c
        for u in indices[indptr[best]:indptr[best + 1]]:
            if u != best:
                seen[u].add(c)
    return colour


def degree_colouring(conf, n):
    """Plain highest-degree-first greedy (the baseline the saturation order improves on)."""
    colour = np.full(n, -1, dtype=np.int64)
    for i in np.argsort(-np.diff(conf.indptr), kind="stable"):
        used = set(colour[conf.indices[conf.indptr[i]:conf.indptr[i + 1]]].tolist())
        c = 0
        while c in used:
            c += 1
        colour[i] = c
    return colour


def collision_free(conf, colour):
    coo = conf.tocoo()
    off = coo.row != coo.col
    return not np.any(colour[coo.row[off]] == colour[coo.col[off]])
