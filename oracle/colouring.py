"""CPU restatement of the cell-block sparsity pattern and the saturation-degree colouring used for
compressed Jacobian assembly (test infrastructure only).

Follows ``aquaflux/solve/sparse_jacobian.py``: the pattern at ``reach`` is the ``reach``-th power
of the cell adjacency with self loops (``block_stencil_colouring``, ``:338-406``), the conflict
graph is ``pattern^T pattern``, and the colouring picks, each round, the uncoloured vertex whose
neighbours show the most distinct colours, ties by degree (``_saturation_colouring``,
``:409-534``).  :func:`bucket_saturation_colouring` restates the reference's bucket queue step for
step -- the same initial push order (vertex order, each pushed at its bucket's head), the same
last-in-first-out pop, and the same neighbour-update order (the conflict graph's CSR order) -- so
it gives the reference's colouring vertex for vertex.  :func:`saturation_colouring` is the full
scan with the smallest-index tie-break of the reference test's ``_naive_saturation_colouring``
(``tests/unit/test_sparse_jacobian.py:120-156``), kept to pin the bucket queue's ranking: the two
agree on the colour of the first vertex and on the colour counts of graphs with known chromatic
numbers.
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


def bucket_saturation_colouring(conf, n):
    """``_saturation_colouring`` (``sparse_jacobian.py:458-534``): saturation-degree greedy over a
    bucket queue of ranks ``saturation * (max degree + 1) + degree`` (degree = the conflict row's
    entry count, diagonal included).  Vertices are pushed at their bucket's head in vertex order, the
    head of the highest non-empty bucket is popped, and after each colouring the neighbours that gain
    a colour move to the head of their new bucket in the row's CSR order."""
    indptr, indices = conf.indptr, conf.indices
    degree = np.diff(indptr).astype(np.int64).tolist()
    colour = [-1] * n
    max_degree = max(degree) if n else 0
    tie = max_degree + 1
    taken = [0] * n  # bitset of the colours each vertex's neighbours hold
    sat = [0] * n
    head = [-1] * (tie * max_degree + max_degree + 1)
    nxt = [-1] * n
    prv = [-1] * n
    top = 0
    for v in range(n):
        r = degree[v]
        h = head[r]
        nxt[v] = h
        if h >= 0:
            prv[h] = v
        head[r] = v
        top = max(top, r)
    for _ in range(n):
        while head[top] < 0:
            top -= 1
        v = head[top]
        f = nxt[v]
        head[top] = f
        if f >= 0:
            prv[f] = -1
        free = ~taken[v]
        chosen = (free & -free).bit_length() - 1
        colour[v] = chosen
        bit = 1 << chosen
        for u in indices[indptr[v]:indptr[v + 1]].tolist():
            had = taken[u] & bit
            taken[u] |= bit
            if had or colour[u] >= 0:
                continue
            r_old = sat[u] * tie + degree[u]
            sat[u] += 1
            r_new = r_old + tie
            p, f = prv[u], nxt[u]
            if p >= 0:
                nxt[p] = f
            else:
                head[r_old] = f
            if f >= 0:
                prv[f] = p
            h = head[r_new]
            nxt[u] = h
            prv[u] = -1
            if h >= 0:
                prv[h] = u
            head[r_new] = u
            top = max(top, r_new)
    return np.asarray(colour, dtype=np.int64)


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
