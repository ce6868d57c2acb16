"""Compressed (coloured) block-Jacobian assembly on the GPU (``aquaflux/solve/sparse_jacobian.py``).

:func:`block_stencil_colouring` builds the cell-block sparsity pattern at a stencil reach and a
probing colouring of it (``sparse_jacobian.py:338-534``) with the library's host code;
:func:`column_probe_plan` gives each column field its own reach over the widest pattern
(``sparse_jacobian.py:88-336``).  :func:`materialize_block_jacobian` (``sparse_jacobian.py:722-848``)
runs one directional derivative per (colour, field) through the residual's HIP tangent and scatters
each response into block-ELL storage on the device, in FP64 or FP32.  The result is exact for the pattern when the reach covers
the residual's true stencil; :func:`jacobian_relative_error` (``sparse_jacobian.py:864-890``) is the
cheap check that it does.  The assembled matrix offers the SpMV, transposed SpMV and point-block
Jacobi of the Newton-Krylov inner solves.
"""

from __future__ import annotations

import ctypes
import copy
import dataclasses

import numpy as np
import torch

from . import _lib


@dataclasses.dataclass
class BlockColouring:
    """Cell-block sparsity pattern and probing colouring (``sparse_jacobian.py:53-85``).

    ``pattern_ptr`` / ``pattern_cols`` hold the pattern as rows: the row cell first, then the cells
    within ``reach`` steps in ascending order.  ``pattern_rows`` / ``pattern_cols_flat`` give the
    reference's coordinate form.
    """

    n_cells: int
    reach: int
    colour: np.ndarray
    n_colours: int
    pattern_ptr: np.ndarray
    pattern_cols: np.ndarray
    pattern_depth: np.ndarray  # graph distance of each pattern entry from its row cell

    @property
    def pattern_rows(self) -> np.ndarray:
        return np.repeat(np.arange(self.n_cells, dtype=np.int64), np.diff(self.pattern_ptr))

    @property
    def pattern_cols_flat(self) -> np.ndarray:
        return self.pattern_cols


def block_stencil_colouring(owner, nb, n: int, reach: int) -> BlockColouring:
    """The pattern at ``reach`` on the graph of the given interior-face edges and a colouring whose
    same-colour cells no row couples to twice (``sparse_jacobian.py:338-406``)."""
    owner = np.ascontiguousarray(np.asarray(owner), dtype=np.int32)
    nb = np.ascontiguousarray(np.asarray(nb), dtype=np.int32)
    if owner.shape != nb.shape or owner.ndim != 1:
        raise ValueError("block_stencil_colouring: owner and nb must be 1-D arrays of equal length")
    if n < 1:
        raise ValueError(f"block_stencil_colouring: need at least one cell, got n={n}")
    if reach < 1:
        raise ValueError(f"block_stencil_colouring: reach must be >= 1, got {reach}.")
    if owner.size and (owner.min() < 0 or nb.min() < 0 or owner.max() >= n or nb.max() >= n):
        raise ValueError("block_stencil_colouring: edge endpoint out of range")
    L = _lib.lib()
    ptr = np.zeros(n + 1, dtype=np.int64)
    _lib.check(L.afx_block_colouring(n, owner.size, owner.ctypes.data, nb.ctypes.data, reach, ptr.ctypes.data, None,
                                     None, None, None), "afx_block_colouring")
    cols = np.zeros(int(ptr[-1]), dtype=np.int32)
    depth = np.zeros(int(ptr[-1]), dtype=np.int8)
    colour = np.zeros(n, dtype=np.int32)
    nc = ctypes.c_int32()
    _lib.check(L.afx_block_colouring(n, owner.size, owner.ctypes.data, nb.ctypes.data, reach, ptr.ctypes.data,
                                     cols.ctypes.data, depth.ctypes.data, colour.ctypes.data, ctypes.byref(nc)),
               "afx_block_colouring")
    return BlockColouring(n, reach, colour, int(nc.value), ptr, cols, depth)


def mesh_colouring(mesh, reach: int) -> BlockColouring:
    """:func:`block_stencil_colouring` over a mesh's interior faces."""
    inner = mesh.neighbour >= 0
    return block_stencil_colouring(mesh.owner[inner], mesh.neighbour[inner], mesh.n_cells, reach)


@dataclasses.dataclass
class ColumnProbePlan:
    """Per column field, the reach it is probed at, over one assembly pattern of the widest reach
    (``sparse_jacobian.py:88-247``).  A column probed at a shorter reach keeps the pattern entries
    beyond it at zero; that is exact only for a column with no coupling that far out (otherwise its
    far couplings fold onto near entries), so the reach is a property of the discretization."""

    pattern: BlockColouring
    colourings: dict  # reach -> BlockColouring
    reach: tuple

    @property
    def n_cells(self) -> int:
        return self.pattern.n_cells

    @property
    def n_fields(self) -> int:
        return len(self.reach)

    @property
    def n_probes(self) -> int:
        return sum(self.colourings[r].n_colours for r in self.reach)

    @classmethod
    def uniform(cls, owner, nb, n: int, n_fields: int, reach: int) -> "ColumnProbePlan":
        return column_probe_plan(owner, nb, n, (reach,) * n_fields)


def column_probe_plan(owner, nb, n: int, column_reach, pattern_reach: int | None = None) -> ColumnProbePlan:
    """A probing plan giving each column field its own stencil reach (``sparse_jacobian.py:249-336``)."""
    reaches = tuple(int(r) for r in column_reach)
    if not reaches:
        raise ValueError("column_probe_plan: need a reach for at least one column field.")
    widest = max(reaches)
    pattern_reach = widest if pattern_reach is None else int(pattern_reach)
    if pattern_reach < widest:
        raise ValueError(f"column_probe_plan: pattern_reach={pattern_reach} is narrower than the widest column "
                         f"reach {widest}; a column cannot be probed beyond the pattern it is assembled into.")
    by_reach = {r: block_stencil_colouring(owner, nb, n, r) for r in sorted({*reaches, pattern_reach})}
    return ColumnProbePlan(by_reach[pattern_reach], by_reach, reaches)


def mesh_probe_plan(mesh, column_reach, pattern_reach: int | None = None) -> ColumnProbePlan:
    """:func:`column_probe_plan` over a mesh's interior faces."""
    inner = mesh.neighbour >= 0
    return column_probe_plan(mesh.owner[inner], mesh.neighbour[inner], mesh.n_cells, column_reach, pattern_reach)


class ProbedJacobian:
    """A block Jacobian on a probing plan's pattern, stored block-ELL cell-major (slot 0 the
    diagonal block)."""

    def __init__(self, plan: ColumnProbePlan, device, precision: str = "fp64"):
        if precision not in ("fp32", "fp64"):
            raise ValueError(f"precision must be 'fp32' or 'fp64', got {precision!r}")
        n_fields = plan.n_fields
        if not 1 <= n_fields <= 6:
            raise ValueError(f"n_fields must be 1 to 6, got {n_fields}")
        self.plan = plan
        colouring = plan.pattern
        C = colouring.n_cells
        self.C, self.B = C, n_fields
        self.prec = 0 if precision == "fp32" else 1
        self.dtype = torch.float32 if self.prec == 0 else torch.float64
        ptr = colouring.pattern_ptr
        lengths = np.diff(ptr)
        self.nslot = int(lengths.max())
        rows = np.repeat(np.arange(C, dtype=np.int64), lengths)
        slot = np.arange(ptr[-1], dtype=np.int64) - np.repeat(ptr[:-1], lengths)
        ell = np.repeat(np.arange(C, dtype=np.int32)[None, :], self.nslot, axis=0)  # padding: the row itself
        ell[slot, rows] = colouring.pattern_cols
        depth = np.zeros((self.nslot, C), dtype=np.int8)
        depth[slot, rows] = colouring.pattern_depth
        # slot of (row, col) in the column cell's row, for the transposed product
        key = rows * C + colouring.pattern_cols.astype(np.int64)
        order = np.argsort(key)
        tkey = colouring.pattern_cols.astype(np.int64) * C + rows
        pos = order[np.searchsorted(key[order], tkey)]
        tslot = np.zeros((self.nslot, C), dtype=np.int32)
        tslot[slot, rows] = slot[pos]
        self.device = torch.device(device)
        self.cols = torch.from_numpy(ell.reshape(-1)).to(self.device)
        self.depth = torch.from_numpy(depth.reshape(-1)).to(self.device)
        self.tslot = torch.from_numpy(tslot.reshape(-1)).to(self.device)
        self.colours = {r: torch.from_numpy(c.colour).to(self.device) for r, c in plan.colourings.items()}
        self.vals = torch.zeros(self.nslot * self.B * self.B * C, dtype=self.dtype, device=self.device)
        self.inv = torch.empty(self.B * self.B * C, dtype=self.dtype, device=self.device)
        self.singular = False

    # vectors: field-major FP64 (residual layout) <-> cell-major in the Jacobian's precision
    def to_cell_major(self, v):
        return v.reshape(self.B, self.C).T.contiguous().reshape(-1).to(self.dtype)

    def from_cell_major(self, v):
        return v.reshape(self.C, self.B).T.contiguous().reshape(-1).to(torch.float64)

    def matvec(self, x):
        """``y = A x`` (cell-major vectors in the Jacobian's precision)."""
        y = torch.empty_like(x)
        _lib.check(
            _lib.lib().afx_bell_spmv(self.C, self.B, self.nslot, self.prec, self.vals.data_ptr(), self.cols.data_ptr(),
                                     x.data_ptr(), y.data_ptr(), _lib.stream_of(x.device)),
            "afx_bell_spmv",
        )
        return y

    def rmatvec(self, x):
        """``y = A^T x``."""
        y = torch.empty_like(x)
        _lib.check(
            _lib.lib().afx_bell_spmv_t(self.C, self.B, self.nslot, self.prec, self.vals.data_ptr(),
                                       self.cols.data_ptr(), self.tslot.data_ptr(), x.data_ptr(), y.data_ptr(),
                                       _lib.stream_of(x.device)),
            "afx_bell_spmv_t",
        )
        return y

    def field_matvec(self, v):
        """``J v`` for a field-major FP64 vector (the residual's layout)."""
        return self.from_cell_major(self.matvec(self.to_cell_major(v)))

    def setup_pbj(self):
        flag = torch.zeros(1, dtype=torch.int32, device=self.device)
        _lib.check(
            _lib.lib().afx_pbj_setup(self.C, self.B, self.prec, self.vals.data_ptr(), self.inv.data_ptr(),
                                     flag.data_ptr(), _lib.stream_of(self.device)),
            "afx_pbj_setup",
        )
        self.singular = bool(flag.item())

    def pbj(self, x):
        """Point-block Jacobi: ``y = D^-1 x`` with D the diagonal blocks."""
        y = torch.empty_like(x)
        _lib.check(
            _lib.lib().afx_pbj_apply(self.C, self.B, self.prec, self.inv.data_ptr(), x.data_ptr(), y.data_ptr(),
                                     _lib.stream_of(x.device)),
            "afx_pbj_apply",
        )
        return y

    def pbj_transpose(self, x):
        """``y = D^-T x`` (the adjoint solve's preconditioner)."""
        inv = self.inv.reshape(self.B, self.B, self.C)  # inv[r, k, c]
        return torch.einsum("rkc,cr->ck", inv, x.reshape(self.C, self.B)).reshape(-1).contiguous()

    def add_diagonal(self, shift):
        """Adds a per-row shift (field-major FP64 vector) to the diagonal blocks
        (``shifted_jacobian``, ``sparse_jacobian.py:893-927``) and refreshes the Jacobi inverses."""
        B, C = self.B, self.C
        sc = shift.reshape(B, C).to(self.dtype)
        diag = self.vals[: B * B * C].view(B, B, C)
        for r in range(B):
            diag[r, r] += sc[r]
        self.setup_pbj()

    def to_dense(self) -> torch.Tensor:
        """The field-major dense matrix (for small meshes)."""
        B, C, S = self.B, self.C, self.nslot
        vals = self.vals.double().reshape(S, B, B, C).cpu()
        cols = self.cols.reshape(S, C).cpu().long()
        J = torch.zeros(B * C, B * C, dtype=torch.float64)
        rows = torch.arange(C)
        for s in range(S):
            keep = (cols[s] != rows) | (s == 0)
            r, j = rows[keep], cols[s][keep]
            for a in range(B):
                for b in range(B):
                    J[a * C + r, b * C + j] = vals[s, a, b][keep]
        return J


def shifted_jacobian(jacobian: ProbedJacobian, shift_diagonal) -> ProbedJacobian:
    """A copy of ``jacobian`` with the pseudo-transient shift ``shift_diagonal`` (a field-major FP64
    vector, one entry per row) added to its diagonal (``solve/sparse_jacobian.py:893-927``); the
    point-block Jacobi inverses of the copy are refreshed.  The input is left unchanged."""
    out = copy.copy(jacobian)
    out.vals = jacobian.vals.clone()
    out.inv = jacobian.inv.clone()
    out.add_diagonal(torch.as_tensor(shift_diagonal, dtype=torch.float64, device=jacobian.device))
    return out


def materialize_block_jacobian(matvec, plan, n_fields: int | None = None, *, device=None,
                               precision: str = "fp64") -> ProbedJacobian:
    """The block Jacobian of ``matvec`` (field-major ``v -> J v``, FP64 device tensors) by coloured
    probing: one directional derivative per (colour, column field) (``sparse_jacobian.py:722-848``).

    ``plan`` is a :class:`ColumnProbePlan`, or a :class:`BlockColouring` used for all ``n_fields``
    columns.  Every column cell of a colour lies more than ``2 reach`` steps from the others, so a
    probe's response at row cell ``i`` holds exactly the coupling to the one probed cell within
    ``reach`` of ``i``.  Couplings beyond a column's reach are folded onto nearer entries, not
    dropped: choose the reaches from the discretization and confirm with
    :func:`jacobian_relative_error`.
    """
    if isinstance(plan, BlockColouring):
        if n_fields is None:
            raise ValueError("materialize_block_jacobian: a single colouring needs n_fields")
        plan = ColumnProbePlan(plan, {plan.reach: plan}, (plan.reach,) * n_fields)
    dev = torch.device(device if device is not None else "cuda")
    J = ProbedJacobian(plan, dev, precision)
    C, B = J.C, J.B
    seed = torch.empty(B * C, dtype=torch.float64, device=dev)
    L = _lib.lib()
    st = _lib.stream_of(dev)
    for b in range(B):
        r = plan.reach[b]
        colour = J.colours[r]
        for k in range(plan.colourings[r].n_colours):
            _lib.check(L.afx_probe_seed(C, B, colour.data_ptr(), k, b, seed.data_ptr(), st), "afx_probe_seed")
            resp = matvec(seed).contiguous()
            if resp.dtype != torch.float64 or resp.numel() != B * C:
                raise ValueError("matvec must return a float64 vector of n_fields * n_cells entries")
            _lib.check(L.afx_probe_scatter(C, B, J.nslot, J.cols.data_ptr(), J.depth.data_ptr(), r,
                                           colour.data_ptr(), k, b, resp.data_ptr(), J.prec, J.vals.data_ptr(), st),
                       "afx_probe_scatter")
    J.setup_pbj()
    return J


def jacobian_relative_error(jacobian: ProbedJacobian, matvec, seed: int = 0) -> float:
    """``||J v - matvec(v)|| / ||matvec(v)||`` for a random ``v`` (``sparse_jacobian.py:864-890``)."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    v = torch.randn(jacobian.B * jacobian.C, generator=g, dtype=torch.float64).to(jacobian.device)
    ref = matvec(v)
    got = jacobian.field_matvec(v)
    return float((got - ref).norm() / (ref.norm() + 1e-300))
