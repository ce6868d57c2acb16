"""Block (SIMPLE-family) preconditioner of the coupled-flow Newton-Krylov solve
(``aquaflux/flow/block_preconditioner.py:1007-1460``, ``flow/preconditioner.py:43-81``).

The saddle ``J = [[F, G], [B, C]]`` (velocity rows / columns first, then pressure) is read off
the assembled block Jacobian of the Newton step -- the compact reach-1 matrix or the probed exact
one, FP32 or FP64, block-ELL -- instead of being linearised matrix-free:

* ``F`` (velocity-velocity) is copied out of the ``(d+1) x (d+1)`` blocks into its own ``d x d``
  block-ELL (same columns) for the velocity solve;
* the divergence ``B du`` and the pressure column ``(G dp, C dp)`` are one product of the whole
  matrix with a zero-padded vector (``FlowBlocks.velocity_column`` / ``pressure_column`` of the
  reference), so they stay exact for either Jacobian.

The pressure Schur is the compact pressure Laplacian of SIMPLE, face coefficient
``rho_f (V/a_P)_f A_f / (d.n)_f`` (``preconditioner.py:43-81``) plus the outlet stiffness
``rho_P (V/a_P)_P A / (d.n)`` of pressure-fixing boundaries (``flow/boundary.py:437-441``), assembled
per Newton step as a scalar ELL on the interior-face graph; ``"msimple"`` replaces ``a_P`` by the
mass diagonal ``rho V / k``, ``k = mean(rho V / a_P)`` (``block_preconditioner.py:1400-1420``).  A
pinned pressure cell is decoupled (identity row and column).

Inner solves run on the GPU: the velocity block by GMRES with its point-block Jacobi (or by that
Jacobi alone), the Schur by a fixed number of conjugate-gradient steps with no host
synchronisation, preconditioned by a V-cycle of an aggregation multigrid (``AggregationHierarchy``)
or by Jacobi.  Neither inner solve is a fixed linear map, so the outer solve must be flexible
GMRES, which the Newton step uses.  The three compositions
(triangular / simple / simpler) follow Klaij & Vuik (2013) as the reference states them
(``block_preconditioner.py:1069-1176``).
"""

from __future__ import annotations

import dataclasses

import numpy as np
import torch

from . import _lib

VELOCITY_SOLVES = ("gmres", "pbj")
SCHUR_SOLVERS = ("amg", "jacobi")
SCHUR_SCALINGS = ("simple", "msimple")
COMPOSITIONS = ("triangular", "simple", "simpler")


# --- compositions (block_preconditioner.py:1007-1176) -----------------------------------------


def compose(composition, velocity_solve, schur_solve, divergence, gradient, pressure_column, inverse_diagonal):
    """The residual-to-correction map ``(r_u, r_p) -> (du, dp)`` of one composition.

    ``velocity_solve(r_u) ~ F^-1 r_u``, ``schur_solve(r_p) ~ S^-1 r_p``; ``divergence(du) = B du``,
    ``gradient(dp) = G dp``, ``pressure_column(dp) = (G dp, C dp)``; ``inverse_diagonal`` is
    ``F~^-1`` per cell (the diagonal the Schur was assembled from).  Velocities are ``(n, d)``,
    pressures ``(n,)``.
    """
    if composition not in COMPOSITIONS:
        raise ValueError(f"unknown composition {composition!r}; use one of {COMPOSITIONS}")
    inv = inverse_diagonal[:, None]

    def triangular(ru, rp):
        du = velocity_solve(ru)
        return du, schur_solve(rp - divergence(du))

    def simple(ru, rp):
        pred = velocity_solve(ru)
        dp = schur_solve(rp - divergence(pred))
        return pred - inv * gradient(dp), dp

    def simpler(ru, rp):
        predicted = schur_solve(-divergence(inv * ru))
        grad, coupling = pressure_column(predicted)
        pred = velocity_solve(ru - grad)
        corr = schur_solve(rp - divergence(pred) - coupling)
        return pred - inv * gradient(corr), predicted + corr

    return {"triangular": triangular, "simple": simple, "simpler": simpler}[composition]


def conjugate_gradient(matvec, b, preconditioner, iterations: int, dot=torch.dot):
    """``iterations`` preconditioned CG steps from zero, with no host synchronisation (the step
    sizes stay on the device; a zero curvature or residual leaves the iterate unchanged).
    ``preconditioner`` is a symmetric positive map or, as a tensor, a Jacobi inverse diagonal;
    ``dot`` returns a 0-d device tensor."""
    precond = preconditioner if callable(preconditioner) else (lambda v: v * preconditioner)
    # no zero-filled allocations (torch fills them with a memory clear, which a graph recording must
    # not contain): the zero iterate is b * 0 and the guards take Python scalars
    x = b * 0.0
    r = b.clone()
    z = precond(r)
    p = z.clone()
    rz = dot(r, z)
    for _ in range(int(iterations)):
        Ap = matvec(p)
        pAp = dot(p, Ap)
        alpha = torch.where(pAp > 0, rz / torch.where(pAp > 0, pAp, 1.0), 0.0)
        x.add_(alpha * p)
        r.sub_(alpha * Ap)
        z = precond(r)
        rz_new = dot(r, z)
        beta = torch.where(rz > 0, rz_new / torch.where(rz > 0, rz, 1.0), 0.0)
        p = z + beta * p
        rz = rz_new
    return x


def captured(fn, n, dtype, device):
    """``fn`` (a map of one length-``n`` vector with no host synchronisation) recorded as a HIP graph
    on a static input buffer: the returned callable copies its argument in, replays the graph and
    returns a copy of the static output.  One warm-up call on a side stream first, so that lazy
    initialisation stays out of the recording.  ``fn`` must make no call that records a memory
    clear (library reductions do), since such nodes are unreliable on replay."""
    static_in = torch.zeros(n, dtype=dtype, device=device)
    side = torch.cuda.Stream(device)
    side.wait_stream(torch.cuda.current_stream(device))
    with torch.cuda.stream(side):
        fn(static_in)
    torch.cuda.current_stream(device).wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        static_out = fn(static_in)

    def run(x):
        static_in.copy_(x)
        graph.replay()
        return static_out.clone()

    # the recording reads the operators ``fn`` closes over (the Schur, the level operators, the
    # coarse inverse) at their addresses: keep ``fn`` alive as long as the graph
    run.graph, run.recorded = graph, fn
    return run


# --- aggregation multigrid for the Schur -------------------------------------------------------


@dataclasses.dataclass
class AggregationLevel:
    """One level of the hierarchy: its operator's scalar-ELL columns, the aggregate (next-level
    row) of each row, and the flat next-level value slot each of its entries sums into."""

    n: int
    nslot: int
    cols: torch.Tensor  # (nslot, n) int32
    agg: torch.Tensor | None  # (n,) int64; None on the coarsest level
    gmap: torch.Tensor | None  # (nslot * n,) int64 into the next level's (nslot', n') values
    agg32: torch.Tensor | None = None  # (n,) int32, the kernels' copy of agg
    members: torch.Tensor | None = None  # (max_members, n') int32 rows of each aggregate, -1 padded
    # (nslot' * n', k) int64: the entries of this level that sum into each next-level entry, in
    # ascending order, padded with nslot * n (a zero appended to the values)
    gsrc: torch.Tensor | None = None


class AggregationHierarchy:
    """Unsmoothed aggregation multigrid on a scalar ELL operator, for the pressure Schur.

    The reference builds smoothed-aggregation hierarchies with scipy on the host
    (``block_preconditioner.py:266-343``); here the aggregates are geometric -- cells binned on a
    grid of cell size doubling per level, about ``2^d`` cells per aggregate on a regular mesh, a
    pinned cell kept a singleton -- and built once per mesh on the host, while the coarse
    operators are Galerkin products ``P^T A P`` of the piecewise-constant prolongation, summed on the
    device per Newton step.  The cycle is a symmetric V-cycle (damped Jacobi, ``sweeps`` pre- and
    post-smoothing steps, over-corrected coarse-grid corrections, a dense pseudo-inverse on the
    coarsest level), used as the preconditioner of the Schur's conjugate gradient.
    """

    def __init__(self, levels, coarse_rows, coarse_cols):
        self.levels = levels
        self.coarse_rows = coarse_rows
        self.coarse_cols = coarse_cols

    @classmethod
    def build(cls, cols, centroids, pin=None, device="cpu", coarsest: int = 1000, max_levels: int = 16):
        """``cols`` (nslot, n) the fine operator's ELL columns (padding = the row itself),
        ``centroids`` (n, d) the cells' positions."""
        cols = np.asarray(cols, dtype=np.int64)
        x = np.asarray(centroids, dtype=np.float64)
        n, dim = x.shape[0], x.shape[1]
        lo, hi = x.min(axis=0), x.max(axis=0)
        # bin width per axis: twice the median spacing of the graph edges that run mostly along that
        # axis (a sample of rows on a large mesh), so anisotropic cells still pair up per axis
        rows = np.arange(0, n, max(1, n // 200000))
        dx = x[cols[1:, rows]] - x[rows][None, :, :]
        dx = dx.reshape(-1, dim)
        norm = np.linalg.norm(dx, axis=1)
        dx, norm = dx[norm > 0], norm[norm > 0]
        iso = (float(np.prod(np.maximum(hi - lo, 1e-300))) / n) ** (1.0 / dim)
        h = np.empty(dim)
        for j in range(dim):
            along = np.abs(dx[:, j])[np.abs(dx[:, j]) >= 0.7 * norm]
            h[j] = 2.0 * (float(np.median(along)) if along.size else iso)
        T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
        levels = []
        tries = 0
        while n > coarsest and len(levels) < max_levels:
            key = np.floor((x - lo) / h + 0.25).astype(np.int64)  # off the bin edges of a regular grid
            span = key.max(axis=0) + 2
            lin = key[:, 0].copy()
            for j in range(1, dim):
                lin = lin * span[j] + key[:, j]
            if pin is not None:
                lin[pin] = -1  # the pinned cell stays a singleton (it holds an identity row)
            _, agg = np.unique(lin, return_inverse=True)
            nc = int(agg.max()) + 1
            target = n / 2.0**dim
            if nc > 2.0 * target and tries < 4:
                # bins too fine for this level (cells clustered closer than their edge spacing, as
                # the tetrahedra of one hexahedron are): widen towards 2^d cells a bin and retry
                h = h * (nc / target) ** (1.0 / dim)
                tries += 1
                continue
            tries = 0
            nslot = cols.shape[0]
            I = np.broadcast_to(agg[None, :], cols.shape)
            J = agg[cols]
            pair = (I * nc + J).ravel()
            up, inv = np.unique(pair, return_inverse=True)
            rc, cc = up // nc, up % nc
            # slot of each coarse entry: the diagonal first, then the rest in column order
            order = np.lexsort((cc, rc != cc, rc))
            counts = np.bincount(rc, minlength=nc)
            start = np.concatenate([[0], np.cumsum(counts)[:-1]])
            slot_sorted = np.arange(up.size) - np.repeat(start, counts)
            slot = np.empty(up.size, dtype=np.int64)
            slot[order] = slot_sorted
            ns_c = int(counts.max())
            cols_c = np.repeat(np.arange(nc, dtype=np.int64)[None, :], ns_c, axis=0)
            cols_c[slot, rc] = cc
            gmap = slot[inv] * nc + rc[inv]
            order = np.argsort(agg, kind="stable")
            size = np.bincount(agg, minlength=nc)
            rank = np.arange(n) - np.repeat(np.concatenate([[0], np.cumsum(size)[:-1]]), size)
            members = np.full((int(size.max()), nc), -1, dtype=np.int32)
            members[rank, agg[order]] = order
            # Galerkin sums as gathers: each next-level entry lists its entries of this level
            src_order = np.argsort(gmap, kind="stable")
            src_count = np.bincount(gmap, minlength=ns_c * nc)
            src_rank = np.arange(gmap.size) - np.repeat(np.concatenate([[0], np.cumsum(src_count)[:-1]]), src_count)
            gsrc = np.full((ns_c * nc, max(1, int(src_count.max()))), gmap.size, dtype=np.int64)
            gsrc[gmap[src_order], src_rank] = src_order
            levels.append(AggregationLevel(n, nslot, T(cols.astype(np.int32)), T(agg.astype(np.int64)), T(gmap),
                                           T(agg.astype(np.int32)), T(members), T(gsrc)))
            sums = np.zeros((nc, dim))
            np.add.at(sums, agg, x)
            x = sums / np.bincount(agg, minlength=nc)[:, None]
            pin = None if pin is None else int(agg[pin])
            cols, n, h = cols_c, nc, 2.0 * h
        levels.append(AggregationLevel(n, cols.shape[0], T(cols.astype(np.int32)), None, None))
        rows = np.broadcast_to(np.arange(n)[None, :], cols.shape)
        return cls(levels, T(rows.ravel().copy()), T(cols.ravel().copy()))

    def operators(self, vals):
        """Every level's ELL values from the fine level's ``(nslot, n)`` values, and the coarsest
        level's dense pseudo-inverse."""
        out = [vals.contiguous()]
        for lv, nx in zip(self.levels[:-1], self.levels[1:]):
            # a gather of each coarse entry's fine entries and a sum along that short axis: the same
            # summation order on every call (no atomics), so the cycle is bitwise reproducible
            flat = torch.cat([out[-1].reshape(-1), out[-1].new_zeros(1)])
            vc = flat[lv.gsrc].sum(dim=1)
            out.append(vc.view(nx.nslot, nx.n))
        last = self.levels[-1]
        A = torch.zeros(last.n, last.n, dtype=torch.float64, device=vals.device)
        # duplicate positions here are the ELL's padding slots (column = row, value exactly 0), so the
        # accumulation order cannot change the sum
        A.index_put_((self.coarse_rows, self.coarse_cols.long()), out[-1].reshape(-1).double(), accumulate=True)
        return out, torch.linalg.pinv(A, hermitian=True).to(vals.dtype).contiguous()

    def cycle(self, spmv, ops, coarse_inverse, b, sweeps: int = 2, omega: float = 2.0 / 3.0,
              correction: float = 1.8):
        """One symmetric V-cycle for ``A x = b`` from zero; ``spmv(level, vals, x)`` is the ELL
        product of a level's operator.  ``correction`` over-corrects each coarse-grid correction:
        a piecewise-constant prolongation's Galerkin operator is too stiff by about the aggregate
        width, and 1.8 gave the fastest conjugate-gradient convergence on 2D and 3D Laplacians with
        2^d-cell aggregates (12 steps to ~1e-9 relative at 64 x 32 .. 256 x 128 in 2D, against
        ~1e-2 without)."""
        dinv = [1.0 / v[0] for v in ops]

        def level(k, rhs):
            if k == len(self.levels) - 1:
                return coarse_inverse @ rhs
            lv = self.levels[k]
            x = omega * dinv[k] * rhs
            for _ in range(sweeps - 1):
                x = x + omega * dinv[k] * (rhs - spmv(k, ops[k], x))
            r = rhs - spmv(k, ops[k], x)
            rc = torch.zeros(self.levels[k + 1].n, dtype=rhs.dtype, device=rhs.device).index_add_(0, lv.agg, r)
            x = x + correction * level(k + 1, rc)[lv.agg]
            for _ in range(sweeps):
                x = x + omega * dinv[k] * (rhs - spmv(k, ops[k], x))
            return x

        return level(0, b)

    def native_cycle(self, ops, coarse_inverse, b, sweeps: int = 2, omega: float = 2.0 / 3.0,
                     correction: float = 1.8):
        """:meth:`cycle` on the library's kernels: a fused Jacobi sweep per smoothing step, a fused
        residual, a member-table restriction (fixed summation order) and a fused prolongation."""
        L = _lib.lib()
        st = _lib.stream_of(b.device)
        prec = 0 if b.dtype == torch.float32 else 1

        def jacobi(k, rhs, x):
            lv = self.levels[k]
            y = torch.empty_like(rhs)
            _lib.check(L.afx_ell_jacobi(lv.n, lv.nslot, prec, ops[k].data_ptr(), lv.cols.data_ptr(), rhs.data_ptr(),
                                        None if x is None else x.data_ptr(), y.data_ptr(), omega, st), "afx_ell_jacobi")
            return y

        def level(k, rhs):
            if k == len(self.levels) - 1:
                y = torch.empty_like(rhs)
                _lib.check(L.afx_dense_matvec(rhs.numel(), prec, coarse_inverse.data_ptr(), rhs.data_ptr(),
                                              y.data_ptr(), st), "afx_dense_matvec")
                return y
            lv, nx = self.levels[k], self.levels[k + 1]
            x = jacobi(k, rhs, None)
            for _ in range(sweeps - 1):
                x = jacobi(k, rhs, x)
            r = torch.empty_like(rhs)
            _lib.check(L.afx_ell_residual(lv.n, lv.nslot, prec, ops[k].data_ptr(), lv.cols.data_ptr(), rhs.data_ptr(),
                                          x.data_ptr(), r.data_ptr(), st), "afx_ell_residual")
            rc = torch.empty(nx.n, dtype=rhs.dtype, device=rhs.device)
            _lib.check(L.afx_aggregate_restrict(nx.n, lv.members.shape[0], prec, lv.members.data_ptr(), r.data_ptr(),
                                                rc.data_ptr(), st), "afx_aggregate_restrict")
            xc = level(k + 1, rc).contiguous()
            _lib.check(L.afx_aggregate_prolong(lv.n, prec, lv.agg32.data_ptr(), xc.data_ptr(), correction,
                                               x.data_ptr(), st), "afx_aggregate_prolong")
            for _ in range(sweeps):
                x = jacobi(k, rhs, x)
            return x

        return level(0, b.contiguous())


# --- the pressure Schur (preconditioner.py:43-81, block_preconditioner.py:123-210) -------------


@dataclasses.dataclass
class SchurGraph:
    """The interior-face graph of a mesh as a scalar ELL (slot 0 the row itself, then one slot per
    interior face of the cell), with the face geometry the Schur coefficient needs."""

    n_cells: int
    nslot: int
    cols: torch.Tensor  # (nslot, C) int32, padding = the row itself
    face: torch.Tensor  # (nslot, C) int64 face of each off-diagonal slot, -1 on the diagonal / padding
    owner: torch.Tensor  # (F,) int64
    neighbour: torch.Tensor  # (F,) int64, -1 on boundary faces
    g: torch.Tensor  # (F,) interpolation factor, 0 on boundary faces
    normal_distance: torch.Tensor  # (F,) (x_N - x_P).n interior, (x_f - x_P).n boundary
    area: torch.Tensor  # (F,)
    volume: torch.Tensor  # (C,)
    outlet: torch.Tensor  # (n_outlet,) int64 faces of pressure-fixing boundaries

    @classmethod
    def of(cls, op) -> "SchurGraph":
        from .flow import PressureOutlet

        mesh, geo = op.mesh, op.geometry
        dev = geo.face.area.device
        C, F = mesh.n_cells, mesh.n_faces
        own = np.asarray(mesh.owner, dtype=np.int64)
        nb = np.asarray(mesh.neighbour, dtype=np.int64)
        inner = np.flatnonzero(nb >= 0)
        rows = np.concatenate([own[inner], nb[inner]])
        other = np.concatenate([nb[inner], own[inner]])
        faces = np.concatenate([inner, inner])
        order = np.argsort(rows, kind="stable")
        rows, other, faces = rows[order], other[order], faces[order]
        counts = np.bincount(rows, minlength=C)
        start = np.concatenate([[0], np.cumsum(counts)[:-1]])
        slot = 1 + np.arange(rows.size) - np.repeat(start, counts)
        nslot = 1 + int(counts.max(initial=0))
        cols = np.repeat(np.arange(C, dtype=np.int32)[None, :], nslot, axis=0)
        face = np.full((nslot, C), -1, dtype=np.int64)
        cols[slot, rows] = other
        face[slot, rows] = faces
        outlet = [faces_ for _, bc, faces_ in op.boundary.items() if isinstance(bc, PressureOutlet)]
        outlet = np.concatenate(outlet).astype(np.int64) if outlet else np.zeros(0, dtype=np.int64)

        T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        owner_t, nb_t = T(own), T(nb)
        cc = geo.cell.centroid_soa.detach()
        fc = geo.face.centroid_soa.detach()
        n = geo.face.normal_soa.detach()
        xP = cc[:, owner_t]
        xN = torch.where(nb_t >= 0, cc[:, nb_t.clamp(min=0)], xP)
        dvec = xN - xP
        interior = nb_t >= 0
        dd = torch.where(interior, (dvec * dvec).sum(0), torch.ones_like(geo.face.area))
        g = torch.where(interior, ((fc - xP) * dvec).sum(0) / dd, torch.zeros_like(dd))
        dist = torch.where(interior, (dvec * n).sum(0), ((fc - xP) * n).sum(0))
        return cls(C, nslot, T(cols), T(face), owner_t, nb_t, g.contiguous(), dist.contiguous(),
                   geo.face.area.detach(), geo.cell.volume.detach(), T(outlet))

    def coefficient(self, d_coeff, rho):
        """Per-face ``rho_f d_f A / (d.n)`` with ``d = V / a_P`` blended ``(1 - g) P + g N``
        (``preconditioner.py:76-81``); zero on boundary faces."""
        interior = self.neighbour >= 0
        N = self.neighbour.clamp(min=0)
        blend = lambda x: (1.0 - self.g) * x[self.owner] + self.g * x[N]  # noqa: E731
        c = blend(rho) * blend(d_coeff) * self.area / self.normal_distance
        return torch.where(interior, c, torch.zeros_like(c))

    def boundary_diagonal(self, d_coeff, rho):
        """Outlet stiffness ``rho_P d_P A / (d.n)`` summed onto each owner (``flow/boundary.py:437-441``)."""
        out = torch.zeros(self.n_cells, dtype=torch.float64, device=self.area.device)
        if self.outlet.numel():
            P = self.owner[self.outlet]
            out.index_add_(0, P, rho[P] * d_coeff[P] * self.area[self.outlet] / self.normal_distance[self.outlet])
        return out

    def laplacian(self, d_coeff, rho, pin=None, dtype=torch.float64):
        """The Schur as scalar-ELL values ``(nslot, C)``: diagonal = sum of face coefficients plus
        outlet stiffness, off-diagonals = minus the face coefficient; a pinned cell's row and
        column are the identity's."""
        c = self.coefficient(d_coeff, rho)
        off = torch.where(self.face >= 0, -c[self.face.clamp(min=0)], torch.zeros((), dtype=c.dtype, device=c.device))
        off[0] = 0.0
        vals = off.clone()
        vals[0] = -off[1:].sum(0) + self.boundary_diagonal(d_coeff, rho)
        if pin is not None:  # decouple the pinned cell (solve/frozen_operator.py:274-299)
            column = self.cols == pin
            column[0] = False
            vals[column] = 0.0
            vals[1:, pin] = 0.0
            vals[0, pin] = 1.0
        return vals.to(dtype).contiguous()


# --- the preconditioner (block_preconditioner.py:1192-1460) -------------------------------------


class BlockPreconditioner:
    """SIMPLE-family block preconditioner of a :class:`MomentumContinuity` Newton step.

    ``velocity``: ``"gmres"`` (GMRES on the velocity block with its point-block Jacobi,
    ``velocity_iters`` iterations to ``velocity_rtol``) or ``"pbj"`` (the point-block Jacobi
    alone).  ``schur_scaling``: ``"simple"`` (``V / a_P``) or ``"msimple"`` (mass diagonal
    ``rho V / k``, ``k`` from ``mass_scale`` or ``mean(rho V / a_P)``).  ``composition``:
    ``"triangular"``, ``"simple"`` or ``"simpler"``.  ``schur_solver``: ``"amg"`` (conjugate
    gradient preconditioned by one aggregation-multigrid V-cycle with ``schur_sweeps`` Jacobi
    sweeps each way) or ``"jacobi"``; ``schur_iters`` conjugate-gradient steps per Schur solve.
    ``capture``: record the Schur solve (its conjugate gradient and V-cycles, launch-bound on the
    coarse levels) once per Newton step as a HIP graph and replay it (:func:`captured`); off by
    default, as replaying the ~120-node graph measured slower than launching its kernels on the
    2M-cell box (FP32 Newton 5.75 s against 4.94 s).
    """

    def __init__(self, assembler, graph, velocity, schur_scaling, composition, mass_scale, velocity_iters,
                 velocity_rtol, schur_iters, schur_solver="jacobi", schur_sweeps=2, hierarchy=None,
                 capture=False):
        self.assembler = assembler
        self.graph = graph
        self.velocity = velocity
        self.schur_scaling = schur_scaling
        self.composition = composition
        self.mass_scale = mass_scale
        self.velocity_iters = int(velocity_iters)
        self.velocity_rtol = float(velocity_rtol)
        self.schur_iters = int(schur_iters)
        self.schur_solver = schur_solver
        self.schur_sweeps = int(schur_sweeps)
        self.hierarchy = hierarchy
        self.capture = bool(capture)

    @classmethod
    def build(cls, assembler, *, velocity: str = "gmres", schur_scaling: str = "simple",
              composition: str = "triangular", mass_scale: float | None = None, velocity_iters: int = 10,
              velocity_rtol: float = 1e-2, schur_solver: str = "amg", schur_iters: int | None = None,
              schur_sweeps: int = 2, capture: bool = False) -> "BlockPreconditioner":
        if velocity not in VELOCITY_SOLVES:
            raise ValueError(f"unknown velocity block {velocity!r}; use one of {VELOCITY_SOLVES} (the reference's "
                             "multigrid velocity hierarchies are replaced by GPU Krylov inner solves)")
        if schur_scaling not in SCHUR_SCALINGS:
            raise ValueError(f"unknown schur_scaling {schur_scaling!r}; use one of {SCHUR_SCALINGS}")
        if composition not in COMPOSITIONS:
            raise ValueError(f"unknown composition {composition!r}; use one of {COMPOSITIONS}")
        if schur_solver not in SCHUR_SOLVERS:
            raise ValueError(f"unknown schur_solver {schur_solver!r}; use one of {SCHUR_SOLVERS}")
        if not hasattr(assembler, "pressure_pin") or not hasattr(assembler, "fields"):
            raise ValueError("BlockPreconditioner needs a MomentumContinuity assembler")
        graph = SchurGraph.of(assembler)
        hierarchy = None
        if schur_solver == "amg":
            cc = assembler.geometry.cell.centroid_soa.detach().T.cpu().numpy()
            hierarchy = AggregationHierarchy.build(graph.cols.cpu().numpy(), cc, assembler.pressure_pin,
                                                   graph.cols.device)
        if schur_iters is None:  # two multigrid-preconditioned steps measured fastest on 2M cells
            schur_iters = 2 if schur_solver == "amg" else 30
        return cls(assembler, graph, velocity, schur_scaling, composition, mass_scale, velocity_iters,
                   velocity_rtol, schur_iters, schur_solver, schur_sweeps, hierarchy, capture)

    def frozen_momentum_diagonal(self, state):
        """``a_P`` per cell at ``state`` (``block_preconditioner.py:1379-1389``)."""
        return self.assembler.fields(state)[1].detach()

    def schur_diagonal(self, state, a_p):
        """The diagonal the Schur is assembled from: ``a_p`` (SIMPLE) or ``rho V / k`` (mass-scaled)."""
        if self.schur_scaling == "simple":
            return a_p
        rho = self.assembler.density.detach()
        mass = rho * self.graph.volume
        k = self.mass_scale
        if k is None:
            k = torch.mean(mass / self.frozen_momentum_diagonal(state))
        return mass / k

    def apply_at(self, state, jacobian, a_p=None):
        """The preconditioner ``v -> M v`` (field-major FP64 vectors) at ``state``, with the saddle
        blocks read off ``jacobian`` (a compact or probed block Jacobian of this assembler, possibly
        with a pseudo-time shift on its diagonal) and the Schur built at ``a_p`` (default: the bare
        momentum diagonal at ``state``)."""
        op = self.assembler
        d = op.mesh.dim
        Bk, C = d + 1, op.mesh.n_cells
        if jacobian.B != Bk or jacobian.C != C:
            raise ValueError(f"jacobian blocks {jacobian.B}x{jacobian.B} over {jacobian.C} cells do not match the "
                             f"flow state ({Bk} fields over {C} cells)")
        dt, dev = jacobian.dtype, jacobian.vals.device
        prec, ns = jacobian.prec, jacobian.nslot
        L = _lib.lib()
        st = _lib.stream_of(dev)
        if a_p is None:
            a_p = self.frozen_momentum_diagonal(state)
        schur_ap = self.schur_diagonal(state, a_p)
        rho = op.density.detach()
        S = self.graph.laplacian(self.graph.volume / schur_ap, rho, op.pressure_pin, dt)
        S_inv_diag = (1.0 / S[0]).contiguous()
        S_cols = self.graph.cols
        inv_diag = (1.0 / schur_ap).to(dt)

        def spmv(block, nslot, vals, cols, x, n_rows=C):
            # the stream is read per launch, so that a graph capture records these launches too
            y = torch.empty_like(x)
            _lib.check(L.afx_bell_spmv(n_rows, block, nslot, prec, vals.data_ptr(), cols.data_ptr(), x.data_ptr(),
                                       y.data_ptr(), _lib.stream_of(x.device)), "afx_bell_spmv")
            return y

        schur_pre = S_inv_diag
        if self.hierarchy is not None:
            H = self.hierarchy
            ops, coarse_inverse = H.operators(S)

            def schur_pre(r):
                return H.native_cycle(ops, coarse_inverse, r, sweeps=self.schur_sweeps)

        def lib_dot(a, b):
            partial = torch.empty(256, dtype=torch.float64, device=a.device)
            out = torch.empty((), dtype=a.dtype, device=a.device)
            _lib.check(L.afx_dot(a.numel(), prec, a.data_ptr(), b.data_ptr(), partial.data_ptr(), out.data_ptr(),
                                 _lib.stream_of(a.device)), "afx_dot")
            return out

        def schur_solve(rp):
            # recorded: the library's dot (no library reduction that records a memory clear)
            return conjugate_gradient(lambda x: spmv(1, self.graph.nslot, S, S_cols, x), rp.contiguous(), schur_pre,
                                      self.schur_iters, dot=lib_dot if self.capture else torch.dot)

        if self.capture:
            schur_solve = captured(schur_solve, C, dt, dev)

        full = jacobian.vals.view(ns, Bk, Bk, C)
        Fv = full[:, :d, :d, :].contiguous()
        Finv = torch.empty(d * d * C, dtype=dt, device=dev)
        flag = torch.zeros(1, dtype=torch.int32, device=dev)
        _lib.check(L.afx_pbj_setup(C, d, prec, Fv.data_ptr(), Finv.data_ptr(), flag.data_ptr(), st), "afx_pbj_setup")

        def f_pbj(x):
            y = torch.empty_like(x)
            _lib.check(L.afx_pbj_apply(C, d, prec, Finv.data_ptr(), x.data_ptr(), y.data_ptr(), st), "afx_pbj_apply")
            return y

        if self.velocity == "pbj":
            def velocity_solve(ru):
                return f_pbj(ru.contiguous().reshape(-1)).view(C, d)
        else:
            from .solve import gmres

            def velocity_solve(ru):
                x, _ = gmres(lambda v: spmv(d, ns, Fv, jacobian.cols, v), ru.contiguous().reshape(-1),
                             rtol=self.velocity_rtol, restart=self.velocity_iters, max_iter=self.velocity_iters,
                             precond=f_pbj, check_every=None)
                return x.view(C, d)

        def column(x):
            return jacobian.matvec(x.reshape(-1)).view(C, Bk)

        def velocity_column(du):
            x = torch.zeros((C, Bk), dtype=dt, device=dev)
            x[:, :d] = du
            return column(x)

        def pressure_column(dp):
            x = torch.zeros((C, Bk), dtype=dt, device=dev)
            x[:, d] = dp
            y = column(x)
            return y[:, :d], y[:, d]

        solve = compose(self.composition, velocity_solve, schur_solve, lambda du: velocity_column(du)[:, d],
                        lambda dp: pressure_column(dp)[0], pressure_column, inv_diag)

        def apply(v):
            x = v.reshape(Bk, C).T.to(dt)
            du, dp = solve(x[:, :d].contiguous(), x[:, d].contiguous())
            return torch.cat([du.T, dp[None, :]]).reshape(-1).to(torch.float64)

        return apply
