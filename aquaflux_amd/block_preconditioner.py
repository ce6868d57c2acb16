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

Inner solves are GPU Krylov iterations where the reference builds smoothed-aggregation multigrid
hierarchies on the host: the velocity block is solved by GMRES with its point-block Jacobi (or by
that Jacobi alone), the Schur by a fixed number of Jacobi-preconditioned conjugate-gradient steps
with no host synchronisation.  Both inner solves are therefore not fixed linear maps: the outer
solve must be flexible GMRES, which the Newton step uses.  The three compositions
(triangular / simple / simpler) follow Klaij & Vuik (2013) as the reference states them
(``block_preconditioner.py:1069-1176``).
"""

from __future__ import annotations

import dataclasses

import numpy as np
import torch

from . import _lib

VELOCITY_SOLVES = ("gmres", "pbj")
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


def conjugate_gradient(matvec, b, inverse_diagonal, iterations: int):
    """``iterations`` Jacobi-preconditioned CG steps from zero, with no host synchronisation (the
    step sizes stay on the device; a zero curvature or residual leaves the iterate unchanged)."""
    x = torch.zeros_like(b)
    r = b.clone()
    z = r * inverse_diagonal
    p = z.clone()
    rz = torch.dot(r, z)
    zero = torch.zeros((), dtype=b.dtype, device=b.device)
    for _ in range(int(iterations)):
        Ap = matvec(p)
        pAp = torch.dot(p, Ap)
        alpha = torch.where(pAp > 0, rz / torch.where(pAp > 0, pAp, torch.ones_like(pAp)), zero)
        x.add_(alpha * p)
        r.sub_(alpha * Ap)
        z = r * inverse_diagonal
        rz_new = torch.dot(r, z)
        beta = torch.where(rz > 0, rz_new / torch.where(rz > 0, rz, torch.ones_like(rz)), zero)
        p = z + beta * p
        rz = rz_new
    return x


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
    ``"triangular"``, ``"simple"`` or ``"simpler"``.  ``schur_iters`` conjugate-gradient steps
    per Schur solve.
    """

    def __init__(self, assembler, graph, velocity, schur_scaling, composition, mass_scale, velocity_iters,
                 velocity_rtol, schur_iters):
        self.assembler = assembler
        self.graph = graph
        self.velocity = velocity
        self.schur_scaling = schur_scaling
        self.composition = composition
        self.mass_scale = mass_scale
        self.velocity_iters = int(velocity_iters)
        self.velocity_rtol = float(velocity_rtol)
        self.schur_iters = int(schur_iters)

    @classmethod
    def build(cls, assembler, *, velocity: str = "gmres", schur_scaling: str = "simple",
              composition: str = "simple", mass_scale: float | None = None, velocity_iters: int = 10,
              velocity_rtol: float = 1e-2, schur_iters: int = 30) -> "BlockPreconditioner":
        if velocity not in VELOCITY_SOLVES:
            raise ValueError(f"unknown velocity block {velocity!r}; use one of {VELOCITY_SOLVES} (the reference's "
                             "multigrid velocity hierarchies are replaced by GPU Krylov inner solves)")
        if schur_scaling not in SCHUR_SCALINGS:
            raise ValueError(f"unknown schur_scaling {schur_scaling!r}; use one of {SCHUR_SCALINGS}")
        if composition not in COMPOSITIONS:
            raise ValueError(f"unknown composition {composition!r}; use one of {COMPOSITIONS}")
        if not hasattr(assembler, "pressure_pin") or not hasattr(assembler, "fields"):
            raise ValueError("BlockPreconditioner needs a MomentumContinuity assembler")
        return cls(assembler, SchurGraph.of(assembler), velocity, schur_scaling, composition, mass_scale,
                   velocity_iters, velocity_rtol, schur_iters)

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

        def spmv(block, nslot, vals, cols, x):
            y = torch.empty_like(x)
            _lib.check(L.afx_bell_spmv(C, block, nslot, prec, vals.data_ptr(), cols.data_ptr(), x.data_ptr(),
                                       y.data_ptr(), st), "afx_bell_spmv")
            return y

        def schur_solve(rp):
            return conjugate_gradient(lambda x: spmv(1, self.graph.nslot, S, S_cols, x), rp.contiguous(), S_inv_diag,
                                      self.schur_iters)

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
                             precond=f_pbj)
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
