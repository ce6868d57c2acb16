"""Numerical strategy objects: gradient reconstruction and slope limiting (``aquaflux/schemes``).

The residual kernels dispatch on these objects' codes.  Each object also keeps the reference's
stand-alone seam -- ``GradientScheme.gradients`` (``schemes/gradient.py:98-126``) and
``Limiter.limit`` (``schemes/limiter.py:78-117``) -- evaluated by the same HIP kernels through the
C ABI (``afx_scalar_gradient`` / ``afx_scalar_gradient_vjp``, ``afx_venkat_limiter``).  Gradients
are differentiable in both modes (they are linear in the field and the boundary values, and the
reverse pass is the hand-written transpose); the limiter in forward mode.
"""

from __future__ import annotations

import ctypes
import dataclasses

import torch

from . import _lib

_BC_NONE, _BC_DIRICHLET = 0, 1


def _boundary_kind(mesh, device):
    """int8 per face: Dirichlet on boundary faces (their values are given), none on interior faces."""
    dm = mesh.device_mesh(device)
    kind = getattr(dm, "_gradient_kind", None)
    if kind is None:
        kind = torch.where(dm.neighbour < 0, _BC_DIRICHLET, _BC_NONE).to(torch.int8).contiguous()
        dm._gradient_kind = kind
    return kind


class _GradientFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, scheme, mesh, geometry, field, bvals):
        ctx.scheme, ctx.mesh, ctx.geometry = scheme, mesh, geometry
        return scheme._apply(mesh, geometry, field, bvals)

    @staticmethod
    def backward(ctx, gbar):
        phi_bar, value_bar = ctx.scheme._transpose(ctx.mesh, ctx.geometry, gbar)
        return None, None, None, phi_bar, value_bar

    @staticmethod
    def jvp(ctx, _s, _m, _g, dfield, dbvals):
        mesh, geometry = ctx.mesh, ctx.geometry
        dev = geometry.face.area.device
        df = dfield if dfield is not None else torch.zeros(mesh.n_cells, dtype=torch.float64, device=dev)
        db = dbvals if dbvals is not None else torch.zeros(mesh.n_faces, dtype=torch.float64, device=dev)
        return ctx.scheme._apply(mesh, geometry, df, db)


class GradientScheme:
    """Strategy interface (``aquaflux/schemes/gradient.py:94-126``)."""

    code: int = -1

    def _params(self):
        return int(getattr(self, "sweeps", 4)), float(getattr(self, "relaxation", 1.0))

    def gradients(self, field, mesh, geometry, boundary_values, *, operator_hook=None):
        """Cell gradients ``(n_cells, dim)`` of ``field`` ``(n_cells,)`` with the face values
        ``boundary_values`` ``(n_faces,)`` on the boundary faces (interior entries are ignored), as
        ``GradientScheme.gradients`` (``schemes/gradient.py:98-126``).  ``operator_hook`` (the
        reference's matrix-free operator override) is not supported and must be None."""
        if operator_hook is not None:
            raise ValueError("gradients: operator_hook is not supported by the GPU gradient")
        _lib.require_cuda(field.contiguous(), "field")
        _lib.require_cuda(boundary_values.contiguous(), "boundary_values")
        if field.shape != (mesh.n_cells,) or boundary_values.shape != (mesh.n_faces,):
            raise ValueError(f"gradients: field must be ({mesh.n_cells},) and boundary_values ({mesh.n_faces},)")
        return _GradientFn.apply(self, mesh, geometry, field.contiguous(), boundary_values.contiguous())

    def _apply(self, mesh, geometry, field, bvals):
        dev = field.device
        dm = mesh.device_mesh(dev)
        G = torch.empty((mesh.dim, mesh.n_cells), dtype=torch.float64, device=dev)
        sweeps, relax = self._params()
        gs = geometry.c_struct()
        _lib.check(_lib.lib().afx_scalar_gradient(dm.handle, ctypes.byref(gs), int(self.code), sweeps, relax,
                                                  _boundary_kind(mesh, dev).data_ptr(), bvals.contiguous().data_ptr(),
                                                  field.contiguous().data_ptr(), G.data_ptr(), _lib.stream_of(dev)),
                   "afx_scalar_gradient")
        return G.T

    def _transpose(self, mesh, geometry, gbar):
        dev = gbar.device
        dm = mesh.device_mesh(dev)
        gb = gbar.T.contiguous()
        phi_bar = torch.empty(mesh.n_cells, dtype=torch.float64, device=dev)
        value_bar = torch.empty(mesh.n_faces, dtype=torch.float64, device=dev)
        sweeps, relax = self._params()
        gs = geometry.c_struct()
        _lib.check(_lib.lib().afx_scalar_gradient_vjp(dm.handle, ctypes.byref(gs), int(self.code), sweeps, relax,
                                                      _boundary_kind(mesh, dev).data_ptr(), gb.data_ptr(),
                                                      phi_bar.data_ptr(), value_bar.data_ptr(), _lib.stream_of(dev)),
                   "afx_scalar_gradient_vjp")
        return phi_bar, value_bar


class CompactGreenGauss(GradientScheme):
    """One-shot Green-Gauss with linearly interpolated face values (``gradient.py:129-153``)."""

    code = 0


@dataclasses.dataclass(frozen=True)
class SweptGradientSolve:
    """Fixed preconditioned-Richardson sweeps ``x <- x + r V^-1 (b - A_g x)`` from the peeled
    ``x0 = r V^-1 b`` (``aquaflux/schemes/gradient.py:546-705``).  ``warn_tol`` is accepted for API
    compatibility; the GPU path does not emit the under-resolution warning."""

    sweeps: int = 4
    relaxation: float = 1.0
    warn_tol: float | None = None

    def __post_init__(self):
        if self.relaxation <= 0.0:
            raise ValueError("SweptGradientSolve: relaxation must be positive")


@dataclasses.dataclass(frozen=True)
class CorrectedGreenGauss(GradientScheme):
    """Green-Gauss with the non-orthogonal skewness correction, ``A_g G = B phi`` solved by swept
    Richardson sweeps preconditioned by ``1/V`` (``aquaflux/schemes/gradient.py:1122-1337``).  Only
    :class:`SweptGradientSolve` and the inverse-volume preconditioner are provided on the GPU."""

    solver: SweptGradientSolve = dataclasses.field(default_factory=SweptGradientSolve)
    code = 3

    @property
    def sweeps(self) -> int:
        return int(self.solver.sweeps)

    @property
    def relaxation(self) -> float:
        return float(self.solver.relaxation)


class LeastSquaresGradient(GradientScheme):
    """Inverse-distance-squared weighted least squares over the face neighbours.

    Not present in the reference (its schemes are Green-Gauss variants); exact for linear fields on
    any mesh.  Boundary faces contribute the face-centroid offset and the boundary value.
    """

    code = 1


class _LimitFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, limiter, mesh, geometry, field, grad):
        ctx.limiter, ctx.mesh, ctx.geometry = limiter, mesh, geometry
        ctx.save_for_forward(field, grad)
        return limiter._apply(mesh, geometry, field, grad)[0]

    @staticmethod
    def backward(ctx, g):
        raise NotImplementedError("Limiter.limit has a forward-mode derivative only; the residual adjoints "
                                  "differentiate the limiter inside their reverse passes")

    @staticmethod
    def jvp(ctx, _l, _m, _g, dfield, dgrad):
        field, grad = ctx.saved_tensors
        return ctx.limiter._apply(ctx.mesh, ctx.geometry, field, grad, dfield, dgrad)[1]


class Limiter:
    """Slope-limiter interface (``aquaflux/schemes/limiter.py:40-63``)."""

    def limit(self, field, gradient, face_cells, geometry):
        """Per-cell limiter ``psi`` ``(n_cells,)`` in [0, 1] of ``field`` with its ``gradient``
        ``(n_cells, dim)``; ``face_cells`` is the mesh whose face-cell connectivity the stencil
        follows (``Limiter.limit``, ``schemes/limiter.py:78-117``)."""
        _lib.require_cuda(field.contiguous(), "field")
        _lib.require_cuda(gradient.contiguous(), "gradient")
        mesh = face_cells
        if field.shape != (mesh.n_cells,) or gradient.shape != (mesh.n_cells, mesh.dim):
            raise ValueError(f"limit: field must be ({mesh.n_cells},) and gradient ({mesh.n_cells}, {mesh.dim})")
        return _LimitFn.apply(self, mesh, geometry, field.contiguous(), gradient)

    def _apply(self, mesh, geometry, field, grad, dfield=None, dgrad=None):
        raise NotImplementedError


@dataclasses.dataclass(frozen=True)
class VenkatakrishnanLimiter(Limiter):
    """Venkatakrishnan (1993) limiter, ``eps^2 = V K^3`` (``aquaflux/schemes/limiter.py:66-117``)."""

    k: float = 5.0

    def _apply(self, mesh, geometry, field, grad, dfield=None, dgrad=None):
        dev = field.device
        dm = mesh.device_mesh(dev)
        gsoa = grad.T.contiguous()
        psi = torch.empty(mesh.n_cells, dtype=torch.float64, device=dev)
        tangent = dfield is not None or dgrad is not None
        psi_t = torch.empty_like(psi) if tangent else None
        ft = dfield.contiguous() if dfield is not None else None
        gt = dgrad.T.contiguous() if dgrad is not None else None
        gs = geometry.c_struct()
        _lib.check(_lib.lib().afx_venkat_limiter(dm.handle, ctypes.byref(gs), float(self.k), field.data_ptr(),
                                                 gsoa.data_ptr(), _lib.ptr(ft), _lib.ptr(gt), psi.data_ptr(),
                                                 _lib.ptr(psi_t), _lib.stream_of(dev)),
                   "afx_venkat_limiter")
        return psi, psi_t
