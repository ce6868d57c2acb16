"""Scalar transport on a given face flux (``aquaflux/transport/scalar.py``) on the GPU.

:class:`ScalarTransport` keeps the reference's ``build(mesh, geometry, diffusivity, boundary,
advection_scheme, *, gradient_scheme, sources, transient)``, ``with_diffusivity`` and
``residual(flux) -> (phi -> R)`` (``transport/scalar.py:134-223``).  The residual is a
``torch.autograd.Function`` whose backward is the hand-written HIP adjoint with respect to phi, the
diffusivity, the face flux, the boundary values and -- when the geometry came from node
coordinates that require grad -- the five geometry arrays, which the geometry's own HIP adjoint
carries on to the node coordinates.  The forward derivative (torch forward AD) is the dual-number
HIP tangent with respect to phi, the diffusivity, the face flux and the boundary values.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .boundary import BC_DIRICHLET, BoundaryConditions, Convective, Dirichlet, Neumann, ZeroGradient
from .discretization import AdvectionScheme, LimitedUpwind
from .properties import FieldProperty

F64 = torch.float64
DIFFUSIVITY = "diffusivity"


def effective_diffusivity(molecular, eddy_viscosity=None, turbulent_number: float = 0.7) -> FieldProperty:
    """``Gamma = D + nu_t / turbulent_number`` per cell (``transport/scalar.py:54-84``)."""
    if eddy_viscosity is None:
        return FieldProperty(molecular)
    return FieldProperty(molecular + eddy_viscosity / turbulent_number)


def face_mass_flux(geometry, velocity) -> torch.Tensor:
    """Owner-outward ``(u . n) A`` of a prescribed uniform or per-face velocity (the verification
    flux builder of the reference's ``tests/support/fields.py:16-43``); differentiable through the
    geometry."""
    n = geometry.face.normal_soa
    u = torch.as_tensor(velocity, dtype=F64, device=n.device)
    if u.dim() == 1:
        un = sum(u[j] * n[j] for j in range(n.shape[0]))
    else:
        un = sum(u[:, j] * n[j] for j in range(n.shape[0]))
    return un * geometry.face.area


def _boundary_arrays(bound: BoundaryConditions, mesh, device):
    """Per-face closure kind, value (Dirichlet value / Neumann q / convective h) and t_inf; tensor
    values stay differentiable."""
    F = mesh.n_faces
    kind = np.zeros(F, dtype=np.int8)
    val = torch.zeros(F, dtype=F64, device=device)
    tinf = torch.zeros(F, dtype=F64, device=device)
    for name, bc, faces in bound.items():
        if np.any(mesh.neighbour[faces] >= 0):
            raise ValueError(f"patch '{name}' names interior faces")
        if not isinstance(bc, (Dirichlet, ZeroGradient, Neumann, Convective)):
            raise ValueError(f"patch '{name}': {type(bc).__name__} is not a scalar boundary condition")
        kind[faces] = bc.kind
        idx = torch.from_numpy(np.asarray(faces, dtype=np.int64)).to(device)
        v, t = bc.values()
        for arr_name, x in (("val", v), ("tinf", t)):
            xt = x if isinstance(x, torch.Tensor) else torch.tensor(float(x), dtype=F64)
            xt = xt.to(device=device, dtype=F64).reshape(()).expand(idx.numel())
            if arr_name == "val":
                val = val.index_put((idx,), xt)
            else:
                tinf = tinf.index_put((idx,), xt)
    if np.any((mesh.neighbour < 0) & (kind == 0)):
        raise ValueError("every boundary face needs a scalar closure")
    return torch.from_numpy(kind).to(device), val, tinf


class _ScalarRaw:
    """Binds the C ABI calls of one configured scalar equation."""

    def __init__(self, mesh, geometry, cfg, kind):
        self.mesh = mesh
        self.dm = mesh.device_mesh(geometry.face.area.device)
        self.cfg = cfg
        self.kind = kind
        L = _lib.lib()
        self.work_bytes = []
        for mode in range(3):
            n = ctypes.c_int64()
            _lib.check(L.afx_scalar_workspace_bytes(self.dm.handle, mode, ctypes.byref(n)), "afx_scalar_workspace_bytes")
            self.work_bytes.append(int(n.value))

    @staticmethod
    def _gs(geo):
        return _lib.AfxGeometry(*[t.data_ptr() for t in geo])

    def residual(self, phi, gam, flux, val, tinf, geo):
        R = torch.empty_like(phi)
        w = self.dm.workspace("scalar0", self.work_bytes[0])
        gs = self._gs(geo)
        ins = _lib.AfxScalarInputs(*[_lib.ptr(t) for t in (phi, gam, flux, val, tinf)])
        _lib.check(
            _lib.lib().afx_scalar_residual(self.dm.handle, ctypes.byref(gs), ctypes.byref(self.cfg), ctypes.byref(ins),
                                           R.data_ptr(), w.data_ptr(), w.numel(), _lib.stream_of(self.dm.device)),
            "afx_scalar_residual",
        )
        return R

    def jvp(self, prim, tans, geo):
        R = torch.empty_like(prim[0])
        dR = torch.empty_like(prim[0])
        w = self.dm.workspace("scalar1", self.work_bytes[1])
        gs = self._gs(geo)
        ins = _lib.AfxScalarInputs(*[_lib.ptr(t) for t in prim])
        tns = _lib.AfxScalarInputs(*[_lib.ptr(t) for t in tans])
        _lib.check(
            _lib.lib().afx_scalar_jvp(self.dm.handle, ctypes.byref(gs), ctypes.byref(self.cfg), ctypes.byref(ins),
                                      ctypes.byref(tns), R.data_ptr(), dR.data_ptr(), w.data_ptr(), w.numel(),
                                      _lib.stream_of(self.dm.device)),
            "afx_scalar_jvp",
        )
        return R, dR

    def vjp(self, prim, cot, geo, with_geometry):
        bars = [torch.empty_like(t) for t in prim]
        gbars = [torch.empty_like(t) for t in geo] if with_geometry else [None] * 5
        w = self.dm.workspace("scalar2", self.work_bytes[2])
        gs = self._gs(geo)
        ins = _lib.AfxScalarInputs(*[_lib.ptr(t) for t in prim])
        outs = _lib.AfxScalarCotangents(*[b.data_ptr() for b in bars],
                                        _lib.AfxGeometryCotangents(*[_lib.ptr(b) for b in gbars]))
        _lib.check(
            _lib.lib().afx_scalar_vjp(self.dm.handle, ctypes.byref(gs), ctypes.byref(self.cfg), ctypes.byref(ins),
                                      cot.data_ptr(), ctypes.byref(outs), w.data_ptr(), w.numel(),
                                      _lib.stream_of(self.dm.device)),
            "afx_scalar_vjp",
        )
        return bars, gbars


class _ScalarResidualFn(torch.autograd.Function):
    """R(phi, gamma, flux, value, t_inf; area, normal, face centroid, volume, cell centroid)."""

    @staticmethod
    def forward(ctx, raw, phi, gam, flux, val, tinf, area, nrm, fc, vol, cc):
        ctx.raw = raw
        prim = (phi, gam, flux, val, tinf)
        geo = (area, nrm, fc, vol, cc)
        ctx.save_for_backward(*prim, *geo)
        ctx.save_for_forward(*prim, *geo)
        return raw.residual(*prim, geo)

    @staticmethod
    def backward(ctx, gR):
        saved = ctx.saved_tensors
        prim, geo = saved[:5], saved[5:]
        with_geo = any(ctx.needs_input_grad[6:11])
        bars, gbars = ctx.raw.vjp(prim, gR.contiguous(), geo, with_geo)
        return (None, *bars, *gbars)

    @staticmethod
    def jvp(ctx, _raw_t, dphi, dgam, dflux, dval, dtinf, *dgeo):
        if any(t is not None for t in dgeo):
            raise NotImplementedError("forward-mode derivatives with respect to the geometry are not provided")
        saved = ctx.saved_tensors
        prim, geo = saved[:5], saved[5:]
        tans = [None if t is None else t.contiguous() for t in (dphi, dgam, dflux, dval, dtinf)]
        _, dR = ctx.raw.jvp(prim, tans, geo)
        return dR


class ScalarTransport:
    """A configured scalar transport equation (``aquaflux/transport/scalar.py:86-223``)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    @classmethod
    def build(
        cls,
        mesh,
        geometry,
        diffusivity,
        boundary: BoundaryConditions,
        advection_scheme: AdvectionScheme | None,
        *,
        gradient_scheme=None,
        sources: tuple = (),
        transient=None,
    ) -> "ScalarTransport":
        if sources:
            raise ValueError("volume sources are not supported by the GPU scalar residual")
        if transient is not None:
            raise ValueError("transient terms are not supported by the GPU scalar residual")
        dev = geometry.face.area.device
        bound = boundary.resolve(mesh.face_patches)
        kind, val, tinf = _boundary_arrays(bound, mesh, dev)
        cfg = _lib.AfxScalarConfig()
        cfg.advection = 0 if advection_scheme is None else int(advection_scheme.code)
        cfg.gradient = 2 if gradient_scheme is None else int(gradient_scheme.code)
        if cfg.gradient not in (0, 1, 2, 3):
            raise ValueError(f"unsupported gradient scheme {type(gradient_scheme).__name__}")
        cfg.gradient_sweeps = int(getattr(gradient_scheme, "sweeps", 4))
        cfg.gradient_relaxation = float(getattr(gradient_scheme, "relaxation", 1.0))
        cfg.limiter_k = advection_scheme.limiter_k() if isinstance(advection_scheme, LimitedUpwind) else 5.0
        cfg.bc_kind = kind.data_ptr()
        return cls(mesh=mesh, geometry=geometry, diffusivity=diffusivity, boundary=bound,
                   advection_scheme=advection_scheme, gradient_scheme=gradient_scheme, sources=(), transient=None,
                   _raw=_ScalarRaw(mesh, geometry, cfg, kind), _val=val, _tinf=tinf)

    def with_diffusivity(self, diffusivity) -> "ScalarTransport":
        new = ScalarTransport(**self.__dict__)
        new.diffusivity = diffusivity
        return new

    def _gamma(self):
        g = self.diffusivity.evaluate(self.mesh.n_cells, self.geometry.face.area.device)
        return g.to(F64).contiguous()

    def _geometry(self):
        g = self.geometry
        return (g.face.area, g.face.normal_soa, g.face.centroid_soa, g.cell.volume, g.cell.centroid_soa)

    def residual(self, flux):
        """The residual function ``phi -> R(phi)`` for the owner-outward volumetric face flux ``flux``."""
        _lib.require_cuda(flux, "flux")
        if flux.shape != (self.mesh.n_faces,):
            raise ValueError(f"flux must have shape ({self.mesh.n_faces},), got {tuple(flux.shape)}")

        def fn(phi):
            _lib.require_cuda(phi, "phi")
            return _ScalarResidualFn.apply(self._raw, phi.contiguous(), self._gamma(), flux.contiguous(),
                                           self._val, self._tinf, *self._geometry())

        return fn

    def residual_jvp(self, flux, phi, tangent):
        """(R, J tangent) with respect to phi only (the Newton-Krylov matvec)."""
        prim = (phi.contiguous(), self._gamma().detach(), flux.detach().contiguous(), self._val.detach(),
                self._tinf.detach())
        geo = tuple(t.detach() for t in self._geometry())
        return self._raw.jvp(prim, (tangent.contiguous(), None, None, None, None), geo)

    def residual_vjp(self, flux, phi, cotangent, with_geometry=False):
        """All cotangents of the residual: ([phi, gamma, flux, value, t_inf], [geometry arrays])."""
        prim = (phi.contiguous(), self._gamma().detach(), flux.detach().contiguous(), self._val.detach(),
                self._tinf.detach())
        geo = tuple(t.detach() for t in self._geometry())
        return self._raw.vjp(prim, cotangent.contiguous(), geo, with_geometry)
