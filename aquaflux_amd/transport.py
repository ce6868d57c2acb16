"""Scalar transport on a given face flux (``aquaflux/transport/scalar.py``) on the GPU.

:class:`ScalarTransport` keeps the reference's ``build(mesh, geometry, diffusivity, boundary,
advection_scheme, *, gradient_scheme, sources, transient)``, ``with_diffusivity`` and
``residual(flux) -> (phi -> R)`` (``transport/scalar.py:134-223``).  The residual is a
``torch.autograd.Function`` whose backward is the hand-written HIP adjoint with respect to phi, the
diffusivity, the face flux, the boundary values and -- when the geometry came from node
coordinates that require grad -- the five geometry arrays, which the geometry's own HIP adjoint
carries on to the node coordinates.  The forward derivative (torch forward AD) is the dual-number
HIP tangent with respect to phi, the diffusivity, the face flux, the boundary values and -- for
Green-Gauss, least-squares or no gradient reconstruction -- the geometry arrays, whose own tangent
comes from the dual-number geometry kernels when the node coordinates carry one.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch
import torch.autograd.forward_ad as fwAD

from . import _lib
from .boundary import BC_DIRICHLET, BoundaryConditions, Convective, Dirichlet, DirichletField, Neumann, ZeroGradient
from .discretization import AdvectionScheme, LimitedUpwind, TransientTerm
from .properties import FieldProperty

F64 = torch.float64
DIFFUSIVITY = "diffusivity"


def effective_diffusivity(molecular, eddy_viscosity=None, turbulent_number: float = 0.7) -> FieldProperty:
    """``Gamma = D + nu_t / turbulent_number`` per cell (``transport/scalar.py:54-84``)."""
    if eddy_viscosity is None:
        return FieldProperty(molecular)
    return FieldProperty(molecular + eddy_viscosity / turbulent_number)


class _FaceFlux(torch.autograd.Function):
    """(u . n) A over the faces with the hand-written transpose (afx_face_flux / _vjp); the forward
    derivative (u . dn) A + (u . n) dA + (du . n) A is three passes of the same kernel."""

    @staticmethod
    def forward(ctx, n, area, u):
        d, F = n.shape
        per_face = int(u.dim() == 2)
        n, area, u = n.contiguous(), area.contiguous(), u.contiguous()
        out = torch.empty(F, dtype=F64, device=n.device)
        _lib.check(_lib.lib().afx_face_flux(F, d, per_face, u.data_ptr(), n.data_ptr(), area.data_ptr(),
                                            out.data_ptr(), _lib.stream_of(n.device)), "afx_face_flux")
        ctx.save_for_backward(n, area, u)
        ctx.save_for_forward(n, area, u)
        return out

    @staticmethod
    def _flux(u, n, area):
        d, F = n.shape
        out = torch.empty(F, dtype=F64, device=n.device)
        _lib.check(_lib.lib().afx_face_flux(F, d, int(u.dim() == 2), u.data_ptr(), n.data_ptr(), area.data_ptr(),
                                            out.data_ptr(), _lib.stream_of(n.device)), "afx_face_flux")
        return out

    @staticmethod
    def jvp(ctx, dn, darea, du):
        n, area, u = ctx.saved_tensors
        dm = torch.zeros(n.shape[1], dtype=F64, device=n.device)
        if dn is not None:
            dm += _FaceFlux._flux(u, dn.contiguous(), area)
        if darea is not None:
            dm += _FaceFlux._flux(u, n, darea.contiguous())
        if du is not None:
            dm += _FaceFlux._flux(du.contiguous(), n, area)
        return dm

    @staticmethod
    def backward(ctx, gm):
        n, area, u = ctx.saved_tensors
        d, F = n.shape
        per_face = int(u.dim() == 2)
        gm = gm.contiguous()
        nb = torch.empty_like(n) if ctx.needs_input_grad[0] else None
        ab = torch.empty_like(area) if ctx.needs_input_grad[1] else None
        ub = torch.empty_like(u) if (ctx.needs_input_grad[2] and per_face) else None
        _lib.check(_lib.lib().afx_face_flux_vjp(F, d, per_face, u.data_ptr(), n.data_ptr(), area.data_ptr(),
                                                gm.data_ptr(), None if nb is None else nb.data_ptr(),
                                                None if ab is None else ab.data_ptr(),
                                                None if ub is None else ub.data_ptr(), _lib.stream_of(n.device)),
                   "afx_face_flux_vjp")
        if ctx.needs_input_grad[2] and not per_face:  # a uniform velocity: its cotangent sums over faces
            ub = (n * area * gm).sum(dim=1)
        return nb, ab, ub


def face_mass_flux(geometry, velocity) -> torch.Tensor:
    """Owner-outward ``(u . n) A`` of a prescribed uniform or per-face velocity (the verification
    flux builder of the reference's ``tests/support/fields.py:16-43``); differentiable through the
    geometry (and the velocity) by its hand-written transpose.  A per-face velocity is ``(F, d)``."""
    n = geometry.face.normal_soa
    u = torch.as_tensor(velocity, dtype=F64, device=n.device)
    if not n.is_cuda:
        raise RuntimeError("face_mass_flux needs the geometry on the GPU")
    if u.dim() == 2:
        u = u.T
    return _FaceFlux.apply(n, geometry.face.area, u)


def _boundary_arrays(bound: BoundaryConditions, mesh, device, face_centroid=None):
    """Per-face closure kind, value (Dirichlet value / Neumann q / convective h) and t_inf; tensor
    values stay differentiable.  Position-dependent closures are evaluated at ``face_centroid``."""
    F = mesh.n_faces
    kind = np.zeros(F, dtype=np.int8)
    val = torch.zeros(F, dtype=F64, device=device)
    tinf = torch.zeros(F, dtype=F64, device=device)
    for name, bc, faces in bound.items():
        if np.any(mesh.neighbour[faces] >= 0):
            raise ValueError(f"patch '{name}' names interior faces")
        if not isinstance(bc, (Dirichlet, DirichletField, ZeroGradient, Neumann, Convective)):
            raise ValueError(f"patch '{name}': {type(bc).__name__} is not a scalar boundary condition")
        kind[faces] = bc.kind
        idx = torch.from_numpy(np.asarray(faces, dtype=np.int64)).to(device)
        if face_centroid is not None:
            v, t = bc.face_values(face_centroid[idx])
        else:
            v, t = bc.face_values(torch.zeros((idx.numel(), mesh.dim), dtype=F64, device=device))
        val = val.index_put((idx,), v)
        tinf = tinf.index_put((idx,), t)
    if np.any((mesh.neighbour < 0) & (kind == 0)):
        raise ValueError("every boundary face needs a scalar closure")
    return torch.from_numpy(kind).to(device), val, tinf


class _ScalarRaw:
    """Binds the C ABI calls of one configured scalar equation."""

    def __init__(self, mesh, geometry, cfg, kind):
        self.mesh = mesh
        self.geometry = geometry
        self.dm = mesh.device_mesh(geometry.face.area.device)
        self.cfg = cfg
        self.kind = kind
        L = _lib.lib()
        self.work_bytes = []
        for mode in range(3):
            n = ctypes.c_int64()
            _lib.check(L.afx_scalar_workspace_bytes(self.dm.handle, mode, ctypes.byref(n)), "afx_scalar_workspace_bytes")
            self.work_bytes.append(int(n.value))

    def _gs(self, geo):
        """The C geometry of `geo`; the internal-order face pack comes from the bound geometry when
        `geo` holds its face arrays (otherwise the library gathers one per call)."""
        g = self.geometry
        same = all(a.data_ptr() == b.data_ptr() for a, b in zip(geo[:3], (g.face.area, g.face.normal_soa,
                                                                          g.face.centroid_soa)))
        return _lib.AfxGeometry(*[t.data_ptr() for t in geo], _lib.ptr(g.face_pack() if same else None))

    def _cfg(self, step):
        """The configuration for one call: ``step`` is None (steady) or (bdf order, dt)."""
        if step is None:
            return self.cfg
        cfg = _lib.AfxScalarConfig.from_buffer_copy(self.cfg)
        cfg.transient, cfg.dt = step
        return cfg

    def residual(self, prim, geo, step=None):
        """prim = (phi, gamma, flux, value, t_inf, phi_old, phi_older); the last two may be None."""
        R = torch.empty_like(prim[0])
        w = self.dm.workspace("scalar0", self.work_bytes[0])
        gs = self._gs(geo)
        ins = _lib.AfxScalarInputs(*[_lib.ptr(t) for t in prim])
        _lib.check(
            _lib.lib().afx_scalar_residual(self.dm.handle, ctypes.byref(gs), ctypes.byref(self._cfg(step)),
                                           ctypes.byref(ins), R.data_ptr(), w.data_ptr(), w.numel(),
                                           _lib.stream_of(self.dm.device)),
            "afx_scalar_residual",
        )
        return R

    @property
    def records(self) -> bool:
        """Whether the residual can keep a forward record for its VJP (not for corrected Green-Gauss)."""
        return self.cfg.gradient != 3

    def residual_record(self, prim, geo, step=None):
        """(R, record): the residual and the gradient pass's outputs that vjp(record=...) reuses."""
        R = torch.empty_like(prim[0])
        n = ctypes.c_int64()
        _lib.check(_lib.lib().afx_scalar_record_bytes(self.dm.handle, ctypes.byref(n)), "afx_scalar_record_bytes")
        rec = torch.empty(int(n.value) // 8, dtype=F64, device=self.dm.device)
        gs = self._gs(geo)
        ins = _lib.AfxScalarInputs(*[_lib.ptr(t) for t in prim])
        _lib.check(
            _lib.lib().afx_scalar_residual_record(self.dm.handle, ctypes.byref(gs), ctypes.byref(self._cfg(step)),
                                                  ctypes.byref(ins), R.data_ptr(), rec.data_ptr(), rec.numel() * 8,
                                                  _lib.stream_of(self.dm.device)),
            "afx_scalar_residual_record",
        )
        return R, rec

    def jvp(self, prim, tans, geo, step=None):
        R = torch.empty_like(prim[0])
        dR = torch.empty_like(prim[0])
        w = self.dm.workspace("scalar1", self.work_bytes[1])
        gs = self._gs(geo)
        ins = _lib.AfxScalarInputs(*[_lib.ptr(t) for t in prim])
        tns = _lib.AfxScalarInputs(*[_lib.ptr(t) for t in tans])
        _lib.check(
            _lib.lib().afx_scalar_jvp(self.dm.handle, ctypes.byref(gs), ctypes.byref(self._cfg(step)),
                                      ctypes.byref(ins), ctypes.byref(tns), R.data_ptr(), dR.data_ptr(), w.data_ptr(),
                                      w.numel(), _lib.stream_of(self.dm.device)),
            "afx_scalar_jvp",
        )
        return R, dR

    def jvp_geometry(self, prim, tans, geo, dgeo, step=None):
        """(R, dR) along tangents of the inputs and of the five geometry arrays together."""
        R = torch.empty_like(prim[0])
        dR = torch.empty_like(prim[0])
        w = self.dm.workspace("scalar1", self.work_bytes[1])
        gs = self._gs(geo)
        dgeo = [t.contiguous() for t in dgeo]
        pack = None
        if self.dm.face_permuted:  # the face tangents in the kernels' internal face order
            pack = torch.empty((2 * self.dm.dim + 1) * self.dm.n_faces, dtype=F64, device=self.dm.device)
            fs = _lib.AfxGeometry(*[t.data_ptr() for t in dgeo[:3]], None, None, None)
            _lib.check(_lib.lib().afx_geometry_pack(self.dm.handle, ctypes.byref(fs), pack.data_ptr(),
                                                    _lib.stream_of(self.dm.device)), "afx_geometry_pack")
        gt = _lib.AfxGeometry(*[t.data_ptr() for t in dgeo], _lib.ptr(pack))
        ins = _lib.AfxScalarInputs(*[_lib.ptr(t) for t in prim])
        tns = _lib.AfxScalarInputs(*[_lib.ptr(t) for t in tans])
        _lib.check(
            _lib.lib().afx_scalar_jvp_geometry(self.dm.handle, ctypes.byref(gs), ctypes.byref(gt),
                                               ctypes.byref(self._cfg(step)), ctypes.byref(ins), ctypes.byref(tns),
                                               R.data_ptr(), dR.data_ptr(), w.data_ptr(), w.numel(),
                                               _lib.stream_of(self.dm.device)),
            "afx_scalar_jvp_geometry",
        )
        return R, dR

    def vjp(self, prim, cot, geo, with_geometry, step=None, record=None):
        """Cotangents of the inputs (and of the geometry); ``record`` is what residual_record kept at
        these inputs (the gradient pass is then not run again)."""
        bars = [None if t is None else torch.empty_like(t) for t in prim]
        gbars = [torch.empty_like(t) for t in geo] if with_geometry else [None] * 5
        w = self.dm.workspace("scalar2", self.work_bytes[2])
        gs = self._gs(geo)
        ins = _lib.AfxScalarInputs(*[_lib.ptr(t) for t in prim])
        outs = _lib.AfxScalarCotangents(*[_lib.ptr(b) for b in bars[:5]],
                                        _lib.AfxGeometryCotangents(*[_lib.ptr(b) for b in gbars]),
                                        *[_lib.ptr(b) for b in bars[5:]])
        L, cfg, st = _lib.lib(), ctypes.byref(self._cfg(step)), _lib.stream_of(self.dm.device)
        if record is None:
            rc = L.afx_scalar_vjp(self.dm.handle, ctypes.byref(gs), cfg, ctypes.byref(ins), cot.data_ptr(),
                                  ctypes.byref(outs), w.data_ptr(), w.numel(), st)
        else:
            rc = L.afx_scalar_vjp_record(self.dm.handle, ctypes.byref(gs), cfg, ctypes.byref(ins), cot.data_ptr(),
                                         record.data_ptr(), record.numel() * 8, ctypes.byref(outs), w.data_ptr(),
                                         w.numel(), st)
        _lib.check(rc, "afx_scalar_vjp" if record is None else "afx_scalar_vjp_record")
        return bars, gbars


class _ScalarResidualFn(torch.autograd.Function):
    """R(phi, gamma, flux, value, t_inf, phi_old, phi_older; area, normal, face centroid, volume,
    cell centroid) at the time step ``step`` (None steady, else (bdf order, dt)); ``geo_tan`` says
    whether the geometry arrays carry forward-mode tangents."""

    @staticmethod
    def forward(ctx, raw, step, geo_tan, phi, gam, flux, val, tinf, phi_old, phi_older, area, nrm, fc, vol, cc):
        ctx.raw, ctx.step, ctx.geo_tan = raw, step, geo_tan
        prim = (phi, gam, flux, val, tinf, phi_old, phi_older)
        geo = (area, nrm, fc, vol, cc)
        ctx.has_old = (phi_old is not None, phi_older is not None)
        saved = [t for t in prim if t is not None] + list(geo)
        ctx.save_for_backward(*saved)
        ctx.save_for_forward(*saved)
        ctx.record = None
        if any(ctx.needs_input_grad) and raw.records:  # keep the gradient pass for the reverse pass
            R, ctx.record = raw.residual_record(prim, geo, step)
            return R
        return raw.residual(prim, geo, step)

    @staticmethod
    def _unpack(ctx):
        saved = list(ctx.saved_tensors)
        prim = saved[:5]
        rest = saved[5:]
        for present in ctx.has_old:
            prim.append(rest.pop(0) if present else None)
        return tuple(prim), tuple(rest)

    @staticmethod
    def backward(ctx, gR):
        prim, geo = _ScalarResidualFn._unpack(ctx)
        with_geo = any(ctx.needs_input_grad[10:15])
        bars, gbars = ctx.raw.vjp(prim, gR.contiguous(), geo, with_geo, ctx.step, ctx.record)
        return (None, None, None, *bars, *gbars)

    @staticmethod
    def jvp(ctx, _raw_t, _step_t, _geo_tan_t, dphi, dgam, dflux, dval, dtinf, dold, dolder, *dgeo):
        prim, geo = _ScalarResidualFn._unpack(ctx)
        tans = [None if t is None else t.contiguous() for t in (dphi, dgam, dflux, dval, dtinf, dold, dolder)]
        if ctx.geo_tan:
            dgeo = [torch.zeros_like(g) if t is None else t for g, t in zip(geo, dgeo)]
            _, dR = ctx.raw.jvp_geometry(prim, tans, geo, dgeo, ctx.step)
        else:  # the geometry is constant: torch fills its tangents with zeros, which are not read
            _, dR = ctx.raw.jvp(prim, tans, geo, ctx.step)
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
        if transient is not None and not isinstance(transient, TransientTerm):
            raise ValueError(f"unsupported transient term {type(transient).__name__}")
        dev = geometry.face.area.device
        bound = boundary.resolve(mesh.face_patches)
        kind, val, tinf = _boundary_arrays(bound, mesh, dev, geometry.face.centroid.detach())
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
                   advection_scheme=advection_scheme, gradient_scheme=gradient_scheme, sources=(), transient=transient,
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

    def _time_level(self, phi, phi_old, phi_older, dt, first_step):
        """(step, phi_old, phi_older) of one call: the accumulation term of ``residual.py:141-147``
        when a transient term was injected, nothing otherwise (history arguments are then ignored,
        as in the reference)."""
        if self.transient is None:
            return None, None, None
        if phi_old is None or dt is None:
            raise ValueError("a transient residual needs phi_old and dt")
        if isinstance(dt, torch.Tensor):
            if dt.requires_grad:
                raise ValueError("derivatives with respect to dt are not provided; pass dt as a float")
            dt = float(dt)
        dt = float(dt)
        if not dt > 0.0:
            raise ValueError(f"dt must be positive, got {dt}")
        tensors = [phi_old] if first_step else [phi_old, phi_older]
        for name, t in zip(("phi_old", "phi_older"), tensors):
            if t is None:
                raise ValueError(f"the BDF2 step needs {name}")
            _lib.require_cuda(t, name)
            if t.shape != phi.shape:
                raise ValueError(f"{name} must have shape {tuple(phi.shape)}, got {tuple(t.shape)}")
        if first_step:
            return (1, dt), phi_old.contiguous(), None
        return (2, dt), phi_old.contiguous(), phi_older.contiguous()

    def residual(self, flux):
        """The residual function ``(phi, phi_old=None, phi_older=None, dt=None, first_step=False) ->
        R`` for the owner-outward volumetric face flux ``flux`` (``residual.py:358-412``)."""
        _lib.require_cuda(flux, "flux")
        if flux.shape != (self.mesh.n_faces,):
            raise ValueError(f"flux must have shape ({self.mesh.n_faces},), got {tuple(flux.shape)}")

        def fn(phi, phi_old=None, phi_older=None, dt=None, first_step=False):
            _lib.require_cuda(phi, "phi")
            step, old, older = self._time_level(phi, phi_old, phi_older, dt, first_step)
            geo = self._geometry()
            geo_tan = any(fwAD.unpack_dual(t).tangent is not None for t in geo)
            return _ScalarResidualFn.apply(self._raw, step, geo_tan, phi.contiguous(), self._gamma(),
                                           flux.contiguous(), self._val, self._tinf, old, older, *geo)

        return fn

    def _prim(self, flux, phi, old, older):
        return (phi.contiguous(), self._gamma().detach(), flux.detach().contiguous(), self._val.detach(),
                self._tinf.detach(), None if old is None else old.detach(), None if older is None else older.detach())

    def residual_jvp(self, flux, phi, tangent, phi_old=None, phi_older=None, dt=None, first_step=False):
        """(R, J tangent) with respect to phi only (the Newton-Krylov matvec)."""
        step, old, older = self._time_level(phi, phi_old, phi_older, dt, first_step)
        geo = tuple(t.detach() for t in self._geometry())
        return self._raw.jvp(self._prim(flux, phi, old, older), (tangent.contiguous(),) + (None,) * 6, geo, step)

    def residual_vjp(self, flux, phi, cotangent, with_geometry=False, phi_old=None, phi_older=None, dt=None,
                     first_step=False):
        """All cotangents of the residual: ([phi, gamma, flux, value, t_inf, phi_old, phi_older],
        [geometry arrays]); the history cotangents are None for a steady equation (phi_older also on
        the first step)."""
        step, old, older = self._time_level(phi, phi_old, phi_older, dt, first_step)
        geo = tuple(t.detach() for t in self._geometry())
        return self._raw.vjp(self._prim(flux, phi, old, older), cotangent.contiguous(), geo, with_geometry, step)
