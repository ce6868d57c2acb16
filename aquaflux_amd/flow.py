"""Coupled momentum + Rhie-Chow continuity residual on the GPU (``aquaflux/flow``).

:class:`MomentumContinuity` keeps the reference's constructor and methods
(``aquaflux/flow/momentum.py:233-924``): ``build(mesh, geometry, properties, gradient_scheme,
boundary, *, advection_scheme, pressure_pin, pressure_pin_value, body_force, sources)``,
``residual(state)``, ``unpack``/``pack``/``initial_state``, ``with_eddy_viscosity``.  The residual
is a ``torch.autograd.Function`` whose backward is the hand-written HIP adjoint and whose forward
derivative is the dual-number HIP tangent, so ``torch.autograd`` (reverse) and
``torch.autograd.forward_ad`` (forward) both differentiate through it with respect to the state,
the per-cell viscosity / density, and the boundary velocity / pressure.
"""

from __future__ import annotations

import ctypes
from typing import NamedTuple

import numpy as np
import torch
import torch.autograd.forward_ad as fwAD

from . import _lib
from .boundary import BoundaryConditions
from .discretization import AdvectionScheme, LimitedUpwind

FLOW_NONE, FLOW_WALL, FLOW_INLET, FLOW_OUTLET = 0, 1, 2, 3


# --- flow boundary bundles (flow/boundary.py:123-446) ----------------------------------------------


class FlowBoundary:
    kind = FLOW_NONE

    def prescribed_velocity(self, dim):
        return None

    def face_velocity(self, dim, centroids):
        """Prescribed velocity on faces with these centroids ``(n, d)``: ``(n, d)``, or None."""
        v = self.prescribed_velocity(dim)
        return None if v is None else v.to(centroids.device)[None, :].expand(centroids.shape[0], dim)

    def prescribed_pressure(self):
        return None

    def shears_flow(self) -> bool:
        return False


class NoSlipWall(FlowBoundary):
    """Zero velocity, zero-gradient pressure, no through-flow (``flow/boundary.py:278-305``)."""

    kind = FLOW_WALL

    def prescribed_velocity(self, dim):
        return torch.zeros(dim, dtype=torch.float64)

    def shears_flow(self):
        return True


class MovingWall(FlowBoundary):
    """Prescribed tangential velocity, no through-flow (``flow/boundary.py:308-346``)."""

    kind = FLOW_WALL

    def __init__(self, velocity):
        self.velocity = velocity

    def prescribed_velocity(self, dim):
        return _as_vector(self.velocity, dim)

    def shears_flow(self):
        return True


class VelocityInlet(FlowBoundary):
    """Prescribed velocity, zero-gradient pressure, ``mdot = rho (u_in . n) A``
    (``flow/boundary.py:360-397``).  ``velocity`` is a constant vector (a tensor may require grad)
    or a profile: a callable mapping face centroids ``(n, d)`` to velocities ``(n, d)``, evaluated
    at the patch's face centroids (``flow/boundary.py:372-376``)."""

    kind = FLOW_INLET

    def __init__(self, velocity):
        self.velocity = velocity

    def prescribed_velocity(self, dim):
        if _is_profile(self.velocity):
            raise ValueError("this VelocityInlet carries a profile; its velocity depends on the face centroid")
        return _as_vector(self.velocity, dim)

    def face_velocity(self, dim, centroids):
        if not _is_profile(self.velocity):
            return super().face_velocity(dim, centroids)
        v = self.velocity(centroids)
        v = v if isinstance(v, torch.Tensor) else torch.as_tensor(np.asarray(v), dtype=torch.float64)
        v = v.to(device=centroids.device, dtype=torch.float64)
        if v.shape != (centroids.shape[0], dim):
            raise ValueError(f"velocity profile must return shape ({centroids.shape[0]}, {dim}), got {tuple(v.shape)}")
        return v

    def face_velocity_function(self, face_centroid, velocity_owner, normal):
        """``FlowBoundary.velocity_face`` (``flow/boundary.py:372-376``) of this inlet."""
        return self.face_velocity(face_centroid.shape[1], face_centroid)


class PressureOutlet(FlowBoundary):
    """Prescribed pressure, zero-gradient velocity, Rhie-Chow mass flux toward ``p_b``
    (``flow/boundary.py:387-446``)."""

    kind = FLOW_OUTLET

    def __init__(self, pressure=0.0):
        self.pressure = pressure

    def prescribed_pressure(self):
        return self.pressure


def _is_profile(v):
    return callable(v) and not isinstance(v, torch.Tensor)


def _as_vector(v, dim):
    if _is_profile(v):
        raise ValueError("only a VelocityInlet takes a velocity profile; pass a constant vector")
    t = v if isinstance(v, torch.Tensor) else torch.tensor(list(v), dtype=torch.float64)
    t = t.to(torch.float64)
    if t.shape != (dim,):
        raise ValueError(f"velocity must have shape ({dim},), got {tuple(t.shape)}")
    return t


def _boundary_data(boundary, d, F, device, face_centroid=None):
    """Prescribed face velocity ``(d, F)`` and pressure ``(F,)`` folded patch by patch
    (``boundary/collection.py:97-137``); tensor-valued closures stay differentiable.  Profiles are
    evaluated at ``face_centroid`` ``(F, d)``."""
    ub = torch.zeros((F, d), dtype=torch.float64, device=device)
    pb = torch.zeros(F, dtype=torch.float64, device=device)
    for _, bc, faces in boundary.items():
        idx = torch.from_numpy(faces).to(device)
        if face_centroid is not None:
            v = bc.face_velocity(d, face_centroid[idx])
        else:
            v0 = bc.prescribed_velocity(d)
            v = None if v0 is None else v0.to(device)[None, :].expand(idx.numel(), d)
        if v is not None:
            ub = ub.index_put((idx,), v.to(device))
        p = bc.prescribed_pressure()
        if p is not None:
            pv = p if isinstance(p, torch.Tensor) else torch.tensor(float(p), dtype=torch.float64)
            pb = pb.index_put((idx,), pv.to(device).reshape(()).expand(idx.numel()))
    return ub.T.contiguous(), pb.contiguous()


# --- the residual operator ---------------------------------------------------------------------


class _Raw:
    """Binds the C ABI calls for one configured flow problem."""

    def __init__(self, mesh, geometry, cfg: _lib.AfxFlowConfig, kind: torch.Tensor):
        self.mesh = mesh
        self.dm = mesh.device_mesh(geometry.face.area.device)
        self.geometry = geometry
        self.gs = geometry.c_struct()
        self.cfg = cfg
        self.kind = kind  # keeps the device buffer alive
        L = _lib.lib()
        self.work_bytes = []
        for mode in range(4):
            n = ctypes.c_int64()
            _lib.check(L.afx_flow_workspace_bytes(self.dm.handle, mode, ctypes.byref(n)), "afx_flow_workspace_bytes")
            self.work_bytes.append(int(n.value))

    def inputs(self, state, mu, rho, ub, pb, gb=None):
        return _lib.AfxFlowInputs(
            _lib.ptr(state), _lib.ptr(mu), _lib.ptr(rho), _lib.ptr(ub), _lib.ptr(pb), _lib.ptr(gb)
        )

    def residual(self, state, mu, rho, ub, pb, gb=None, out=None):
        R = torch.empty_like(state) if out is None else out
        w = self.dm.workspace("flow0", self.work_bytes[0])
        ins = self.inputs(state, mu, rho, ub, pb, gb)
        _lib.check(
            _lib.lib().afx_flow_residual(
                self.dm.handle, ctypes.byref(self.gs), ctypes.byref(self.cfg), ctypes.byref(ins),
                R.data_ptr(), w.data_ptr(), w.numel(), _lib.stream_of(self.dm.device),
            ),
            "afx_flow_residual",
        )
        return R

    def jvp(self, primals, tangents):
        state = primals[0]
        R = torch.empty_like(state)
        dR = torch.empty_like(state)
        w = self.dm.workspace("flow1", self.work_bytes[1])
        ins = self.inputs(*primals)
        tns = self.inputs(*tangents)
        _lib.check(
            _lib.lib().afx_flow_jvp(
                self.dm.handle, ctypes.byref(self.gs), ctypes.byref(self.cfg), ctypes.byref(ins), ctypes.byref(tns),
                R.data_ptr(), dR.data_ptr(), w.data_ptr(), w.numel(), _lib.stream_of(self.dm.device),
            ),
            "afx_flow_jvp",
        )
        return R, dR

    def vjp(self, primals, cot):
        state, mu, rho, ub, pb = primals[:5]
        gb = primals[5] if len(primals) > 5 else None
        bars = [torch.empty_like(state), torch.empty_like(mu), torch.empty_like(rho), torch.empty_like(ub), torch.empty_like(pb)]
        gbb = torch.empty_like(gb) if gb is not None else None
        w = self.dm.workspace("flow2", self.work_bytes[2])
        ins = self.inputs(state, mu, rho, ub, pb, gb)
        outs = _lib.AfxFlowCotangents(*[b.data_ptr() for b in bars], _lib.ptr(gbb))
        _lib.check(
            _lib.lib().afx_flow_vjp(
                self.dm.handle, ctypes.byref(self.gs), ctypes.byref(self.cfg), ctypes.byref(ins), cot.data_ptr(),
                ctypes.byref(outs), w.data_ptr(), w.numel(), _lib.stream_of(self.dm.device),
            ),
            "afx_flow_vjp",
        )
        return bars + ([gbb] if gb is not None else [])

    def jvp_geometry(self, primals, tangents, dgeo):
        """(R, dR) along tangents of the inputs and of the five geometry arrays together."""
        state = primals[0]
        R = torch.empty_like(state)
        dR = torch.empty_like(state)
        w = self.dm.workspace("flow1", self.work_bytes[1])
        ins = self.inputs(*primals)
        tns = self.inputs(*tangents)
        gt = _geometry_tangent(self.dm, dgeo)
        _lib.check(
            _lib.lib().afx_flow_jvp_geometry(
                self.dm.handle, ctypes.byref(self.gs), ctypes.byref(gt[0]), ctypes.byref(self.cfg), ctypes.byref(ins),
                ctypes.byref(tns), R.data_ptr(), dR.data_ptr(), w.data_ptr(), w.numel(),
                _lib.stream_of(self.dm.device),
            ),
            "afx_flow_jvp_geometry",
        )
        return R, dR

    def vjp_geometry(self, primals, cot):
        """The cotangents of afx_flow_vjp plus those of the five geometry arrays (area, normal,
        face centroid, volume, cell centroid)."""
        state, mu, rho, ub, pb = primals[:5]
        gb = primals[5] if len(primals) > 5 else None
        bars = [torch.empty_like(state), torch.empty_like(mu), torch.empty_like(rho), torch.empty_like(ub), torch.empty_like(pb)]
        gbb = torch.empty_like(gb) if gb is not None else None
        g = self.geometry
        gbars = [torch.empty_like(t) for t in (g.face.area, g.face.normal_soa, g.face.centroid_soa, g.cell.volume,
                                                g.cell.centroid_soa)]
        w = self.dm.workspace("flow3", self.work_bytes[3])
        ins = self.inputs(state, mu, rho, ub, pb, gb)
        outs = _lib.AfxFlowCotangents(*[b.data_ptr() for b in bars], _lib.ptr(gbb))
        gouts = _lib.AfxGeometryCotangents(*[b.data_ptr() for b in gbars])
        _lib.check(
            _lib.lib().afx_flow_vjp_geometry(
                self.dm.handle, ctypes.byref(self.gs), ctypes.byref(self.cfg), ctypes.byref(ins), cot.data_ptr(),
                ctypes.byref(outs), ctypes.byref(gouts), w.data_ptr(), w.numel(), _lib.stream_of(self.dm.device),
            ),
            "afx_flow_vjp_geometry",
        )
        return bars + ([gbb] if gb is not None else []), gbars

    def mass_flux(self, primals, tangents=None):
        """Face mass flux ``(F,)`` (caller face order) and, with tangents, its tangent."""
        state = primals[0]
        F = self.mesh.n_faces
        mdot = torch.empty(F, dtype=torch.float64, device=state.device)
        dm = torch.empty_like(mdot) if tangents is not None else None
        mode = 0 if tangents is None else 1
        w = self.dm.workspace(f"flow{mode}", self.work_bytes[mode])
        ins = self.inputs(*primals)
        tns = self.inputs(*tangents) if tangents is not None else None
        _lib.check(
            _lib.lib().afx_flow_mass_flux(
                self.dm.handle, ctypes.byref(self.gs), ctypes.byref(self.cfg), ctypes.byref(ins),
                ctypes.byref(tns) if tns is not None else None, mdot.data_ptr(), _lib.ptr(dm), w.data_ptr(),
                w.numel(), _lib.stream_of(self.dm.device),
            ),
            "afx_flow_mass_flux",
        )
        return mdot, dm

    def fields(self, state, mu, rho, ub, pb, gb=None):
        d, C = self.mesh.dim, self.mesh.n_cells
        dev = state.device
        G = torch.empty(((d + 1) * d, C), dtype=torch.float64, device=dev)
        aP = torch.empty(C, dtype=torch.float64, device=dev)
        psi = torch.empty((d, C), dtype=torch.float64, device=dev) if self.cfg.advection == 2 else None
        ins = self.inputs(state, mu, rho, ub, pb, gb)
        _lib.check(
            _lib.lib().afx_flow_fields(
                self.dm.handle, ctypes.byref(self.gs), ctypes.byref(self.cfg), ctypes.byref(ins),
                G.data_ptr(), aP.data_ptr(), _lib.ptr(psi), _lib.stream_of(dev),
            ),
            "afx_flow_fields",
        )
        return G, aP, psi


class _FlowResidualFn(torch.autograd.Function):
    """R(state, mu, rho, u_b, p_b[, gamma_b]; area, normal, face centroid, volume, cell centroid)
    with HIP forward, tangent and adjoint, including those with respect to the geometry arrays
    (forward mode when ``geo_tan``; compact Green-Gauss or least squares), which the geometry's own
    derivatives carry on to the node coordinates."""

    @staticmethod
    def forward(ctx, raw, geo_tan, state, mu, rho, ub, pb, gb, area, nrm, fc, vol, cc):
        ctx.raw = raw
        ctx.has_gb = gb is not None
        ctx.geo_tan = geo_tan
        prim = (state, mu, rho, ub, pb) + ((gb,) if gb is not None else ())
        ctx.save_for_backward(*prim)
        ctx.save_for_forward(*prim, area, nrm, fc, vol, cc)
        return raw.residual(state, mu, rho, ub, pb, gb)

    @staticmethod
    def backward(ctx, gR):
        prim = ctx.saved_tensors
        if any(ctx.needs_input_grad[8:13]):
            bars, gbars = ctx.raw.vjp_geometry(prim, gR.contiguous())
        else:
            bars, gbars = ctx.raw.vjp(prim, gR.contiguous()), [None] * 5
        gbb = bars[5] if ctx.has_gb else None
        return (None, None, bars[0], bars[1], bars[2], bars[3], bars[4], gbb, *gbars)

    @staticmethod
    def jvp(ctx, _raw_t, _geo_tan_t, dstate, dmu, drho, dub, dpb, dgb, *dgeo):
        saved = ctx.saved_tensors
        prim, geo = saved[:-5], saved[-5:]
        tans = [t.contiguous() if t is not None else None for t in (dstate, dmu, drho, dub, dpb)]
        if ctx.has_gb:
            tans.append(dgb.contiguous() if dgb is not None else None)
        if ctx.geo_tan:
            dgeo = [torch.zeros_like(g) if t is None else t for g, t in zip(geo, dgeo)]
            _, dR = ctx.raw.jvp_geometry(prim, tans, dgeo)
        else:  # constant geometry: torch fills its tangents with zeros, which are not read
            _, dR = ctx.raw.jvp(prim, tans)
        return dR


def _geometry_tangent(dm, dgeo):
    """The C struct of geometry tangents (area, normal, face centroid, volume, cell centroid), with
    the face tangents packed in the kernels' internal face order when the mesh reorders faces.
    Returns (struct, tensors kept alive)."""
    dgeo = [t.contiguous() for t in dgeo]
    pack = None
    if dm.face_permuted:
        pack = torch.empty((2 * dm.dim + 1) * dm.n_faces, dtype=torch.float64, device=dm.device)
        fs = _lib.AfxGeometry(*[t.data_ptr() for t in dgeo[:3]], None, None, None)
        _lib.check(_lib.lib().afx_geometry_pack(dm.handle, ctypes.byref(fs), pack.data_ptr(),
                                                _lib.stream_of(dm.device)), "afx_geometry_pack")
    return _lib.AfxGeometry(*[t.data_ptr() for t in dgeo], _lib.ptr(pack)), (dgeo, pack)


class _MassFluxFn(torch.autograd.Function):
    """mdot(state, mu, rho, u_b, p_b[, gamma_b]): HIP forward and dual-number tangent."""

    @staticmethod
    def forward(ctx, raw, state, mu, rho, ub, pb, gb):
        ctx.raw = raw
        ctx.has_gb = gb is not None
        prim = (state, mu, rho, ub, pb) + ((gb,) if gb is not None else ())
        ctx.save_for_forward(*prim)
        return raw.mass_flux(prim)[0]

    @staticmethod
    def backward(ctx, g):
        raise NotImplementedError("mass_flux has a forward-mode derivative only; differentiate the residual, whose "
                                  "adjoint carries mdot's, for reverse mode")

    @staticmethod
    def jvp(ctx, _raw_t, dstate, dmu, drho, dub, dpb, dgb):
        prim = ctx.saved_tensors
        tans = [t.contiguous() if t is not None else None for t in (dstate, dmu, drho, dub, dpb)]
        if ctx.has_gb:
            tans.append(dgb.contiguous() if dgb is not None else None)
        return ctx.raw.mass_flux(prim, tans)[1]


class FlowFields(NamedTuple):
    """``FlowFields`` (``flow/momentum.py:92-109``) with the velocity fields flattened in: velocity
    ``(C, d)``, boundary velocity ``(F, d)``, velocity gradient ``(C, d, d)`` (``[c, i, j] = du_i /
    dx_j``), pressure ``(C,)``, boundary pressure ``(F,)`` (the owner value on walls and inlets),
    pressure gradient ``(C, d)``, mass flux ``(F,)``; plus the momentum diagonal ``a_P`` ``(C,)``."""

    velocity: torch.Tensor
    boundary_velocity: torch.Tensor
    grad_velocity: torch.Tensor
    pressure: torch.Tensor
    boundary_pressure: torch.Tensor
    grad_pressure: torch.Tensor
    mdot: torch.Tensor
    a_p: torch.Tensor


class MomentumContinuity:
    """Coupled momentum + Rhie-Chow continuity residual (``aquaflux/flow/momentum.py:158-924``)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    @classmethod
    def build(
        cls,
        mesh,
        geometry,
        properties,
        gradient_scheme,
        boundary: BoundaryConditions,
        *,
        advection_scheme: AdvectionScheme | None = None,
        pressure_pin: int | None = None,
        pressure_pin_value: float = 0.0,
        body_force=None,
        sources: tuple = (),
    ) -> "MomentumContinuity":
        properties.require("viscosity", "density")
        if sources:
            raise ValueError("momentum sources are not supported by the GPU residual")
        d = mesh.dim
        dev = geometry.face.area.device
        bound = boundary.resolve(mesh.face_patches)
        kind = np.zeros(mesh.n_faces, dtype=np.int8)
        for name, bc, faces in bound.items():
            if not isinstance(bc, FlowBoundary):
                raise ValueError(f"patch '{name}': {type(bc).__name__} is not a flow boundary")
            if np.any(mesh.neighbour[faces] >= 0):
                raise ValueError(f"patch '{name}' names interior faces; baffles are not supported")
            kind[faces] = bc.kind
        cfg = _lib.AfxFlowConfig()
        cfg.advection = 0 if advection_scheme is None else int(advection_scheme.code)
        cfg.gradient = int(gradient_scheme.code)
        if cfg.gradient not in (0, 1, 3):
            raise ValueError(f"unsupported gradient scheme {type(gradient_scheme).__name__}")
        cfg.gradient_sweeps = int(getattr(gradient_scheme, "sweeps", 4))
        cfg.gradient_relaxation = float(getattr(gradient_scheme, "relaxation", 1.0))
        cfg.pressure_pin = -1 if pressure_pin is None else int(pressure_pin)
        cfg.pressure_pin_value = float(pressure_pin_value)
        bf = [0.0, 0.0, 0.0]
        if body_force is not None:
            for i, v in enumerate(np.asarray(body_force, dtype=float).ravel()[:d]):
                bf[i] = float(v)
        cfg.body_force = (ctypes.c_double * 3)(*bf)
        cfg.limiter_k = advection_scheme.limiter_k() if isinstance(advection_scheme, LimitedUpwind) else 5.0
        kind_dev = torch.from_numpy(kind).to(dev)
        cfg.face_kind = kind_dev.data_ptr()
        raw = _Raw(mesh, geometry, cfg, kind_dev)
        return cls(
            mesh=mesh,
            geometry=geometry,
            properties=properties,
            gradient_scheme=gradient_scheme,
            advection_scheme=advection_scheme,
            boundary=bound,
            pressure_pin=pressure_pin,
            pressure_pin_value=pressure_pin_value,
            body_force=bf[:d],
            eddy_viscosity=None,
            wall_eddy_viscosity=None,
            _raw=raw,
        )

    # --- layout (flow/state.py:17-73) ------------------------------------------------------------
    @property
    def dim(self):
        return self.mesh.dim

    @property
    def device(self):
        return self.geometry.face.area.device

    def unpack(self, state):
        blocks = state.reshape(self.dim + 1, self.mesh.n_cells)
        return blocks[: self.dim].T, blocks[self.dim]

    def pack(self, velocity, pressure):
        return torch.cat([velocity.T, pressure[None, :]]).reshape(-1)

    def initial_state(self):
        return torch.zeros((self.dim + 1) * self.mesh.n_cells, dtype=torch.float64, device=self.device)

    # --- properties ------------------------------------------------------------------------------
    def with_eddy_viscosity(self, eddy_viscosity, wall_eddy_viscosity=None) -> "MomentumContinuity":
        """Copy carrying a kinematic eddy viscosity (``flow/momentum.py:314-349``)."""
        new = MomentumContinuity(**self.__dict__)
        new.eddy_viscosity = eddy_viscosity
        new.wall_eddy_viscosity = wall_eddy_viscosity
        return new

    def material(self):
        props = self.properties.evaluate(self.mesh.n_cells, self.device)
        return props["viscosity"], props["density"]

    @property
    def viscosity(self):
        mu, rho = self.material()
        return mu if self.eddy_viscosity is None else mu + rho * self.eddy_viscosity

    @property
    def density(self):
        return self.material()[1]

    def shear_mask(self):
        mask = torch.zeros(self.mesh.n_faces, dtype=torch.bool)
        for _, bc, faces in self.boundary.items():
            if bc.shears_flow():
                mask[torch.from_numpy(faces)] = True
        return mask.to(self.device)

    def boundary_data(self):
        """Per-face prescribed velocity ``(d, F)`` and outlet pressure ``(F,)`` (differentiable)."""
        return _boundary_data(self.boundary, self.dim, self.mesh.n_faces, self.device,
                              self.geometry.face.centroid.detach())

    def _inputs(self):
        mu, rho = self.material()
        mu_eff = mu if self.eddy_viscosity is None else mu + rho * self.eddy_viscosity
        ub, pb = self.boundary_data()
        gb = None
        if self.wall_eddy_viscosity is not None:
            own = self.mesh.device_mesh(self.device).owner
            wall_mu = mu[own] + rho[own] * self.wall_eddy_viscosity
            gb = torch.where(self.shear_mask(), wall_mu, mu_eff[own]).contiguous()
        return mu_eff.contiguous(), rho.contiguous(), ub, pb, gb

    # --- residual --------------------------------------------------------------------------------
    def _geometry(self):
        g = self.geometry
        return (g.face.area, g.face.normal_soa, g.face.centroid_soa, g.cell.volume, g.cell.centroid_soa)

    def residual(self, state):
        """Coupled residual, same shape as ``state`` (``flow/momentum.py:880-882``).  Forward-mode
        derivatives also follow geometry tangents (node coordinates that carry one) for compact
        Green-Gauss or least-squares gradients."""
        _lib.require_cuda(state, "state")
        mu, rho, ub, pb, gb = self._inputs()
        geo = self._geometry()
        geo_tan = any(fwAD.unpack_dual(t).tangent is not None for t in geo)
        return _FlowResidualFn.apply(self._raw, geo_tan, state.contiguous(), mu, rho, ub, pb, gb, *geo)

    __call__ = residual

    def residual_raw(self, state):
        """R(state) with no autograd bookkeeping (Newton iterations, line searches)."""
        mu, rho, ub, pb, gb = self._inputs()
        return self._raw.residual(state.contiguous(), mu.detach(), rho.detach(), ub.detach(), pb.detach(),
                                  None if gb is None else gb.detach())

    def residual_jvp(self, state, tangent):
        """(R, J tangent) with respect to the state only (the Newton-Krylov matvec)."""
        mu, rho, ub, pb, gb = self._inputs()
        prim = [state.contiguous(), mu.detach(), rho.detach(), ub.detach(), pb.detach()]
        if gb is not None:
            prim.append(gb.detach())
        tans = [tangent.contiguous(), None, None, None, None] + ([None] if gb is not None else [])
        return self._raw.jvp(prim, tans)

    def residual_vjp(self, state, cotangent):
        """J^T cotangent with respect to the state (the adjoint-Krylov matvec)."""
        mu, rho, ub, pb, gb = self._inputs()
        prim = [state.contiguous(), mu.detach(), rho.detach(), ub.detach(), pb.detach()]
        if gb is not None:
            prim.append(gb.detach())
        return self._raw.vjp(prim, cotangent.contiguous())[0]

    def mass_flux(self, state):
        """The Rhie-Chow face mass flux ``mdot`` ``(F,)`` that continuity sums and the advection
        fluxes carry (``flow/momentum.py:884-899``); differentiable in forward mode."""
        _lib.require_cuda(state, "state")
        mu, rho, ub, pb, gb = self._inputs()
        return _MassFluxFn.apply(self._raw, state.contiguous(), mu, rho, ub, pb, gb)

    def flow_fields(self, state):
        """The shared flow quantities at ``state`` (``flow/momentum.py:822-857``): boundary fields,
        gradients, ``a_P`` and the mass flux, as a :class:`FlowFields` (values, no autograd)."""
        _lib.require_cuda(state, "state")
        d, C = self.dim, self.mesh.n_cells
        with torch.no_grad():
            G, aP, _ = self.fields(state)
            vel, p = self.unpack(state)
            ub, pb = self._inputs()[2:4]
            own = self.mesh.device_mesh(self.device).owner
            kinds = self._face_kinds()
            bvel = torch.where((kinds == FLOW_OUTLET)[:, None], vel[own], ub.T)
            bp = torch.where(kinds == FLOW_OUTLET, pb, torch.where(kinds == FLOW_NONE, torch.zeros_like(pb), p[own]))
            Gr = G.reshape(d + 1, d, C)
            return FlowFields(vel, bvel, Gr[:d].permute(2, 0, 1), p, bp, Gr[d].T, self.mass_flux(state).detach(), aP)

    def _face_kinds(self):
        return self._raw.kind.to(torch.int64)

    def fields(self, state):
        """Gradients ``((d+1) d, C)``, ``a_P`` ``(C,)`` and limiter ``psi`` for a state."""
        mu, rho, ub, pb, gb = self._inputs()
        return self._raw.fields(state.contiguous(), mu.detach(), rho.detach(), ub.detach(), pb.detach(),
                                None if gb is None else gb.detach())
