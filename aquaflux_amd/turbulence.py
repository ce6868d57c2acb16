"""k-omega SST turbulence and the coupled RANS residual on the GPU (``aquaflux/turbulence``).

:class:`SSTModel` carries the reference's constants (``turbulence/sst.py:37-105``).
:class:`SSTTurbulence` keeps ``SSTTurbulence.build(model, mesh, geometry, gradient_scheme,
advection_scheme, density, molecular_viscosity, wall_patches, k_boundary, omega_boundary)``
(``turbulence/transport.py:259-313``): the wall distance (nearest wall-face centroid through a k-d
tree, as ``mesh/distance.py:52-113``), the wall-adjacent cell set and the wall-distance gradient
are set-up products.  :class:`CoupledRANS` keeps ``CoupledRANS.build(momentum, turbulence)`` and
``residual(state)`` (``turbulence/coupled.py:306-473``) on the state ``[u, v, (w), p, k, omega]``;
the residual is a ``torch.autograd.Function`` whose backward is the hand-written HIP adjoint and
whose forward derivative is the dual-number HIP tangent, with respect to the state, the molecular
viscosity (dynamic ``mu`` of the momentum block and kinematic ``nu`` of the SST block), the
density and the boundary velocity / pressure.
"""

from __future__ import annotations

import ctypes
import dataclasses
import math

import numpy as np
import torch
import torch.autograd.forward_ad as fwAD

from . import _lib
from .boundary import BC_DIRICHLET, BC_ZERO_GRADIENT, BoundaryConditions, Dirichlet, ZeroGradient
from .discretization import AdvectionScheme, DifferenceRow, LimitedUpwind, LogRatioRow

F64 = torch.float64


@dataclasses.dataclass(frozen=True)
class SSTModel:
    """The k-omega SST constants (``turbulence/sst.py:37-82``)."""

    a1: float = 0.31
    beta_star: float = 0.09
    alpha_1: float = 5.0 / 9.0
    beta_1: float = 3.0 / 40.0
    sigma_k1: float = 0.85
    sigma_omega1: float = 0.5
    alpha_2: float = 0.44
    beta_2: float = 0.0828
    sigma_k2: float = 1.0
    sigma_omega2: float = 0.856
    kappa: float = 0.41
    e_wall: float = 9.8
    wall_omega_exponent: float = 2.0
    wall_omega_viscous_coeff: float = 1.0

    @property
    def wall_y_star_lam(self) -> float:
        """Root of ``y = ln(E y) / kappa`` by 10 fixed-point steps from 11 (``sst.py:84-105``)."""
        y = 11.0
        for _ in range(10):
            y = math.log(self.e_wall * y) / self.kappa
        return y

    def blend(self, f1, inner, outer):
        return f1 * inner + (1.0 - f1) * outer

    def c_constants(self):
        return [
            self.a1, self.beta_star, self.alpha_1, self.beta_1, self.sigma_k1, self.sigma_omega1,
            self.alpha_2, self.beta_2, self.sigma_k2, self.sigma_omega2, self.kappa, self.e_wall,
            self.wall_omega_exponent, self.wall_omega_viscous_coeff, self.beta_star**0.25,
            self.wall_y_star_lam,
        ]


class ScalarVariableTransform:
    """Map between a solved turbulence unknown ``w`` and its physical field
    (``turbulence/coupled.py:111-152``).  ``code`` is the C ABI's AFX_SCALAR_* value: the residual
    kernels apply the map themselves, so these methods serve callers that build or read states."""

    code = -1


class DirectScalars(ScalarVariableTransform):
    """The identity parametrization (``turbulence/coupled.py:155-175``)."""

    code = 0

    def to_physical(self, w):
        return w

    def to_solved(self, phi):
        return phi

    def jacobian_scale(self, phi):
        return torch.ones_like(phi)

    def fixation_row(self):
        return DifferenceRow()


class LogScalars(ScalarVariableTransform):
    """``phi = e^w`` (``turbulence/coupled.py:178-211``): positive under any Newton step; the wall
    fixation becomes the log-ratio row."""

    code = 1

    def to_physical(self, w):
        return torch.exp(w)

    def to_solved(self, phi):
        return torch.log(phi)

    def jacobian_scale(self, phi):
        return phi

    def fixation_row(self):
        return LogRatioRow()


def _scalar_face_data(bc: BoundaryConditions, mesh, what):
    """Per-face closure kind and Dirichlet value of a k / omega boundary (host arrays)."""
    kind = np.zeros(mesh.n_faces, dtype=np.int8)
    value = np.zeros(mesh.n_faces, dtype=np.float64)
    for name, c, faces in bc.items():
        if np.any(mesh.neighbour[faces] >= 0):
            raise ValueError(f"{what} patch '{name}' names interior faces")
        if isinstance(c, Dirichlet):
            kind[faces] = BC_DIRICHLET
            value[faces] = float(c.value)
        elif isinstance(c, ZeroGradient):
            kind[faces] = BC_ZERO_GRADIENT
        else:
            raise ValueError(f"{what} patch '{name}': only Dirichlet / ZeroGradient closures are supported")
    return kind, value


def wall_distance(mesh, geometry, wall_faces, with_nearest=False):
    """Distance from each cell centroid to the nearest wall-face centroid (``mesh/distance.py:52-113``).

    A set-up product computed once per build on the host with a k-d tree, as the reference does for
    concrete inputs; the result is ``sqrt(distance**2)`` like the reference's squared-distance path.
    With ``with_nearest`` also the nearest wall face of each cell (mesh face index), along which a
    node-coordinate tangent moves the distance (the reference's traced path differentiates the
    same nearest-centroid distance, ``mesh/distance.py:115-140``)."""
    from scipy.spatial import cKDTree

    if len(wall_faces) == 0:
        raise ValueError("wall_distance: the wall patches contain no faces")
    fc = geometry.face.centroid_soa.detach().cpu().numpy().T[wall_faces]
    cc = geometry.cell.centroid_soa.detach().cpu().numpy().T
    dist, idx = cKDTree(fc).query(cc)
    d2 = torch.tensor(dist, dtype=F64) ** 2
    d = torch.sqrt(d2).to(geometry.face.area.device)
    if with_nearest:
        return d, torch.from_numpy(np.asarray(wall_faces)[idx]).to(geometry.face.area.device)
    return d


def _nearest_tables(nearest, n_cells, device):
    """For each wall face that is some cell's nearest: the face (``_near_faces``) and its cells,
    padded with ``n_cells`` (``_near_table``), so that the distance cotangents reach the face
    centroids as fixed-order sums rather than atomic scatters."""
    near = nearest.cpu().numpy()
    order = np.argsort(near, kind="stable")
    faces, start, count = np.unique(near[order], return_index=True, return_counts=True)
    table = np.full((faces.size, int(count.max()) if faces.size else 1), n_cells, dtype=np.int64)
    for i, (s0, n) in enumerate(zip(start, count)):
        table[i, :n] = order[s0:s0 + n]
    return {"_near_faces": torch.from_numpy(faces).to(device), "_near_table": torch.from_numpy(table).to(device)}


def scalar_gradient(mesh, geometry, gradient_scheme, phi, kind, value):
    """Cell gradient ``(d, C)`` of ``phi`` with Dirichlet / zero-gradient closures (HIP kernel)."""
    dev = geometry.face.area.device
    dm = mesh.device_mesh(dev)
    G = torch.empty((mesh.dim, mesh.n_cells), dtype=F64, device=dev)
    gs = geometry.c_struct()
    _lib.check(
        _lib.lib().afx_scalar_gradient(
            dm.handle, ctypes.byref(gs), int(gradient_scheme.code), int(getattr(gradient_scheme, "sweeps", 4)),
            float(getattr(gradient_scheme, "relaxation", 1.0)), kind.data_ptr(), value.data_ptr(),
            phi.contiguous().data_ptr(), G.data_ptr(), _lib.stream_of(dev),
        ),
        "afx_scalar_gradient",
    )
    return G


class SSTTurbulence:
    """The k-omega SST closure and equation assembler (``turbulence/transport.py:180-718``)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    @classmethod
    def build(
        cls,
        model: SSTModel,
        mesh,
        geometry,
        gradient_scheme,
        advection_scheme: AdvectionScheme,
        density: float,
        molecular_viscosity,
        wall_patches,
        k_boundary: BoundaryConditions,
        omega_boundary: BoundaryConditions,
        *,
        explicit_production_limiter: bool = False,
    ) -> "SSTTurbulence":
        if explicit_production_limiter:
            raise ValueError("explicit_production_limiter (a segregated-solve linearization) is not supported")
        if not wall_patches:
            raise ValueError("SSTTurbulence.build: no wall patches given")
        dev = geometry.face.area.device
        wall_faces = np.concatenate([np.asarray(mesh.face_patches.indices(p), dtype=np.int64) for p in wall_patches])
        if np.any(mesh.neighbour[wall_faces] >= 0):
            raise ValueError("wall patches must be boundary faces")
        wall_cells = np.unique(mesh.owner[wall_faces])
        d, nearest = wall_distance(mesh, geometry, wall_faces, with_nearest=True)
        # wall-distance gradient: Dirichlet 0 on the walls, zero gradient elsewhere (transport.py:86-110)
        dkind = np.where(mesh.neighbour < 0, BC_ZERO_GRADIENT, 0).astype(np.int8)
        dkind[wall_faces] = BC_DIRICHLET
        dkind_t = torch.from_numpy(dkind).to(dev)
        dval = torch.zeros(mesh.n_faces, dtype=F64, device=dev)
        dgrad = scalar_gradient(mesh, geometry, gradient_scheme, d, dkind_t, dval)
        kb = k_boundary.resolve(mesh.face_patches)
        ob = omega_boundary.resolve(mesh.face_patches)
        kk, kv = _scalar_face_data(kb, mesh, "k")
        ok, ov = _scalar_face_data(ob, mesh, "omega")
        wall_face = np.zeros(mesh.n_faces, dtype=np.int8)
        wall_face[wall_faces] = 1
        wall_cell = np.zeros(mesh.n_cells, dtype=np.int8)
        wall_cell[wall_cells] = 1
        nu = molecular_viscosity
        if not isinstance(nu, torch.Tensor):
            nu = torch.as_tensor(nu, dtype=F64)
        nu = nu.to(dev, F64)
        if nu.dim() == 0:
            nu = nu.expand(mesh.n_cells).contiguous()
        if advection_scheme is not None and isinstance(advection_scheme, LimitedUpwind):
            advection_scheme.limiter_k()
        return cls(
            model=model,
            mesh=mesh,
            geometry=geometry,
            gradient_scheme=gradient_scheme,
            advection_scheme=advection_scheme,
            density=float(density),
            molecular_viscosity=nu,
            wall_distance=d,
            wall_distance_gradient=dgrad.contiguous(),
            wall_cells=torch.from_numpy(wall_cells).to(dev),
            wall_faces=torch.from_numpy(wall_faces).to(dev),
            k_boundary=kb,
            omega_boundary=ob,
            _k_kind=torch.from_numpy(kk).to(dev),
            _k_value=torch.from_numpy(kv).to(dev),
            _omega_kind=torch.from_numpy(ok).to(dev),
            _omega_value=torch.from_numpy(ov).to(dev),
            _wall_face=torch.from_numpy(wall_face).to(dev),
            _wall_nearest=nearest,
            **_nearest_tables(nearest, mesh.n_cells, dev),
            _dkind=dkind_t,
            _dval=dval,
            _wall_cell=torch.from_numpy(wall_cell).to(dev),
        )

    def with_scaled_molecular_viscosity(self, factor) -> "SSTTurbulence":
        new = SSTTurbulence(**self.__dict__)
        new.molecular_viscosity = self.molecular_viscosity * factor
        return new


class _RansRaw:
    """Binds the C ABI calls of one configured coupled RANS problem."""

    def __init__(self, momentum, turbulence: SSTTurbulence, k_transform=None, omega_transform=None):
        self.mesh = momentum.mesh
        self.dm = momentum._raw.dm
        self.gs = momentum._raw.gs
        t = turbulence
        cfg = _lib.AfxRansConfig()
        fc = momentum._raw.cfg
        for name, _ in _lib.AfxFlowConfig._fields_:
            setattr(cfg.flow, name, getattr(fc, name))
        adv = t.advection_scheme
        cfg.scalar_advection = 0 if adv is None else int(adv.code)
        if cfg.scalar_advection == 2 and cfg.flow.advection == 2:
            if adv.limiter_k() != cfg.flow.limiter_k:
                raise ValueError("the k/omega and momentum limiters must share one Venkatakrishnan K")
        elif cfg.scalar_advection == 2:
            cfg.flow.limiter_k = adv.limiter_k()
        if cfg.scalar_advection not in (1, 2) or cfg.flow.advection not in (1, 2):
            raise ValueError("the coupled RANS residual needs upwind or limited advection in both blocks")
        cfg.density = t.density
        cfg.sst = (ctypes.c_double * 16)(*t.model.c_constants())
        cfg.k_kind = t._k_kind.data_ptr()
        cfg.omega_kind = t._omega_kind.data_ptr()
        cfg.wall_face = t._wall_face.data_ptr()
        cfg.wall_cell = t._wall_cell.data_ptr()
        cfg.wall_distance = t.wall_distance.data_ptr()
        cfg.wall_distance_gradient = t.wall_distance_gradient.data_ptr()
        cfg.k_transform = (k_transform or DirectScalars()).code
        cfg.omega_transform = (omega_transform or DirectScalars()).code
        self.cfg = cfg
        self.keep = (t, momentum._raw)
        self.k_value = t._k_value
        self.omega_value = t._omega_value
        L = _lib.lib()
        self.work_bytes = []
        for mode in range(4):
            n = ctypes.c_int64()
            _lib.check(L.afx_rans_workspace_bytes(self.dm.handle, mode, ctypes.byref(n)), "afx_rans_workspace_bytes")
            self.work_bytes.append(int(n.value))

    def inputs(self, state, mu, rho, nu, ub, pb, kv=None, wv=None):
        return _lib.AfxRansInputs(
            _lib.ptr(state), _lib.ptr(mu), _lib.ptr(rho), _lib.ptr(nu), _lib.ptr(ub), _lib.ptr(pb),
            _lib.ptr(self.k_value if kv is None else kv), _lib.ptr(self.omega_value if wv is None else wv),
        )

    def residual(self, state, mu, rho, nu, ub, pb):
        R = torch.empty_like(state)
        w = self.dm.workspace("rans0", self.work_bytes[0])
        ins = self.inputs(state, mu, rho, nu, ub, pb)
        _lib.check(
            _lib.lib().afx_rans_residual(
                self.dm.handle, ctypes.byref(self.gs), ctypes.byref(self.cfg), ctypes.byref(ins),
                R.data_ptr(), w.data_ptr(), w.numel(), _lib.stream_of(self.dm.device),
            ),
            "afx_rans_residual",
        )
        return R

    def jvp(self, primals, tangents):
        state = primals[0]
        R = torch.empty_like(state)
        dR = torch.empty_like(state)
        w = self.dm.workspace("rans1", self.work_bytes[1])
        ins = self.inputs(*primals)
        tns = _lib.AfxRansInputs(*[_lib.ptr(t) for t in tangents], None, None)
        _lib.check(
            _lib.lib().afx_rans_jvp(
                self.dm.handle, ctypes.byref(self.gs), ctypes.byref(self.cfg), ctypes.byref(ins), ctypes.byref(tns),
                R.data_ptr(), dR.data_ptr(), w.data_ptr(), w.numel(), _lib.stream_of(self.dm.device),
            ),
            "afx_rans_jvp",
        )
        return R, dR

    def wall_distance_tangent(self, dgeo, gt):
        """Tangents of the wall distance and of its gradient along the geometry tangents `dgeo`
        (C struct `gt`): d = |x_c - x_f*| to the nearest wall-face centroid f*, and the gradient
        scheme's tangent with the geometry and d moving together."""
        t = self.keep[0]
        g = t.geometry
        near = t._wall_nearest
        d = t.wall_distance
        cc = fwAD.unpack_dual(g.cell.centroid_soa).primal
        fc = fwAD.unpack_dual(g.face.centroid_soa).primal
        diff = cc - fc[:, near]
        ddiff = dgeo[4] - dgeo[2][:, near]
        dd = ((diff * ddiff).sum(dim=0) / d).contiguous()
        dgrad = torch.empty_like(t.wall_distance_gradient)
        dgrad_t = torch.empty_like(dgrad)
        _lib.check(
            _lib.lib().afx_scalar_gradient_jvp_geometry(
                self.dm.handle, ctypes.byref(self.gs), ctypes.byref(gt), int(self.cfg.flow.gradient),
                int(self.cfg.flow.gradient_sweeps), float(self.cfg.flow.gradient_relaxation), t._dkind.data_ptr(),
                t._dval.data_ptr(), d.data_ptr(), dd.data_ptr(), dgrad.data_ptr(), dgrad_t.data_ptr(),
                _lib.stream_of(self.dm.device),
            ),
            "afx_scalar_gradient_jvp_geometry",
        )
        return dd, dgrad_t

    def jvp_geometry(self, primals, tangents, dgeo):
        """(R, dR) along tangents of the inputs and of the five geometry arrays together; the wall
        distance and its gradient move with the geometry."""
        from .flow import _geometry_tangent

        state = primals[0]
        R = torch.empty_like(state)
        dR = torch.empty_like(state)
        w = self.dm.workspace("rans1", self.work_bytes[1])
        ins = self.inputs(*primals)
        tns = _lib.AfxRansInputs(*[_lib.ptr(t) for t in tangents], None, None)
        gt = _geometry_tangent(self.dm, dgeo)
        dd, dgrad_t = self.wall_distance_tangent(gt[1][0], gt[0])
        _lib.check(
            _lib.lib().afx_rans_jvp_geometry(
                self.dm.handle, ctypes.byref(self.gs), ctypes.byref(gt[0]), dd.data_ptr(), dgrad_t.data_ptr(),
                ctypes.byref(self.cfg), ctypes.byref(ins), ctypes.byref(tns), R.data_ptr(), dR.data_ptr(),
                w.data_ptr(), w.numel(), _lib.stream_of(self.dm.device),
            ),
            "afx_rans_jvp_geometry",
        )
        return R, dR

    def vjp_geometry(self, primals, cot):
        """The cotangents of afx_rans_vjp plus those of the five geometry arrays, with the wall
        distance and its gradient moving with the geometry (a differentiated build)."""
        t = self.keep[0]
        L = _lib.lib()
        stream = _lib.stream_of(self.dm.device)
        bars = [torch.empty_like(p) for p in primals]
        g = t.geometry
        geo = [fwAD.unpack_dual(x).primal for x in (g.face.area, g.face.normal_soa, g.face.centroid_soa, g.cell.volume,
                                                    g.cell.centroid_soa)]
        gbars = [torch.empty_like(x) for x in geo]
        d = t.wall_distance
        db = torch.empty_like(d)
        dgb = torch.empty_like(t.wall_distance_gradient)
        w = self.dm.workspace("rans3", self.work_bytes[3])
        ins = self.inputs(*primals)
        outs = _lib.AfxRansCotangents(*[b.data_ptr() for b in bars])
        gouts = _lib.AfxGeometryCotangents(*[b.data_ptr() for b in gbars])
        _lib.check(
            L.afx_rans_vjp_geometry(
                self.dm.handle, ctypes.byref(self.gs), ctypes.byref(self.cfg), ctypes.byref(ins), cot.data_ptr(),
                ctypes.byref(outs), ctypes.byref(gouts), db.data_ptr(), dgb.data_ptr(), w.data_ptr(), w.numel(), stream,
            ),
            "afx_rans_vjp_geometry",
        )
        # the wall-distance gradient: d's share and the geometry's
        gcode = int(self.cfg.flow.gradient)
        sweeps, relax = int(self.cfg.flow.gradient_sweeps), float(self.cfg.flow.gradient_relaxation)
        d_extra = torch.empty_like(d)
        _lib.check(L.afx_scalar_gradient_vjp(self.dm.handle, ctypes.byref(self.gs), gcode, sweeps, relax,
                                             t._dkind.data_ptr(), dgb.data_ptr(), d_extra.data_ptr(), None, stream),
                   "afx_scalar_gradient_vjp")
        db = db + d_extra
        g2 = [torch.empty_like(x) for x in geo]
        gouts2 = _lib.AfxGeometryCotangents(*[b.data_ptr() for b in g2])
        _lib.check(L.afx_scalar_gradient_vjp_geometry(self.dm.handle, ctypes.byref(self.gs), gcode, sweeps, relax,
                                                      t._dkind.data_ptr(), t._dval.data_ptr(), d.data_ptr(),
                                                      dgb.data_ptr(), ctypes.byref(gouts2), stream),
                   "afx_scalar_gradient_vjp_geometry")
        gbars = [a + b for a, b in zip(gbars, g2)]
        # d = |x_c - x_f*| to the nearest wall-face centroid f*
        near, faces, table = t._wall_nearest, t._near_faces, t._near_table
        unit = (geo[4] - geo[2][:, near]) * (db / d)
        pad = torch.cat([unit, torch.zeros((unit.shape[0], 1), dtype=unit.dtype, device=unit.device)], dim=1)
        gbars[4] = gbars[4] + unit
        fcb = gbars[2].clone()
        fcb[:, faces] = fcb[:, faces] - pad[:, table].sum(dim=2)  # fixed-order sums over each face's cells
        gbars[2] = fcb
        return bars, gbars

    def vjp(self, primals, cot):
        bars = [torch.empty_like(p) for p in primals]
        w = self.dm.workspace("rans2", self.work_bytes[2])
        ins = self.inputs(*primals)
        outs = _lib.AfxRansCotangents(*[b.data_ptr() for b in bars])
        _lib.check(
            _lib.lib().afx_rans_vjp(
                self.dm.handle, ctypes.byref(self.gs), ctypes.byref(self.cfg), ctypes.byref(ins), cot.data_ptr(),
                ctypes.byref(outs), w.data_ptr(), w.numel(), _lib.stream_of(self.dm.device),
            ),
            "afx_rans_vjp",
        )
        return bars

    def residual_vjp(self, primals, cot):
        """(R, cotangents) with the forward passes shared between the two."""
        R = torch.empty_like(primals[0])
        bars = [torch.empty_like(p) for p in primals]
        w = self.dm.workspace("rans2", self.work_bytes[2])
        ins = self.inputs(*primals)
        outs = _lib.AfxRansCotangents(*[b.data_ptr() for b in bars])
        _lib.check(
            _lib.lib().afx_rans_residual_vjp(
                self.dm.handle, ctypes.byref(self.gs), ctypes.byref(self.cfg), ctypes.byref(ins), cot.data_ptr(),
                R.data_ptr(), ctypes.byref(outs), w.data_ptr(), w.numel(), _lib.stream_of(self.dm.device),
            ),
            "afx_rans_residual_vjp",
        )
        return R, bars

    def host_line(self, primals, cot):
        """The OpenMP CPU line (afx_rans_residual_vjp_host) on host copies of every array, made once:
        returns run(threads) -> (R, cotangents) as host tensors, which runs the same cell bodies as
        the device kernels in the caller's face order (threads <= 0: OpenMP's default)."""
        t, fraw = self.keep
        mesh = self.mesh
        held = []

        def hp(x):
            h = torch.as_tensor(x).detach().to("cpu").contiguous()
            held.append(h)
            return h.data_ptr()

        g = fraw.geometry
        geo = _lib.AfxGeometry(hp(g.face.area), hp(g.face.normal_soa), hp(g.face.centroid_soa), hp(g.cell.volume),
                               hp(g.cell.centroid_soa), None)
        cfg = _lib.AfxRansConfig.from_buffer_copy(self.cfg)
        cfg.flow.face_kind = hp(fraw.kind)
        cfg.k_kind = hp(t._k_kind)
        cfg.omega_kind = hp(t._omega_kind)
        cfg.wall_face = hp(t._wall_face)
        cfg.wall_cell = hp(t._wall_cell)
        cfg.wall_distance = hp(t.wall_distance)
        cfg.wall_distance_gradient = hp(t.wall_distance_gradient)
        ins = _lib.AfxRansInputs(*[hp(x) for x in primals], hp(self.k_value), hp(self.omega_value))
        lam = hp(cot)
        R = torch.empty(primals[0].shape, dtype=F64)
        bars = [torch.empty(p.shape, dtype=F64) for p in primals]
        outs = _lib.AfxRansCotangents(*[b.data_ptr() for b in bars])
        owner = np.ascontiguousarray(mesh.owner, dtype=np.int32)
        nb = np.ascontiguousarray(mesh.neighbour, dtype=np.int32)
        L = _lib.lib()

        def run(threads=0):
            _lib.check(
                L.afx_rans_residual_vjp_host(mesh.dim, mesh.n_cells, mesh.n_faces, owner.ctypes.data, nb.ctypes.data,
                                             ctypes.byref(geo), ctypes.byref(cfg), ctypes.byref(ins), lam,
                                             R.data_ptr(), ctypes.byref(outs), int(threads)),
                "afx_rans_residual_vjp_host",
            )
            _ = held  # the host copies live as long as run does
            return R, bars

        return run

    def residual_vjp_host(self, primals, cot, threads=0):
        """(R, cotangents) of residual_vjp from the OpenMP CPU line, as host tensors."""
        R, bars = self.host_line(primals, cot)(threads)
        return R.clone(), [b.clone() for b in bars]

    def fields(self, state, mu, rho, nu, ub, pb):
        d, C = self.mesh.dim, self.mesh.n_cells
        n = ((d + 3) * d + (d + 2) + 1 + 7) * C
        buf = torch.empty(n, dtype=F64, device=state.device)
        ins = self.inputs(state, mu, rho, nu, ub, pb)
        _lib.check(
            _lib.lib().afx_rans_fields(
                self.dm.handle, ctypes.byref(self.gs), ctypes.byref(self.cfg), ctypes.byref(ins), buf.data_ptr(),
                _lib.stream_of(self.dm.device),
            ),
            "afx_rans_fields",
        )
        o = 0
        out = {}
        for name, rows in (("gradients", (d + 3) * d), ("limiters", d + 2), ("a_p", 1), ("closure", 7)):
            out[name] = buf[o : o + rows * C].view(rows, C)
            o += rows * C
        names = ("nu_t", "f1", "strain_rate", "cross_diffusion", "wall_nu_t", "wall_shear_rate", "mu_eff")
        for i, nm in enumerate(names):
            out[nm] = out["closure"][i]
        return out


class _RansResidualFn(torch.autograd.Function):
    """R(state, mu, rho, nu, u_b, p_b; area, normal, face centroid, volume, cell centroid) with HIP
    forward, tangent and adjoint, in both modes also through the geometry arrays (compact
    Green-Gauss or least squares), with the wall distance and its gradient moving with them."""

    @staticmethod
    def forward(ctx, raw, geo_tan, state, mu, rho, nu, ub, pb, *geo):
        ctx.raw = raw
        ctx.geo_tan = geo_tan
        prim = (state, mu, rho, nu, ub, pb)
        ctx.save_for_backward(*prim)
        ctx.save_for_forward(*prim, *geo)
        return raw.residual(*prim)

    @staticmethod
    def backward(ctx, gR):
        if any(ctx.needs_input_grad[8:]):
            bars, gbars = ctx.raw.vjp_geometry(ctx.saved_tensors, gR.contiguous())
            return (None, None, *bars, *gbars)
        bars = ctx.raw.vjp(ctx.saved_tensors, gR.contiguous())
        return (None, None, *bars) + (None,) * (len(ctx.needs_input_grad) - 8)

    @staticmethod
    def jvp(ctx, _raw_t, _geo_tan_t, *tans):
        saved = ctx.saved_tensors
        prim, geo = saved[:6], saved[6:]
        dgeo = tans[6:]
        tans = [t.contiguous() if t is not None else None for t in tans[:6]]
        if ctx.geo_tan:
            dgeo = [torch.zeros_like(g) if t is None else t for g, t in zip(geo, dgeo)]
            _, dR = ctx.raw.jvp_geometry(prim, tans, dgeo)
        else:  # constant geometry: torch fills its tangents with zeros, which are not read
            _, dR = ctx.raw.jvp(prim, tans)
        return dR


class CoupledRANS:
    """Monolithic ``(u, v[, w], p, k, omega)`` residual (``turbulence/coupled.py:253-473``)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    @classmethod
    def build(cls, momentum, turbulence: SSTTurbulence, k_transform=None, omega_transform=None) -> "CoupledRANS":
        """``CoupledRANS.build`` (``turbulence/coupled.py:306-355``); ``omega_transform=LogScalars()``
        solves for ``log(omega)``."""
        for name, tr in (("k_transform", k_transform), ("omega_transform", omega_transform)):
            if tr is not None and not isinstance(tr, (DirectScalars, LogScalars)):
                raise ValueError(f"{name}: {type(tr).__name__} is not a supported ScalarVariableTransform")
        rho = momentum.density
        if not bool(torch.all(torch.isclose(rho, torch.tensor(turbulence.density, dtype=F64, device=rho.device)))):
            raise ValueError(
                f"SSTTurbulence.density ({turbulence.density}) does not match the flow assembler's density "
                f"(range [{float(rho.min())}, {float(rho.max())}]); pass the same density to both"
            )
        if momentum.eddy_viscosity is not None:
            raise ValueError("CoupledRANS.build expects the momentum assembler without an eddy viscosity")
        k_transform = k_transform or DirectScalars()
        omega_transform = omega_transform or DirectScalars()
        return cls(momentum=momentum, turbulence=turbulence, k_transform=k_transform,
                   omega_transform=omega_transform,
                   _raw=_RansRaw(momentum, turbulence, k_transform, omega_transform))

    # --- layout (turbulence/coupled.py:56-120) ---------------------------------------------------
    @property
    def mesh(self):
        return self.momentum.mesh

    @property
    def dim(self):
        return self.momentum.mesh.dim

    @property
    def n_cells(self):
        return self.momentum.mesh.n_cells

    @property
    def device(self):
        return self.momentum.device

    def pack_state(self, flow, k, omega):
        return torch.cat([flow.reshape(-1), k.reshape(-1), omega.reshape(-1)])

    def unpack(self, state):
        C, d = self.n_cells, self.dim
        return state[: (d + 1) * C], state[(d + 1) * C : (d + 2) * C], state[(d + 2) * C :]

    def physical_fields(self, state):
        """Flow sub-vector and physical ``k``, ``omega`` (``turbulence/coupled.py:379-392``)."""
        flow, k, omega = self.unpack(state)
        return flow, self.k_transform.to_physical(k), self.omega_transform.to_physical(omega)

    def state_from_physical(self, flow, k, omega):
        """Inverse of :meth:`physical_fields` (``turbulence/coupled.py:394-404``)."""
        return self.pack_state(flow, self.k_transform.to_solved(k), self.omega_transform.to_solved(omega))

    def with_scaled_molecular_viscosity(self, factor) -> "CoupledRANS":
        mom = MomentumScaled.scale(self.momentum, factor)
        turb = self.turbulence.with_scaled_molecular_viscosity(factor)
        return CoupledRANS.build(mom, turb, self.k_transform, self.omega_transform)

    # --- inputs ----------------------------------------------------------------------------------
    def _inputs(self):
        mu, rho = self.momentum.material()
        ub, pb = self.momentum.boundary_data()
        nu = self.turbulence.molecular_viscosity
        return mu.contiguous(), rho.contiguous(), nu.contiguous(), ub, pb

    # --- residual --------------------------------------------------------------------------------
    def residual(self, state):
        """``R(u, p, k, omega)``, same shape as ``state`` (``turbulence/coupled.py:437-473``)."""
        _lib.require_cuda(state, "state")
        mu, rho, nu, ub, pb = self._inputs()
        g = self.momentum.geometry
        geo = (g.face.area, g.face.normal_soa, g.face.centroid_soa, g.cell.volume, g.cell.centroid_soa)
        geo_tan = any(fwAD.unpack_dual(t).tangent is not None for t in geo)
        return _RansResidualFn.apply(self._raw, geo_tan, state.contiguous(), mu, rho, nu, ub, pb, *geo)

    __call__ = residual

    def _detached(self, state):
        return [state.contiguous()] + [t.detach() for t in self._inputs()]

    def residual_raw(self, state):
        return self._raw.residual(*self._detached(state))

    def residual_jvp(self, state, tangent):
        prim = self._detached(state)
        return self._raw.jvp(prim, [tangent.contiguous(), None, None, None, None, None])

    def residual_vjp(self, state, cotangent):
        return self._raw.vjp(self._detached(state), cotangent.contiguous())[0]

    def closure_fields(self, state):
        """Gradients, limiters, ``a_P`` and the SST closure fields at ``state``."""
        return self._raw.fields(*self._detached(state))

    def eddy_viscosity(self, state):
        return self.closure_fields(state)["nu_t"]

    def pseudo_time_diagonal(self, state):
        """The pseudo-transient shift basis at ``state`` (``turbulence/coupled.py:607-640``): the
        momentum diagonal a_P on every velocity row, zero on continuity, and on the k and omega rows
        the same transport diagonal per unit density, a_P / rho, times the coordinate factor
        d(phi)/d(w) of their transform (phi for LogScalars, 1 for DirectScalars)."""
        f = self.closure_fields(state)
        d, C = self.dim, self.n_cells
        aP = f["a_p"][0]
        _, rho = self.momentum.material()
        t = aP / rho.detach()
        blocks = state.reshape(d + 3, C)
        diag = torch.zeros_like(blocks)
        diag[:d] = aP[None, :]
        phys = self.physical_fields(state)
        for row, transform, val in ((d + 1, self.k_transform, phys[1]), (d + 2, self.omega_transform, phys[2])):
            diag[row] = t * (val if isinstance(transform, LogScalars) else 1.0)
        return diag.reshape(-1)


class MomentumScaled:
    """Copy of a momentum assembler with its molecular viscosity scaled (``flow/momentum.py``)."""

    @staticmethod
    def scale(momentum, factor):
        from .flow import MomentumContinuity

        new = MomentumContinuity(**momentum.__dict__)
        new.properties = momentum.properties.with_scaled("viscosity", factor)
        return new
