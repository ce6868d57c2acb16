"""Oracle: k-omega SST closure, sources, wall treatment and the coupled RANS residual (torch FP64).

Test infrastructure only (see ``oracle/__init__.py``).  Restates, function by function:
  ``aquaflux/turbulence/sst.py:37-239``        SSTModel (blend, F1, F2, eddy viscosity, y*_lam)
  ``aquaflux/turbulence/strain.py:32-66``      safe_sqrt, strain-rate magnitude
  ``aquaflux/turbulence/boundary.py:45-589``   wall closures (omega_wall, nut_wall, weights, ...)
  ``aquaflux/turbulence/sources.py:60-412``    k / omega sources and the near-wall k closure
  ``aquaflux/turbulence/transport.py:113-718`` SSTTurbulence (closure fields, k / omega residuals)
  ``aquaflux/turbulence/coupled.py:306-473``   CoupledRANS.residual (direct scalars)
Pinned by the closed forms of the reference's ``tests/unit/test_sst.py``, ``test_sst_sources.py``,
``test_turbulence_boundary.py``, ``test_strain.py`` (CPU tests) and by central finite differences.
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np
import torch

from . import fv

F64 = torch.float64
CROSS_DIFFUSION_FLOOR = 1e-10
PRODUCTION_LIMIT_RATIO = 10.0
EDDY_VISCOSITY_FLOOR = 1e-30


@dataclasses.dataclass(frozen=True)
class SSTModel:
    """``aquaflux/turbulence/sst.py:37-82`` constants."""

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
        """``sst.py:84-105``: 10 fixed-point iterations of ``y <- ln(E y)/kappa`` from 11."""
        y = 11.0
        for _ in range(10):
            y = math.log(self.e_wall * y) / self.kappa
        return y

    def blend(self, f1, inner, outer):
        return f1 * inner + (1.0 - f1) * outer

    def f2(self, k, omega, nu, d):
        """``sst.py:125-150``."""
        arg2 = fv.maximum(
            2.0 * safe_sqrt(fv.maximum(k, 0.0)) / (self.beta_star * omega * d),
            500.0 * nu / (d**2 * omega),
        )
        return torch.tanh(arg2**2)

    def f1(self, k, omega, nu, d, grad_k, grad_omega):
        """``sst.py:152-195``."""
        cd = fv.maximum(
            2.0 * self.sigma_omega2 * fv.dot(grad_k, grad_omega) / omega,
            torch.full_like(omega, CROSS_DIFFUSION_FLOOR),
        )
        pk = fv.maximum(k, 0.0)
        arg1 = fv.minimum(
            fv.maximum(safe_sqrt(pk) / (self.beta_star * omega * d), 500.0 * nu / (d**2 * omega)),
            4.0 * self.sigma_omega2 * pk / (cd * d**2),
        )
        return torch.tanh(arg1**4)

    def eddy_viscosity(self, k, omega, strain, nu, d):
        """``sst.py:197-239``: ``a1 max(k,0) / max(a1 omega, S F2)``."""
        return self.a1 * fv.maximum(k, 0.0) / fv.maximum(
            self.a1 * omega, strain * self.f2(k, omega, nu, d)
        )


def safe_sqrt(x):
    """``strain.py:32-43``: zero value and derivative at 0."""
    pos = x > 0.0
    return torch.where(pos, torch.sqrt(torch.where(pos, x, torch.ones_like(x))), torch.zeros_like(x))


def strain_rate_magnitude(grad):
    """``strain.py:46-66``: ``sqrt(2 S_ij S_ij)``."""
    s = 0.5 * (grad + grad.transpose(-1, -2))
    return safe_sqrt(2.0 * torch.sum(s * s, dim=(-2, -1)))


# --- wall closures (turbulence/boundary.py) ---------------------------------------------------------


def omega_wall_value(nu, d, model):
    return 6.0 * nu / (model.beta_1 * d**2)


def _power_mean(a, b, p):
    m = fv.maximum(a, b)
    ra, rb = a / m, b / m
    pos = rb > 0.0
    rbs = torch.where(pos, rb, torch.ones_like(rb))
    bt = torch.where(pos, rbs**p, torch.zeros_like(rb))
    return m * (ra**p + bt) ** (1.0 / p)


def omega_wall(nu, d, k, model):
    vis = model.wall_omega_viscous_coeff * omega_wall_value(nu, d, model)
    log = safe_sqrt(fv.maximum(k, 0.0)) / (model.beta_star**0.25 * model.kappa * d)
    return _power_mean(vis, log, model.wall_omega_exponent)


def wall_y_star(nu, d, k, model):
    return model.beta_star**0.25 * safe_sqrt(fv.maximum(k, 0.0)) * d / nu


def nut_wall(nu, d, k, model):
    ys = wall_y_star(nu, d, k, model)
    yl = model.wall_y_star_lam
    arg = model.e_wall * fv.maximum(ys, yl)
    r = ys * model.kappa / torch.log(arg) - 1.0
    return nu * torch.where(ys > yl, r, torch.zeros_like(r))


def wall_function_weight(nu, d, k, model):
    return torch.tanh((wall_y_star(nu, d, k, model) / model.wall_y_star_lam) ** 4)


def log_layer_shear_rate(d, k, model):
    return model.beta_star**0.25 * safe_sqrt(fv.maximum(k, 0.0)) / (model.kappa * d)


def k_wall_production(nu, d, k, shear_rate, model):
    return (nu + nut_wall(nu, d, k, model)) * shear_rate * log_layer_shear_rate(d, k, model)


def omega_wall_gradient(nu, d, k, dgrad, kgrad, model):
    """``boundary.py:560-589``: ``d omega_wall/dd grad d + d omega_wall/dk grad k``."""
    dod = torch.func.grad(lambda dd: omega_wall(nu, dd, k, model).sum())(d)
    dok = torch.func.grad(lambda kk: omega_wall(nu, d, kk, model).sum())(k)
    return fv.scale(dgrad, dod) + fv.scale(kgrad, dok)


# --- volume sources (turbulence/sources.py) -----------------------------------------------------------
# Each returns the source per unit volume; the balance multiplies by the cell volume.


def production_and_limit(nu_t, strain, omega, k, model):
    """``sources.py:165-202``: ``(nu_t S^2, 10 beta* max(k, 0) omega)``."""
    production = nu_t * strain**2
    limit = PRODUCTION_LIMIT_RATIO * model.beta_star * fv.maximum(k, 0.0) * omega
    return production, limit


def near_wall_production(resolved, k, cells, nu, d, shear_rate, model):
    """``NearWallKClosure.production`` (``sources.py:109-139``): the wall cells' resolved production
    blended toward ``k_wall_production`` by ``wall_function_weight`` at the live ``k``."""
    kw = k[cells]
    w = wall_function_weight(nu, d, kw, model)
    wall_value = k_wall_production(nu, d, kw, shear_rate, model)
    out = resolved.clone()
    out[cells] = (1.0 - w) * resolved[cells] + w * wall_value
    return out


def near_wall_dissipation_rate(omega, k, cells, nu, d, model):
    """``NearWallKClosure.dissipation_rate`` (``sources.py:141-162``): ``omega_wall(k)`` on the wall
    cells, ``omega`` elsewhere."""
    out = omega.clone()
    out[cells] = omega_wall(nu, d, k[cells], model)
    return out


def k_production(nu_t, strain, omega, k, model, near_wall=None):
    """``KProduction.source / V`` (``sources.py:205-262``): ``min(nu_t S^2, 10 beta* max(k,0) omega)``,
    then the near-wall blend; ``near_wall`` = (cells, nu, d, shear_rate) or None."""
    production, limit = production_and_limit(nu_t, strain, omega, k, model)
    limited = fv.minimum(production, limit)
    if near_wall is not None:
        limited = near_wall_production(limited, k, *near_wall, model)
    return limited


def k_destruction(k, omega, model, near_wall=None):
    """``KDestruction.source / V`` (``sources.py:265-293``): ``-beta* k omega`` with the wall cells'
    ``omega`` read live as ``omega_wall(k)``; ``near_wall`` = (cells, nu, d) or None."""
    if near_wall is not None:
        omega = near_wall_dissipation_rate(omega, k, *near_wall, model)
    return -model.beta_star * k * omega


def omega_production(strain, nu_t, k, omega, f1, model):
    """``OmegaProduction.source / V`` (``sources.py:296-355``): ``alpha(F1) min(S^2, 10 beta* max(k,0)
    omega / max(nu_t, 1e-30))``."""
    alpha = model.blend(f1, model.alpha_1, model.alpha_2)
    cap = (PRODUCTION_LIMIT_RATIO * model.beta_star * fv.maximum(k, 0.0) * omega
           / fv.maximum(nu_t, EDDY_VISCOSITY_FLOOR))
    return alpha * fv.minimum(strain**2, cap)


def omega_destruction(omega, f1, model):
    """``OmegaDestruction.source / V`` (``sources.py:358-376``): ``-beta(F1) omega^2``."""
    beta = model.blend(f1, model.beta_1, model.beta_2)
    return -beta * omega**2


def omega_cross_diffusion(omega, grad_k, grad_omega, f1, model):
    """``OmegaCrossDiffusion.source / V`` (``sources.py:379-412``): ``2 (1 - F1) sigma_w2 (grad k . grad omega) / omega``."""
    return 2.0 * (1.0 - f1) * model.sigma_omega2 * fv.dot(grad_k, grad_omega) / omega


# --- SST turbulence (turbulence/transport.py) --------------------------------------------------------


@dataclasses.dataclass
class SSTSetup:
    model: SSTModel
    nu: torch.Tensor  # (C,) kinematic molecular viscosity
    density: float
    wall_faces: torch.Tensor  # (nw,) face indices of the wall patches
    wall_cells: torch.Tensor  # (nwc,) unique owners of wall faces
    wall_distance: torch.Tensor  # (C,)
    wall_distance_gradient: torch.Tensor  # (C, d)
    k_bc: fv.ScalarSetup
    omega_bc: fv.ScalarSetup


def wall_distance(m, g, wall_faces):
    """``mesh/distance.py:52-113``: nearest wall-face centroid (cKDTree)."""
    from scipy.spatial import cKDTree

    tree = cKDTree(g.face_centroid[wall_faces].detach().numpy())
    dist, _ = tree.query(g.cell_centroid.detach().numpy())
    return torch.sqrt(torch.tensor(dist, dtype=F64) ** 2)


def wall_distance_traced(m, g, wall_faces):
    """``mesh/distance.py:87-140`` (the traced path): sqrt of the smallest squared distance from each
    cell centroid to a wall-face centroid, as a differentiable search (brute force; small meshes)."""
    diff = g.cell_centroid[:, None, :] - g.face_centroid[wall_faces][None, :, :]
    d2 = (diff * diff).sum(dim=-1)
    idx = torch.argmin(d2.detach(), dim=1, keepdim=True)
    return torch.sqrt(torch.gather(d2, 1, idx).squeeze(1))


def gradient_spec(s: fv.FlowSetup):
    """The flow setup's gradient scheme as scalar_gradient takes it."""
    if s.gradient == "corrected":
        return ("corrected", s.gradient_sweeps, s.gradient_relaxation)
    return s.gradient


def scalar_gradient(m, g, gradient, phi, bc: fv.ScalarSetup):
    """``ResidualAssembler.gradient`` for Dirichlet / ZeroGradient closures (leading-order values)."""
    zero = torch.zeros((m.n_cells, m.dim), dtype=F64)
    gamma = torch.zeros(m.n_cells, dtype=F64)
    lead = fv.scalar_boundary_values(m, g, bc, phi, zero, gamma)
    if gradient == "least_squares":
        return fv.least_squares_gradient(m, g, phi, lead)
    if isinstance(gradient, tuple) and gradient[0] == "corrected":
        return fv.corrected_green_gauss(m, g, phi, lead, gradient[1], gradient[2])
    return fv.compact_green_gauss(m, g, phi, lead)


def build_sst_setup(m, g, gradient, model, nu, density, wall_faces, k_bc, omega_bc, traced=False):
    """``SSTTurbulence.build`` (``transport.py:259-313``); ``traced`` computes the wall distance
    the way a differentiated build does (``mesh/distance.py:87-140``)."""
    wall_faces = torch.as_tensor(wall_faces, dtype=torch.long)
    wall_cells = torch.unique(m.owner[wall_faces])
    d = wall_distance_traced(m, g, wall_faces) if traced else wall_distance(m, g, wall_faces)
    kind = torch.full((m.n_faces,), fv.BC_ZERO_GRADIENT, dtype=torch.long)
    kind[wall_faces] = fv.BC_DIRICHLET
    dbc = fv.ScalarSetup(kind=kind, value=torch.zeros(m.n_faces, dtype=F64), t_inf=torch.zeros(m.n_faces, dtype=F64))
    dgrad = scalar_gradient(m, g, gradient, d, dbc)
    return SSTSetup(model, nu, density, wall_faces, wall_cells, d, dgrad, k_bc, omega_bc)


def closure_fields(m, g, gradient, t: SSTSetup, vel, ub, grad_vel, k, omega):
    """``SSTTurbulence.closure_fields`` (``transport.py:508-546``)."""
    model = t.model
    wall = t.wall_cells
    dw = t.wall_distance
    # strain_rate with the near-wall log-layer blend (transport.py:369-390)
    strain = strain_rate_magnitude(grad_vel)
    w = wall_function_weight(t.nu[wall], dw[wall], k[wall], model)
    strain = strain.clone()
    strain[wall] = (1.0 - w) * strain[wall] + w * log_layer_shear_rate(dw[wall], k[wall], model)
    grad_k = scalar_gradient(m, g, gradient, k, t.k_bc)
    grad_w = scalar_gradient(m, g, gradient, omega, t.omega_bc).clone()
    grad_w[wall] = omega_wall_gradient(t.nu[wall], dw[wall], k[wall], t.wall_distance_gradient[wall], grad_k[wall], model)
    f1 = model.f1(k, omega, t.nu, dw, grad_k, grad_w)
    nu_t = model.eddy_viscosity(k, omega, strain, t.nu, dw)
    # wall_shear_rate (transport.py:508-531)
    own = m.owner[t.wall_faces]
    slip = vel[own] - ub[t.wall_faces]
    rate = safe_sqrt(fv.dot(slip, slip)) / dw[own]
    F = m.n_faces
    zeros = torch.zeros(F, dtype=F64)
    area = zeros.index_put((t.wall_faces,), g.area[t.wall_faces])
    weighted = zeros.index_put((t.wall_faces,), rate) * area
    tau = (fv.scatter(m, weighted, zeros)[wall] / fv.scatter(m, area, zeros)[wall])
    return dict(nu_t=nu_t, strain=strain, f1=f1, grad_k=grad_k, grad_omega=grad_w, omega=omega, k=k, tau=tau)


def wall_face_eddy_viscosity(m, t: SSTSetup, k):
    """``transport.py:434-458``."""
    own = m.owner[t.wall_faces]
    vals = nut_wall(t.nu[own], t.wall_distance[own], k[own], t.model)
    return torch.zeros(m.n_faces, dtype=F64).index_put((t.wall_faces,), vals)


def _scalar_balance(m, g, gradient, bc, phi, Q, gamma, boundary_gamma, advection, limiter_k):
    """``ResidualAssembler.residual`` flux part for (Advection(Q), Diffusion) (``residual.py:358-412``)."""
    grad = scalar_gradient(m, g, gradient, phi, bc)
    bvals = fv.scalar_boundary_values(m, g, bc, phi, grad, gamma)
    face = torch.zeros(m.n_faces, dtype=F64)
    if advection == "upwind":
        face = face + Q * fv.upwind_face_value(m, phi, Q, bvals)
    elif advection == "limited":
        psi = fv.venkatakrishnan(m, g, phi, grad, limiter_k)
        face = face + Q * fv.limited_face_value(m, g, phi, grad, psi, Q, bvals)
    face = face + fv.diffusion_flux(m, g, phi, grad, gamma, bvals, boundary_gamma)
    return fv.scatter_conservative(m, face)


def rans_residual(m, g, s: fv.FlowSetup, t: SSTSetup, state, mu, rho, scalar_advection="limited", k_log=False,
                  omega_log=False):
    """``CoupledRANS.residual`` (``turbulence/coupled.py:437-473``); ``k_log`` / ``omega_log``
    select ``LogScalars`` for that unknown (``coupled.py:178-211``): the physics runs on ``e^w`` and
    the omega wall fixation becomes ``LogRatioRow`` (``discretization/fixed_value.py:96-124``)."""
    d, C = m.dim, m.n_cells
    model = t.model
    gspec = gradient_spec(s)
    blocks = state.reshape(d + 3, C)
    flow = blocks[: d + 1].reshape(-1)
    k, omega = blocks[d + 1], blocks[d + 2]
    if k_log:
        k = torch.exp(k)
    if omega_log:
        omega = torch.exp(omega)
    vel, p = blocks[:d].T, blocks[d]
    ub, _ = fv.flow_boundary_fields(m, g, s, vel, p)
    grad_vel = torch.stack([fv._gradient(m, g, s, vel[:, i], ub[:, i]) for i in range(d)], dim=1)
    cl = closure_fields(m, g, gspec, t, vel, ub, grad_vel, k, omega)
    wall_nut = wall_face_eddy_viscosity(m, t, k)
    shear = s.kind == fv.FLOW_WALL
    f = fv.flow_fields(m, g, s, flow, mu, rho, nu_t=cl["nu_t"], wall_nu_t=wall_nut, shear_mask=shear)
    psi = None
    if s.advection == "limited":
        psi = [fv.venkatakrishnan(m, g, f["vel"][:, i], f["grad_vel"][:, i], s.limiter_k) for i in range(d)]
    flow_res = fv.flow_residual_from_fields(m, g, s, f, psi)
    Q = f["mdot"] / t.density
    V = g.volume
    wall = t.wall_cells
    dw = t.wall_distance
    # k residual (transport.py:581-623)
    gk = t.nu + model.blend(cl["f1"], model.sigma_k1, model.sigma_k2) * cl["nu_t"]
    own = m.owner
    face_gamma = gk[own].clone()
    wo = own[t.wall_faces]
    face_gamma[t.wall_faces] = (1.0 - wall_function_weight(t.nu[wo], dw[wo], k[wo], model)) * face_gamma[t.wall_faces]
    Rk = _scalar_balance(m, g, gspec, t.k_bc, k, Q, gk, face_gamma, scalar_advection, s.limiter_k)
    kw = k[wall]
    near = (wall, t.nu[wall], dw[wall])
    Rk = Rk - k_production(cl["nu_t"], cl["strain"], omega, k, model, near_wall=near + (cl["tau"],)) * V
    Rk = Rk - k_destruction(k, omega, model, near_wall=near) * V
    # omega residual (transport.py:655-718)
    gw = t.nu + model.blend(cl["f1"], model.sigma_omega1, model.sigma_omega2) * cl["nu_t"]
    Rw = _scalar_balance(m, g, gspec, t.omega_bc, omega, Q, gw, None, scalar_advection, s.limiter_k)
    Rw = Rw - omega_production(cl["strain"], cl["nu_t"], k, omega, cl["f1"], model) * V
    Rw = Rw - omega_destruction(omega, cl["f1"], model) * V
    Rw = Rw - omega_cross_diffusion(omega, cl["grad_k"], cl["grad_omega"], cl["f1"], model) * V
    Rw = Rw.clone()
    target = omega_wall(t.nu[wall], dw[wall], kw, model)
    Rw[wall] = torch.log(omega[wall] / target) if omega_log else omega[wall] - target
    return torch.cat([flow_res, Rk, Rw])
