"""Oracle: geometry, schemes, face fluxes and the coupled-flow / scalar residuals (torch FP64, CPU).

Test infrastructure only (see ``oracle/__init__.py``).  Every function restates one reference
function; the citation is in its docstring.  Arrays are torch float64 on the CPU; index arrays are
int64.  Layouts follow the reference: per-face vectors ``(F, d)``, per-cell vectors ``(C, d)``,
vector gradients ``(C, d, d)`` with ``[c, i, j] = d u_i / d x_j``.
"""

from __future__ import annotations

import dataclasses

import numpy as np
import torch

F64 = torch.float64

# Boundary-face kinds for the coupled flow (one per face; 0 = interior or unlisted).
FLOW_NONE, FLOW_WALL, FLOW_INLET, FLOW_OUTLET = 0, 1, 2, 3
# Scalar closures.
BC_NONE, BC_DIRICHLET, BC_ZERO_GRADIENT, BC_NEUMANN, BC_CONVECTIVE = 0, 1, 2, 3, 4


def dot(a, b):
    """Left-to-right component sum of ``a * b`` over the last axis (``aquaflux/vectors.py:28-65``)."""
    p = a * b
    out = p[..., 0]
    for i in range(1, p.shape[-1]):
        out = out + p[..., i]
    return out


def scale(v, s):
    """``s[..., None] * v`` (``aquaflux/vectors.py:88-107``)."""
    return s[..., None] * v


def segment_sum(values, index, n):
    out = torch.zeros((n,) + tuple(values.shape[1:]), dtype=values.dtype)
    return out.index_add(0, index, values)


@dataclasses.dataclass
class OMesh:
    """Plain-array mesh view for the oracle (owner/neighbour, face rings, patches)."""

    node_coords: torch.Tensor
    owner: torch.Tensor
    neighbour: torch.Tensor
    offsets: np.ndarray
    indices: np.ndarray
    n_cells: int
    patches: dict

    @classmethod
    def from_mesh(cls, mesh, node_coords=None):
        nc = mesh.node_coords if node_coords is None else node_coords
        if not isinstance(nc, torch.Tensor):
            nc = torch.tensor(np.asarray(nc), dtype=F64)
        patches = {
            name: torch.tensor(mesh.face_patches.indices(name), dtype=torch.long)
            for name in mesh.face_patches.names
        }
        return cls(
            node_coords=nc,
            owner=torch.tensor(mesh.owner.astype(np.int64)),
            neighbour=torch.tensor(mesh.neighbour.astype(np.int64)),
            offsets=np.asarray(mesh.face_node_offsets, dtype=np.int64),
            indices=np.asarray(mesh.face_node_indices, dtype=np.int64),
            n_cells=int(mesh.n_cells),
            patches=patches,
        )

    @property
    def dim(self):
        return int(self.node_coords.shape[1])

    @property
    def n_faces(self):
        return int(self.owner.shape[0])

    @property
    def interior(self):
        return self.neighbour >= 0

    @property
    def safe_nb(self):
        """``connectivity.py:126-136``: boundary faces read their owner as neighbour."""
        return torch.where(self.interior, self.neighbour, self.owner)

    @property
    def nb_scatter(self):
        """Trash-row scatter index: ``n_cells`` on boundary faces (``connectivity.py:131-136``)."""
        return torch.where(self.interior, self.neighbour, torch.full_like(self.neighbour, self.n_cells))


# --- max / min with JAX's derivative rule ---------------------------------------------------


def maximum(a, b):
    """``jnp.maximum``: value ``max(a, b)``; the derivative takes the larger operand's (half of each
    on a tie).  Written with ``torch.where`` because torch's own forward-mode rule for
    ``torch.maximum`` (``t_b + mask * (t_a - t_b)``) loses the selected tangent to cancellation when
    the other operand's tangent is many orders larger, which JAX's per-operand weighting does not."""
    a, b = torch.broadcast_tensors(torch.as_tensor(a, dtype=F64), torch.as_tensor(b, dtype=F64))
    return torch.where(a > b, a, torch.where(a < b, b, 0.5 * (a + b)))


def minimum(a, b):
    """``jnp.minimum`` with the same derivative rule as :func:`maximum`."""
    a, b = torch.broadcast_tensors(torch.as_tensor(a, dtype=F64), torch.as_tensor(b, dtype=F64))
    return torch.where(a < b, a, torch.where(a > b, b, 0.5 * (a + b)))


# --- scatter primitives (aquaflux/mesh/connectivity.py:236-342) ------------------------------


def scatter(m: OMesh, own, nb):
    """Owner ``segment_sum`` plus neighbour ``segment_sum`` with the trash row dropped (``:236-260``)."""
    return segment_sum(own, m.owner, m.n_cells) + segment_sum(nb, m.nb_scatter, m.n_cells + 1)[: m.n_cells]


def scatter_conservative(m: OMesh, flux):
    """Owner ``+flux``, neighbour ``-flux`` (``connectivity.py:262-280``)."""
    return scatter(m, flux, -flux)


def _scatter_ext(m: OMesh, own, nb, how):
    n = m.n_cells
    fill = -torch.inf if how == "amax" else torch.inf
    a = torch.full((n,), fill, dtype=own.dtype).scatter_reduce(0, m.owner, own, how, include_self=True)
    b = torch.full((n + 1,), fill, dtype=own.dtype).scatter_reduce(
        0, m.nb_scatter, nb, how, include_self=True
    )[:n]
    return torch.maximum(a, b) if how == "amax" else torch.minimum(a, b)


def scatter_max(m, own, nb):
    """Per-cell max of face contributions (``connectivity.py:299-318``)."""
    return _scatter_ext(m, own, nb, "amax")


def scatter_min(m, own, nb):
    """Per-cell min of face contributions (``connectivity.py:320-342``)."""
    return _scatter_ext(m, own, nb, "amin")


# --- geometry (aquaflux/mesh/face.py, cell.py, mesh.py) ----------------------------------------


def _safe_magnitude(v):
    """``face.py:70-88``: zero-safe magnitude (finite gradient at the zero vector)."""
    sq = dot(v, v)
    return torch.where(sq > 0.0, torch.sqrt(torch.where(sq > 0.0, sq, torch.ones_like(sq))), torch.zeros_like(sq))


def _safe_unit(v):
    """``face.py:91-111``: zero-safe unit vectors."""
    sq = torch.sum(v * v, dim=-1, keepdim=True)
    return torch.where(sq > 0.0, v / torch.sqrt(torch.where(sq > 0.0, sq, torch.ones_like(sq))), torch.zeros_like(v))


def _incidence_maps(m: OMesh):
    counts = np.diff(m.offsets)
    inc_face = np.repeat(np.arange(m.n_faces), counts)
    starts = m.offsets[:-1][inc_face]
    local = np.arange(m.indices.size) - starts
    nxt = starts + (local + 1) % counts[inc_face]
    return torch.tensor(inc_face), torch.tensor(nxt), torch.tensor(counts, dtype=F64)


@dataclasses.dataclass
class OGeometry:
    area: torch.Tensor
    face_centroid: torch.Tensor
    normal: torch.Tensor
    volume: torch.Tensor
    cell_centroid: torch.Tensor


def geometry(m: OMesh) -> OGeometry:
    """``Mesh.geometry`` / ``_fused_geometry`` (``aquaflux/mesh/mesh.py:341-387``)."""
    nodes = m.node_coords
    idx = torch.tensor(m.indices)
    if m.dim == 2:
        # EdgeFaceGeometry.unoriented_geometry (face.py:217-230)
        verts = nodes[idx].reshape(m.n_faces, 2, -1)
        p0, p1 = verts[:, 0], verts[:, 1]
        d = p1 - p0
        area = _safe_magnitude(d)
        centroid = 0.5 * (p0 + p1)
        normal = _safe_unit(torch.stack([d[:, 1], -d[:, 0]], dim=1))
    else:
        # PolygonFaceGeometry.centre_fan (face.py:292-349)
        inc_face, nxt, counts = _incidence_maps(m)
        pb = nodes[idx]
        pc = pb[nxt]
        centre = scale(segment_sum(pb, inc_face, m.n_faces), 1.0 / counts)
        pa = centre[inc_face]
        directed = 0.5 * torch.linalg.cross(pb - pa, pc - pa, dim=-1)
        tri_centroid = (pa + pb + pc) / 3.0
        vector_area = segment_sum(directed, inc_face, m.n_faces)
        area = _safe_magnitude(vector_area)
        normal = _safe_unit(vector_area)
        signed = dot(directed, normal[inc_face])
        safe_area = torch.where(area > 0.0, area, torch.ones_like(area))[:, None]
        centroid = segment_sum(scale(tri_centroid, signed), inc_face, m.n_faces) / safe_area
    # CellGeometry.approx_centroids (cell.py:84-112)
    ones = torch.ones(m.n_faces, dtype=F64)
    count = scatter(m, ones, ones)
    approx = scale(scatter(m, centroid, centroid), 1.0 / count)
    # orient_owner_outward (face.py:196-211)
    outward = centroid - approx[m.owner]
    projection = torch.sum(outward * normal, dim=1, keepdim=True)
    normal = normal * torch.where(projection < 0.0, -1.0, 1.0)
    # CellGeometry.from_faces (cell.py:123-170)
    own_local = centroid - approx[m.owner]
    nb_local = centroid - approx[m.safe_nb]
    flux = dot(own_local, normal) * area
    nb_flux = dot(nb_local, normal) * area
    volume = scatter(m, flux, -nb_flux) / m.dim
    csum = scatter(m, scale(own_local, flux), -scale(nb_local, nb_flux))
    cell_centroid = scale(csum, 1.0 / ((m.dim + 1) * volume)) + approx
    return OGeometry(area, centroid, normal, volume, cell_centroid)


# --- interpolation & gradients (aquaflux/schemes) -----------------------------------------------


def interpolation_factor(m: OMesh, g: OGeometry):
    """``schemes/interpolation.py:27-55``."""
    x_p = g.cell_centroid[m.owner]
    d = g.cell_centroid[m.safe_nb] - x_p
    dd = torch.where(m.interior, dot(d, d), torch.ones(m.n_faces, dtype=F64))
    return torch.where(m.interior, dot(g.face_centroid - x_p, d) / dd, torch.zeros(m.n_faces, dtype=F64))


def blend(m: OMesh, own_field, nb_field, factor):
    """``blend_owner_neighbour`` (``interpolation.py:84-122``)."""
    gg = factor.reshape(factor.shape + (1,) * (own_field.dim() - 1))
    return (1.0 - gg) * own_field[m.owner] + gg * nb_field[m.safe_nb]


def interpolate_to_face(m: OMesh, g: OGeometry, field, grad, factor):
    """Blend plus skewness correction ``grad_f . (x_ip - x_g)`` (``interpolation.py:125-176``)."""
    b = blend(m, field, field, factor)
    grad_face = blend(m, grad, grad, factor)
    x_p = g.cell_centroid[m.owner]
    d = g.cell_centroid[m.safe_nb] - x_p
    skew = g.face_centroid - (x_p + scale(d, factor))
    if grad_face.dim() == 2:
        corr = dot(grad_face, skew)
    else:
        corr = dot(grad_face, skew[:, None, :])
    return b + corr


def compact_green_gauss(m: OMesh, g: OGeometry, field, bvals):
    """``CompactGreenGauss.gradients`` (``schemes/gradient.py:129-153``)."""
    fac = interpolation_factor(m, g)
    phi_int = blend(m, field, field, fac)
    phi_face = torch.where(m.interior, phi_int, bvals)
    area_vector = scale(g.normal, g.area)
    return scale(scatter_conservative(m, scale(area_vector, phi_face)), 1.0 / g.volume)


def corrected_terms(m: OMesh, g: OGeometry):
    """``CorrectedGreenGauss.terms`` (``schemes/gradient.py:1160-1185``): interpolation factor,
    skewness offset x_ip - x_g, owner-outward area vector, volume."""
    fac = interpolation_factor(m, g)
    x_p = g.cell_centroid[m.owner]
    d = g.cell_centroid[m.safe_nb] - x_p
    skew = g.face_centroid - (x_p + scale(d, fac))
    return fac, skew, scale(g.normal, g.area), g.volume


def corrected_operator(m: OMesh, g: OGeometry, x):
    """``A_g x = V x - correction(x)`` (``schemes/gradient.py:1187-1204``)."""
    fac, skew, S, V = corrected_terms(m, g)
    w = (1.0 - fac) * dot(skew, x[m.owner]) + fac * dot(skew, x[m.safe_nb])
    corr = scatter_conservative(m, torch.where(m.interior[:, None], scale(S, w), torch.zeros_like(S)))
    return scale(x, V) - corr


def corrected_rhs(m: OMesh, g: OGeometry, field, bvals):
    """``B phi`` (``schemes/gradient.py:1311-1319``): Green-Gauss sum of interpolated face values."""
    fac, _, S, _ = corrected_terms(m, g)
    phi_face = torch.where(m.interior, blend(m, field, field, fac), bvals)
    return scatter_conservative(m, scale(S, phi_face))


def corrected_green_gauss(m: OMesh, g: OGeometry, field, bvals, sweeps=4, relaxation=1.0):
    """``CorrectedGreenGauss`` with ``SweptGradientSolve`` (``schemes/gradient.py:618-705``):
    x0 = r V^-1 b (the peeled first sweep), then x <- x + r V^-1 (b - A_g x), ``sweeps - 1`` times."""
    b = corrected_rhs(m, g, field, bvals)
    inv = 1.0 / g.volume
    if sweeps <= 0:
        return torch.zeros_like(b)
    x = relaxation * scale(b, inv)
    for _ in range(sweeps - 1):
        x = x + relaxation * scale(b - corrected_operator(m, g, x), inv)
    return x


def least_squares_gradient(m: OMesh, g: OGeometry, field, bvals):
    """Weighted least-squares cell gradient (no reference counterpart: a new ``GradientScheme``).

    For cell ``c`` minimise ``sum_f w_f (dphi_f - G . dx_f)^2`` over its faces, with
    ``dx_f = x_other - x_c`` on interior faces and ``x_ip - x_c`` (``dphi = phi_b - phi_c``) on
    boundary faces, and inverse-distance-squared weights ``w_f = 1 / |dx_f|^2``.  Exact for linear
    fields whenever the stencil spans the space.  Parity: pinned by linear exactness and by a
    brute-force ``numpy.linalg.lstsq`` per cell in the tests.
    """
    d = m.dim
    cc = g.cell_centroid
    x_own = cc[m.owner]
    x_other_own = torch.where(m.interior[:, None], cc[m.safe_nb], g.face_centroid)
    dx_own = x_other_own - x_own  # as seen from the owner
    dphi_own = torch.where(m.interior, field[m.safe_nb], bvals) - field[m.owner]
    w = 1.0 / dot(dx_own, dx_own)
    outer = dx_own[:, :, None] * dx_own[:, None, :]
    # Neighbour side sees -dx, -dphi: the outer product and dx*dphi are unchanged.
    M = scatter(m, w[:, None, None] * outer, w[:, None, None] * outer)
    b = scatter(m, scale(dx_own, w * dphi_own), scale(dx_own, w * dphi_own))
    del d
    return torch.linalg.solve(M, b.unsqueeze(-1)).squeeze(-1)


def venkatakrishnan(m: OMesh, g: OGeometry, phi, grad, k=5.0):
    """``VenkatakrishnanLimiter.limit`` (``schemes/limiter.py:78-117``)."""
    own, nb = m.owner, m.safe_nb
    phi_max = torch.maximum(phi, scatter_max(m, phi[nb], phi[own]))
    phi_min = torch.minimum(phi, scatter_min(m, phi[nb], phi[own]))
    eps2 = g.volume * k**3

    def face_limiter(cell):
        dm = dot(grad[cell], g.face_centroid - g.cell_centroid[cell])
        sign = torch.where(dm >= 0.0, 1.0, -1.0).to(F64)
        dm = sign * (torch.abs(dm) + 1e-12)
        head = torch.where(dm > 0.0, phi_max[cell] - phi[cell], phi_min[cell] - phi[cell])
        e2 = eps2[cell]
        num = (head**2 + e2) * dm + 2.0 * dm**2 * head
        den = (head**2 + 2.0 * dm**2 + head * dm + e2) * dm
        return num / den

    psi = scatter_min(m, face_limiter(own), face_limiter(nb))
    # jnp.clip(x, 0, 1) == minimum(maximum(x, 0), 1): ties split the derivative
    return minimum(maximum(psi, 0.0), 1.0)


# --- face fluxes (aquaflux/discretization) --------------------------------------------------------


def diffusion_flux(m: OMesh, g: OGeometry, phi, grad, gamma, bvals, boundary_gamma=None):
    """``DiffusionFlux.face_flux`` (``discretization/diffusion.py:163-199``)."""
    own, nb = m.owner, m.safe_nb
    n, area, x_ip = g.normal, g.area, g.face_centroid
    d_p = x_ip - g.cell_centroid[own]
    d_n = x_ip - g.cell_centroid[nb]
    dpn = dot(d_p, n)
    dnn = dot(d_n, n)
    corr_p = dot(grad[own], d_p - scale(n, dpn))
    corr_n = dot(grad[nb], d_n - scale(n, dnn))
    g_own, g_nb = gamma[own], gamma[nb]
    denom = dpn - (g_own / g_nb) * dnn
    denom_safe = torch.where(m.interior, denom, torch.ones_like(denom))
    ng_int = ((phi[nb] - phi[own]) + corr_n - corr_p) / denom_safe
    ng_bnd = (bvals - phi[own] - corr_p) / dpn
    ng = torch.where(m.interior, ng_int, ng_bnd)
    gamma_face = g_own if boundary_gamma is None else torch.where(m.interior, g_own, boundary_gamma)
    return -gamma_face * ng * area


def conductance(m: OMesh, g: OGeometry, gamma):
    """``flux_continuous_conductance`` (``discretization/diffusion.py:85-134``)."""
    own, nb = m.owner, m.safe_nb
    n = g.normal
    dpn = dot(g.face_centroid - g.cell_centroid[own], n)
    dnn = dot(g.face_centroid - g.cell_centroid[nb], n)
    denom = torch.where(m.interior, dpn - (gamma[own] / gamma[nb]) * dnn, dpn)
    return gamma[own] * g.area / denom


def upwind_face_value(m: OMesh, phi, mdot, bvals):
    """``FirstOrderUpwind.face_value`` (``discretization/advection.py:96-110``)."""
    outflow = mdot >= 0.0
    interior_value = torch.where(outflow, phi[m.owner], phi[m.safe_nb])
    boundary_value = torch.where(outflow, phi[m.owner], bvals)
    return torch.where(m.interior, interior_value, boundary_value)


def limited_face_value(m: OMesh, g: OGeometry, phi, grad, psi, mdot, bvals):
    """``LimitedUpwind.face_value`` (``discretization/advection.py:113-158``)."""
    outflow = mdot >= 0.0
    own, nb = m.owner, m.safe_nb
    phi_up = torch.where(outflow, phi[own], phi[nb])
    psi_up = torch.where(outflow, psi[own], psi[nb])
    grad_up = torch.where(outflow[:, None], grad[own], grad[nb])
    x_up = torch.where(outflow[:, None], g.cell_centroid[own], g.cell_centroid[nb])
    rec = phi_up + psi_up * dot(grad_up, g.face_centroid - x_up)
    return torch.where(m.interior | outflow, rec, bvals)


# --- coupled flow (aquaflux/flow) -------------------------------------------------------------------


@dataclasses.dataclass
class FlowSetup:
    """Resolved flow boundary data: per-face kind + prescribed values, and the options."""

    kind: torch.Tensor  # (F,) int64, FLOW_*
    u_b: torch.Tensor  # (F, d) prescribed velocity (walls, inlets)
    p_b: torch.Tensor  # (F,) prescribed outlet pressure
    advection: str = "upwind"  # "none" | "upwind" | "limited"
    gradient: str = "green_gauss"  # "green_gauss" | "least_squares" | "corrected"
    limiter_k: float = 5.0
    pressure_pin: int | None = None
    pressure_pin_value: float = 0.0
    body_force: tuple = (0.0, 0.0, 0.0)
    gradient_sweeps: int = 4
    gradient_relaxation: float = 1.0


def flow_boundary_fields(m, g, setup: FlowSetup, vel, p):
    """``_boundary_fields`` (``flow/momentum.py:445-470``) with the closures of ``flow/boundary.py``."""
    own = m.owner
    k = setup.kind
    zero = torch.zeros_like(setup.u_b)
    ub = torch.where((k == FLOW_WALL)[:, None] | (k == FLOW_INLET)[:, None], setup.u_b, zero)
    ub = torch.where((k == FLOW_OUTLET)[:, None], vel[own], ub)
    pz = torch.zeros(m.n_faces, dtype=F64)
    pb = torch.where((k == FLOW_WALL) | (k == FLOW_INLET), p[own], pz)
    pb = torch.where(k == FLOW_OUTLET, setup.p_b, pb)
    return ub, pb


def _gradient(m, g, setup, field, bvals):
    if setup.gradient == "least_squares":
        return least_squares_gradient(m, g, field, bvals)
    if setup.gradient == "corrected":
        return corrected_green_gauss(m, g, field, bvals, setup.gradient_sweeps, setup.gradient_relaxation)
    return compact_green_gauss(m, g, field, bvals)


def normal_distance(m, g):
    """``MomentumContinuity.build`` normal distance (``flow/momentum.py:268-277``)."""
    x_p = g.cell_centroid[m.owner]
    d = g.cell_centroid[m.safe_nb] - x_p
    return torch.where(m.interior, dot(d, g.normal), dot(g.face_centroid - x_p, g.normal))


def schur_face_coefficient(m, g, a_p, rho):
    """Per-face SIMPLE Schur coefficient ``rho_f (V/a_P)_f A / (d.n)``, blended ``(1 - g) P + g N``,
    zero on boundary faces (``flow/preconditioner.py:43-81``)."""
    fac = interpolation_factor(m, g)
    d = g.volume / a_p
    rho_f = blend(m, rho, rho, fac)
    d_f = blend(m, d, d, fac)
    c = rho_f * d_f * g.area / normal_distance(m, g)
    return torch.where(m.interior, c, torch.zeros_like(c))


def pressure_schur_matrix(m, g, a_p, rho, outlet_faces=(), pin=None):
    """Dense pressure Schur: the graph Laplacian of the interior face coefficients
    (``solve/frozen_operator.py:214-271`` without a flux) plus the outlet stiffness
    ``rho_P (V/a_P)_P A / (d.n)`` on each pressure-fixing boundary face's owner
    (``flow/boundary.py:437-441``, ``flow/block_preconditioner.py:191-209``), with a pinned cell
    decoupled (``solve/frozen_operator.py:274-299``)."""
    n = m.n_cells
    c = schur_face_coefficient(m, g, a_p, rho)
    S = torch.zeros((n, n), dtype=F64)
    for f in torch.nonzero(m.interior).flatten().tolist():
        P, N = int(m.owner[f]), int(m.neighbour[f])
        S[P, P] += c[f]
        S[N, N] += c[f]
        S[P, N] -= c[f]
        S[N, P] -= c[f]
    nd = normal_distance(m, g)
    d = g.volume / a_p
    for f in outlet_faces:
        P = int(m.owner[f])
        S[P, P] += rho[P] * d[P] * g.area[f] / nd[f]
    if pin is not None:
        S[pin, :] = 0.0
        S[:, pin] = 0.0
        S[pin, pin] = 1.0
    return S


def advective_momentum_flux(m, g, vel, rho, fac, grad_vel):
    """``flow/rhie_chow.py:220-271``: ``(rho u)_f . n`` reconstructed to the integration point."""
    mom = scale(vel, rho)
    if grad_vel is None:
        mf = blend(m, mom, mom, fac)
    else:
        mf = interpolate_to_face(m, g, mom, rho[:, None, None] * grad_vel, fac)
    return dot(mf, g.normal)


def momentum_diagonal_coefficient(kind, viscous, convective):
    """``FlowBoundary.momentum_diagonal_coefficient`` per patch (``flow/boundary.py:216-238, 302-308,
    349-355, 443-446``): a wall adds the viscous ``mu A / (d.n)`` only, a pressure outlet the
    convective ``max(mdot, 0)`` only, a velocity inlet both; other faces nothing."""
    out = torch.where(kind == FLOW_WALL, viscous, torch.zeros_like(viscous))
    out = torch.where(kind == FLOW_INLET, viscous + convective, out)
    return torch.where(kind == FLOW_OUTLET, convective, out)


def momentum_diagonal(m, g, setup: FlowSetup, mu_eff, boundary_mu, mdot_est):
    """``momentum_matrix_diagonal`` (``flow/momentum.py:596-635``) with ``momentum_diagonal``
    (``flow/rhie_chow.py:81-168``): per cell, the interior conductances plus ``max(+-mdot_est, 0)``
    and each boundary face's per-patch term."""
    dist = normal_distance(m, g)
    visc_b = boundary_mu * g.area / dist
    conv_b = maximum(mdot_est, 0.0) if mdot_est is not None else torch.zeros(m.n_faces, dtype=F64)
    bcoef = momentum_diagonal_coefficient(setup.kind, visc_b, conv_b)
    visc = conductance(m, g, mu_eff)
    oc, nc = visc, visc
    if mdot_est is not None:
        oc = oc + maximum(mdot_est, 0.0)
        nc = nc + maximum(-mdot_est, 0.0)
    oc = torch.where(m.interior, oc, bcoef)
    return scatter(m, oc, nc)


def interior_mass_flux(m, g, vel, grad_vel, p, grad_p, d_coeff, rho, fac, dist):
    """``interior_mass_flux`` (``flow/rhie_chow.py:274-353``): ``[(rho u)_f . n - rho_f d_hat ((p_N -
    p_P) - grad p_f . d) / (d.n)] A`` with ``d_hat = sum_i n_i^2 (V/a_P)_{f,i}``."""
    own = m.owner
    mom_n = advective_momentum_flux(m, g, vel, rho, fac, grad_vel)
    rho_f = blend(m, rho, rho, fac)
    gp_f = blend(m, grad_p, grad_p, fac)
    dc_f = blend(m, d_coeff, d_coeff, fac)
    d_hat = dot(g.normal * g.normal, dc_f)
    dvec = g.cell_centroid[m.safe_nb] - g.cell_centroid[own]
    dp = p[m.safe_nb] - p[own]
    damping = (dp - dot(gp_f, dvec)) / dist
    return (mom_n - rho_f * d_hat * damping) * g.area


def inlet_mass_flux(rho_owner, u_b, normal, area):
    """``VelocityInlet.mass_flux`` (``flow/boundary.py:381-397``): ``rho (u_in . n) A``."""
    return rho_owner * dot(u_b, normal) * area


def outlet_mass_flux(vel_owner, p_owner, grad_p_owner, d_coeff_owner, normal, area, dist, rho_owner, p_b):
    """``PressureOutlet.mass_flux`` (``flow/boundary.py:418-434``): ``rho (u_P . n - d_hat ((p_b -
    p_P) / (d.n) - grad p_P . n)) A``."""
    compact = (p_b - p_owner) / dist
    interp = dot(grad_p_owner, normal)
    dh_own = dot(normal * normal, d_coeff_owner)
    return rho_owner * (dot(vel_owner, normal) - dh_own * (compact - interp)) * area


def pressure_force_flux(p_face, normal, area, component):
    """``PressureForce.face_flux`` (``flow/momentum.py:112-155``): ``p_f n_i A``."""
    return p_face * normal[:, component] * area


def flow_fields(m, g, setup: FlowSetup, state, mu, rho, nu_t=None, wall_nu_t=None, shear_mask=None, frozen=None):
    """``MomentumContinuity.flow_fields`` (``flow/momentum.py:822-857``) -- returns a dict.

    ``frozen`` (a dict from an earlier call) substitutes its gradients and ``a_p``: the reach-1
    linearisation the block Jacobian of the Newton preconditioner is built from."""
    d, C = m.dim, m.n_cells
    blocks = state.reshape(d + 1, C)
    vel, p = blocks[:d].T, blocks[d]
    mu_eff = mu if nu_t is None else mu + rho * nu_t
    ub, pb = flow_boundary_fields(m, g, setup, vel, p)
    grad_vel = torch.stack([_gradient(m, g, setup, vel[:, i], ub[:, i]) for i in range(d)], dim=1)
    grad_p = _gradient(m, g, setup, p, pb)
    if frozen is not None:
        grad_vel, grad_p = frozen["grad_vel"].detach(), frozen["grad_p"].detach()
    fac = interpolation_factor(m, g)
    dist = normal_distance(m, g)
    own = m.owner
    # boundary face viscosity (momentum.py:_boundary_face_viscosity / _wall_boundary_viscosity)
    if wall_nu_t is not None:
        wall_mu = mu[own] + rho[own] * wall_nu_t
        bvisc = torch.where(shear_mask, wall_mu, mu_eff[own])
        boundary_gamma = bvisc
    else:
        bvisc = mu_eff[own]
        boundary_gamma = None
    # momentum_matrix_diagonal (momentum.py:596-635) + momentum_diagonal (rhie_chow.py:81-175)
    mdot_est = None
    if setup.advection != "none":
        mdot_est = advective_momentum_flux(m, g, vel, rho, fac, grad_vel) * g.area
    a_p = momentum_diagonal(m, g, setup, mu_eff, bvisc, mdot_est)
    if frozen is not None:
        a_p = frozen["a_p"].detach()
    d_coeff = (g.volume / a_p)[:, None].expand(C, d)
    # _mass_flux (momentum.py:531-554) + interior_mass_flux (rhie_chow.py:274-353)
    mdot_int = interior_mass_flux(m, g, vel, grad_vel, p, grad_p, d_coeff, rho, fac, dist)
    mdot = torch.where(m.interior, mdot_int, torch.zeros_like(mdot_int))
    # boundary mass-flux closures (flow/boundary.py:278-446)
    k = setup.kind
    n = g.normal
    mdot_in = inlet_mass_flux(rho[own], ub, n, g.area)
    mdot_out = outlet_mass_flux(vel[own], p[own], grad_p[own], d_coeff[own], n, g.area, dist, rho[own], setup.p_b)
    mdot = torch.where(k == FLOW_WALL, torch.zeros_like(mdot), mdot)
    mdot = torch.where(k == FLOW_INLET, mdot_in, mdot)
    mdot = torch.where(k == FLOW_OUTLET, mdot_out, mdot)
    return dict(
        vel=vel, p=p, ub=ub, pb=pb, grad_vel=grad_vel, grad_p=grad_p, a_p=a_p, mdot=mdot,
        mu_eff=mu_eff, boundary_gamma=boundary_gamma, fac=fac, mdot_est=mdot_est,
    )


def flow_residual_from_fields(m, g, setup: FlowSetup, f, psi_vel=None):
    """``residual_from_fields`` (``flow/momentum.py:859-878``)."""
    d = m.dim
    own = m.owner
    # _face_pressure (momentum.py:498-513)
    pf_int = interpolate_to_face(m, g, f["p"], f["grad_p"], f["fac"])
    p_face = torch.where(m.interior, pf_int, f["pb"])
    mdot = f["mdot"]
    cols = []
    for i in range(d):
        phi = f["vel"][:, i]
        grad = f["grad_vel"][:, i]
        bv = f["ub"][:, i]
        flux = torch.zeros(m.n_faces, dtype=F64)
        flux = flux + diffusion_flux(m, g, phi, grad, f["mu_eff"], bv, f["boundary_gamma"])
        flux = flux + pressure_force_flux(p_face, g.normal, g.area, i)
        if setup.advection == "upwind":
            flux = flux + mdot * upwind_face_value(m, phi, mdot, bv)
        elif setup.advection == "limited":
            psi = psi_vel[i]
            flux = flux + mdot * limited_face_value(m, g, phi, grad, psi, mdot, bv)
        cols.append(scatter_conservative(m, flux) - setup.body_force[i] * g.volume)
    cont = scatter_conservative(m, mdot)
    if setup.pressure_pin is not None:
        cont = cont.clone()
        cont[setup.pressure_pin] = f["p"][setup.pressure_pin] - setup.pressure_pin_value
    return torch.cat(cols + [cont])


def flow_residual(m, g, setup: FlowSetup, state, mu, rho, frozen_state=None):
    """``MomentumContinuity.residual`` (``flow/momentum.py:880-882``).

    With ``frozen_state`` the gradients, ``a_p`` and limiter are those of ``frozen_state`` (held
    constant): differentiating that at ``state == frozen_state`` gives the reach-1 block Jacobian."""
    frozen = None
    if frozen_state is not None:
        with torch.no_grad():
            frozen = flow_fields(m, g, setup, frozen_state, mu, rho)
    f = flow_fields(m, g, setup, state, mu, rho, frozen=frozen)
    psi = None
    if setup.advection == "limited":
        src = frozen if frozen is not None else f
        psi = [venkatakrishnan(m, g, src["vel"][:, i], src["grad_vel"][:, i], setup.limiter_k) for i in range(m.dim)]
        if frozen is not None:
            psi = [p.detach() for p in psi]
    return flow_residual_from_fields(m, g, setup, f, psi)


# --- scalar transport (aquaflux/discretization/residual.py, transport/scalar.py) -----------------


@dataclasses.dataclass
class ScalarSetup:
    kind: torch.Tensor  # (F,) BC_*
    value: torch.Tensor  # (F,) Dirichlet value / Neumann flux q / convective h
    t_inf: torch.Tensor  # (F,) convective ambient value
    advection: str = "limited"  # "none" | "upwind" | "limited"
    gradient: str = "green_gauss"  # "none" | "green_gauss" | "least_squares"
    limiter: bool = True
    limiter_k: float = 5.0
    gradient_sweeps: int = 4
    gradient_relaxation: float = 1.0


def tangential_correction(grad_owner, d, normal):
    """``boundary/conditions.py`` ``_tangential_correction``: ``grad_P . (d - (d.n) n)``."""
    return dot(grad_owner, d - scale(normal, dot(d, normal)))


def zero_gradient_face_value(phi_owner, grad_owner, d, normal):
    """``ZeroGradient.face_value`` (``boundary/conditions.py:145-153``): ``phi_P + corr``."""
    return phi_owner + tangential_correction(grad_owner, d, normal)


def neumann_face_value(phi_owner, grad_owner, d, normal, gamma_owner, q):
    """``Neumann.face_value`` (``boundary/conditions.py:156-173``): ``phi_P + corr - (q / Gamma_P)(d.n)``."""
    return phi_owner + tangential_correction(grad_owner, d, normal) - q / gamma_owner * dot(d, normal)


def convective_face_value(phi_owner, grad_owner, d, normal, gamma_owner, h, t_inf):
    """``Convective.face_value`` (``boundary/conditions.py:176-201``):
    ``(phi_P + corr + beta T_inf) / (1 + beta)`` with ``beta = (h / Gamma_P)(d.n)``."""
    beta = h / gamma_owner * dot(d, normal)
    return (phi_owner + tangential_correction(grad_owner, d, normal) + beta * t_inf) / (1.0 + beta)


def scalar_boundary_values(m, g, s: ScalarSetup, phi, grad, gamma):
    """``ResidualAssembler.boundary_values`` (``discretization/residual.py:298-356``): each boundary
    face's closure (``boundary/conditions.py:111-201``) evaluated from its owner."""
    own = m.owner
    dvec = g.face_centroid - g.cell_centroid[own]
    n = g.normal
    k = s.kind
    phi_o, grad_o = phi[own], grad[own]
    out = torch.zeros(m.n_faces, dtype=F64)
    out = torch.where(k == BC_DIRICHLET, s.value, out)
    out = torch.where(k == BC_ZERO_GRADIENT, zero_gradient_face_value(phi_o, grad_o, dvec, n), out)
    # the Neumann / Convective branches divide by the owner diffusivity; faces of other kinds read a
    # unit stand-in there so that no unselected 0/0 reaches the reverse pass of torch.where
    robin = (k == BC_NEUMANN) | (k == BC_CONVECTIVE)
    gam = torch.where(robin, gamma[own], torch.ones_like(gamma[own]))
    out = torch.where(k == BC_NEUMANN, neumann_face_value(phi_o, grad_o, dvec, n, gam, s.value), out)
    out = torch.where(k == BC_CONVECTIVE, convective_face_value(phi_o, grad_o, dvec, n, gam, s.value, s.t_inf), out)
    return out


def transient_term(phi, phi_old, phi_older, dt, first_step, volume):
    """``TransientTerm.residual`` (``discretization/transient.py:32-62``): BDF1 on the first step,
    BDF2 afterwards."""
    if first_step:
        return volume * (phi - phi_old) / dt
    return volume * (1.5 * phi - 2.0 * phi_old + 0.5 * phi_older) / dt


def scalar_residual(m, g, s: ScalarSetup, phi, flux, gamma, phi_old=None, phi_older=None, dt=None,
                    first_step=False):
    """``ResidualAssembler.residual`` for ``ScalarTransport`` (Advection, Diffusion) order
    (``transport/scalar.py:185-208``, ``discretization/residual.py:308-412``); with ``phi_old`` the
    accumulation term is added after the flux scatter (``residual.py:141-147``)."""
    C, d = m.n_cells, m.dim
    if s.gradient == "none":
        grad = torch.zeros((C, d), dtype=F64)
        bvals = scalar_boundary_values(m, g, s, phi, grad, gamma)
    else:
        zero = torch.zeros((C, d), dtype=F64)
        lead = scalar_boundary_values(m, g, s, phi, zero, gamma)
        if s.gradient == "least_squares":
            grad = least_squares_gradient(m, g, phi, lead)
        elif s.gradient == "corrected":
            grad = corrected_green_gauss(m, g, phi, lead, s.gradient_sweeps, s.gradient_relaxation)
        else:
            grad = compact_green_gauss(m, g, phi, lead)
        bvals = scalar_boundary_values(m, g, s, phi, grad, gamma)
    face = torch.zeros(m.n_faces, dtype=F64)
    if s.advection == "upwind":
        face = face + flux * upwind_face_value(m, phi, flux, bvals)
    elif s.advection == "limited":
        psi = venkatakrishnan(m, g, phi, grad, s.limiter_k) if s.limiter else torch.ones(C, dtype=F64)
        face = face + flux * limited_face_value(m, g, phi, grad, psi, flux, bvals)
    face = face + diffusion_flux(m, g, phi, grad, gamma, bvals)
    R = scatter_conservative(m, face)
    if phi_old is not None:
        R = R + transient_term(phi, phi_old, phi_older, dt, first_step, g.volume)
    return R
