"""Pins the oracle's source terms, per-patch momentum diagonal, mass fluxes, pressure force and
scalar boundary closures against the closed forms of the reference's unit tests (CPU only).

Each test cites the reference test it restates.  Every function pinned here is the one the
oracle's whole residual calls (``oracle/sst.py:rans_residual``, ``oracle/fv.py:flow_fields`` /
``flow_residual_from_fields`` / ``scalar_boundary_values``), so a wrong blend, sign or operand in
any of them fails a test here before it can reach a GPU parity comparison.
"""

import math

import numpy as np
import pytest
import torch

from aquaflux_amd.mesh import structured_grid_2d
from oracle import fv
from oracle import sst as S

F64 = torch.float64
M = S.SSTModel()


def c(*v):
    return torch.tensor([float(x) for x in v], dtype=F64)


# --- SST sources: tests/unit/test_sst_sources.py ------------------------------------------------


def test_k_production_unlimited_and_limited():
    """test_sst_sources.py:49-64: nu_t S^2 below the cap; capped at 10 beta* k omega above it."""
    one = c(1.0, 1.0)
    assert torch.allclose(S.k_production(c(0.01, 0.01), one, one, one, M), c(0.01, 0.01), rtol=0, atol=1e-15)
    cap = 10.0 * M.beta_star
    assert torch.allclose(S.k_production(c(10.0, 10.0), one, one, one, M), c(cap, cap), rtol=0, atol=1e-15)


def test_k_destruction_is_a_negative_sink():
    """test_sst_sources.py:85-88."""
    got = S.k_destruction(c(2.0, 2.0), c(3.0, 3.0), M)
    assert torch.allclose(got, c(-M.beta_star * 6.0, -M.beta_star * 6.0), rtol=0, atol=1e-15)


def _near(cells, d, nu=1e-5, shear=900.0):
    n = len(cells)
    return (torch.tensor(cells), torch.full((n,), nu, dtype=F64), torch.full((n,), d, dtype=F64),
            torch.full((n,), shear, dtype=F64))


def test_near_wall_closure_touches_only_its_own_cells():
    """test_sst_sources.py:108-114."""
    cells, nu, d, shear = _near([0], d=3e-3)
    resolved, omega, k = c(0.5, 0.5), c(7.0, 7.0), c(0.3, 0.3)
    assert float(S.near_wall_production(resolved, k, cells, nu, d, shear, M)[1]) == 0.5
    assert float(S.near_wall_dissipation_rate(omega, k, cells, nu, d, M)[1]) == 7.0


def test_near_wall_production_resolved_in_the_sublayer_and_log_law_outside():
    """test_sst_sources.py:117-136: weight ~0 deep in the sublayer, ~1 in the log layer."""
    resolved = c(0.5, 0.5)
    cells, nu, d, shear = _near([0, 1], d=1e-5)
    got = S.near_wall_production(resolved, c(0.01, 0.01), cells, nu, d, shear, M)
    assert torch.allclose(got, resolved, rtol=1e-6, atol=0)
    cells, nu, d, shear = _near([0, 1], d=3e-3)
    k = c(0.3, 0.3)
    got = S.near_wall_production(resolved, k, cells, nu, d, shear, M)
    expected = S.k_wall_production(nu, d, k, shear, M)
    assert torch.allclose(got, expected, rtol=1e-6, atol=0)
    # and the weight really is a blend: the log-law value differs from the resolved one by far more
    assert float((expected - resolved).abs().min()) > 1e-2


def test_near_wall_dissipation_rate_is_the_wall_omega_at_the_live_k():
    """test_sst_sources.py:139-145."""
    cells, nu, d, _ = _near([0, 1], d=3e-3)
    k = c(0.3, 0.3)
    got = S.near_wall_dissipation_rate(c(7.0, 7.0), k, cells, nu, d, M)
    assert torch.allclose(got, S.omega_wall(nu, d, k, M), rtol=1e-15, atol=0)


def test_near_wall_destruction_grows_faster_than_linearly_in_k():
    """test_sst_sources.py:148-168: frozen omega gives exactly 4x for 4x k; the live wall omega >6x."""
    cells, nu, d, _ = _near([0, 1], d=3e-3)
    small, large = c(0.3, 0.3), c(1.2, 0.3)
    om = c(7.0, 7.0)

    def ratio(near):
        return float(S.k_destruction(large, om, M, near)[0] / S.k_destruction(small, om, M, near)[0])

    assert abs(ratio(None) - 4.0) < 1e-9
    assert ratio((cells, nu, d)) > 6.0


def test_k_production_with_near_wall_uses_the_live_k_in_the_wall_rows():
    """test_sst_sources.py:117-136 through ``k_production``: the near-wall blend applies after the cap."""
    cells, nu, d, shear = _near([0], d=3e-3)
    one = c(1.0, 1.0)
    k = c(0.3, 0.3)
    limited = S.k_production(c(0.01, 0.01), one, one, k, M)
    blended = S.k_production(c(0.01, 0.01), one, one, k, M, near_wall=(cells, nu, d, shear))
    assert float(blended[1]) == float(limited[1])
    w = S.wall_function_weight(nu, d, k[:1], M)
    want = (1.0 - w) * limited[:1] + w * S.k_wall_production(nu, d, k[:1], shear, M)
    assert torch.allclose(blended[:1], want, rtol=1e-15, atol=0)


def test_omega_production_blends_alpha_and_caps():
    """test_sst_sources.py:171-204: alpha_1 S^2 at F1 = 1, alpha_2 S^2 at F1 = 0 (cap inactive); the
    cap 10 beta* k omega / nu_t times alpha when S^2 exceeds it."""
    one = c(1.0, 1.0)
    s2 = c(2.0, 2.0)
    nu_t = c(0.01, 0.01)
    assert torch.allclose(S.omega_production(s2, nu_t, one, one, one, M), M.alpha_1 * 4.0 * one, rtol=1e-15, atol=0)
    assert torch.allclose(S.omega_production(s2, nu_t, one, one, 0 * one, M), M.alpha_2 * 4.0 * one, rtol=1e-15, atol=0)
    cap = 10.0 * M.beta_star
    assert torch.allclose(S.omega_production(c(3.0, 3.0), one, one, one, one, M), M.alpha_1 * cap * one, rtol=1e-15,
                          atol=0)


def test_omega_destruction_blends_beta():
    """test_sst_sources.py:217-220: F1 = 1 gives -beta_1 omega^2; F1 = 0 gives -beta_2 omega^2."""
    assert torch.allclose(S.omega_destruction(c(2.0, 2.0), c(1.0, 1.0), M), c(-4 * M.beta_1, -4 * M.beta_1))
    assert torch.allclose(S.omega_destruction(c(2.0), c(0.0), M), c(-4 * M.beta_2))


def test_omega_cross_diffusion_value_and_vanishes_at_wall():
    """test_sst_sources.py:223-237."""
    gk = torch.tensor([[1.0, 0.0], [1.0, 0.0]], dtype=F64)
    gw = torch.tensor([[2.0, 0.0], [2.0, 0.0]], dtype=F64)
    one = c(1.0, 1.0)
    outer = S.omega_cross_diffusion(one, gk, gw, 0 * one, M)
    assert torch.allclose(outer, 2.0 * M.sigma_omega2 * 2.0 * one, rtol=1e-15, atol=0)
    assert torch.allclose(S.omega_cross_diffusion(one, gk, gw, one, M), 0 * one, rtol=0, atol=0)


def test_omega_production_cap_on_a_negative_k():
    """test_sst_sources.py:256-284: finite and zero at k < 0 with a negative nu_t."""
    src = S.omega_production(c(2.0, 2.0), c(-1e-9, -1e-9), c(-2.9e-13, -2.9e-13), c(4e5, 4e5), c(1.0, 1.0), M)
    assert bool(torch.isfinite(src).all())
    assert torch.allclose(src, torch.zeros(2, dtype=F64), rtol=0, atol=0)


def test_omega_production_cap_unchanged_for_a_positive_k():
    """test_sst_sources.py:287-297: exactly alpha_1 min(S^2, 10 beta* k omega / nu_t)."""
    k, om, nu_t, s = c(1.0, 1.0), c(1.0, 1.0), c(0.01, 0.01), c(2.0, 2.0)
    got = S.omega_production(s, nu_t, k, om, c(1.0, 1.0), M)
    want = M.alpha_1 * torch.minimum(s**2, 10.0 * M.beta_star * k * om / nu_t)
    assert torch.equal(got, want)


def test_k_production_never_a_sink_and_unchanged_for_positive_k():
    """test_sst_sources.py:300-340: nu_t >= 0 and production >= 0 at a negative k; bitwise the
    unguarded closed form at k > 0 on both cap branches; eddy viscosity unchanged for k > 0."""
    neg = c(-2.9e-13, -2.9e-13)
    om, s, d, nu = c(4e5, 4e5), c(2.0, 2.0), c(1.0, 1.0), c(1e-5, 1e-5)
    nu_t = M.eddy_viscosity(neg, om, s, nu, d)
    assert bool((nu_t >= 0).all())
    src = S.k_production(nu_t, s, om, neg, M)
    assert bool(torch.isfinite(src).all()) and bool((src >= 0).all())
    om = c(1.0, 1.0)
    for k in (c(1.0, 1.0), c(1e-4, 1e-4)):
        nu_t = M.eddy_viscosity(k, om, s, nu, d)
        got = S.k_production(nu_t, s, om, k, M)
        assert torch.equal(got, torch.minimum(nu_t * s**2, 10.0 * M.beta_star * k * om))
    k, om, s, d, nu = c(0.81, 1.0), c(100.0, 1.0), c(2.0, 2.0), c(1.0, 1.0), c(0.06, 1e-5)
    want = M.a1 * k / torch.maximum(M.a1 * om, s * M.f2(k, om, nu, d))
    assert torch.equal(M.eddy_viscosity(k, om, s, nu, d), want)


# --- flow boundaries: tests/unit/test_flow_boundary.py -------------------------------------------

VEL = torch.tensor([[2.0, 1.0]], dtype=F64)
NORMAL = torch.tensor([[1.0, 0.0]], dtype=F64)
AREA = c(2.0)
DN = c(0.5)
GRADP = torch.tensor([[3.0, 0.0]], dtype=F64)
DCOEFF = torch.tensor([[0.4, 0.4]], dtype=F64)
P = c(1.5)


def test_momentum_diagonal_coefficient_per_patch():
    """test_flow_boundary.py:30-57: wall viscous only, outlet convective only, inlet both."""
    kinds = torch.tensor([fv.FLOW_WALL, fv.FLOW_OUTLET, fv.FLOW_INLET, fv.FLOW_NONE])
    got = fv.momentum_diagonal_coefficient(kinds, torch.full((4,), 7.0, dtype=F64), torch.full((4,), 5.0, dtype=F64))
    assert got.tolist() == [7.0, 5.0, 12.0, 0.0]


def test_inlet_and_outlet_mass_flux():
    """test_flow_boundary.py:63-91: inlet rho (u_in . n) A = 8; outlet rho (u.n - d_hat((p_b - p)/dn -
    grad p . n)) A."""
    assert abs(float(fv.inlet_mass_flux(c(1.0), torch.tensor([[4.0, 0.0]], dtype=F64), NORMAL, AREA)[0]) - 8.0) < 1e-12
    expected = 1.0 * (2.0 - 0.4 * ((0.0 - 1.5) / 0.5 - 3.0)) * 2.0
    got = float(fv.outlet_mass_flux(VEL, P, GRADP, DCOEFF, NORMAL, AREA, DN, c(1.0), c(0.0))[0])
    assert abs(got - expected) < 1e-12


def _single_face(p_owner=0.0, p_neighbour=0.0, grad=(0.0, 0.0), d_coeff=0.5):
    """test_rhie_chow.py:19-47: one interior face between cells at (0,0) and (1,0); centroid (0.5,0),
    n = x, A = 2, d.n = 1; velocities (3,1) and (5,1); zero velocity gradient."""
    m = fv.OMesh(node_coords=torch.zeros((2, 2), dtype=F64), owner=torch.tensor([0]), neighbour=torch.tensor([1]),
                 offsets=np.array([0, 2]), indices=np.array([0, 1]), n_cells=2, patches={})
    g = fv.OGeometry(area=c(2.0), face_centroid=torch.tensor([[0.5, 0.0]], dtype=F64), normal=NORMAL.clone(),
                     volume=c(1.0, 1.0), cell_centroid=torch.tensor([[0.0, 0.0], [1.0, 0.0]], dtype=F64))
    vel = torch.tensor([[3.0, 1.0], [5.0, 1.0]], dtype=F64)
    gp = torch.tensor([list(grad), list(grad)], dtype=F64)
    return fv.interior_mass_flux(m, g, vel, torch.zeros((2, 2, 2), dtype=F64), c(p_owner, p_neighbour), gp,
                                 torch.full((2, 2), d_coeff, dtype=F64), c(1.0, 1.0), c(0.5), c(1.0))


def test_interior_mass_flux_closed_forms():
    """test_rhie_chow.py:50-69: uniform pressure gives rho u_f . n A = 8; an adverse pressure jump
    reduces the flux; a compact jump equal to the reconstructed one cancels the correction."""
    base = float(_single_face()[0])
    assert abs(base - 8.0) < 1e-12
    adverse = float(_single_face(p_neighbour=1.0)[0])
    assert adverse < base
    assert abs(adverse - (8.0 - 1.0 * 0.5 * 1.0 * 2.0)) < 1e-12  # rho_f d_hat (dp / d.n) A = 1
    assert abs(float(_single_face(p_neighbour=1.0, grad=(1.0, 0.0))[0]) - base) < 1e-12


def _grid(n=4):
    mesh = structured_grid_2d(n, n, named_boundaries=True)
    m = fv.OMesh.from_mesh(mesh)
    return mesh, m, fv.geometry(m)


def _setup(m, kind, advection="none"):
    F, d = m.n_faces, m.dim
    return fv.FlowSetup(kind=kind, u_b=torch.zeros((F, d), dtype=F64), p_b=torch.zeros(F, dtype=F64),
                        advection=advection)


def test_momentum_diagonal_closed_form_on_a_uniform_grid():
    """test_rhie_chow.py:72-80: positive and linear in mu; on the uniform 4x4 unit grid with walls all
    round, a_P = mu (interior faces) + 2 mu (wall faces: A / (h/2)) per face."""
    mesh, m, g = _grid(4)
    kind = torch.where(m.interior, torch.tensor(fv.FLOW_NONE), torch.tensor(fv.FLOW_WALL))
    s = _setup(m, kind)
    mu = torch.full((m.n_cells,), 0.3, dtype=F64)
    a1 = fv.momentum_diagonal(m, g, s, mu, mu[m.owner], None)
    a2 = fv.momentum_diagonal(m, g, s, 2 * mu, 2 * mu[m.owner], None)
    n_int = fv.scatter(m, m.interior.to(F64), m.interior.to(F64))
    n_bnd = fv.scatter(m, (~m.interior).to(F64), torch.zeros(m.n_faces, dtype=F64))
    assert torch.allclose(a1, 0.3 * (n_int + 2.0 * n_bnd), rtol=1e-14, atol=0)
    assert bool((a1 > 0).all()) and torch.allclose(a2, 2 * a1, rtol=1e-15, atol=0)


def test_flow_fields_a_p_and_boundary_mdot_closed_form():
    """The a_P and mass flux that ``flow_fields`` hands the residual, on the 4x4 unit grid with a uniform
    flow u = (1, 0): inlet left, outlet right, walls top/bottom, upwind.  Each cell's a_P is its
    viscous diagonal plus the outflow max(mdot_est, 0) = h through its +x face (interior or outlet;
    the inlet adds no convective term since its mdot_est is inflow); wall mdot is 0, inlet mdot is
    -h, outlet mdot is +h (no pressure gradient)."""
    mesh, m, g = _grid(4)
    h = 0.25
    kind = torch.full((m.n_faces,), fv.FLOW_NONE)
    pk = {"left": fv.FLOW_INLET, "right": fv.FLOW_OUTLET, "bottom": fv.FLOW_WALL, "top": fv.FLOW_WALL}
    for name, k in pk.items():
        kind[m.patches[name]] = k
    s = _setup(m, kind, "upwind")
    s.u_b[m.patches["left"], 0] = 1.0
    C = m.n_cells
    state = torch.cat([torch.ones(C, dtype=F64), torch.zeros(C, dtype=F64), torch.zeros(C, dtype=F64)])
    mu = torch.full((C,), 0.3, dtype=F64)
    f = fv.flow_fields(m, g, s, state, mu, torch.ones(C, dtype=F64))
    n_int = fv.scatter(m, m.interior.to(F64), m.interior.to(F64))
    dirichlet = ((kind == fv.FLOW_WALL) | (kind == fv.FLOW_INLET)).to(F64)
    n_dir = fv.scatter(m, dirichlet, torch.zeros(m.n_faces, dtype=F64))
    assert torch.allclose(f["a_p"], 0.3 * (n_int + 2.0 * n_dir) + h, rtol=1e-14, atol=0)
    mdot = f["mdot"]
    assert float(mdot[m.patches["bottom"]].abs().max()) == 0.0 and float(mdot[m.patches["top"]].abs().max()) == 0.0
    assert torch.allclose(mdot[m.patches["left"]], torch.full((4,), -h, dtype=F64), rtol=1e-14, atol=0)
    assert torch.allclose(mdot[m.patches["right"]], torch.full((4,), h, dtype=F64), rtol=1e-14, atol=0)


# --- pressure force: tests/unit/test_pressure_force.py --------------------------------------------


@pytest.mark.parametrize("component", [0, 1])
def test_pressure_force_closed_form_and_no_net_uniform_force(component):
    """test_pressure_force.py:38-79: p_f n_i A face by face; a uniform face pressure scatters to zero;
    the derivative in the face pressure is n_i A."""
    mesh = structured_grid_2d(2, 2, named_boundaries=True)
    m = fv.OMesh.from_mesh(mesh)
    g = fv.geometry(m)
    pf = torch.linspace(0.5, 3.0, m.n_faces, dtype=F64)
    assert torch.equal(fv.pressure_force_flux(pf, g.normal, g.area, component), pf * g.normal[:, component] * g.area)
    uni = fv.pressure_force_flux(torch.full((m.n_faces,), 2.5, dtype=F64), g.normal, g.area, component)
    assert float(fv.scatter_conservative(m, uni).abs().max()) < 1e-12
    x = pf.clone().requires_grad_(True)
    (gx,) = torch.autograd.grad(fv.pressure_force_flux(x, g.normal, g.area, component).sum(), (x,))
    assert torch.allclose(gx, g.normal[:, component] * g.area, rtol=1e-15, atol=0)


def test_flow_residual_pressure_term_is_the_pressure_force():
    """The momentum rows of the whole flow residual at zero velocity (no-slip walls, so no viscous
    flux) are the scattered pressure force ``sum_f p_f n_i A`` with p_f the face pressure (interior: interpolated,
    boundary: the owner value), i.e. -grad p V for a linear pressure on interior cells."""
    mesh, m, g = _grid(4)
    kind = torch.where(m.interior, torch.tensor(fv.FLOW_NONE), torch.tensor(fv.FLOW_WALL))
    s = _setup(m, kind)
    C = m.n_cells
    x = g.cell_centroid
    p = 2.0 * x[:, 0] - 0.5 * x[:, 1]
    state = torch.cat([torch.zeros(C, dtype=F64), torch.zeros(C, dtype=F64), p])
    R = fv.flow_residual(m, g, s, state, torch.ones(C, dtype=F64), torch.ones(C, dtype=F64))
    interior_cells = fv.scatter(m, (~m.interior).to(F64), torch.zeros(m.n_faces, dtype=F64)) == 0
    V = g.volume
    assert torch.allclose(R[:C][interior_cells], (2.0 * V)[interior_cells], rtol=1e-12, atol=1e-14)
    assert torch.allclose(R[C:2 * C][interior_cells], (-0.5 * V)[interior_cells], rtol=1e-12, atol=1e-14)


# --- scalar boundary closures: tests/unit/test_boundary.py ----------------------------------------

N = torch.tensor([[1.0, 0.0]], dtype=F64)
D = torch.tensor([[0.5, 0.3]], dtype=F64)
GRAD = torch.tensor([[1.0, 4.0]], dtype=F64)
PHI = c(3.0)
GAMMA = c(2.0)
CORR = 1.2
DNORM = 0.5
D_ORTHO = torch.tensor([[0.5, 0.0]], dtype=F64)


def test_zero_gradient_closure():
    """test_boundary.py:46-52."""
    assert abs(float(fv.zero_gradient_face_value(PHI, GRAD, D, N)[0]) - (3.0 + CORR)) < 1e-14
    assert float(fv.zero_gradient_face_value(PHI, GRAD, D_ORTHO, N)[0]) == 3.0


def test_neumann_closure():
    """test_boundary.py:55-58: phi_P + corr - (q / Gamma_P)(d.n)."""
    got = float(fv.neumann_face_value(PHI, GRAD, D, N, GAMMA, c(1.5))[0])
    assert abs(got - (3.0 + CORR - 1.5 / 2.0 * DNORM)) < 1e-14


def test_convective_closure_robin_balance_and_high_biot():
    """test_boundary.py:61-80: closed form, Gamma dphi/dn = h (T_inf - phi_ip), and the Dirichlet limit."""
    h, t_inf = 4.0, 1.0
    beta = h / 2.0 * DNORM
    got = float(fv.convective_face_value(PHI, GRAD, D, N, GAMMA, c(h), c(t_inf))[0])
    assert abs(got - (3.0 + CORR + beta * t_inf) / (1.0 + beta)) < 1e-12
    dphi_dn = (got - 3.0 - CORR) / DNORM
    assert abs(2.0 * dphi_dn - h * (t_inf - got)) < 1e-12
    hi = float(fv.convective_face_value(PHI, GRAD, D_ORTHO, N, GAMMA, c(1e6), c(0.25))[0])
    assert abs(hi - 0.25) < 1e-4


def test_scalar_boundary_values_dispatch_each_closure():
    """``scalar_boundary_values`` applies each face's own closure: on a 3x2 grid with Dirichlet left,
    Neumann right, Convective top and ZeroGradient bottom, every boundary value equals its closure
    evaluated from the owner (each a closed form pinned above)."""
    mesh = structured_grid_2d(3, 2, named_boundaries=True)
    m = fv.OMesh.from_mesh(mesh)
    g = fv.geometry(m)
    F = m.n_faces
    kind = torch.zeros(F, dtype=torch.long)
    value = torch.zeros(F, dtype=F64)
    t_inf = torch.zeros(F, dtype=F64)
    for name, k, v in (("left", fv.BC_DIRICHLET, 2.0), ("right", fv.BC_NEUMANN, 0.7), ("top", fv.BC_CONVECTIVE, 3.0),
                       ("bottom", fv.BC_ZERO_GRADIENT, 0.0)):
        kind[m.patches[name]] = k
        value[m.patches[name]] = v
    t_inf[m.patches["top"]] = -1.0
    s = fv.ScalarSetup(kind=kind, value=value, t_inf=t_inf)
    rng = np.random.default_rng(0)
    phi = torch.tensor(rng.normal(size=m.n_cells), dtype=F64)
    grad = torch.tensor(rng.normal(size=(m.n_cells, 2)), dtype=F64)
    gamma = torch.tensor(rng.uniform(0.5, 2.0, size=m.n_cells), dtype=F64)
    bv = fv.scalar_boundary_values(m, g, s, phi, grad, gamma)
    own = m.owner
    dvec = g.face_centroid - g.cell_centroid[own]
    for f in range(F):
        P = int(own[f])
        dn = float(fv.dot(dvec[f:f + 1], g.normal[f:f + 1])[0])
        corr = float(fv.dot(grad[P:P + 1], dvec[f:f + 1] - dn * g.normal[f:f + 1])[0])
        k = int(kind[f])
        if k == fv.BC_DIRICHLET:
            want = float(value[f])
        elif k == fv.BC_ZERO_GRADIENT:
            want = float(phi[P]) + corr
        elif k == fv.BC_NEUMANN:
            want = float(phi[P]) + corr - float(value[f]) / float(gamma[P]) * dn
        elif k == fv.BC_CONVECTIVE:
            b = float(value[f]) / float(gamma[P]) * dn
            want = (float(phi[P]) + corr + b * float(t_inf[f])) / (1.0 + b)
        else:
            want = 0.0
        assert abs(float(bv[f]) - want) < 1e-13, (f, k)
