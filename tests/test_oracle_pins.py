"""Pins the CPU oracle against the reference's own closed-form unit-test facts (CPU only).

Each test cites the reference test it restates.  These run without a GPU.
"""

import numpy as np
import pytest
import torch

from aquaflux_amd.mesh import Mesh, perturbed_grid_2d, perturbed_grid_3d, structured_grid_2d, structured_grid_3d, tet_grid_3d
from oracle import fv
from oracle.connectivity import structured_2d, structured_3d

F64 = torch.float64


def _geom(mesh):
    m = fv.OMesh.from_mesh(mesh)
    return m, fv.geometry(m)


# --- geometry: tests/unit/test_mesh.py closed forms --------------------------------------------


def test_two_unit_squares():
    """test_mesh.py:276-305: unit edges, midpoints, +x shared normal, unit volumes/centroids."""
    nodes = [[0, 0], [1, 0], [2, 0], [0, 1], [1, 1], [2, 1]]
    faces = [[1, 4], [0, 1], [4, 3], [3, 0], [1, 2], [2, 5], [5, 4]]
    mesh = Mesh.from_faces(nodes, faces, [0, 0, 0, 0, 1, 1, 1], [1, -1, -1, -1, -1, -1, -1], n_cells=2)
    m, g = _geom(mesh)
    assert torch.allclose(g.area, torch.ones(7, dtype=F64))
    assert torch.allclose(g.face_centroid[0], torch.tensor([1.0, 0.5], dtype=F64))
    assert torch.allclose(g.normal[0], torch.tensor([1.0, 0.0], dtype=F64))
    assert torch.allclose(g.volume, torch.ones(2, dtype=F64))
    assert torch.allclose(g.cell_centroid, torch.tensor([[0.5, 0.5], [1.5, 0.5]], dtype=F64))


def test_triangle_and_parallelogram():
    """test_mesh.py:307-321."""
    tri = Mesh.from_faces([[0, 0], [1, 0], [0, 1]], [[0, 1], [1, 2], [2, 0]], [0, 0, 0], [-1, -1, -1], n_cells=1)
    _, g = _geom(tri)
    assert torch.allclose(g.volume, torch.tensor([0.5], dtype=F64))
    assert torch.allclose(g.cell_centroid, torch.tensor([[1 / 3, 1 / 3]], dtype=F64))
    par = Mesh.from_faces([[0, 0], [2, 0], [3, 1], [1, 1]], [[0, 1], [1, 2], [2, 3], [3, 0]], [0] * 4, [-1] * 4, n_cells=1)
    _, g = _geom(par)
    assert torch.allclose(g.volume, torch.tensor([2.0], dtype=F64))
    assert torch.allclose(g.cell_centroid, torch.tensor([[1.5, 0.5]], dtype=F64))


CUBE_NODES = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]]
CUBE_FACES = [[0, 1, 2, 3], [4, 5, 6, 7], [0, 1, 5, 4], [3, 2, 6, 7], [0, 3, 7, 4], [1, 2, 6, 5]]


def test_cube_tet_box():
    """test_mesh.py:324-362: unit cube, tetrahedron, 2x3x4 box."""
    cube = Mesh.from_faces(CUBE_NODES, CUBE_FACES, [0] * 6, [-1] * 6, n_cells=1)
    _, g = _geom(cube)
    assert torch.allclose(g.area, torch.ones(6, dtype=F64))
    assert torch.allclose(g.volume, torch.tensor([1.0], dtype=F64))
    assert torch.allclose(g.cell_centroid, torch.tensor([[0.5, 0.5, 0.5]], dtype=F64))
    out = g.face_centroid - g.cell_centroid[0]
    assert bool(torch.all((out * g.normal).sum(1) > 0))
    tet = Mesh.from_faces(
        [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]], [0] * 4, [-1] * 4, n_cells=1
    )
    _, g = _geom(tet)
    assert torch.allclose(g.volume, torch.tensor([1 / 6], dtype=F64))
    assert torch.allclose(g.cell_centroid, torch.tensor([[0.25, 0.25, 0.25]], dtype=F64))
    box = Mesh.from_faces(np.array(CUBE_NODES) * [2.0, 3.0, 4.0], CUBE_FACES, [0] * 6, [-1] * 6, n_cells=1)
    _, g = _geom(box)
    assert torch.allclose(g.volume, torch.tensor([24.0], dtype=F64))
    assert torch.allclose(g.cell_centroid, torch.tensor([[1.0, 1.5, 2.0]], dtype=F64))
    assert torch.allclose(torch.sort(g.area).values, torch.tensor([6.0, 6.0, 8.0, 8.0, 12.0, 12.0], dtype=F64))


def test_prism_mixed_faces():
    """test_mesh.py:375-394: triangular prism, volume 1/2, centroid (1/3, 1/3, 1/2)."""
    nodes = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]]
    faces = [[0, 1, 2], [3, 4, 5], [0, 1, 4, 3], [1, 2, 5, 4], [2, 0, 3, 5]]
    mesh = Mesh.from_faces(nodes, faces, [0] * 5, [-1] * 5, n_cells=1)
    _, g = _geom(mesh)
    assert torch.allclose(g.area[:2], torch.tensor([0.5, 0.5], dtype=F64))
    assert torch.allclose(g.volume, torch.tensor([0.5], dtype=F64))
    assert torch.allclose(g.cell_centroid, torch.tensor([[1 / 3, 1 / 3, 0.5]], dtype=F64))


def test_far_from_origin_centroid():
    """test_mesh.py:663-683: recentred accumulation keeps a tiny far cell's centroid accurate."""
    s = 1e-3
    off = np.array([1.7e6, -5.1e5, 3.2e5])
    mesh = Mesh.from_faces(np.array(CUBE_NODES) * s + off, CUBE_FACES, [0] * 6, [-1] * 6, n_cells=1)
    _, g = _geom(mesh)
    assert abs(float(g.volume[0]) - s**3) <= 1e-6 * s**3
    err = np.linalg.norm(g.cell_centroid[0].numpy() - (off + 0.5 * s)) / s
    assert err < 1e-3


def test_structured_volumes_sum_and_tets():
    """Volumes tile the domain; every face normal is unit; tet split volumes are h^3/6."""
    for mesh, total in (
        (perturbed_grid_2d(8, 5, 2.0, 1.0, perturb=0.2, seed=0), 2.0),
        (perturbed_grid_3d(4, 3, 3, perturb=0.2, seed=1), 1.0),
        (tet_grid_3d(3, perturb=0.15, seed=3), 1.0),
    ):
        _, g = _geom(mesh)
        assert abs(float(g.volume.sum()) - total) < 1e-12
        assert torch.allclose(g.normal.norm(dim=1), torch.ones(mesh.n_faces, dtype=F64))
        assert bool(torch.all(g.volume > 0))
    _, g = _geom(tet_grid_3d(2))
    assert torch.allclose(g.volume, torch.full((48,), 1 / 48, dtype=F64))


def test_structured_connectivity_bit_exact():
    """Vectorised generators == loop restatement of mesh/structured.py (faces, owner, nb, patches)."""
    m = structured_grid_2d(6, 4, named_boundaries=True)
    r, o, n, p = structured_2d(6, 4)
    assert np.array_equal(m.face_node_indices.reshape(-1, 2), np.array(r))
    assert np.array_equal(m.owner, o) and np.array_equal(m.neighbour, n)
    for k, v in p.items():
        assert np.array_equal(m.face_patches.indices(k), v)
    m = structured_grid_3d(4, 3, 2, named_boundaries=True)
    r, o, n, p = structured_3d(4, 3, 2)
    assert np.array_equal(m.face_node_indices.reshape(-1, 4), np.array(r))
    assert np.array_equal(m.owner, o) and np.array_equal(m.neighbour, n)
    for k, v in p.items():
        assert np.array_equal(m.face_patches.indices(k), v)
    # sizes of the survey's configurations (face counts of the scope table)
    assert structured_grid_2d(64, 64).n_faces == 8320
    assert structured_grid_2d(512, 256).n_faces == 262912


# --- fluxes: tests/unit/test_diffusion.py, test_advection.py closed forms ---------------------------


def _one_face(phi_o, phi_n, *, grad=(0.0, 0.0), interior=True, bval=0.0, gamma=(1.0, 1.0)):
    """test_diffusion.py:34-74: owner (0.5,0), neighbour (1.5,0), face (1,0), n=x, A=0.5."""
    nc = 2 if interior else 1
    m = fv.OMesh(
        node_coords=torch.zeros((2, 2), dtype=F64),
        owner=torch.tensor([0]),
        neighbour=torch.tensor([1 if interior else -1]),
        offsets=np.array([0, 2]),
        indices=np.array([0, 1]),
        n_cells=nc,
        patches={},
    )
    g = fv.OGeometry(
        area=torch.tensor([0.5], dtype=F64),
        face_centroid=torch.tensor([[1.0, 0.0]], dtype=F64),
        normal=torch.tensor([[1.0, 0.0]], dtype=F64),
        volume=torch.ones(nc, dtype=F64),
        cell_centroid=torch.tensor([[0.5, 0.0], [1.5, 0.0]], dtype=F64)[:nc],
    )
    phi = torch.tensor([phi_o, phi_n][:nc], dtype=F64)
    grad_t = torch.tensor([list(grad)] * nc, dtype=F64)
    gam = torch.tensor(list(gamma)[:nc], dtype=F64)
    return m, g, phi, grad_t, gam, torch.tensor([bval], dtype=F64)


def test_diffusion_closed_forms():
    m, g, phi, grad, gam, bv = _one_face(1.0, 3.0)
    assert abs(float(fv.diffusion_flux(m, g, phi, grad, gam, bv)[0]) + 0.5 * 2.0) < 1e-13  # :77-82
    m, g, phi, grad, gam, bv = _one_face(1.0, 0.0, interior=False, bval=2.0)
    assert abs(float(fv.diffusion_flux(m, g, phi, grad, gam, bv)[0]) + 1.0 * (2.0 - 1.0) / 0.5 * 0.5) < 1e-13  # :118-122
    m, g, phi, grad, gam, bv = _one_face(0.0, 0.0, gamma=(2.0, 8.0))
    cond = fv.conductance(m, g, gam)
    assert abs(float(cond[0]) - 2 * 2 * 8 / 10 * 0.5) < 1e-13  # :85-97 harmonic mean
    # conductance == d(owner flux)/d(phi_P)  (:100-115)
    phi = torch.tensor([1.0, 0.0], dtype=F64, requires_grad=True)
    flux = fv.diffusion_flux(m, g, phi, grad, gam, bv)[0]
    (dphi,) = torch.autograd.grad(flux, phi)
    assert abs(float(dphi[0]) - float(cond[0])) < 1e-13


def test_upwind_selection():
    """test_advection.py:26-60: owner value on outflow, neighbour on inflow, weak value at inlets."""
    m, g, phi, grad, gam, bv = _one_face(1.0, 3.0)
    assert float(fv.upwind_face_value(m, phi, torch.tensor([2.0], dtype=F64), bv)[0]) == 1.0
    assert float(fv.upwind_face_value(m, phi, torch.tensor([-2.0], dtype=F64), bv)[0]) == 3.0
    m, g, phi, grad, gam, bv = _one_face(1.0, 0.0, interior=False, bval=7.0)
    assert float(fv.upwind_face_value(m, phi, torch.tensor([-1.0], dtype=F64), bv)[0]) == 7.0
    assert float(fv.upwind_face_value(m, phi, torch.tensor([1.0], dtype=F64), bv)[0]) == 1.0


def test_limited_reconstruction_closed_form():
    """test_advection.py:63-100: phi_C + psi grad.(x_f - x_C) = 3 + 2 * 0.5 with psi = 1."""
    m, g, phi, grad, gam, bv = _one_face(3.0, 0.0, grad=(2.0, 0.0))
    psi = torch.ones(2, dtype=F64)
    v = fv.limited_face_value(m, g, phi, grad, psi, torch.tensor([1.0], dtype=F64), bv)
    assert abs(float(v[0]) - 4.0) < 1e-13


# --- gradients / limiter -------------------------------------------------------------------------


def _linear(g, c=(0.3, -1.2, 0.7)):
    x = g.cell_centroid
    return 1.5 + sum(c[j] * x[:, j] for j in range(x.shape[1]))


@pytest.mark.parametrize("scheme", ["gg", "lsq"])
def test_gradient_linear_exactness(scheme):
    """test_gradient.py:57-110: linear fields are reconstructed exactly (GG: orthogonal grid)."""
    mesh = structured_grid_2d(6, 5, 1.0, 1.0) if scheme == "gg" else perturbed_grid_2d(6, 5, perturb=0.2, seed=0)
    m, g = _geom(mesh)
    c = (0.3, -1.2)
    phi = _linear(g, c)
    bv = 1.5 + c[0] * g.face_centroid[:, 0] + c[1] * g.face_centroid[:, 1]
    grad = fv.compact_green_gauss(m, g, phi, bv) if scheme == "gg" else fv.least_squares_gradient(m, g, phi, bv)
    interior_cells = torch.ones(m.n_cells, dtype=torch.bool)
    if scheme == "gg":  # orthogonal uniform grid: exact in every cell
        pass
    assert torch.allclose(grad[interior_cells], torch.tensor(c, dtype=F64).expand(m.n_cells, 2), atol=1e-10)


def test_lsq_matches_bruteforce_lstsq():
    """LSQ (no reference counterpart) == numpy.linalg.lstsq per cell on a skewed 3D mesh."""
    mesh = perturbed_grid_3d(3, 3, 2, perturb=0.2, seed=4)
    m, g = _geom(mesh)
    rng = np.random.default_rng(0)
    phi = torch.tensor(rng.normal(size=m.n_cells), dtype=F64)
    bv = torch.tensor(rng.normal(size=m.n_faces), dtype=F64)
    grad = fv.least_squares_gradient(m, g, phi, bv).numpy()
    own, nb = mesh.owner, mesh.neighbour
    cc, fc = g.cell_centroid.numpy(), g.face_centroid.numpy()
    for c in range(m.n_cells):
        rows, rhs = [], []
        for f in np.nonzero((own == c) | (nb == c))[0]:
            o = nb[f] if own[f] == c else own[f]
            dx = (cc[o] if o >= 0 else fc[f]) - cc[c]
            dphi = (phi[o] if o >= 0 else bv[f]) - phi[c]
            w = 1.0 / np.sqrt(dx @ dx)
            rows.append(w * dx)
            rhs.append(w * float(dphi))
        ref = np.linalg.lstsq(np.array(rows), np.array(rhs), rcond=None)[0]
        assert np.allclose(grad[c], ref, atol=1e-10)


def _psi_24(field):
    """test_limiter.py:10-25: 24x24 grid, Green-Gauss gradient (orthogonal grid), K = 5."""
    mesh = structured_grid_2d(24, 24)
    m, g = _geom(mesh)
    phi, bv = field(g.cell_centroid), field(g.face_centroid)
    grad = fv.compact_green_gauss(m, g, phi, bv)
    psi = fv.venkatakrishnan(m, g, phi, grad, 5.0)
    bcells = set(mesh.owner[mesh.neighbour < 0].tolist())
    interior = torch.tensor([c not in bcells for c in range(m.n_cells)])
    return psi, interior, g.cell_centroid


def test_limiter_linear_field_is_unlimited():
    """test_limiter.py:31-34: psi > 0.999 in the interior on a linear field."""
    psi, interior, _ = _psi_24(lambda x: 2.0 * x[:, 0] - 3.0 * x[:, 1] + 1.0)
    assert float(psi[interior].min()) > 0.999


def test_limiter_activates_near_step():
    """test_limiter.py:37-53: psi < 0.9 near a step, > 0.99 far from it, and within [0, 1]."""
    step = lambda x: torch.where(x[:, 0] < 0.5, 1.0, 0.0).to(F64)  # noqa: E731
    psi, _, cc = _psi_24(step)
    near = (cc[:, 0] - 0.5).abs() < 0.1
    far = (cc[:, 0] - 0.5).abs() > 0.3
    assert float(psi[near].min()) < 0.9
    assert float(psi[far].min()) > 0.99
    assert float(psi.min()) >= 0.0 and float(psi.max()) <= 1.0 + 1e-12


# --- coupled flow: conservation and Rhie-Chow properties ---------------------------------------


def _cavity_setup(m, n_faces, top_faces):
    kind = torch.full((n_faces,), fv.FLOW_NONE, dtype=torch.long)
    kind[~m.interior] = fv.FLOW_WALL
    ub = torch.zeros((n_faces, m.dim), dtype=F64)
    ub[top_faces, 0] = 1.0
    return fv.FlowSetup(kind=kind, u_b=ub, p_b=torch.zeros(n_faces, dtype=F64), advection="upwind", pressure_pin=0)


def test_flow_continuity_sums_to_zero_on_closed_domain():
    """Sum of continuity rows (excluding the pin) == -(pinned cell's flux); mdot = 0 on walls."""
    mesh = structured_grid_2d(8, 8, named_boundaries=True)
    m, g = _geom(mesh)
    s = _cavity_setup(m, mesh.n_faces, torch.tensor(mesh.face_patches.indices("top")))
    s.pressure_pin = None
    rng = np.random.default_rng(0)
    state = torch.tensor(rng.uniform(-1, 1, 3 * m.n_cells), dtype=F64)
    mu = torch.full((m.n_cells,), 0.01, dtype=F64)
    rho = torch.ones(m.n_cells, dtype=F64)
    R = fv.flow_residual(m, g, s, state, mu, rho)
    assert abs(float(R[2 * m.n_cells:].sum())) < 1e-12


def test_rhie_chow_cancels_on_linear_pressure():
    """test_rhie_chow.py:19-69: with a linear p and uniform u, mdot = rho u.n A on interior faces."""
    mesh = perturbed_grid_2d(6, 6, perturb=0.2, seed=2, named_boundaries=True)
    m, g = _geom(mesh)
    s = _cavity_setup(m, mesh.n_faces, torch.tensor([], dtype=torch.long))
    s.pressure_pin = None
    s.gradient = "least_squares"
    C = m.n_cells
    x = g.cell_centroid
    u = torch.stack([torch.full((C,), 0.7, dtype=F64), torch.full((C,), -0.2, dtype=F64)], 1)
    p = 0.4 * x[:, 0] - 1.1 * x[:, 1]
    # impose the exact linear boundary pressure through an outlet-like closure on every face
    s.kind[~m.interior] = fv.FLOW_OUTLET
    s.p_b = 0.4 * g.face_centroid[:, 0] - 1.1 * g.face_centroid[:, 1]
    state = torch.cat([u[:, 0], u[:, 1], p])
    f = fv.flow_fields(m, g, s, state, torch.full((C,), 0.01, dtype=F64), torch.ones(C, dtype=F64))
    ref = (u[m.owner] * g.normal).sum(1) * g.area
    assert torch.allclose(f["mdot"][m.interior], ref[m.interior], atol=1e-12)


def test_geometry_gradient_pins():
    """tests/unit/test_mesh.py:730-748 (d total volume / d node vs central FD), plus invariants the
    node-coordinate adjoint must honour: the total volume does not change when an interior node
    moves, and a random functional of all five geometry arrays matches FD on every coordinate of a
    small jittered tet mesh."""
    from aquaflux_amd.mesh import structured_grid_3d, tet_grid_3d

    mesh = structured_grid_3d(2, 3, 4, lx=2.0, ly=3.0, lz=4.0)
    X = torch.tensor(mesh.node_coords, dtype=F64, requires_grad=True)
    total = fv.geometry(fv.OMesh.from_mesh(mesh, node_coords=X)).volume.sum()
    (gx,) = torch.autograd.grad(total, X)

    def vol(Xn):
        return float(fv.geometry(fv.OMesh.from_mesh(mesh, node_coords=Xn)).volume.sum())

    h = 1e-6
    Xp, Xm = X.detach().clone(), X.detach().clone()
    Xp[6, 0] += h
    Xm[6, 0] -= h
    assert abs(float(gx[6, 0]) - (vol(Xp) - vol(Xm)) / (2 * h)) < 1e-5
    tm = tet_grid_3d(2, perturb=0.15, seed=3)
    Xt = torch.tensor(tm.node_coords, dtype=F64, requires_grad=True)
    g = fv.geometry(fv.OMesh.from_mesh(tm, node_coords=Xt))
    (gt,) = torch.autograd.grad(g.volume.sum(), Xt)
    on_b = np.zeros(tm.n_nodes, dtype=bool)
    for f in np.nonzero(tm.neighbour < 0)[0]:
        on_b[tm.face_node_indices[tm.face_node_offsets[f]:tm.face_node_offsets[f + 1]]] = True
    assert float(gt[torch.from_numpy(~on_b)].abs().max()) < 1e-13
    rng = np.random.default_rng(0)
    arrays = (g.area, g.normal, g.face_centroid, g.volume, g.cell_centroid)
    w = [torch.tensor(rng.normal(size=a.shape)) for a in arrays]

    def functional(Xn):
        gg = fv.geometry(fv.OMesh.from_mesh(tm, node_coords=Xn))
        return sum((a * b).sum() for a, b in zip((gg.area, gg.normal, gg.face_centroid, gg.volume,
                                                  gg.cell_centroid), w))

    (ga,) = torch.autograd.grad(functional(Xt), Xt)
    Xd = Xt.detach()
    for node in range(0, tm.n_nodes, 5):
        for ax in range(3):
            e = torch.zeros_like(Xd)
            e[node, ax] = 1.0
            fd = (4 * (functional(Xd + 5e-6 * e) - functional(Xd - 5e-6 * e)) / 1e-5
                  - (functional(Xd + 1e-5 * e) - functional(Xd - 1e-5 * e)) / 2e-5) / 3
            assert abs(float(fd) - float(ga[node, ax])) < 1e-7 * max(1.0, abs(float(ga[node, ax])))


def test_scalar_residual_jvp_matches_richardson_fd():
    """The oracle's scalar-transport residual tangent (phi, diffusivity, flux) against Richardson-
    extrapolated central differences on a jittered tet mesh with limited advection."""
    from aquaflux_amd.mesh import tet_grid_3d

    mesh = tet_grid_3d(3, perturb=0.15, seed=3)
    m = fv.OMesh.from_mesh(mesh)
    g = fv.geometry(m)
    F, C = mesh.n_faces, mesh.n_cells
    kind = torch.where(m.interior, 0, fv.BC_ZERO_GRADIENT)
    left = torch.tensor(mesh.face_patches.indices("left"), dtype=torch.long)
    kind[left] = fv.BC_DIRICHLET
    value = torch.zeros(F, dtype=F64)
    value[left] = 1.0
    s = fv.ScalarSetup(kind=kind, value=value, t_inf=torch.zeros(F, dtype=F64), advection="limited",
                       gradient="green_gauss")
    rng = np.random.default_rng(1)
    phi = torch.tensor(rng.uniform(size=C))
    gam = torch.full((C,), 0.01, dtype=F64)
    flux = g.normal[:, 0] * g.area
    # no flux tangent where the flux is exactly zero (wall faces parallel to the flow): the upwind
    # switch is not differentiable there
    tan = (torch.tensor(rng.normal(size=C)), torch.tensor(rng.normal(size=C)) * 1e-3,
           torch.tensor(rng.normal(size=F)) * 1e-2 * (flux.abs() > 1e-12))

    def f(p, ga, fl):
        return fv.scalar_residual(m, g, s, p, fl, ga)

    _, jv = torch.func.jvp(f, (phi, gam, flux), tan)

    def cd(e):
        return (f(phi + e * tan[0], gam + e * tan[1], flux + e * tan[2])
                - f(phi - e * tan[0], gam - e * tan[1], flux - e * tan[2])) / (2 * e)

    rich = (4 * cd(5e-6) - cd(1e-5)) / 3
    assert float((jv - rich).abs().max() / jv.abs().max()) < 1e-8


def test_corrected_green_gauss_pins():
    """tests/unit/test_gradient.py:163-229: the swept corrected gradient equals compact Green-Gauss on
    an orthogonal grid and after one sweep anywhere; with enough sweeps it is linear-exact on an
    irregular grid (where compact is not) and equals the direct solve of A_g G = B phi assembled
    densely by probing the operator."""
    from aquaflux_amd.mesh import perturbed_grid_2d, structured_grid_2d

    def lin(x):
        return 1.5 * x[:, 0] - 0.7 * x[:, 1] + 0.3

    orth = structured_grid_2d(12, 12)
    mo = fv.OMesh.from_mesh(orth)
    go = fv.geometry(mo)
    rng = np.random.default_rng(0)
    phi = torch.tensor(rng.normal(size=orth.n_cells))
    bv = torch.tensor(rng.normal(size=orth.n_faces))
    cg = fv.corrected_green_gauss(mo, go, phi, bv, sweeps=4)
    assert float((cg - fv.compact_green_gauss(mo, go, phi, bv)).abs().max()) < 1e-12
    pm = perturbed_grid_2d(16, 16, perturb=0.2, seed=0)
    m = fv.OMesh.from_mesh(pm)
    g = fv.geometry(m)
    phi = lin(g.cell_centroid)
    bv = lin(g.face_centroid)
    assert torch.equal(fv.corrected_green_gauss(m, g, phi, bv, sweeps=1), fv.compact_green_gauss(m, g, phi, bv))
    exact = torch.tensor([1.5, -0.7], dtype=F64)
    interior = torch.ones(pm.n_cells, dtype=torch.bool)
    interior[m.owner[~m.interior]] = False
    err_c = float((fv.compact_green_gauss(m, g, phi, bv)[interior] - exact).abs().max())
    err_s = float((fv.corrected_green_gauss(m, g, phi, bv, sweeps=40)[interior] - exact).abs().max())
    assert err_c > 1e-3 and err_s < 1e-10
    # direct solve on a small irregular grid
    sm = perturbed_grid_2d(5, 4, perturb=0.2, seed=2)
    ms = fv.OMesh.from_mesh(sm)
    gs = fv.geometry(ms)
    C = sm.n_cells
    cols = []
    for i in range(2 * C):
        e = torch.zeros(2 * C, dtype=F64)
        e[i] = 1.0
        cols.append(fv.corrected_operator(ms, gs, e.reshape(C, 2)).reshape(-1))
    A = torch.stack(cols, dim=1)
    phis = torch.tensor(rng.normal(size=C))
    bvs = torch.tensor(rng.normal(size=sm.n_faces))
    direct = torch.linalg.solve(A, fv.corrected_rhs(ms, gs, phis, bvs).reshape(-1)).reshape(C, 2)
    swept = fv.corrected_green_gauss(ms, gs, phis, bvs, sweeps=60)
    assert float((swept - direct).abs().max() / direct.abs().max()) < 1e-12


# --- transient accumulation: tests/unit/test_transient.py -------------------------------------


def test_transient_term_known_answers():
    """test_transient.py:9-43 inputs: V = (2, 4), phi = (1, 2), phi_old = (0.5, 1), phi_older =
    (0, 0.5), dt = 0.1 give BDF1 (10, 40) and BDF2 (10, 50); BDF1 ignores phi_older; a constant
    history vanishes."""
    V = torch.tensor([2.0, 4.0], dtype=F64)
    phi, old, older = (torch.tensor(x, dtype=F64) for x in ([1.0, 2.0], [0.5, 1.0], [0.0, 0.5]))
    assert torch.allclose(fv.transient_term(phi, old, older, 0.1, True, V), torch.tensor([10.0, 40.0], dtype=F64))
    assert torch.allclose(fv.transient_term(phi, old, older, 0.1, False, V), torch.tensor([10.0, 50.0], dtype=F64))
    assert torch.equal(fv.transient_term(phi, old, older, 0.1, True, V),
                       fv.transient_term(phi, old, 99.0 * older, 0.1, True, V))
    c = torch.full((2,), 3.0, dtype=F64)
    assert torch.all(fv.transient_term(c, c, c, 0.1, False, V) == 0.0)


def _order(errors):
    return min(float(np.log2(errors[i] / errors[i + 1])) for i in range(len(errors) - 1))


def test_transient_term_orders_of_accuracy():
    """test_transient.py:46-101: the stencils estimate d/dt exp(t) at t = 0 to first (BDF1) and
    second (BDF2) order; a BDF2 branch replaced by BDF1 fails the second bound."""
    one = torch.ones(1, dtype=F64)

    def err(dt, first):
        r = fv.transient_term(one, np.exp(-dt) * one, np.exp(-2 * dt) * one, dt, first, one)
        return abs(float(r[0]) - 1.0)

    dts = [0.1, 0.05, 0.025, 0.0125]
    assert 0.9 < _order([err(d, True) for d in dts]) < 1.1
    assert _order([err(d, False) for d in dts]) > 1.9


def test_transient_decay_march_is_second_order():
    """test_transient.py:104-134: dphi/dt = -2 phi over [0, 1] marched BDF1 then BDF2 (each step
    linear in phi_new, solved from the stencil's own slope and offset) converges at second order."""
    rate, one = 2.0, torch.ones(1, dtype=F64)

    def march(n):
        dt = 1.0 / n
        phi, older = one.clone(), one.clone()
        for k in range(n):
            first = k == 0
            rest = fv.transient_term(0 * one, phi, older, dt, first, one)
            slope = fv.transient_term(one, phi, older, dt, first, one) - rest
            phi, older = -rest / (slope + rate), phi
        return float(phi[0])

    errors = [abs(march(n) - np.exp(-rate)) for n in (10, 20, 40, 80)]
    assert _order(errors) > 1.9


def test_scalar_transient_residual_adds_accumulation():
    """residual.py:141-147: R_transient = R_steady + V (BDF stencil) / dt on a perturbed quad mesh,
    and the accumulation cotangent with respect to phi is lambda V c0 / dt."""
    from tests import scalar_cases as SC
    mesh, u, bcs = SC.make("quad_mixed")
    m, g = _geom(mesh)
    s = SC.oracle_setup(mesh, bcs, "limited", "gg")
    phi, gam = (torch.tensor(x) for x in SC.seeded(mesh, 0))
    old, older = (torch.tensor(x) for x in SC.history(mesh, 0))
    flux = SC.oracle_flux(g, u)
    steady = fv.scalar_residual(m, g, s, phi, flux, gam)
    for first, c0 in ((True, 1.0), (False, 1.5)):
        R = fv.scalar_residual(m, g, s, phi, flux, gam, old, older, 0.25, first)
        acc = (R - steady) * 0.25 / g.volume
        want = phi - old if first else 1.5 * phi - 2.0 * old + 0.5 * older
        assert torch.allclose(acc, want, rtol=0, atol=1e-12)
        lam = torch.randn(mesh.n_cells, dtype=F64, generator=torch.Generator().manual_seed(1))
        p = phi.clone().requires_grad_(True)
        (gs,) = torch.autograd.grad(fv.scalar_residual(m, g, s, p, flux, gam), p, lam)
        p = phi.clone().requires_grad_(True)
        (gt,) = torch.autograd.grad(fv.scalar_residual(m, g, s, p, flux, gam, old, older, 0.25, first), p, lam)
        assert torch.allclose(gt - gs, lam * g.volume * c0 / 0.25, rtol=0, atol=1e-12)
