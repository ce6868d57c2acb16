"""Host logic of the SIMPLE-family block preconditioner (CPU): the three compositions against what
exact inner solves make of them, the synchronisation-free conjugate gradient, and the assembled
pressure Schur against the oracle, which is itself pinned by closed forms on a uniform grid."""

import types

import numpy as np
import pytest
import torch

from aquaflux_amd import block_preconditioner as BP
from aquaflux_amd.boundary import BoundaryConditions
from aquaflux_amd.flow import NoSlipWall, PressureOutlet, VelocityInlet
from aquaflux_amd.geometry import CellGeometry, FaceGeometry
from aquaflux_amd.mesh import perturbed_grid_2d, perturbed_grid_3d, structured_grid_2d
from oracle import fv

F64 = torch.float64


def _saddle(n=7, d=2, seed=0):
    """A random saddle ``[[F, G], [B, C]]`` with a diagonal positive F."""
    g = torch.Generator().manual_seed(seed)
    Fd = 1.0 + torch.rand(n * d, generator=g, dtype=F64)
    G = torch.randn(n * d, n, generator=g, dtype=F64)
    B = torch.randn(n, n * d, generator=g, dtype=F64)
    Cp = 0.3 * torch.randn(n, n, generator=g, dtype=F64)
    J = torch.zeros(n * (d + 1), n * (d + 1), dtype=F64)
    J[: n * d, : n * d] = torch.diag(Fd)
    J[: n * d, n * d:] = G
    J[n * d:, : n * d] = B
    J[n * d:, n * d:] = Cp
    return J, Fd, G, B, Cp


def _isotropic_saddle(n=7, d=2, seed=0):
    """A random saddle whose velocity block is ``a_P I`` per cell (isotropic, as the SIMPLE Schur
    assumes), so F~^-1 = 1 / a_P exactly."""
    J, _, G, B, Cp = _saddle(n, d, seed)
    g = torch.Generator().manual_seed(seed + 100)
    aP = 1.0 + torch.rand(n, generator=g, dtype=F64)
    J[: n * d, : n * d] = torch.diag(aP.repeat_interleave(d))
    return J, aP, G, B, Cp


def _composition_matrix(composition, n=7, d=2, seed=0):
    J, aP, G, B, Cp = _isotropic_saddle(n, d, seed)
    S = Cp - B @ torch.diag((1.0 / aP).repeat_interleave(d)) @ G
    Sinv = torch.linalg.inv(S)
    fn = BP.compose(composition, lambda ru: ru / aP[:, None], lambda rp: Sinv @ rp, lambda du: B @ du.reshape(-1),
                    lambda dp: (G @ dp).view(n, d), lambda dp: ((G @ dp).view(n, d), Cp @ dp), 1.0 / aP)
    N = n * (d + 1)
    M = torch.zeros(N, N, dtype=F64)
    for k in range(N):
        e = torch.zeros(N, dtype=F64)
        e[k] = 1.0
        du, dp = fn(e[: n * d].view(n, d), e[n * d:])
        M[:, k] = torch.cat([du.reshape(-1), dp])
    return M, J


@pytest.mark.parametrize("composition", ["simple", "simpler"])
def test_exact_inner_solves_invert_the_saddle(composition):
    """With F~ = F and the exact Schur, SIMPLE and SIMPLER are the block LU solve: M J = I."""
    M, J = _composition_matrix(composition)
    I = torch.eye(J.shape[0], dtype=F64)
    assert torch.allclose(M @ J, I, atol=1e-11)


def test_triangular_is_unipotent():
    """The block-triangular pass with exact inner solves leaves M J = [[I, F^-1 G], [0, I]]: not
    the identity, but (M J - I)^2 = 0, so a Krylov method finishes in two iterations."""
    M, J = _composition_matrix("triangular")
    E = M @ J - torch.eye(J.shape[0], dtype=F64)
    assert float(E.abs().max()) > 1e-3
    assert float((E @ E).abs().max()) < 1e-11


def test_triangular_and_simple_differ_only_by_the_velocity_update():
    """Dropping SIMPLE's closing update is the only difference: their pressures agree exactly."""
    Mt, J = _composition_matrix("triangular", seed=3)
    Ms, _ = _composition_matrix("simple", seed=3)
    n, d = 7, 2
    assert torch.allclose(Mt[n * d:], Ms[n * d:], atol=1e-13)
    assert float((Mt[: n * d] - Ms[: n * d]).abs().max()) > 1e-3


def test_unknown_composition_is_refused():
    with pytest.raises(ValueError, match="composition"):
        BP.compose("uzawa", None, None, None, None, None, torch.ones(1))


def test_conjugate_gradient_solves_spd():
    g = torch.Generator().manual_seed(1)
    A0 = torch.randn(20, 20, generator=g, dtype=F64)
    A = A0 @ A0.T + 20.0 * torch.eye(20, dtype=F64)
    b = torch.randn(20, generator=g, dtype=F64)
    x = BP.conjugate_gradient(lambda v: A @ v, b, 1.0 / torch.diagonal(A), 40)
    assert float(torch.linalg.vector_norm(A @ x - b)) < 1e-10 * float(torch.linalg.vector_norm(b))
    z = BP.conjugate_gradient(lambda v: A @ v, torch.zeros(20, dtype=F64), 1.0 / torch.diagonal(A), 5)
    assert torch.equal(z, torch.zeros(20, dtype=F64))


# --- the assembled Schur -----------------------------------------------------------------------


def _fake_op(mesh, outlet=True, pin=None):
    m = fv.OMesh.from_mesh(mesh)
    g = fv.geometry(m)
    d = mesh.dim
    bcs = {"left": VelocityInlet((1.0,) + (0.0,) * (d - 1)), "bottom": NoSlipWall(), "top": NoSlipWall(),
           "right": PressureOutlet(0.0) if outlet else NoSlipWall()}
    if d == 3:
        bcs.update({"back": NoSlipWall(), "front": NoSlipWall()})
    bound = BoundaryConditions(bcs).resolve(mesh.face_patches)
    geo = types.SimpleNamespace(
        face=FaceGeometry(area=g.area, centroid_soa=g.face_centroid.T.contiguous(), normal_soa=g.normal.T.contiguous()),
        cell=CellGeometry(volume=g.volume, centroid_soa=g.cell_centroid.T.contiguous()))
    op = types.SimpleNamespace(mesh=mesh, geometry=geo, boundary=bound, pressure_pin=pin)
    outlet_faces = mesh.face_patches.indices("right") if outlet else ()
    return op, m, g, outlet_faces


def _dense(graph, vals):
    C = graph.n_cells
    S = torch.zeros(C, C, dtype=F64)
    cols = graph.cols.long()
    for s in range(graph.nslot):
        S.index_put_((torch.arange(C), cols[s]), vals[s].to(F64), accumulate=True)
    return S


@pytest.mark.parametrize("case", ["perturbed2d", "perturbed3d", "pinned"])
def test_schur_laplacian_matches_oracle(case):
    if case == "perturbed3d":
        mesh = perturbed_grid_3d(5, 4, 3, lx=2.0, perturb=0.15, seed=1, named_boundaries=True)
    else:
        mesh = perturbed_grid_2d(9, 6, lx=3.0, perturb=0.2, seed=2, named_boundaries=True)
    pin = 5 if case == "pinned" else None
    op, m, g, outlet = _fake_op(mesh, outlet=case != "pinned", pin=pin)
    gen = torch.Generator().manual_seed(4)
    aP = 0.5 + torch.rand(mesh.n_cells, generator=gen, dtype=F64)
    rho = 0.8 + 0.4 * torch.rand(mesh.n_cells, generator=gen, dtype=F64)
    graph = BP.SchurGraph.of(op)
    got = _dense(graph, graph.laplacian(graph.volume / aP, rho, pin))
    ref = fv.pressure_schur_matrix(m, g, aP, rho, outlet, pin)
    assert float((got - ref).abs().max()) <= 1e-13 * float(ref.abs().max())
    assert torch.allclose(got, got.T, atol=1e-14)


def test_schur_oracle_closed_form_on_uniform_grid():
    """On a uniform nx x ny grid with constant a_P and rho the face coefficient is
    rho (V/a_P) A / h: the five-point Laplacian, symmetric, zero row sums on cells without an
    outlet face, a linear field annihilated on interior rows, and the outlet stiffness
    rho (V/a_P) A / (h/2) on the outlet column."""
    nx, ny, lx, ly = 6, 4, 3.0, 1.0
    hx, hy = lx / nx, ly / ny
    mesh = structured_grid_2d(nx, ny, lx, ly, named_boundaries=True)
    m = fv.OMesh.from_mesh(mesh)
    g = fv.geometry(m)
    aP, rho = 2.0, 1.5
    C = nx * ny
    S = fv.pressure_schur_matrix(m, g, torch.full((C,), aP, dtype=F64), torch.full((C,), rho, dtype=F64),
                                 mesh.face_patches.indices("right"))
    V = hx * hy
    cx, cy = rho * V / aP * hy / hx, rho * V / aP * hx / hy
    idx = lambda i, j: j * nx + i  # noqa: E731  (x-fastest cell numbering of the generator)
    for j in range(ny):
        for i in range(nx):
            c = idx(i, j)
            if i + 1 < nx:
                assert abs(float(S[c, idx(i + 1, j)]) + cx) < 1e-13
            if j + 1 < ny:
                assert abs(float(S[c, idx(i, j + 1)]) + cy) < 1e-13
            row = float(S[c].sum())
            expect = rho * V / aP * hy / (hx / 2) if i == nx - 1 else 0.0
            assert abs(row - expect) < 1e-12
    x = g.cell_centroid[:, 0]
    Lx = S @ x
    for j in range(1, ny - 1):
        for i in range(1, nx - 1):
            assert abs(float(Lx[idx(i, j)])) < 1e-12
    assert torch.allclose(S, S.T, atol=1e-15)


def test_build_refuses_unknown_options():
    op = types.SimpleNamespace(pressure_pin=None, fields=None)
    with pytest.raises(ValueError, match="velocity"):
        BP.BlockPreconditioner.build(op, velocity="smoothed")
    with pytest.raises(ValueError, match="schur_scaling"):
        BP.BlockPreconditioner.build(op, schur_scaling="lsc")
    with pytest.raises(ValueError, match="composition"):
        BP.BlockPreconditioner.build(op, composition="uzawa")


# --- the aggregation multigrid -----------------------------------------------------------------


def _torch_spmv(H):
    return lambda k, vals, x: (vals * x[H.levels[k].cols.long()]).sum(0)


def _grid_schur(nx, ny, pin=None):
    mesh = structured_grid_2d(nx, ny, 2.0, 1.0, named_boundaries=True)
    op, m, g, outlet = _fake_op(mesh, outlet=pin is None, pin=pin)
    graph = BP.SchurGraph.of(op)
    C = mesh.n_cells
    vals = graph.laplacian(graph.volume / torch.full((C,), 2.0, dtype=F64), torch.ones(C, dtype=F64), pin)
    H = BP.AggregationHierarchy.build(graph.cols.numpy(), g.cell_centroid.numpy(), pin, coarsest=12)
    return mesh, graph, vals, H


def test_aggregates_are_two_by_two_blocks_on_a_regular_grid():
    nx, ny = 16, 12
    _, _, _, H = _grid_schur(nx, ny)
    agg = H.levels[0].agg.numpy()
    assert H.levels[1].n == (nx // 2) * (ny // 2)
    for a in range(H.levels[1].n):
        cells = np.flatnonzero(agg == a)
        i, j = cells % nx, cells // nx
        assert cells.size == 4 and i.max() - i.min() == 1 and j.max() - j.min() == 1 and i.min() % 2 == 0


def test_coarse_operators_are_galerkin_products():
    _, graph, vals, H = _grid_schur(16, 12)
    ops, cinv = H.operators(vals)
    A = _dense(graph, vals)
    for k in range(len(H.levels) - 1):
        lv, nx_ = H.levels[k], H.levels[k + 1]
        P = torch.zeros(lv.n, nx_.n, dtype=F64)
        P[torch.arange(lv.n), lv.agg] = 1.0
        Ac = torch.zeros(nx_.n, nx_.n, dtype=F64)
        Ac.index_put_((torch.arange(nx_.n).repeat(nx_.nslot), nx_.cols.long().reshape(-1)), ops[k + 1].reshape(-1),
                      accumulate=True)
        assert torch.allclose(Ac, P.T @ A @ P, atol=1e-13)
        A = Ac
    assert torch.allclose(cinv @ A, torch.eye(A.shape[0], dtype=F64), atol=1e-9)


def test_pinned_cell_stays_a_singleton():
    _, _, _, H = _grid_schur(16, 12, pin=37)
    p = 37
    for lv in H.levels[:-1]:
        agg = lv.agg.numpy()
        assert int((agg == agg[p]).sum()) == 1
        p = int(agg[p])


def test_multigrid_cg_beats_jacobi_cg():
    """On a 64 x 32 Laplacian with an outlet, 12 V-cycle-preconditioned CG steps reach 1e-8; as
    many Jacobi-preconditioned steps do not reach 1e-2."""
    _, graph, vals, H = _grid_schur(64, 32)
    A = _dense(graph, vals)
    b = torch.randn(A.shape[0], generator=torch.Generator().manual_seed(5), dtype=F64)
    ops, cinv = H.operators(vals)
    spmv = _torch_spmv(H)
    x = BP.conjugate_gradient(lambda v: A @ v, b, lambda r: H.cycle(spmv, ops, cinv, r), 12)
    rel = float(torch.linalg.vector_norm(A @ x - b) / torch.linalg.vector_norm(b))
    xj = BP.conjugate_gradient(lambda v: A @ v, b, 1.0 / torch.diagonal(A), 12)
    relj = float(torch.linalg.vector_norm(A @ xj - b) / torch.linalg.vector_norm(b))
    assert rel < 1e-8 and relj > 1e-2, (rel, relj)


def test_v_cycle_is_symmetric():
    """CG needs a symmetric preconditioner: <u, M v> = <M u, v>."""
    _, _, vals, H = _grid_schur(16, 12)
    ops, cinv = H.operators(vals)
    spmv = _torch_spmv(H)
    g = torch.Generator().manual_seed(6)
    u, v = torch.randn(192, generator=g, dtype=F64), torch.randn(192, generator=g, dtype=F64)
    a = float(torch.dot(u, H.cycle(spmv, ops, cinv, v)))
    b = float(torch.dot(H.cycle(spmv, ops, cinv, u), v))
    assert abs(a - b) < 1e-12 * max(abs(a), 1.0)


def test_member_tables_list_each_aggregate():
    _, _, _, H = _grid_schur(16, 12, pin=37)
    for lv in H.levels[:-1]:
        agg = lv.agg.numpy()
        mem = lv.members.numpy()
        assert np.array_equal(lv.agg32.numpy(), agg)
        for a in range(mem.shape[1]):
            got = sorted(int(m) for m in mem[:, a] if m >= 0)
            assert got == sorted(np.flatnonzero(agg == a).tolist())


def test_multigrid_on_jittered_tetrahedra():
    """Geometric aggregation on an unstructured mesh (10 x 8 x 8 Kuhn tetrahedra, nodes jittered):
    the hierarchy coarsens by at least 4 per level, its coarse operators stay Galerkin products, and
    10 multigrid-preconditioned CG steps beat 10 Jacobi-preconditioned ones by 100x in residual."""
    from aquaflux_amd.mesh import tet_grid_3d

    mesh = tet_grid_3d(10, 8, 8, perturb=0.15, seed=3)
    op, m, g, outlet = _fake_op(mesh, outlet=True)
    graph = BP.SchurGraph.of(op)
    C = mesh.n_cells
    aP = 1.0 + torch.rand(C, generator=torch.Generator().manual_seed(2), dtype=F64)
    vals = graph.laplacian(graph.volume / aP, torch.ones(C, dtype=F64))
    H = BP.AggregationHierarchy.build(graph.cols.numpy(), g.cell_centroid.numpy(), None, coarsest=200)
    sizes = [lv.n for lv in H.levels]
    assert len(sizes) >= 3 and all(a >= 4 * b for a, b in zip(sizes, sizes[1:])), sizes
    ops, cinv = H.operators(vals)
    A = _dense(graph, vals)
    lv, nx_ = H.levels[0], H.levels[1]
    P = torch.zeros(lv.n, nx_.n, dtype=F64)
    P[torch.arange(lv.n), lv.agg] = 1.0
    Ac = torch.zeros(nx_.n, nx_.n, dtype=F64)
    Ac.index_put_((torch.arange(nx_.n).repeat(nx_.nslot), nx_.cols.long().reshape(-1)), ops[1].reshape(-1),
                  accumulate=True)
    assert torch.allclose(Ac, P.T @ A @ P, atol=1e-12)
    b = torch.randn(C, generator=torch.Generator().manual_seed(5), dtype=F64)
    spmv = _torch_spmv(H)
    x = BP.conjugate_gradient(lambda v: A @ v, b, lambda r: H.cycle(spmv, ops, cinv, r), 10)
    xj = BP.conjugate_gradient(lambda v: A @ v, b, 1.0 / torch.diagonal(A), 10)
    res = lambda y: float(torch.linalg.vector_norm(A @ y - b) / torch.linalg.vector_norm(b))  # noqa: E731
    assert res(x) < 1e-2 * res(xj), (res(x), res(xj))


def test_gmres_with_an_fp32_basis():
    """An FP32 Krylov basis (products, residuals and the update in FP64) still reaches a few-digit
    tolerance, in about the iterations of the FP64 basis, flexible or not."""
    from aquaflux_amd.solve import gmres

    g = torch.Generator().manual_seed(11)
    n = 300
    A = torch.eye(n, dtype=F64) * 4.0 + torch.randn(n, n, generator=g, dtype=F64) / n**0.5
    b = torch.randn(n, generator=g, dtype=F64)
    dinv = 1.0 / torch.diagonal(A)
    for flexible in (False, True):
        x64, i64 = gmres(lambda v: A @ v, b, rtol=1e-5, restart=20, max_iter=200, precond=lambda v: v * dinv,
                         flexible=flexible)
        x32, i32 = gmres(lambda v: A @ v, b, rtol=1e-5, restart=20, max_iter=200, precond=lambda v: v * dinv,
                         flexible=flexible, basis_dtype=torch.float32)
        assert x32.dtype == F64 and i32.converged and abs(i32.iterations - i64.iterations) <= 2
        assert float(torch.linalg.vector_norm(A @ x32 - b)) <= 1.01e-5 * float(torch.linalg.vector_norm(b))
