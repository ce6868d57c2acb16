"""GPU Newton-Krylov: block Jacobian parity, Newton convergence (FP32 vs FP64 inner), IFT adjoint."""

import numpy as np
import pytest
import torch

from aquaflux_amd.solve import BlockJacobian, DampedNewtonStep, ImplicitNewtonSolver, gmres
from oracle import fv
from tests.flow_cases import build, oracle_setup, rel_err, seeded_state

pytestmark = pytest.mark.gpu
F64 = torch.float64


@pytest.mark.parametrize("name,scheme,gradient", [("cavity", "upwind", "gg"), ("channel2d", "limited", "lsq"),
                                                  ("channel3d", "upwind", "lsq"), ("channel2d", "limited", "cgg")])
def test_block_jacobian_matches_frozen_oracle(cuda, name, scheme, gradient):
    """The block-ELL Jacobian equals d/dx of the residual with gradients, a_P, psi frozen."""
    mesh, geom, mc = build(name, scheme, gradient, device=cuda)
    kind = "cavity" if name.startswith("cavity") else "channel"
    x = seeded_state(mesh, 3, kind)
    J = BlockJacobian(mc, torch.tensor(x, dtype=F64, device=cuda), "fp64")
    n = x.size
    eye = torch.eye(n, dtype=F64, device=cuda)
    dense = torch.stack([J.from_cell_major(J.matvec(J.to_cell_major(eye[j]))) for j in range(n)], 1).cpu()
    m = fv.OMesh.from_mesh(mesh)
    g = fv.geometry(m)
    s = oracle_setup(mesh, mc, scheme, gradient)
    C = mesh.n_cells
    mu, rho = torch.full((C,), 0.01, dtype=F64), torch.ones(C, dtype=F64)
    x0 = torch.tensor(x, dtype=F64)
    ref = torch.func.jacfwd(lambda z: fv.flow_residual(m, g, s, z, mu, rho, frozen_state=x0))(x0)
    assert rel_err(dense, ref) < 1e-12


def test_spmv_fp32_matches_fp64(cuda):
    mesh, geom, mc = build("channel3d", "upwind", "lsq", device=cuda)
    x = torch.tensor(seeded_state(mesh, 1, "channel"), dtype=F64, device=cuda)
    J64, J32 = BlockJacobian(mc, x, "fp64"), BlockJacobian(mc, x, "fp32")
    v = torch.randn(x.numel(), dtype=F64, device=cuda)
    a = J64.from_cell_major(J64.matvec(J64.to_cell_major(v)))
    b = J32.from_cell_major(J32.matvec(J32.to_cell_major(v)))
    assert rel_err(b, a) < 1e-5
    # point-block Jacobi really inverts the diagonal blocks
    e = J64.to_cell_major(v)
    diag_only = torch.zeros_like(J64.vals)
    B, C = J64.B, J64.C
    diag_only[: B * B * C] = J64.vals[: B * B * C]
    J64.vals, keep = diag_only, J64.vals
    back = J64.matvec(J64.pbj(e))
    J64.vals = keep
    assert rel_err(back, e) < 1e-12
    # transposed product: <A v, w> == <v, A^T w>
    a_, b_ = J64.to_cell_major(v), J64.to_cell_major(torch.randn_like(v))
    lhs = float(torch.dot(J64.matvec(a_), b_))
    rhs = float(torch.dot(a_, J64.rmatvec(b_)))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_cavity_newton_converges(cuda, precision):
    """test_cavity.py:59-70 at 64x64 (the Re=100 cavity): Newton to 1e-10 relative."""
    mesh, geom, mc = build("cavity64", "upwind", "gg", device=cuda)
    solver = ImplicitNewtonSolver(forward_step=DampedNewtonStep(preconditioner="inner", precision=precision))
    phi, rep = solver.forward(mc, mc.initial_state())
    assert rep.converged and rep.residual_norms[-1] <= 1e-12 + 1e-10 * rep.residual_norms[0]
    # the converged state is a root of the oracle's residual too
    m = fv.OMesh.from_mesh(mesh)
    g = fv.geometry(m)
    s = oracle_setup(mesh, mc, "upwind", "gg")
    C = mesh.n_cells
    R = fv.flow_residual(m, g, s, phi.cpu(), torch.full((C,), 0.01, dtype=F64), torch.ones(C, dtype=F64))
    r0 = fv.flow_residual(m, g, s, torch.zeros_like(phi.cpu()), torch.full((C,), 0.01, dtype=F64), torch.ones(C, dtype=F64))
    assert float(R.norm()) <= 1e-9 * float(r0.norm())


def test_mixed_precision_matches_fp64_root(cuda):
    mesh, geom, mc = build("cavity64", "upwind", "gg", device=cuda)
    roots = {}
    for prec in ("fp32", "fp64"):
        solver = ImplicitNewtonSolver(forward_step=DampedNewtonStep(preconditioner="inner", precision=prec))
        roots[prec], _ = solver.forward(mc, mc.initial_state())
    d = float((roots["fp32"] - roots["fp64"]).abs().max() / roots["fp64"].abs().max())
    assert d <= 1e-9


def test_ghia_centrelines(cuda):
    """test_cavity.py:98-132: Re=100 centreline extrema on a 48x48 mesh."""
    from aquaflux_amd.mesh import structured_grid_2d
    from aquaflux_amd.boundary import BoundaryConditions
    from aquaflux_amd.discretization import FirstOrderUpwind
    from aquaflux_amd.flow import MomentumContinuity, MovingWall, NoSlipWall
    from aquaflux_amd.properties import Constant, PropertyModel
    from aquaflux_amd.schemes import CompactGreenGauss

    n = 48
    mesh = structured_grid_2d(n, n, named_boundaries=True)
    geom = mesh.geometry(device=cuda)
    mc = MomentumContinuity.build(
        mesh, geom, PropertyModel({"viscosity": Constant(0.01), "density": Constant(1.0)}), CompactGreenGauss(),
        BoundaryConditions({"top": MovingWall((1.0, 0.0)), "bottom": NoSlipWall(), "left": NoSlipWall(),
                            "right": NoSlipWall()}),
        advection_scheme=FirstOrderUpwind(), pressure_pin=0,
    )
    phi, rep = ImplicitNewtonSolver(rtol=1e-9).forward(mc, mc.initial_state())
    vel, _ = mc.unpack(phi)
    xc, yc = (geom.cell.centroid[:, 0].cpu().numpy(), geom.cell.centroid[:, 1].cpu().numpy())
    u, v = vel[:, 0].cpu().numpy(), vel[:, 1].cpu().numpy()
    gy = np.array([0.0625, 0.1016, 0.2813, 0.4531, 0.5, 0.6172, 0.7344, 0.8516, 0.9531])
    gu = np.array([-0.0419, -0.0643, -0.1566, -0.2109, -0.2058, -0.1364, 0.0033, 0.2315, 0.6872])
    gx = np.array([0.0625, 0.0938, 0.1563, 0.2266, 0.5, 0.8047, 0.8594, 0.9063, 0.9453])
    gv = np.array([0.0923, 0.1232, 0.1608, 0.1751, 0.0545, -0.2453, -0.2245, -0.1691, -0.1031])
    ux = np.unique(xc)[np.argmin(np.abs(np.unique(xc) - 0.5))]
    col = np.abs(xc - ux) < 1e-9
    o = np.argsort(yc[col])
    u_line = np.interp(gy, yc[col][o], u[col][o])
    vy = np.unique(yc)[np.argmin(np.abs(np.unique(yc) - 0.5))]
    row = np.abs(yc - vy) < 1e-9
    o = np.argsort(xc[row])
    v_line = np.interp(gx, xc[row][o], v[row][o])
    assert abs(u_line.min() - (-0.211)) < 0.03
    assert abs(v_line.min() - (-0.245)) < 0.04
    assert abs(v_line.max() - 0.175) < 0.04
    assert np.max(np.abs(u_line - gu)) < 0.1
    assert np.max(np.abs(v_line - gv)) < 0.04


def test_ift_adjoint_matches_finite_difference(cuda):
    """test_cavity.py:74-95: d mean|u| / d mu through the converged cavity vs central FD (3e-2)."""
    n = 20
    from aquaflux_amd.boundary import BoundaryConditions
    from aquaflux_amd.discretization import FirstOrderUpwind
    from aquaflux_amd.flow import MomentumContinuity, MovingWall, NoSlipWall
    from aquaflux_amd.mesh import structured_grid_2d
    from aquaflux_amd.properties import Constant, PropertyModel
    from aquaflux_amd.schemes import CompactGreenGauss

    mesh = structured_grid_2d(n, n, named_boundaries=True)
    geom = mesh.geometry(device=cuda)
    solver = ImplicitNewtonSolver(forward_step=DampedNewtonStep(preconditioner="inner", precision="fp64"))

    def mean_speed(mu):
        mc = MomentumContinuity.build(
            mesh, geom, PropertyModel({"viscosity": Constant(mu), "density": Constant(1.0)}), CompactGreenGauss(),
            BoundaryConditions({"top": MovingWall((1.0, 0.0)), "bottom": NoSlipWall(), "left": NoSlipWall(),
                                "right": NoSlipWall()}),
            advection_scheme=FirstOrderUpwind(), pressure_pin=0,
        )
        phi = solver.solve(mc, mc.initial_state(), (mu,) if isinstance(mu, torch.Tensor) else ())
        vel, _ = mc.unpack(phi)
        return vel[:, 0].abs().mean()

    mu = torch.tensor(0.01, dtype=F64, device=cuda, requires_grad=True)
    (grad,) = torch.autograd.grad(mean_speed(mu), (mu,))
    h = 1e-5
    fd = float(mean_speed(0.01 + h) - mean_speed(0.01 - h)) / (2 * h)
    assert abs(fd) > 1e-2
    assert abs(float(grad) - fd) <= 3e-2 * abs(fd)


def _hex_box(cuda, nx, ny, nz, mu):
    """The 3D hex box of the mixed-precision Newton workload: inlet u = (1, 0, 0) left, pressure outlet
    right, no-slip walls elsewhere, first-order upwind, least-squares gradients, Re = 1/mu, and the
    seeded state u = 1 + 0.05 N, v, w = 0.05 N, p = N(0, 0.1) (seed 2)."""
    from aquaflux_amd.boundary import BoundaryConditions
    from aquaflux_amd.discretization import FirstOrderUpwind
    from aquaflux_amd.flow import MomentumContinuity, NoSlipWall, PressureOutlet, VelocityInlet
    from aquaflux_amd.mesh import structured_grid_3d
    from aquaflux_amd.properties import Constant, PropertyModel
    from aquaflux_amd.schemes import LeastSquaresGradient

    mesh = structured_grid_3d(nx, ny, nz, named_boundaries=True)
    geom = mesh.geometry(device=cuda)
    bcs = BoundaryConditions({"left": VelocityInlet((1.0, 0.0, 0.0)), "right": PressureOutlet(0.0),
                              "bottom": NoSlipWall(), "top": NoSlipWall(), "back": NoSlipWall(), "front": NoSlipWall()})
    mc = MomentumContinuity.build(mesh, geom, PropertyModel({"viscosity": Constant(mu), "density": Constant(1.0)}),
                                  LeastSquaresGradient(), bcs, advection_scheme=FirstOrderUpwind())
    rng = np.random.default_rng(2)
    C = mesh.n_cells
    x0 = torch.tensor(np.concatenate([1.0 + 0.05 * rng.normal(size=C), 0.05 * rng.normal(size=C),
                                      0.05 * rng.normal(size=C), rng.normal(0.0, 0.1, C)]), dtype=F64, device=cuda)
    return mc, x0


@pytest.mark.parametrize("preconditioner", ["inner", "block"])
def test_mixed_precision_newton_on_the_re1000_hex_box(cuda, preconditioner):
    """The mixed-precision Newton workload on a reduced box (40x20x20, Re = 1000, the same boundary
    conditions and seeded state): FP32 inner solves on the FP32 block Jacobian (GMRES(30) with
    point-block Jacobi, or the block-SIMPLER preconditioner) against the all-FP64 path (FP64 Jacobian,
    FP64 inner solves, FP64 outer basis).  Both reach ||R|| <= 1e-12 + 1e-10 ||R0|| and their roots
    agree to 1e-9."""
    from aquaflux_amd.block_preconditioner import BlockPreconditioner

    print(f"case hex box 40x20x20 Re=1000, {preconditioner}", flush=True)
    mc, x0 = _hex_box(cuda, 40, 20, 20, 1e-3)
    roots = {}
    for prec in ("fp32", "fp64"):
        kw = dict(precision=prec, ptc_cfl0=10.0, restart=60, outer_basis=prec)
        if preconditioner == "inner":
            step = DampedNewtonStep(preconditioner="inner", inner_restart=30, inner_iters=30, **kw)
        else:
            step = DampedNewtonStep(preconditioner="block", block_preconditioner=BlockPreconditioner.build(
                mc, composition="simpler"), **kw)
        phi, rep = ImplicitNewtonSolver(max_steps=60, forward_step=step).forward(mc, x0)
        print(prec, rep.steps, "steps", f"{rep.seconds:.2f} s", rep.residual_norms[-1] / rep.residual_norms[0],
              flush=True)
        assert rep.converged and rep.residual_norms[-1] <= 1e-12 + 1e-10 * rep.residual_norms[0]
        roots[prec] = phi
    d = float((roots["fp32"] - roots["fp64"]).abs().max() / roots["fp64"].abs().max())
    print("root difference", d, flush=True)
    assert d <= 1e-9


def test_fp32_basis_update_keeps_fp64_accuracy(cuda):
    """The GMRES update x + Z y from an FP32 basis widens each entry and sums in FP64
    (afx_basis_combine): equal to the FP64 product of the widened basis to 1e-15; and GMRES with an
    FP32 basis still reaches a 1e-13 residual through its FP64 restarts."""
    from aquaflux_amd import _lib

    g = torch.Generator(device="cpu").manual_seed(3)
    Z = torch.randn(12, 10001, generator=g, dtype=torch.float32).to(cuda)
    y = np.random.default_rng(4).normal(size=9)
    x0 = torch.randn(10001, generator=g, dtype=F64).to(cuda)
    got = x0.clone()
    yd = torch.tensor(y, dtype=F64, device=cuda)
    _lib.check(_lib.lib().afx_basis_combine(got.numel(), 9, 0, Z.data_ptr(), Z.stride(0), yd.data_ptr(), got.data_ptr(),
                                            _lib.stream_of(cuda)), "afx_basis_combine")
    want = x0 + Z[:9].double().T @ torch.tensor(y, dtype=F64, device=cuda)
    assert float((got - want).abs().max()) <= 1e-15 * float(want.abs().max())
    n = 4096
    d = torch.linspace(1.0, 3.0, n, dtype=F64, device=cuda)
    b = torch.sin(torch.arange(n, dtype=F64, device=cuda))
    x, info = gmres(lambda v: d * v, b, rtol=1e-13, restart=20, max_iter=400, basis_dtype=torch.float32)
    assert info.converged
    assert float((d * x - b).norm() / b.norm()) <= 1e-13


def test_rans_adjoint_solve_uses_the_probed_jacobian(cuda):
    """The IFT adjoint solve J^T lambda = g for a coupled RANS operator runs on the forward step's
    probed Jacobian (FP64) and returns lambda with ||J^T lambda - g|| <= 1e-9 ||g|| (checked with the
    HIP adjoint); a compact-Jacobian step refuses the RANS operator with a ValueError."""
    from tests import rans_cases as RC

    print("case channel2d_sst limited lsq, adjoint solve", flush=True)
    mesh, rans, _ = RC.build("channel2d_sst", "limited", "limited", "lsq", nu=1e-3, device=cuda)
    st = torch.tensor(RC.seeded_state(mesh, 5, "channel"), device=cuda)
    g = torch.tensor(np.random.default_rng(7).normal(size=st.numel()), device=cuda)
    step = DampedNewtonStep(preconditioner="inner", precision="fp32", jacobian="probed", column_reach=(3, 3, 2, 2, 2))
    solver = ImplicitNewtonSolver(forward_step=step, adjoint_rtol=1e-11)
    lam = solver.adjoint(rans, st, g)
    res = rans.residual_vjp(st, lam) - g
    print("adjoint residual", float(res.norm() / g.norm()), flush=True)
    assert float(res.norm()) <= 1e-9 * float(g.norm())
    with pytest.raises(ValueError):
        ImplicitNewtonSolver(forward_step=DampedNewtonStep()).adjoint(rans, st, g)
