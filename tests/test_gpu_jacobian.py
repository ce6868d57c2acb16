"""Compressed (coloured) Jacobian assembly on the GPU: the probed block Jacobian equals the matrix
of unit-vector tangents (the HIP JVP, itself pinned to the oracle) when the reach covers the
stencil, folds far couplings when it does not, and its SpMV / transposed SpMV / point-block Jacobi
agree with the dense matrix.  After tests/unit/test_sparse_jacobian.py:230-400."""

import numpy as np
import pytest
import torch

from aquaflux_amd.jacobian import jacobian_relative_error, materialize_block_jacobian, mesh_colouring, mesh_probe_plan
from tests import flow_cases as FC
from tests import rans_cases as RC
from tests import scalar_cases as SC

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _dense_from_jvp(matvec, n, device):
    cols = []
    for i in range(n):
        e = torch.zeros(n, dtype=F64, device=device)
        e[i] = 1.0
        cols.append(matvec(e).cpu())
    return torch.stack(cols, dim=1)


def _rel(a, b):
    return float((a - b).abs().max() / max(float(b.abs().max()), 1e-300))


# Column reaches.  The scalar residual couples cells at most two steps apart (the neighbours'
# gradients and limiters).  In the flow and RANS residuals the velocity columns reach three: the
# Rhie-Chow a_P of a neighbour holds its faces' lagged mass-flux estimates, whose face velocities
# carry a skewness correction with the next ring's velocity gradients (zero on an orthogonal mesh),
# and in RANS its face viscosities carry the next ring's eddy viscosity, a function of that ring's
# velocity gradient; the pressure, k and omega columns reach two.
def _rans(cuda):
    print("case box3d_sst limited lsq", flush=True)
    mesh, rans, _ = RC.build("box3d_sst", "limited", "limited", "lsq", nu=1e-3, device=cuda)
    st = torch.tensor(RC.seeded_state(mesh, 3, "cavity"), device=cuda)
    return mesh, (lambda v: rans.residual_jvp(st, v)[1]), (3, 3, 3, 2, 2, 2)


def _flow(cuda):
    print("case channel2d limited lsq", flush=True)
    mesh, geom, mc = FC.build("channel2d", "limited", "lsq", device=cuda)
    st = torch.tensor(FC.seeded_state(mesh, 2, "channel"), dtype=F64, device=cuda)
    return mesh, (lambda v: mc.residual_jvp(st, v)[1]), (3, 3, 2)


def _scalar(cuda):
    print("case tet3 limited gg", flush=True)
    mesh, u, bcs = SC.make("tet3")
    phi, gamma = SC.seeded(mesh, 4)
    geom, st = SC.build_gpu(mesh, bcs, "limited", "gg", gamma, cuda)
    flux = SC.gpu_flux(geom, u).detach()
    x = torch.tensor(phi, device=cuda)
    return mesh, (lambda v: st.residual_jvp(flux, x, v)[1]), (2,)


@pytest.mark.parametrize("make", [_rans, _flow, _scalar], ids=["rans", "flow", "scalar"])
def test_probed_jacobian_equals_the_tangent_matrix(make, cuda):
    mesh, matvec, reach = make(cuda)
    B = len(reach)
    n = B * mesh.n_cells
    dense = _dense_from_jvp(matvec, n, cuda)
    plan = mesh_probe_plan(mesh, reach)
    J = materialize_block_jacobian(matvec, plan, device=cuda)
    err = _rel(J.to_dense(), dense)
    print("reach", reach, plan.n_probes, "probes, err", err)
    assert err < 1e-13
    assert jacobian_relative_error(J, matvec) < 1e-13
    # every nonzero of the true Jacobian lies in its column's reach
    C = mesh.n_cells
    col = plan.pattern
    depth = torch.full((C, C), 99, dtype=torch.int64)
    depth[torch.from_numpy(col.pattern_rows), torch.from_numpy(col.pattern_cols.astype(np.int64))] = torch.from_numpy(
        col.pattern_depth.astype(np.int64))
    for b in range(B):
        nz = dense.reshape(B, C, B, C)[:, :, b, :].abs().sum(dim=0) > 0
        assert not bool((nz & (depth > reach[b])).any()), b


def test_short_reach_folds_far_couplings(cuda):
    """A probe shorter than the residual's reach is not the Jacobian (far couplings land on near
    entries: ColumnProbePlan, sparse_jacobian.py:88-147) -- the relative-error guard shows it, for
    reach 2 on the velocity columns of the flow (perturbed mesh) and RANS residuals."""
    mesh, matvec, reach = _flow(cuda)
    J1 = materialize_block_jacobian(matvec, mesh_colouring(mesh, reach=2), len(reach), device=cuda)
    assert jacobian_relative_error(J1, matvec) > 1e-6
    mesh, matvec, reach = _rans(cuda)
    J2 = materialize_block_jacobian(matvec, mesh_colouring(mesh, reach=2), len(reach), device=cuda)
    assert jacobian_relative_error(J2, matvec) > 1e-6


def test_spmv_transpose_and_point_block_jacobi(cuda):
    mesh, matvec, reach = _rans(cuda)
    B = len(reach)
    plan = mesh_probe_plan(mesh, reach)
    J = materialize_block_jacobian(matvec, plan, device=cuda)
    Jf = materialize_block_jacobian(matvec, plan, device=cuda, precision="fp32")
    dense = J.to_dense()
    C = mesh.n_cells
    g = torch.Generator(device="cpu").manual_seed(1)
    v = torch.randn(B * C, generator=g, dtype=F64)
    assert _rel(J.field_matvec(v.to(cuda)).cpu(), dense @ v) < 1e-13
    assert _rel(Jf.field_matvec(v.to(cuda)).cpu(), dense @ v) < 1e-5
    # transposed product in the cell-major layout
    y = J.from_cell_major(J.rmatvec(J.to_cell_major(v.to(cuda)))).cpu()
    assert _rel(y, dense.T @ v) < 1e-13
    # point-block Jacobi inverts the diagonal blocks
    D = torch.zeros_like(dense)
    for a in range(B):
        for b in range(B):
            idx = torch.arange(C)
            D[a * C + idx, b * C + idx] = dense[a * C + idx, b * C + idx]
    x = J.from_cell_major(J.pbj(J.to_cell_major((D @ v).to(cuda)))).cpu()
    assert not J.singular
    assert _rel(x, v) < 1e-10


def test_newton_with_the_probed_jacobian_converges_on_the_cavity(cuda):
    """The Re=100 cavity (test_cavity.py:59-70, here 12x10) by Newton whose inner solves run on the
    FP32 probed Jacobian with point-block Jacobi: same root as with the compact Jacobian."""
    from aquaflux_amd.solve import DampedNewtonStep, ImplicitNewtonSolver

    print("case cavity 12x10 upwind gg, probed fp32", flush=True)
    mesh, geom, mc = FC.build("cavity", "upwind", "gg", device=cuda)
    roots = {}
    for jac in ("compact", "probed"):
        step = DampedNewtonStep(preconditioner="inner", precision="fp32", jacobian=jac, column_reach=(3, 3, 2))
        phi, rep = ImplicitNewtonSolver(forward_step=step).forward(mc, mc.initial_state())
        assert rep.converged and rep.residual_norms[-1] <= 1e-12 + 1e-10 * rep.residual_norms[0]
        roots[jac] = (phi, sum(rep.krylov_iterations))
        print(jac, rep.steps, "steps", roots[jac][1], "Krylov iterations", flush=True)
    assert float((roots["probed"][0] - roots["compact"][0]).abs().max()) < 1e-9
    # the exact Jacobian preconditions at least as well as the frozen compact one
    assert roots["probed"][1] <= roots["compact"][1]


def test_probed_jacobian_preconditions_the_rans_newton_step(cuda):
    """With the exact FP64 Jacobian solved tightly inside, the outer FGMRES on the HIP tangent of the
    coupled RANS residual converges in one or two iterations."""
    from aquaflux_amd.solve import DampedNewtonStep

    print("case channel2d_sst limited lsq, probed fp64", flush=True)
    mesh, rans, _ = RC.build("channel2d_sst", "limited", "limited", "lsq", nu=1e-3, device=cuda)
    st = torch.tensor(RC.seeded_state(mesh, 5, "channel"), device=cuda)
    r = rans.residual_raw(st)
    step = DampedNewtonStep(preconditioner="inner", precision="fp64", jacobian="probed", column_reach=(3, 3, 2, 2, 2),
                            rtol=1e-8, inner_rtol=1e-12, inner_restart=200, inner_iters=2000)
    delta, info = step.correction(rans, st, r)
    print("outer iterations", info.iterations, flush=True)
    assert info.converged and info.iterations <= 2
    _, Jd = rans.residual_jvp(st, delta)
    assert float((Jd + r).norm()) <= 1e-8 * float(r.norm())


def test_shifted_jacobian_adds_the_shift_and_leaves_the_input(cuda):
    """shifted_jacobian (sparse_jacobian.py:893-927): J + diag(shift) as a new matrix, J unchanged."""
    from aquaflux_amd.jacobian import shifted_jacobian

    mesh, matvec, reach = _scalar(cuda)
    plan = mesh_probe_plan(mesh, reach)
    J = materialize_block_jacobian(matvec, plan, device=cuda, precision="fp64")
    n = mesh.n_cells
    v = torch.randn(n, dtype=F64, device=cuda)
    before = J.field_matvec(v)
    shift = torch.linspace(0.5, 2.0, n, dtype=F64, device=cuda)
    Js = shifted_jacobian(J, shift)
    assert torch.allclose(Js.field_matvec(v), before + shift * v, rtol=1e-13, atol=1e-13)
    assert torch.equal(J.field_matvec(v), before)
    # the shifted copy's point-block Jacobi inverts its own (shifted) 1x1 diagonal blocks
    d = Js.to_dense().diagonal().to(cuda)
    assert torch.allclose(Js.from_cell_major(Js.pbj(Js.to_cell_major(v))), v / d, rtol=1e-13, atol=0.0)


def _oracle_jacobian(kind):
    """Dense Jacobian of the oracle residual by forward-mode AD (torch.func.jacfwd) at the same state
    and parameters as the matching case above."""
    from oracle import fv

    if kind == "rans":
        mesh, oset = RC.oracle_case("box3d_sst", "limited", "lsq", nu=1e-3)
        st = torch.tensor(RC.seeded_state(mesh, 3, "cavity"))
        C = mesh.n_cells
        mu, rho = torch.full((C,), 1e-3, dtype=F64), torch.ones(C, dtype=F64)
        return torch.func.jacfwd(lambda z: RC.oracle_residual(oset, z, mu, rho, "limited"))(st)
    if kind == "flow":
        mesh, (m, g, s) = FC.oracle_case("channel2d", "limited", "lsq")
        st = torch.tensor(FC.seeded_state(mesh, 2, "channel"), dtype=F64)
        C = mesh.n_cells
        mu, rho = torch.full((C,), 0.01, dtype=F64), torch.ones(C, dtype=F64)
        return torch.func.jacfwd(lambda z: fv.flow_residual(m, g, s, z, mu, rho))(st)
    mesh, u, bcs = SC.make("tet3")
    phi, gamma = SC.seeded(mesh, 4)
    m = fv.OMesh.from_mesh(mesh)
    g = fv.geometry(m)
    s = SC.oracle_setup(mesh, bcs, "limited", "gg")
    flux = SC.oracle_flux(g, u)
    return torch.func.jacfwd(lambda z: fv.scalar_residual(m, g, s, z, flux, torch.tensor(gamma)))(torch.tensor(phi))


@pytest.mark.parametrize("make", [_rans, _flow, _scalar], ids=["rans", "flow", "scalar"])
def test_probed_jacobian_equals_the_oracle_jacobian(make, cuda):
    """materialize_block_jacobian on the GPU against torch.func.jacfwd of the CPU oracle (the stand-in
    for jax.jacfwd of the reference residual), FP64, on the same small cases."""
    mesh, matvec, reach = make(cuda)
    plan = mesh_probe_plan(mesh, reach)
    J = materialize_block_jacobian(matvec, plan, device=cuda).to_dense()
    ref = _oracle_jacobian(make.__name__.lstrip("_"))
    err = _rel(J, ref)
    print("probed vs oracle jacfwd err", err, flush=True)
    assert err < 1e-11
