"""The SIMPLE-family block preconditioner on the GPU: the Schur it assembles from the HIP geometry
and momentum diagonal equals the oracle's, and Newton preconditioned by it (each composition,
both velocity solves, both Schur scalings) reaches the root the point-block inner solve reaches."""

import pytest
import torch

from aquaflux_amd.block_preconditioner import BlockPreconditioner
from aquaflux_amd.solve import DampedNewtonStep, ImplicitNewtonSolver
from oracle import fv
from tests import flow_cases as FC

pytestmark = pytest.mark.gpu
F64 = torch.float64


@pytest.mark.parametrize("case", ["cavity", "channel2d", "channel3d"])
def test_schur_matches_oracle_on_device(case, cuda):
    print("schur", case, flush=True)
    mesh, geom, mc = FC.build(case, "upwind", "gg", device=cuda)
    om, og, _ = FC.oracle_case(case, "upwind", "gg")[1]
    state = torch.tensor(FC.seeded_state(mesh, seed=1, kind="cavity" if case == "cavity" else "channel"), dtype=F64,
                         device=cuda)
    bp = BlockPreconditioner.build(mc)
    aP = bp.frozen_momentum_diagonal(state)
    rho = mc.density.detach()
    g = bp.graph
    vals = g.laplacian(g.volume / aP, rho, mc.pressure_pin).cpu()
    C = mesh.n_cells
    got = torch.zeros(C, C, dtype=F64)
    cols = g.cols.long().cpu()
    for s in range(g.nslot):
        got.index_put_((torch.arange(C), cols[s]), vals[s], accumulate=True)
    outlet = mesh.face_patches.indices("right") if mc.pressure_pin is None else ()
    ref = fv.pressure_schur_matrix(om, og, aP.cpu(), rho.cpu(), outlet, mc.pressure_pin)
    err = float((got - ref).abs().max()) / float(ref.abs().max())
    print("relative error", err, flush=True)
    assert err <= 1e-12


CASES = [
    ("cavity", "triangular", "gmres", "simple"),
    ("cavity", "simple", "pbj", "simple"),
    ("cavity", "simpler", "gmres", "msimple"),
    ("channel2d", "simple", "gmres", "simple"),
    ("channel2d", "simpler", "gmres", "simple"),
    ("channel3d", "simple", "gmres", "msimple"),
]


@pytest.mark.parametrize("case,composition,velocity,scaling", CASES)
def test_block_preconditioned_newton_reaches_the_root(case, composition, velocity, scaling, cuda):
    print("newton", case, composition, velocity, scaling, flush=True)
    mesh, geom, mc = FC.build(case, "upwind", "gg", device=cuda)
    # the cavity from rest; a channel from the seeded near-uniform stream (plain Newton from rest
    # does not reach the 3D channel's root)
    x0 = mc.initial_state() if case == "cavity" else torch.tensor(FC.seeded_state(mesh, seed=0, kind="channel"),
                                                                  dtype=F64, device=cuda)
    ref_step = DampedNewtonStep(preconditioner="inner", precision="fp64")
    ref, rep0 = ImplicitNewtonSolver(forward_step=ref_step).forward(mc, x0)
    bp = BlockPreconditioner.build(mc, composition=composition, velocity=velocity, schur_scaling=scaling)
    step = DampedNewtonStep(preconditioner="block", precision="fp64", block_preconditioner=bp)
    phi, rep = ImplicitNewtonSolver(forward_step=step).forward(mc, x0)
    print("inner", rep0.steps, rep0.krylov_iterations, "block", rep.steps, rep.krylov_iterations, flush=True)
    assert rep.converged and rep.residual_norms[-1] <= 1e-12 + 1e-10 * rep.residual_norms[0]
    assert step.block_preconditioner is bp
    err = float((phi - ref).abs().max()) / float(ref.abs().max())
    print("root difference", err, flush=True)
    assert err <= 1e-8


@pytest.mark.parametrize("case", ["cavity64", "channel3d_mg"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_native_v_cycle_matches_the_torch_cycle(case, dtype, cuda):
    """The library's fused smoothing / residual / restriction / prolongation kernels against the
    same V-cycle written with torch operations over the library's ELL product."""
    from aquaflux_amd import _lib

    print("v-cycle", case, dtype, flush=True)
    mesh, geom, mc = FC.build(case, "upwind", "gg", device=cuda)
    state = torch.tensor(FC.seeded_state(mesh, seed=1, kind="cavity" if case.startswith("cavity") else "channel"),
                         dtype=F64, device=cuda)
    bp = BlockPreconditioner.build(mc)
    g, H = bp.graph, bp.hierarchy
    S = g.laplacian(g.volume / bp.frozen_momentum_diagonal(state), mc.density.detach(), mc.pressure_pin, dtype)
    ops, cinv = H.operators(S)
    prec = 0 if dtype == torch.float32 else 1

    def spmv(k, vals, x):
        lv = H.levels[k]
        y = torch.empty_like(x)
        _lib.check(_lib.lib().afx_bell_spmv(lv.n, 1, lv.nslot, prec, vals.data_ptr(), lv.cols.data_ptr(),
                                            x.data_ptr(), y.data_ptr(), _lib.stream_of(x.device)), "afx_bell_spmv")
        return y

    b = torch.randn(mesh.n_cells, generator=torch.Generator().manual_seed(3), dtype=F64).to(cuda, dtype)
    want = H.cycle(spmv, ops, cinv, b)
    got = H.native_cycle(ops, cinv, b)
    err = float((got - want).abs().max() / want.abs().max())
    print("levels", len(H.levels), "relative difference", err, flush=True)
    assert len(H.levels) >= 3
    assert err <= (1e-12 if dtype == torch.float64 else 1e-5)
    # the hierarchy set-up and the native cycle are bitwise reproducible (Galerkin sums by gather,
    # restriction by member table: no atomics)
    ops2, cinv2 = H.operators(S)
    assert all(torch.equal(a, c) for a, c in zip(ops, ops2)) and torch.equal(cinv, cinv2)
    assert torch.equal(H.native_cycle(ops2, cinv2, b), got)


def test_block_preconditioner_over_the_probed_jacobian(cuda):
    """The saddle blocks read off the probed exact Jacobian (reach 3 / 3 / 2) precondition Newton to
    the same root as the inner solve on the compact Jacobian."""
    print("newton channel2d simpler over the probed Jacobian", flush=True)
    mesh, geom, mc = FC.build("channel2d", "upwind", "gg", device=cuda)
    x0 = torch.tensor(FC.seeded_state(mesh, seed=0, kind="channel"), dtype=F64, device=cuda)
    ref, _ = ImplicitNewtonSolver(forward_step=DampedNewtonStep(preconditioner="inner", precision="fp64")).forward(mc, x0)
    bp = BlockPreconditioner.build(mc, composition="simpler")
    step = DampedNewtonStep(preconditioner="block", precision="fp32", jacobian="probed", column_reach=(3, 3, 2),
                            block_preconditioner=bp)
    phi, rep = ImplicitNewtonSolver(forward_step=step).forward(mc, x0)
    print("steps", rep.steps, "krylov", rep.krylov_iterations, flush=True)
    assert rep.converged
    assert float((phi - ref).abs().max()) / float(ref.abs().max()) <= 1e-8


def test_block_preconditioner_refuses_a_foreign_assembler(cuda):
    _, _, mc = FC.build("cavity", "upwind", "gg", device=cuda)
    _, _, other = FC.build("cavity", "upwind", "gg", device=cuda)
    step = DampedNewtonStep(preconditioner="block", block_preconditioner=BlockPreconditioner.build(mc))
    x0 = other.initial_state()
    with pytest.raises(ValueError, match="another assembler"):
        step.correction(other, x0, other.residual_raw(x0))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_captured_schur_solve_replays_match_eager(precision, cuda):
    """The preconditioner with its Schur solve recorded as a HIP graph against the same
    preconditioner run eagerly, over several replays with eager work and synchronisations between."""
    from aquaflux_amd.solve import BlockJacobian

    print("capture", precision, flush=True)
    mesh, geom, mc = FC.build("channel3d_mg", "upwind", "gg", device=cuda)
    state = torch.tensor(FC.seeded_state(mesh, seed=2, kind="channel"), dtype=F64, device=cuda)
    J = BlockJacobian(mc, state, precision)
    eager = BlockPreconditioner.build(mc, composition="simpler", capture=False).apply_at(state, J)
    graphed = BlockPreconditioner.build(mc, composition="simpler", capture=True).apply_at(state, J)
    gen = torch.Generator().manual_seed(9)
    tol = 1e-12 if precision == "fp64" else 1e-5
    for k in range(4):
        v = torch.randn(state.numel(), generator=gen, dtype=F64).to(cuda)
        want = eager(v)
        torch.cuda.synchronize()
        got = graphed(v)
        torch.cuda.synchronize()
        err = float((got - want).abs().max() / want.abs().max())
        print("replay", k, "relative difference", err, flush=True)
        assert err <= tol


def test_jacobi_schur_solve_reaches_the_root(cuda):
    """The Schur solved by Jacobi-preconditioned CG (no multigrid) still preconditions Newton to the
    inner-solve root."""
    print("newton channel2d simple, jacobi schur", flush=True)
    mesh, geom, mc = FC.build("channel2d", "upwind", "gg", device=cuda)
    x0 = torch.tensor(FC.seeded_state(mesh, seed=0, kind="channel"), dtype=F64, device=cuda)
    ref, _ = ImplicitNewtonSolver(forward_step=DampedNewtonStep(preconditioner="inner", precision="fp64")).forward(mc, x0)
    bp = BlockPreconditioner.build(mc, composition="simple", schur_solver="jacobi")
    assert bp.hierarchy is None and bp.schur_iters == 30
    phi, rep = ImplicitNewtonSolver(forward_step=DampedNewtonStep(preconditioner="block", precision="fp64",
                                                                  block_preconditioner=bp)).forward(mc, x0)
    assert rep.converged
    assert float((phi - ref).abs().max()) / float(ref.abs().max()) <= 1e-8


def test_re1000_cavity_from_rest(cuda):
    """The 64 x 64 lid-driven cavity at Re = 1000 from rest (pseudo-transient continuation from CFL
    1): block-SIMPLER reaches the root the inner point-block solve reaches."""
    print("newton cavity64 Re=1000 from rest, block-simpler", flush=True)
    mesh, geom, mc = FC.build("cavity64", "upwind", "gg", mu=1e-3, device=cuda)
    x0 = mc.initial_state()
    ref, _ = ImplicitNewtonSolver(max_steps=60, forward_step=DampedNewtonStep(
        preconditioner="inner", precision="fp32", ptc_cfl0=1.0, restart=60)).forward(mc, x0)
    bp = BlockPreconditioner.build(mc, composition="simpler")
    phi, rep = ImplicitNewtonSolver(max_steps=60, forward_step=DampedNewtonStep(
        preconditioner="block", precision="fp32", ptc_cfl0=1.0, restart=60, block_preconditioner=bp)).forward(mc, x0)
    print("steps", rep.steps, "krylov", sum(rep.krylov_iterations), flush=True)
    assert rep.converged
    assert float((phi - ref).abs().max()) / float(ref.abs().max()) <= 1e-8


@pytest.mark.parametrize("preconditioner", ["inner", "block"])
def test_fp32_outer_basis_reaches_the_same_root(preconditioner, cuda):
    """Newton whose outer FGMRES keeps its basis in FP32 (residuals, JVPs and updates in FP64)
    reaches the root of the FP64-basis solve, in the same number of Newton steps."""
    print("newton channel3d_mg outer basis fp32,", preconditioner, flush=True)
    mesh, geom, mc = FC.build("channel3d_mg", "upwind", "gg", device=cuda)
    x0 = torch.tensor(FC.seeded_state(mesh, seed=0, kind="channel"), dtype=F64, device=cuda)
    roots = {}
    for basis in ("fp64", "fp32"):
        kw = {"block_preconditioner": BlockPreconditioner.build(mc, composition="simpler")} if preconditioner == "block" else {}
        step = DampedNewtonStep(preconditioner=preconditioner, precision="fp32", outer_basis=basis, **kw)
        phi, rep = ImplicitNewtonSolver(forward_step=step).forward(mc, x0)
        assert rep.converged
        roots[basis] = (phi, rep.steps)
        print(basis, rep.steps, rep.krylov_iterations, flush=True)
    assert roots["fp32"][1] == roots["fp64"][1]
    assert float((roots["fp32"][0] - roots["fp64"][0]).abs().max()) / float(roots["fp64"][0].abs().max()) <= 1e-9
    with pytest.raises(ValueError, match="outer_basis"):
        DampedNewtonStep(outer_basis="bf16").correction(mc, x0, mc.residual_raw(x0))
