"""GPU parity of the coupled-flow residual, tangent and adjoint against the CPU oracle.

Tolerances: the FP64 residual and its derivatives must match the oracle to 1e-12 relative to the
largest entry (summation order differs between the per-cell register sums on the GPU and the
oracle's index_add scatter, so bitwise equality is not expected).
"""

import numpy as np
import pytest
import torch

from oracle import fv
from tests.flow_cases import build, oracle_setup, rel_err, seeded_state

pytestmark = pytest.mark.gpu
F64 = torch.float64
TOL = 1e-12

CASES = [
    ("cavity", "upwind", "gg"),
    ("cavity", "none", "gg"),
    ("cavity", "limited", "lsq"),
    ("channel2d", "upwind", "gg"),
    ("channel2d", "limited", "lsq"),
    ("channel3d", "upwind", "lsq"),
    ("channel3d", "limited", "lsq"),
    ("box3d", "upwind", "gg"),
    ("channel2d", "limited", "cgg"),
    ("channel2d", "upwind", "cgg9r0.8"),
    ("channel3d", "limited", "cgg6"),
    ("cavity", "none", "cgg1"),
    ("brick2d", "limited", "lsq"),
    ("brick3d", "limited", "lsq"),
    ("brick3d", "upwind", "cgg"),
]


@pytest.mark.parametrize("name", ["channel2d", "channel3d"])
def test_geometry_matches_oracle(cuda, name):
    from tests.flow_cases import make_mesh

    mesh = make_mesh(name)
    geom = mesh.geometry(device=cuda)
    og = fv.geometry(fv.OMesh.from_mesh(mesh))
    assert rel_err(geom.face.area, og.area) < 1e-13
    assert rel_err(geom.face.normal, og.normal) < 1e-13
    assert rel_err(geom.face.centroid, og.face_centroid) < 1e-13
    assert rel_err(geom.cell.volume, og.volume) < 1e-12
    assert rel_err(geom.cell.centroid, og.cell_centroid) < 1e-12


def test_tet_geometry_matches_oracle(cuda):
    from aquaflux_amd.mesh import tet_grid_3d

    mesh = tet_grid_3d(4, perturb=0.15, seed=3)
    geom = mesh.geometry(device=cuda)
    og = fv.geometry(fv.OMesh.from_mesh(mesh))
    assert rel_err(geom.face.area, og.area) < 1e-13
    assert rel_err(geom.face.normal, og.normal) < 1e-13
    assert rel_err(geom.cell.volume, og.volume) < 1e-12
    assert rel_err(geom.cell.centroid, og.cell_centroid) < 1e-12


def _oracle_inputs(mesh, mc, scheme, gradient, state):
    m = fv.OMesh.from_mesh(mesh)
    g = fv.geometry(m)
    s = oracle_setup(mesh, mc, scheme, gradient)
    C = mesh.n_cells
    mu = torch.full((C,), 0.01, dtype=F64)
    rho = torch.ones(C, dtype=F64)
    return m, g, s, torch.tensor(state, dtype=F64), mu, rho


@pytest.mark.parametrize("name,scheme,gradient", CASES)
def test_residual_matches_oracle(cuda, name, scheme, gradient):
    mesh, geom, mc = build(name, scheme, gradient, device=cuda)
    kind = "cavity" if name.startswith(("cavity", "box")) else "channel"
    state = seeded_state(mesh, seed=0, kind=kind)
    R = mc.residual(torch.tensor(state, dtype=F64, device=cuda))
    m, g, s, x, mu, rho = _oracle_inputs(mesh, mc, scheme, gradient, state)
    R0 = fv.flow_residual(m, g, s, x, mu, rho)
    assert rel_err(R, R0) < TOL


@pytest.mark.parametrize("name,scheme,gradient", CASES)
def test_jvp_matches_oracle(cuda, name, scheme, gradient):
    mesh, geom, mc = build(name, scheme, gradient, device=cuda)
    kind = "cavity" if name.startswith(("cavity", "box")) else "channel"
    state = seeded_state(mesh, seed=1, kind=kind)
    rng = np.random.default_rng(7)
    v = rng.normal(size=state.size)
    R, dR = mc.residual_jvp(torch.tensor(state, dtype=F64, device=cuda), torch.tensor(v, dtype=F64, device=cuda))
    m, g, s, x, mu, rho = _oracle_inputs(mesh, mc, scheme, gradient, state)
    _, dR0 = torch.func.jvp(lambda z: fv.flow_residual(m, g, s, z, mu, rho), (x,), (torch.tensor(v, dtype=F64),))
    assert rel_err(dR, dR0) < TOL


@pytest.mark.parametrize("name,scheme,gradient", CASES)
def test_vjp_matches_oracle(cuda, name, scheme, gradient):
    mesh, geom, mc = build(name, scheme, gradient, device=cuda)
    kind = "cavity" if name.startswith(("cavity", "box")) else "channel"
    state = seeded_state(mesh, seed=2, kind=kind)
    rng = np.random.default_rng(11)
    w = rng.normal(size=state.size)
    # GPU: reverse through the autograd.Function, w.r.t. state and the viscosity / density leaves
    mu_leaf = torch.tensor(0.01, dtype=F64, device=cuda, requires_grad=True)
    rho_leaf = torch.tensor(1.0, dtype=F64, device=cuda, requires_grad=True)
    from aquaflux_amd.properties import Constant, PropertyModel

    mc.properties = PropertyModel({"viscosity": Constant(mu_leaf), "density": Constant(rho_leaf)})
    xs = torch.tensor(state, dtype=F64, device=cuda, requires_grad=True)
    R = mc.residual(xs)
    gx, gmu, grho = torch.autograd.grad(R, (xs, mu_leaf, rho_leaf), torch.tensor(w, dtype=F64, device=cuda))
    # oracle
    m, g, s, x, mu, rho = _oracle_inputs(mesh, mc, scheme, gradient, state)
    x = x.clone().requires_grad_(True)
    mu0 = torch.tensor(0.01, dtype=F64, requires_grad=True)
    rho0 = torch.tensor(1.0, dtype=F64, requires_grad=True)
    C = mesh.n_cells
    R0 = fv.flow_residual(m, g, s, x, mu0.expand(C), rho0.expand(C))
    gx0, gmu0, grho0 = torch.autograd.grad(R0, (x, mu0, rho0), torch.tensor(w, dtype=F64))
    assert rel_err(gx, gx0) < TOL
    assert abs(float(gmu) - float(gmu0)) <= TOL * max(abs(float(gmu0)), 1.0)
    assert abs(float(grho) - float(grho0)) <= TOL * max(abs(float(grho0)), 1.0)


def test_vjp_inlet_velocity_and_outlet_pressure(cuda):
    """Cotangents w.r.t. the inlet velocity vector and outlet pressure (dynamic leaves here)."""
    uin = torch.tensor([1.0, 0.1], dtype=F64, device=cuda, requires_grad=True)
    mesh, geom, mc = build("channel2d", "limited", "lsq", device=cuda, inlet_velocity=uin)
    state = seeded_state(mesh, seed=3, kind="channel")
    xs = torch.tensor(state, dtype=F64, device=cuda)
    w = torch.tensor(np.random.default_rng(5).normal(size=state.size), dtype=F64, device=cuda)
    (gu,) = torch.autograd.grad(mc.residual(xs), (uin,), w)
    m, g, s, x, mu, rho = _oracle_inputs(mesh, mc, "limited", "lsq", state)
    u0 = torch.tensor([1.0, 0.1], dtype=F64, requires_grad=True)
    inlet = torch.from_numpy(mesh.face_patches.indices("left"))
    ub = s.u_b.clone()
    ub = ub.index_put((inlet,), u0[None, :].expand(inlet.numel(), 2))
    s.u_b = ub
    (gu0,) = torch.autograd.grad(fv.flow_residual(m, g, s, x, mu, rho), (u0,), w.cpu())
    assert rel_err(gu, gu0) < TOL


def test_adjoint_dot_product(cuda):
    """<J v, w> == <v, J^T w> between the dual-number tangent and the hand-written adjoint."""
    mesh, geom, mc = build("channel3d", "limited", "lsq", device=cuda)
    state = torch.tensor(seeded_state(mesh, 4, "channel"), dtype=F64, device=cuda)
    g = torch.Generator(device="cpu").manual_seed(3)
    v = torch.randn(state.numel(), generator=g, dtype=F64).to(cuda)
    w = torch.randn(state.numel(), generator=g, dtype=F64).to(cuda)
    _, Jv = mc.residual_jvp(state, v)
    JTw = mc.residual_vjp(state, w)
    a, b = float(torch.dot(Jv, w)), float(torch.dot(v, JTw))
    assert abs(a - b) <= 1e-12 * max(abs(a), abs(b))


def test_residual_is_deterministic(cuda):
    """Two evaluations are bitwise identical (no atomics in any reduction)."""
    mesh, geom, mc = build("channel3d", "limited", "lsq", device=cuda)
    state = torch.tensor(seeded_state(mesh, 5, "channel"), dtype=F64, device=cuda)
    assert torch.equal(mc.residual(state), mc.residual(state))


def test_internal_face_order_and_face_pack(cuda):
    """The residual kernels' internal face order sorts the caller's faces by (rank among the owner's
    faces, owner); a
    call without the geometry's face pack (the library gathers one) gives the same bits; a mesh
    whose faces already are owner-sorted reports no permutation."""
    import ctypes

    from aquaflux_amd import _lib
    from aquaflux_amd.mesh import tet_grid_3d

    mesh, geom, mc = build("channel3d", "limited", "lsq", device=cuda)
    dm = mesh.device_mesh(cuda)
    assert dm.face_permuted
    fo = dm.face_order().cpu().numpy()
    by_owner = np.argsort(mesh.owner, kind="stable")
    rank = np.empty(mesh.n_faces, dtype=np.int64)
    run = np.r_[0, np.flatnonzero(np.diff(mesh.owner[by_owner])) + 1]
    pos = np.arange(mesh.n_faces) - np.repeat(run, np.diff(np.r_[run, mesh.n_faces]))
    rank[by_owner] = pos
    assert np.array_equal(fo, by_owner[np.argsort(rank[by_owner], kind="stable")])
    state = torch.tensor(seeded_state(mesh, 7, "channel"), dtype=F64, device=cuda)
    raw = mc._raw
    prim = [state] + [t.detach() for t in mc._inputs()[:4]]
    R_pack = raw.residual(*prim)
    gs_pack = raw.gs
    assert gs_pack.face_pack
    raw.gs = _lib.AfxGeometry(*[getattr(gs_pack, n) for n in ("area", "normal", "face_centroid", "volume",
                                                                   "cell_centroid")], None)
    try:
        R_gather = raw.residual(*prim)
    finally:
        raw.gs = gs_pack
    assert torch.equal(R_pack, R_gather)
    tet = tet_grid_3d(3, perturb=0.1, seed=1)
    tdm = tet.device_mesh(cuda)
    assert not tdm.face_permuted and tdm.face_order() is None
    assert tet.geometry(device=cuda).face_pack() is None
    flag = ctypes.c_int32(7)
    _lib.check(_lib.lib().afx_mesh_face_order(tdm.handle, ctypes.byref(flag), None, None), "afx_mesh_face_order")
    assert flag.value == 0


# --- node coordinates -----------------------------------------------------------------------------
GEO_CASES = [
    ("cavity", "upwind", "gg"),
    ("cavity", "limited", "lsq"),
    ("channel2d", "upwind", "gg"),
    ("channel2d", "limited", "lsq"),
    ("channel3d", "limited", "lsq"),
    ("channel3d", "upwind", "gg"),
    ("box3d", "limited", "gg"),
    ("brick2d", "limited", "lsq"),
    ("brick3d", "limited", "lsq"),
    ("channel2d", "limited", "cgg"),
    ("channel2d", "upwind", "cgg9r0.8"),
    ("channel3d", "limited", "cgg6"),
    ("cavity", "upwind", "cgg1"),
    ("brick3d", "upwind", "cgg"),
]


def _kind(name):
    return "cavity" if name in ("cavity", "box3d") else "channel"


@pytest.mark.parametrize("name,scheme,gradient", GEO_CASES)
def test_node_coordinate_tangent_matches_oracle(cuda, name, scheme, gradient):
    """dR/dX . Xdot with the geometry recomputed from the node coordinates: the dual-number
    geometry kernels feeding the flow passes on dual geometry, against torch forward-mode AD of the
    oracle (1e-11 relative to the largest entry)."""
    import torch.autograd.forward_ad as fwAD

    from tests.flow_cases import make_mesh

    print("case", name, scheme, gradient, flush=True)
    mesh = make_mesh(name)
    state = seeded_state(mesh, seed=4, kind=_kind(name))
    dX = np.random.default_rng(5).normal(size=mesh.node_coords.shape) * 0.01
    T = torch.tensor
    with fwAD.dual_level():
        Xd = fwAD.make_dual(T(mesh.node_coords, dtype=F64, device=cuda), T(dX, device=cuda))
        mesh, geom, mc = build(name, scheme, gradient, device=cuda, node_coords=Xd)
        out = fwAD.unpack_dual(mc.residual(T(state, dtype=F64, device=cuda)))
        R, dR = out.primal.clone(), out.tangent.clone()
    s = oracle_setup(mesh, mc, scheme, gradient)
    C = mesh.n_cells
    mu, rho = torch.full((C,), 0.01, dtype=F64), torch.ones(C, dtype=F64)

    def f(X):
        m = fv.OMesh.from_mesh(mesh, node_coords=X)
        return fv.flow_residual(m, fv.geometry(m), s, T(state, dtype=F64), mu, rho)

    R_ref, dR_ref = torch.func.jvp(f, (T(mesh.node_coords, dtype=F64),), (T(dX),))
    assert rel_err(R, R_ref) < TOL
    err = rel_err(dR, dR_ref)
    print("node tangent err", err, float(dR_ref.abs().max()))
    assert err < 1e-11


@pytest.mark.parametrize("name,scheme,gradient", GEO_CASES)
def test_node_coordinate_adjoint_matches_oracle(cuda, name, scheme, gradient):
    """d <lambda, R(state; geometry(X))> / dX through the hand-written flow adjoint with geometry
    cotangents (afx_flow_vjp_geometry) and the geometry adjoint, against torch autograd of the
    oracle (1e-11 relative to the largest entry), and <J Xdot, lambda> against the forward-mode
    tangent of the same HIP path."""
    import torch.autograd.forward_ad as fwAD

    from tests.flow_cases import make_mesh

    print("case", name, scheme, gradient, flush=True)
    mesh = make_mesh(name)
    state = seeded_state(mesh, seed=4, kind=_kind(name))
    rng = np.random.default_rng(6)
    lam = rng.normal(size=state.size)
    T = torch.tensor
    X = T(mesh.node_coords, dtype=F64, device=cuda, requires_grad=True)
    mesh, geom, mc = build(name, scheme, gradient, device=cuda, node_coords=X)
    R = mc.residual(T(state, dtype=F64, device=cuda))
    (got,) = torch.autograd.grad((R * T(lam, device=cuda)).sum(), X)
    s = oracle_setup(mesh, mc, scheme, gradient)
    C = mesh.n_cells
    Xo = T(mesh.node_coords, dtype=F64, requires_grad=True)
    m = fv.OMesh.from_mesh(mesh, node_coords=Xo)
    Ro = fv.flow_residual(m, fv.geometry(m), s, T(state, dtype=F64), torch.full((C,), 0.01, dtype=F64),
                          torch.ones(C, dtype=F64))
    (ref,) = torch.autograd.grad((Ro * T(lam)).sum(), Xo)
    err = rel_err(got, ref)
    print("node adjoint err", err, float(ref.abs().max()))
    assert err < 1e-11
    dX = rng.normal(size=mesh.node_coords.shape) * 0.01
    with fwAD.dual_level():
        Xd = fwAD.make_dual(T(mesh.node_coords, dtype=F64, device=cuda), T(dX, device=cuda))
        _, _, mcd = build(name, scheme, gradient, device=cuda, node_coords=Xd)
        dR = fwAD.unpack_dual(mcd.residual(T(state, dtype=F64, device=cuda))).tangent.clone()
    a = float((dR.cpu() * T(lam)).sum())
    b = float((got.cpu() * T(dX)).sum())
    assert abs(a - b) <= 1e-12 * float(dR.abs().sum()) * float(np.abs(lam).max())
