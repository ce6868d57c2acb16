"""GPU parity of the scalar-transport residual, tangent and adjoint (including the tangent and the
adjoint with respect to mesh node coordinates through the geometry) against the CPU oracle.

Tolerances: residual 1e-12 and derivatives 1e-11 relative to the largest entry of each compared
array (register sums on the GPU and index_add scatters in the oracle associate differently).
"""

import dataclasses

import numpy as np
import pytest
import torch
import torch.autograd.forward_ad as fwAD

from oracle import fv
from tests import scalar_cases as SC

pytestmark = pytest.mark.gpu
F64 = torch.float64
TOL_R = 1e-12
TOL_D = 1e-11

CASES = [
    ("tet3", "limited", "gg"),
    ("tet3", "upwind", "gg"),
    ("tet3", "limited", "lsq"),
    ("tet3", "upwind", "none"),
    ("quad_mixed", "limited", "gg"),
    ("quad_mixed", "upwind", "lsq"),
    ("quad_mixed", "none", "gg"),
    ("quad_mixed", "limited", "none"),
    ("tet3", "limited", "cgg"),
    ("quad_mixed", "limited", "cgg"),
    ("brick3d", "limited", "gg"),
    ("brick3d", "limited", "lsq"),
]
# Node-coordinate derivatives: every scheme.  tet3 with limited least squares is left out: at that
# case's seed two limiter candidates of some cells are exactly equal, where the oracle (like JAX's
# segment_min) splits the derivative between the tied faces and the kernels take the first (a
# documented parity-unpinned point); brick3d and quad_mixed cover the scheme.
GEO_CASES = [c for c in CASES if c != ("tet3", "limited", "lsq")] + [("brick3d", "upwind", "cgg")]


def _ids(c):
    return "-".join(c)


def rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).abs().max() / max(float(b.abs().max()), 1e-300))


def _setup(case, cuda, seed=0):
    name, adv, grad = case
    print(f"case {_ids(case)} seed {seed}", flush=True)
    mesh, u, bcs = SC.make(name)
    phi, gamma = SC.seeded(mesh, seed)
    m = fv.OMesh.from_mesh(mesh)
    g = fv.geometry(m)
    s = SC.oracle_setup(mesh, bcs, adv, grad)
    geom, st = SC.build_gpu(mesh, bcs, adv, grad, gamma, cuda)
    return mesh, u, bcs, phi, gamma, (m, g, s), geom, st


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_scalar_residual_matches_oracle(case, cuda):
    mesh, u, bcs, phi, gamma, (m, g, s), geom, st = _setup(case, cuda)
    ref = fv.scalar_residual(m, g, s, torch.tensor(phi), SC.oracle_flux(g, u), torch.tensor(gamma))
    got = st.residual(SC.gpu_flux(geom, u))(torch.tensor(phi, device=cuda))
    err = rel(got, ref)
    print("residual err", err)
    assert err < TOL_R


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_scalar_jvp_matches_oracle(case, cuda):
    mesh, u, bcs, phi, gamma, (m, g, s), geom, st = _setup(case, cuda, seed=1)
    rng = np.random.default_rng(2)
    C, F = mesh.n_cells, mesh.n_faces
    t_phi = torch.tensor(rng.normal(size=C))
    t_gam = torch.tensor(rng.normal(size=C)) * 0.001
    t_flux = torch.tensor(rng.normal(size=F)) * 0.01
    t_val = torch.tensor(rng.normal(size=F))
    t_tinf = torch.tensor(rng.normal(size=F))
    flux_o = SC.oracle_flux(g, u)

    def f(ph, ga, fl, va, ti):
        return fv.scalar_residual(m, g, dataclasses.replace(s, value=va, t_inf=ti), ph, fl, ga)

    _, ref = torch.func.jvp(f, (torch.tensor(phi), torch.tensor(gamma), flux_o, s.value, s.t_inf),
                            (t_phi, t_gam, t_flux, t_val, t_tinf))
    prim = (torch.tensor(phi, device=cuda), torch.tensor(gamma, device=cuda),
            SC.gpu_flux(geom, u).detach().contiguous(), st._val, st._tinf)
    geo = tuple(t.detach() for t in st._geometry())
    _, dR = st._raw.jvp(prim, tuple(t.to(cuda) for t in (t_phi, t_gam, t_flux, t_val, t_tinf)), geo)
    err = rel(dR, ref)
    print("jvp err", err)
    assert err < TOL_D


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_scalar_vjp_matches_oracle(case, cuda):
    mesh, u, bcs, phi, gamma, (m, g, s), geom, st = _setup(case, cuda, seed=2)
    rng = np.random.default_rng(4)
    lam = torch.tensor(rng.normal(size=mesh.n_cells))
    ph = torch.tensor(phi, requires_grad=True)
    ga = torch.tensor(gamma, requires_grad=True)
    fl = SC.oracle_flux(g, u).detach().clone().requires_grad_(True)
    va = s.value.clone().requires_grad_(True)
    ti = s.t_inf.clone().requires_grad_(True)
    R = fv.scalar_residual(m, g, dataclasses.replace(s, value=va, t_inf=ti), ph, fl, ga)
    refs = torch.autograd.grad(R, (ph, ga, fl, va, ti), lam, allow_unused=True)
    bars, _ = st.residual_vjp(SC.gpu_flux(geom, u), torch.tensor(phi, device=cuda), lam.to(cuda))
    bnd = torch.from_numpy(mesh.neighbour < 0)
    for name, got, ref in zip(("phi", "gamma", "flux", "value", "t_inf"), bars, refs):
        ref = torch.zeros_like(got.cpu()) if ref is None else ref
        if name in ("value", "t_inf"):
            got, ref = got.cpu()[bnd], ref[bnd]
        scale = max(float(ref.abs().max()), 1e-300)
        err = float((got.cpu() - ref).abs().max()) / scale
        print(name, err, scale)
        assert err < TOL_D or float(ref.abs().max()) == 0.0 and float(got.abs().max()) < 1e-14, name


@pytest.mark.parametrize("case", GEO_CASES, ids=_ids)
def test_scalar_node_coordinate_adjoint_matches_oracle(case, cuda):
    """d <lambda, R(phi; geometry(X), flux(geometry(X)))> / dX through the hand-written residual and
    geometry adjoints, against torch autograd of the oracle."""
    name, adv, grad = case
    mesh, u, bcs = SC.make(name)
    phi, gamma = SC.seeded(mesh, 3)
    rng = np.random.default_rng(5)
    lam = torch.tensor(rng.normal(size=mesh.n_cells))
    # oracle
    Xo = torch.tensor(mesh.node_coords, dtype=F64, requires_grad=True)
    m = fv.OMesh.from_mesh(mesh, node_coords=Xo)
    g = fv.geometry(m)
    s = SC.oracle_setup(mesh, bcs, adv, grad)
    R = fv.scalar_residual(m, g, s, torch.tensor(phi), SC.oracle_flux(g, u), torch.tensor(gamma))
    (ref,) = torch.autograd.grad((R * lam).sum(), Xo)
    # HIP path
    X = torch.tensor(mesh.node_coords, dtype=F64, device=cuda, requires_grad=True)
    geom, st = SC.build_gpu(mesh, bcs, adv, grad, gamma, cuda, node_coords=X)
    Rg = st.residual(SC.gpu_flux(geom, u))(torch.tensor(phi, device=cuda))
    (got,) = torch.autograd.grad((Rg * lam.to(cuda)).sum(), X)
    err = rel(got, ref)
    print("node adjoint err", err, float(ref.abs().max()))
    assert err < TOL_D


def test_geometry_adjoint_matches_oracle(cuda):
    """Cotangents of all five geometry arrays back to the nodes (tet and quad meshes)."""
    for name in ("tet3", "quad_mixed"):
        mesh, _, _ = SC.make(name)
        rng = np.random.default_rng(9)
        Xo = torch.tensor(mesh.node_coords, dtype=F64, requires_grad=True)
        g = fv.geometry(fv.OMesh.from_mesh(mesh, node_coords=Xo))
        w = [torch.tensor(rng.normal(size=t.shape)) for t in (g.area, g.normal, g.face_centroid, g.volume,
                                                              g.cell_centroid)]
        L = sum((a * b).sum() for a, b in zip((g.area, g.normal, g.face_centroid, g.volume, g.cell_centroid), w))
        (ref,) = torch.autograd.grad(L, Xo)
        X = torch.tensor(mesh.node_coords, dtype=F64, device=cuda, requires_grad=True)
        geom = mesh.geometry(node_coords=X, device=cuda)
        arrs = (geom.face.area, geom.face.normal, geom.face.centroid, geom.cell.volume, geom.cell.centroid)
        Lg = sum((a * b.to(cuda)).sum() for a, b in zip(arrs, w))
        (got,) = torch.autograd.grad(Lg, X)
        err = rel(got, ref)
        print(name, "geometry adjoint err", err)
        assert err < TOL_D


def test_scalar_node_adjoint_finite_difference(cuda):
    """The node-coordinate adjoint against central differences of the GPU residual itself.  Only
    interior nodes move: the wall faces then stay parallel to the flow (Q = 0 exactly), where the
    upwind switch makes the residual non-differentiable in the node coordinates."""
    mesh, u, bcs = SC.make("tet3")
    phi, gamma = SC.seeded(mesh, 6)
    rng = np.random.default_rng(7)
    lam = torch.tensor(rng.normal(size=mesh.n_cells), device=cuda)
    X0 = torch.tensor(mesh.node_coords, dtype=F64, device=cuda)
    on_boundary = np.zeros(mesh.n_nodes, dtype=bool)
    off = mesh.face_node_offsets
    for f in np.nonzero(mesh.neighbour < 0)[0]:
        on_boundary[mesh.face_node_indices[off[f]:off[f + 1]]] = True
    dX = torch.tensor(rng.normal(size=X0.shape), device=cuda) * 1e-2 / 3.0
    dX[torch.from_numpy(on_boundary).to(cuda)] = 0.0

    def L(X):
        geom, st = SC.build_gpu(mesh, bcs, "limited", "gg", gamma, cuda, node_coords=X)
        R = st.residual(SC.gpu_flux(geom, u))(torch.tensor(phi, device=cuda))
        return (R * lam).sum()

    X = X0.clone().requires_grad_(True)
    (gx,) = torch.autograd.grad(L(X), X)
    an = float((gx * dX).sum())

    def cd(e):
        with torch.no_grad():
            return float((L(X0 + e * dX) - L(X0 - e * dX)) / (2 * e))

    fd = (4.0 * cd(5e-5) - cd(1e-4)) / 3.0
    print("fd", fd, "adjoint", an)
    assert abs(fd - an) <= 1e-7 * max(abs(an), 1e-12)


def _node_tangent_pair(mesh, u, bcs, adv, grad, phi, gamma, dX, cuda, hist=None):
    """(R, dR/dX . dX) of the oracle (torch forward AD) and of the HIP path (dual-number geometry
    and residual kernels under torch forward AD)."""
    T = torch.tensor
    s = SC.oracle_setup(mesh, bcs, adv, grad)
    extra_o = () if hist is None else (T(hist[0]), T(hist[1]), DT, hist[2])

    def f(Xo):
        m = fv.OMesh.from_mesh(mesh, node_coords=Xo)
        g = fv.geometry(m)
        return fv.scalar_residual(m, g, s, T(phi), SC.oracle_flux(g, u), T(gamma), *extra_o)

    ref = torch.func.jvp(f, (T(mesh.node_coords, dtype=F64),), (T(dX),))
    extra = () if hist is None else (T(hist[0], device=cuda), None if hist[2] else T(hist[1], device=cuda), DT,
                                     hist[2])
    with fwAD.dual_level():
        Xd = fwAD.make_dual(T(mesh.node_coords, dtype=F64, device=cuda), T(dX, device=cuda))
        geom, st = SC.build_gpu(mesh, bcs, adv, grad, gamma, cuda, node_coords=Xd, transient=hist is not None)
        out = fwAD.unpack_dual(st.residual(SC.gpu_flux(geom, u))(T(phi, device=cuda), *extra))
        got = (out.primal.clone(), out.tangent.clone())
    return ref, got


@pytest.mark.parametrize("case", GEO_CASES, ids=_ids)
def test_scalar_node_coordinate_tangent_matches_oracle(case, cuda):
    """dR/dX . Xdot of R(phi; geometry(X), flux(geometry(X))) through the dual-number geometry and
    residual kernels, against torch forward-mode AD of the oracle; then <J Xdot, lambda> against
    <Xdot, J^T lambda> from the hand-written adjoints (two independent derivations)."""
    name, adv, grad = case
    mesh, u, bcs = SC.make(name)
    phi, gamma = SC.seeded(mesh, 3)
    rng = np.random.default_rng(11)
    dX = rng.normal(size=mesh.node_coords.shape)
    (R_ref, dR_ref), (R, dR) = _node_tangent_pair(mesh, u, bcs, adv, grad, phi, gamma, dX, cuda)
    assert rel(R, R_ref) < TOL_R
    err = rel(dR, dR_ref)
    print("node tangent err", err, float(dR_ref.abs().max()))
    assert err < TOL_D
    lam = rng.normal(size=mesh.n_cells)
    X = torch.tensor(mesh.node_coords, dtype=F64, device=cuda, requires_grad=True)
    geom, st = SC.build_gpu(mesh, bcs, adv, grad, gamma, cuda, node_coords=X)
    Rg = st.residual(SC.gpu_flux(geom, u))(torch.tensor(phi, device=cuda))
    (gx,) = torch.autograd.grad((Rg * torch.tensor(lam, device=cuda)).sum(), X)
    a = float((dR.cpu() * torch.tensor(lam)).sum())
    b = float((gx.cpu() * torch.tensor(dX)).sum())
    scale = float(dR.abs().sum()) * float(np.abs(lam).max())
    print("duality", a, b)
    assert abs(a - b) <= 1e-12 * scale


@pytest.mark.parametrize("first", [True, False], ids=["bdf1", "bdf2"])
def test_scalar_transient_node_coordinate_tangent_matches_oracle(first, cuda):
    """The node-coordinate tangent with the accumulation term V (phi - history) / dt, whose volume
    moves with the nodes."""
    mesh, u, bcs = SC.make("tet3")
    phi, gamma = SC.seeded(mesh, 4)
    old, older = SC.history(mesh, 4)
    dX = np.random.default_rng(12).normal(size=mesh.node_coords.shape)
    (R_ref, dR_ref), (R, dR) = _node_tangent_pair(mesh, u, bcs, "limited", "gg", phi, gamma, dX, cuda,
                                                  hist=(old, older, first))
    assert rel(R, R_ref) < TOL_R
    err = rel(dR, dR_ref)
    print("transient node tangent err", err)
    assert err < TOL_D


def test_geometry_tangent_matches_oracle(cuda):
    """Tangents of all five geometry arrays along a node-coordinate direction (tet, quad and hex
    meshes) against torch forward-mode AD of the oracle's geometry."""
    for name in ("tet3", "quad_mixed", "brick3d"):
        mesh, _, _ = SC.make(name)
        dX = np.random.default_rng(13).normal(size=mesh.node_coords.shape)

        def f(Xo):
            g = fv.geometry(fv.OMesh.from_mesh(mesh, node_coords=Xo))
            return g.area, g.normal, g.face_centroid, g.volume, g.cell_centroid

        vals, tans = torch.func.jvp(f, (torch.tensor(mesh.node_coords, dtype=F64),), (torch.tensor(dX),))
        with fwAD.dual_level():
            Xd = fwAD.make_dual(torch.tensor(mesh.node_coords, dtype=F64, device=cuda), torch.tensor(dX, device=cuda))
            geom = mesh.geometry(node_coords=Xd, device=cuda)
            arrs = (geom.face.area, geom.face.normal, geom.face.centroid, geom.cell.volume, geom.cell.centroid)
            got = [fwAD.unpack_dual(a) for a in arrs]
            got = [(p.primal.clone(), p.tangent.clone()) for p in got]
        for label, (pv, pt), v, t in zip(("area", "normal", "face centroid", "volume", "cell centroid"), got, vals,
                                         tans):
            ev, et = rel(pv, v), rel(pt, t)
            print(name, label, "value err", ev, "tangent err", et)
            assert ev < TOL_R and et < TOL_D, (name, label)


# --- transient (BDF1 first step, BDF2 afterwards) ----------------------------------------------
TRANSIENT_CASES = [
    ("tet3", "limited", "gg", True),
    ("tet3", "limited", "gg", False),
    ("quad_mixed", "upwind", "lsq", False),
    ("quad_mixed", "limited", "cgg", True),
    ("tet3", "none", "none", False),
    ("brick3d", "limited", "gg", False),
]
DT = 0.037


def _tids(c):
    return "-".join(c[:3]) + ("-bdf1" if c[3] else "-bdf2")


def _transient_setup(case, cuda, seed):
    name, adv, grad, first = case
    print(f"case {_tids(case)} seed {seed}", flush=True)
    mesh, u, bcs = SC.make(name)
    phi, gamma = SC.seeded(mesh, seed)
    old, older = SC.history(mesh, seed)
    m = fv.OMesh.from_mesh(mesh)
    g = fv.geometry(m)
    s = SC.oracle_setup(mesh, bcs, adv, grad)
    geom, st = SC.build_gpu(mesh, bcs, adv, grad, gamma, cuda, transient=True)
    return mesh, u, phi, gamma, old, older, first, (m, g, s), geom, st


@pytest.mark.parametrize("case", TRANSIENT_CASES, ids=_tids)
def test_scalar_transient_residual_matches_oracle(case, cuda):
    mesh, u, phi, gamma, old, older, first, (m, g, s), geom, st = _transient_setup(case, cuda, 0)
    T = torch.tensor
    ref = fv.scalar_residual(m, g, s, T(phi), SC.oracle_flux(g, u), T(gamma), T(old), T(older), DT, first)
    fn = st.residual(SC.gpu_flux(geom, u))
    got = fn(T(phi, device=cuda), T(old, device=cuda), T(older, device=cuda) if not first else None, DT, first)
    err = rel(got, ref)
    print("transient residual err", err)
    assert err < TOL_R
    # the steady residual plus the accumulation term alone
    steady = fv.scalar_residual(m, g, s, T(phi), SC.oracle_flux(g, u), T(gamma))
    acc = fv.transient_term(T(phi), T(old), T(older), DT, first, g.volume)
    assert rel(got, steady + acc) < TOL_R


@pytest.mark.parametrize("case", TRANSIENT_CASES, ids=_tids)
def test_scalar_transient_jvp_matches_oracle(case, cuda):
    """Tangent with respect to phi and both history levels (forward AD through the autograd
    Function)."""
    mesh, u, phi, gamma, old, older, first, (m, g, s), geom, st = _transient_setup(case, cuda, 1)
    rng = np.random.default_rng(11)
    C = mesh.n_cells
    tp, to, too = (torch.tensor(rng.normal(size=C)) for _ in range(3))
    flux_o = SC.oracle_flux(g, u)
    T = torch.tensor

    def f(ph, po, poo):
        return fv.scalar_residual(m, g, s, ph, flux_o, T(gamma), po, poo, DT, first)

    _, ref = torch.func.jvp(f, (T(phi), T(old), T(older)), (tp, to, too if not first else torch.zeros(C)))
    fn = st.residual(SC.gpu_flux(geom, u).detach())
    import torch.autograd.forward_ad as fwAD
    with fwAD.dual_level():
        ph = fwAD.make_dual(T(phi, device=cuda), tp.to(cuda))
        po = fwAD.make_dual(T(old, device=cuda), to.to(cuda))
        poo = None if first else fwAD.make_dual(T(older, device=cuda), too.to(cuda))
        dR = fwAD.unpack_dual(fn(ph, po, poo, DT, first)).tangent
    err = rel(dR, ref)
    print("transient jvp err", err)
    assert err < TOL_D


@pytest.mark.parametrize("case", TRANSIENT_CASES, ids=_tids)
def test_scalar_transient_vjp_matches_oracle(case, cuda):
    mesh, u, phi, gamma, old, older, first, (m, g, s), geom, st = _transient_setup(case, cuda, 2)
    rng = np.random.default_rng(12)
    lam = torch.tensor(rng.normal(size=mesh.n_cells))
    ph, po, poo = (torch.tensor(x, requires_grad=True) for x in (phi, old, older))
    R = fv.scalar_residual(m, g, s, ph, SC.oracle_flux(g, u), torch.tensor(gamma), po, poo, DT, first)
    refs = torch.autograd.grad(R, (ph, po, poo), lam, allow_unused=True)
    T = torch.tensor
    xs = [T(phi, device=cuda, requires_grad=True), T(old, device=cuda, requires_grad=True),
          None if first else T(older, device=cuda, requires_grad=True)]
    Rg = st.residual(SC.gpu_flux(geom, u).detach())(xs[0], xs[1], xs[2], DT, first)
    got = torch.autograd.grad(Rg, [x for x in xs if x is not None], lam.to(cuda))
    for name, gv, ref in zip(("phi", "phi_old", "phi_older"), got, refs):
        err = rel(gv, ref)
        print(name, err)
        assert err < TOL_D, name
    # the raw entry returns the same cotangents (and None for phi_older on the first step)
    bars, _ = st.residual_vjp(SC.gpu_flux(geom, u), xs[0].detach(), lam.to(cuda), phi_old=xs[1].detach(),
                              phi_older=None if first else xs[2].detach(), dt=DT, first_step=first)
    assert rel(bars[0], refs[0]) < TOL_D and rel(bars[5], refs[1]) < TOL_D
    assert (bars[6] is None) == first


@pytest.mark.parametrize("first", [True, False], ids=["bdf1", "bdf2"])
def test_scalar_transient_node_adjoint_matches_oracle(first, cuda):
    """The accumulation term's volume enters the node-coordinate adjoint."""
    mesh, u, bcs = SC.make("tet3")
    phi, gamma = SC.seeded(mesh, 3)
    old, older = SC.history(mesh, 3)
    rng = np.random.default_rng(5)
    lam = torch.tensor(rng.normal(size=mesh.n_cells))
    T = torch.tensor
    Xo = torch.tensor(mesh.node_coords, dtype=F64, requires_grad=True)
    m = fv.OMesh.from_mesh(mesh, node_coords=Xo)
    g = fv.geometry(m)
    s = SC.oracle_setup(mesh, bcs, "limited", "gg")
    R = fv.scalar_residual(m, g, s, T(phi), SC.oracle_flux(g, u), T(gamma), T(old), T(older), DT, first)
    (ref,) = torch.autograd.grad((R * lam).sum(), Xo)
    X = torch.tensor(mesh.node_coords, dtype=F64, device=cuda, requires_grad=True)
    geom, st = SC.build_gpu(mesh, bcs, "limited", "gg", gamma, cuda, node_coords=X, transient=True)
    Rg = st.residual(SC.gpu_flux(geom, u))(T(phi, device=cuda), T(old, device=cuda), T(older, device=cuda), DT,
                                           first)
    (got,) = torch.autograd.grad((Rg * lam.to(cuda)).sum(), X)
    err = rel(got, ref)
    print("transient node adjoint err", err)
    assert err < TOL_D


RECORD_CASES = [c for c in CASES if c[2] != "cgg"]


@pytest.mark.parametrize("with_geometry", [False, True], ids=["inputs", "geometry"])
@pytest.mark.parametrize("case", RECORD_CASES, ids=_ids)
def test_vjp_from_forward_record_is_bitwise_the_recomputing_vjp(case, with_geometry, cuda):
    """afx_scalar_residual_record + afx_scalar_vjp_record (the reverse pass reusing the forward's
    gradient pass) give bit for bit what afx_scalar_residual and afx_scalar_vjp give."""
    mesh, u, bcs, phi, gamma, _, geom, st = _setup(case, cuda, seed=4)
    lam = torch.tensor(np.random.default_rng(6).normal(size=mesh.n_cells), device=cuda)
    raw = st._raw
    prim = st._prim(SC.gpu_flux(geom, u), torch.tensor(phi, device=cuda), None, None)
    geo = tuple(t.detach() for t in st._geometry())
    R0 = raw.residual(prim, geo)
    R1, rec = raw.residual_record(prim, geo)
    assert torch.equal(R0, R1)
    b0, g0 = raw.vjp(prim, lam, geo, with_geometry)
    b1, g1 = raw.vjp(prim, lam, geo, with_geometry, record=rec)
    for x, y in zip(b0 + g0, b1 + g1):
        assert (x is None) == (y is None)
        if x is not None:
            assert torch.equal(x, y)


def test_vjp_record_transient_and_autograd(cuda):
    """The autograd backward takes the record path; its cotangents of phi and phi_old equal the
    recomputing entry's bit for bit (BDF2, limited Green-Gauss on tetrahedra)."""
    mesh, u, phi, gamma, old, older, first, _, geom, st = _transient_setup(("tet3", "limited", "gg", False), cuda, 3)
    lam = torch.tensor(np.random.default_rng(13).normal(size=mesh.n_cells), device=cuda)
    T = torch.tensor
    xs = [T(x, device=cuda, requires_grad=True) for x in (phi, old, older)]
    Rg = st.residual(SC.gpu_flux(geom, u).detach())(xs[0], xs[1], xs[2], DT, False)
    got = torch.autograd.grad(Rg, xs, lam)
    bars, _ = st.residual_vjp(SC.gpu_flux(geom, u), xs[0].detach(), lam, phi_old=xs[1].detach(),
                              phi_older=xs[2].detach(), dt=DT, first_step=False)
    for g, b in zip(got, (bars[0], bars[5], bars[6])):
        assert torch.equal(g, b)


def test_record_refused_for_corrected_green_gauss(cuda):
    mesh, u, bcs, phi, gamma, _, geom, st = _setup(("quad_mixed", "limited", "cgg"), cuda)
    assert not st._raw.records
    prim = st._prim(SC.gpu_flux(geom, u), torch.tensor(phi, device=cuda), None, None)
    geo = tuple(t.detach() for t in st._geometry())
    with pytest.raises(ValueError, match="corrected Green-Gauss"):
        st._raw.residual_record(prim, geo)


def test_geometry_vjp_from_record_is_bitwise_the_recomputing_vjp(cuda):
    """The autograd backward of the geometry reads the forward's approximate centroids
    (afx_geometry_vjp_record); it equals afx_geometry_vjp, which forms them again, bit for bit."""
    import ctypes

    from aquaflux_amd import _lib

    for name in ("tet3", "quad_mixed", "brick3d"):
        mesh, _, _ = SC.make(name)
        X = torch.tensor(mesh.node_coords.T, dtype=F64, device=cuda).contiguous().requires_grad_(True)
        geom = mesh.geometry(node_coords=X.T, device=cuda)
        dm = geom.device_mesh
        soa = (geom.face.area, geom.face.normal_soa, geom.face.centroid_soa, geom.cell.volume,
               geom.cell.centroid_soa)
        rng = np.random.default_rng(10)
        w = [torch.tensor(rng.normal(size=tuple(t.shape)), device=cuda) for t in soa]
        (got,) = torch.autograd.grad(soa, X, w)
        Xs = geom.node_coords_soa.detach()
        gs = _lib.AfxGeometry(*[t.data_ptr() for t in soa])
        gb = _lib.AfxGeometryCotangents(*[t.data_ptr() for t in w])
        n = ctypes.c_int64()
        L = _lib.lib()
        _lib.check(L.afx_geometry_vjp_workspace_bytes(dm.handle, ctypes.byref(n)), "workspace")
        work = torch.empty(int(n.value), dtype=torch.uint8, device=cuda)
        ref = torch.empty_like(Xs)
        _lib.check(L.afx_geometry_vjp(dm.handle, Xs.data_ptr(), ctypes.byref(gs), ctypes.byref(gb), ref.data_ptr(),
                                      work.data_ptr(), work.numel(), _lib.stream_of(cuda)), "afx_geometry_vjp")
        torch.cuda.synchronize()
        print(name, float((got - ref).abs().max()))
        assert torch.equal(got, ref)
