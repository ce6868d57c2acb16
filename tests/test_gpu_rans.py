"""GPU parity of the coupled RANS (SST) residual, tangent and adjoint against the CPU oracle.

Tolerances: each field block (u, v, (w), p, k, omega) is normalised by its own largest entry.  The
FP64 residual matches to 1e-12; the tangent and adjoint to 1e-11 (the closure's derivatives pass
through tanh / log / sqrt chains whose rounding differs between the hand-written dual arithmetic and
torch autograd by a few ulps of intermediate terms).
"""

import numpy as np
import pytest
import torch

from tests import rans_cases as RC

pytestmark = pytest.mark.gpu
F64 = torch.float64
TOL_R = 1e-12
TOL_D = 1e-11

CASES = [
    ("cavity_sst", "upwind", "upwind", "gg", 1e-2),
    ("channel2d_sst", "limited", "limited", "lsq", 1e-5),
    ("channel2d_sst", "upwind", "limited", "lsq", 1e-3),
    ("channel3d_sst", "limited", "limited", "lsq", 1e-4),
    ("channel3d_sst", "limited", "upwind", "lsq", 1e-5),
    ("box3d_sst", "limited", "limited", "gg", 1e-3),
    ("channel2d_sst", "limited", "limited", "cgg", 1e-5),
    ("channel3d_sst", "upwind", "upwind", "cgg6r0.9", 1e-4),
    ("brick2d_sst", "limited", "limited", "lsq", 1e-5),
    ("brick3d_sst", "limited", "limited", "lsq", 1e-4),
]


def _ids(c):
    return f"{c[0]}-{c[1]}-{c[2]}-{c[3]}-nu{c[4]:g}"


def _setup(case, cuda, seed=0):
    name, fs, ss, gr, nu = case
    print(f"case {_ids(case)} seed {seed}", flush=True)
    mesh, rans, oset = RC.build(name, fs, ss, gr, nu=nu, device=cuda)
    kind = "cavity" if name.startswith("cavity") else "channel"
    st = torch.tensor(RC.seeded_state(mesh, seed, kind))
    return mesh, rans, oset, st


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_rans_residual_matches_oracle(case, cuda):
    mesh, rans, oset, st = _setup(case, cuda)
    C, nb = mesh.n_cells, mesh.dim + 3
    mu = torch.full((C,), case[4], dtype=F64)
    rho = torch.ones(C, dtype=F64)
    ref = RC.oracle_residual(oset, st, mu, rho, case[2])
    got = rans.residual(st.to(cuda))
    errs = RC.block_errors(got, ref, nb)
    print("block errors", errs)
    assert max(errs) < TOL_R


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_rans_closure_fields_match_oracle(case, cuda):
    from oracle import fv
    from oracle import sst as S

    mesh, rans, oset, st = _setup(case, cuda)
    m, g, s, t = oset
    d, C = mesh.dim, mesh.n_cells
    blocks = st.reshape(d + 3, C)
    vel, p = blocks[:d].T, blocks[d]
    ub, _ = fv.flow_boundary_fields(m, g, s, vel, p)
    gv = torch.stack([fv._gradient(m, g, s, vel[:, i], ub[:, i]) for i in range(d)], dim=1)
    cl = S.closure_fields(m, g, S.gradient_spec(s), t, vel, ub, gv, blocks[d + 1], blocks[d + 2])
    f = rans.closure_fields(st.to(cuda))
    for key, ref in (("nu_t", cl["nu_t"]), ("f1", cl["f1"]), ("strain_rate", cl["strain"])):
        err = float((f[key].cpu() - ref).abs().max() / ref.abs().max())
        print(key, err)
        assert err < TOL_R
    tau = f["wall_shear_rate"].cpu()[t.wall_cells]
    assert float((tau - cl["tau"]).abs().max() / cl["tau"].abs().max()) < TOL_R
    wn = S.wall_face_eddy_viscosity(m, t, blocks[d + 1])
    got_wn = f["wall_nu_t"].cpu()[m.owner[t.wall_faces]]
    scale = max(float(wn.abs().max()), 1e-300)
    assert float((got_wn - wn[t.wall_faces]).abs().max()) / scale < TOL_R


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_rans_jvp_matches_oracle(case, cuda):
    mesh, rans, oset, st = _setup(case, cuda, seed=1)
    C, nb = mesh.n_cells, mesh.dim + 3
    mu = torch.full((C,), case[4], dtype=F64)
    rho = torch.ones(C, dtype=F64)
    rng = np.random.default_rng(7)
    v = torch.tensor(rng.normal(size=st.numel())) * st.abs().clamp_min(1e-3)
    _, ref = torch.func.jvp(lambda z: RC.oracle_residual(oset, z, mu, rho, case[2]), (st,), (v,))
    R, dR = rans.residual_jvp(st.to(cuda), v.to(cuda))
    errs = RC.block_errors(dR, ref, nb)
    print("jvp block errors", errs)
    assert max(errs) < TOL_D
    assert max(RC.block_errors(R, RC.oracle_residual(oset, st, mu, rho, case[2]), nb)) < TOL_R


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_rans_parameter_jvp_matches_oracle(case, cuda):
    """Tangent with respect to mu, rho, nu and the boundary velocity / pressure together."""
    mesh, rans, oset, st = _setup(case, cuda, seed=2)
    m, g, s, t = oset
    d, C, F, nb = mesh.dim, mesh.n_cells, mesh.n_faces, mesh.dim + 3
    mu = torch.full((C,), case[4], dtype=F64)
    rho = torch.ones(C, dtype=F64)
    nu = t.nu.clone()
    rng = np.random.default_rng(3)
    dmu = torch.tensor(rng.normal(size=C)) * case[4]
    drho = torch.tensor(rng.normal(size=C)) * 0.1
    dnu = torch.tensor(rng.normal(size=C)) * case[4]
    dub = torch.tensor(rng.normal(size=(F, d)))
    dpb = torch.tensor(rng.normal(size=F))

    def f(mu_, rho_, nu_, ub_, pb_):
        return RC.oracle_residual(oset, st, mu_, rho_, case[2], nu=nu_, u_b=ub_, p_b=pb_)

    _, ref = torch.func.jvp(f, (mu, rho, nu, s.u_b, s.p_b), (dmu, drho, dnu, dub, dpb))
    prim = rans._detached(st.to(cuda))
    tans = [torch.zeros_like(prim[0]), dmu.to(cuda), drho.to(cuda), dnu.to(cuda), dub.T.contiguous().to(cuda),
            dpb.to(cuda)]
    _, dR = rans._raw.jvp(prim, tans)
    errs = RC.block_errors(dR, ref, nb)
    print("param jvp block errors", errs)
    assert max(errs) < TOL_D


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_rans_vjp_matches_oracle(case, cuda):
    mesh, rans, oset, st = _setup(case, cuda, seed=3)
    m, g, s, t = oset
    d, C, nb = mesh.dim, mesh.n_cells, mesh.dim + 3
    rng = np.random.default_rng(11)
    lam = torch.tensor(rng.normal(size=st.numel()))
    x = st.clone().requires_grad_(True)
    mu = torch.full((C,), case[4], dtype=F64, requires_grad=True)
    rho = torch.ones(C, dtype=F64, requires_grad=True)
    nu = t.nu.clone().requires_grad_(True)
    ub = s.u_b.clone().requires_grad_(True)
    pb = s.p_b.clone().requires_grad_(True)
    R = RC.oracle_residual(oset, x, mu, rho, case[2], nu=nu, u_b=ub, p_b=pb)
    gx, gmu, grho, gnu, gub, gpb = torch.autograd.grad(R, (x, mu, rho, nu, ub, pb), lam)
    prim = rans._detached(st.to(cuda))
    bars = rans._raw.vjp(prim, lam.to(cuda))
    errs = RC.block_errors(bars[0], gx, nb)
    print("vjp state block errors", errs)
    assert max(errs) < TOL_D
    for name, got, ref in (("mu", bars[1], gmu), ("rho", bars[2], grho), ("nu", bars[3], gnu),
                           ("u_b", bars[4], gub.T), ("p_b", bars[5], gpb)):
        scale = max(float(ref.abs().max()), 1e-300)
        err = float((got.cpu() - ref).abs().max()) / scale
        print(name, err, scale)
        assert err < TOL_D, name


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_rans_fused_residual_vjp_matches_separate_calls(case, cuda):
    """afx_rans_residual_vjp returns the cotangents of afx_rans_vjp bitwise and the residual of
    afx_rans_residual to rounding (its k / omega rows come from the adjoint kernel's face loop)."""
    mesh, rans, oset, st = _setup(case, cuda, seed=8)
    rng = np.random.default_rng(12)
    lam = torch.tensor(rng.normal(size=st.numel()), device=cuda)
    prim = rans._detached(st.to(cuda))
    R0 = rans._raw.residual(*prim)
    b0 = rans._raw.vjp(prim, lam)
    R1, b1 = rans._raw.residual_vjp(prim, lam)
    assert max(RC.block_errors(R1, R0, mesh.dim + 3)) < 1e-14
    for x, y in zip(b0, b1):
        assert torch.equal(x, y)


@pytest.mark.parametrize("case", CASES[:2], ids=_ids)
def test_rans_adjoint_dot_product(case, cuda):
    """<lambda, J v> = <J^T lambda, v> on the GPU alone, to rounding."""
    mesh, rans, oset, st = _setup(case, cuda, seed=4)
    rng = np.random.default_rng(5)
    st_d = st.to(cuda)
    v = torch.tensor(rng.normal(size=st.numel()), device=cuda) * st_d.abs().clamp_min(1e-3)
    lam = torch.tensor(rng.normal(size=st.numel()), device=cuda)
    _, Jv = rans.residual_jvp(st_d, v)
    JTl = rans.residual_vjp(st_d, lam)
    a, b = float(lam @ Jv), float(JTl @ v)
    assert abs(a - b) <= 1e-11 * (lam.abs() @ Jv.abs()).item()


def test_rans_autograd_through_inlet_velocity(cuda):
    """Reverse-mode autograd through CoupledRANS.residual reaches a tensor-valued inlet velocity."""
    from aquaflux_amd.boundary import BoundaryConditions
    from aquaflux_amd.flow import VelocityInlet

    mesh, rans, oset, st = _setup(("channel2d_sst", "limited", "limited", "lsq", 1e-4), cuda, seed=6)
    uin = torch.tensor([1.0, 0.0], dtype=F64, device=cuda, requires_grad=True)
    conds = dict(rans.momentum.boundary.conditions)
    conds["left"] = VelocityInlet(uin)
    mom = type(rans.momentum)(**rans.momentum.__dict__)
    mom.boundary = BoundaryConditions(conds).resolve(mesh.face_patches)
    rans2 = type(rans).build(mom, rans.turbulence)
    R = rans2.residual(st.to(cuda))
    C = mesh.n_cells
    obj = (R[:C] ** 2).sum()
    (g,) = torch.autograd.grad(obj, uin)
    # oracle: same objective through the oracle residual with u_b as a function of the inlet vector
    m, g_, s, t = oset
    u0 = torch.tensor([1.0, 0.0], dtype=F64, requires_grad=True)
    left = torch.from_numpy(np.asarray(mesh.face_patches.indices("left"), dtype=np.int64))
    ub = s.u_b.index_put((left,), u0[None, :].expand(left.numel(), 2))
    Ro = RC.oracle_residual(oset, st, torch.full((C,), 1e-4, dtype=F64), torch.ones(C, dtype=F64), "limited", u_b=ub)
    (go,) = torch.autograd.grad((Ro[:C] ** 2).sum(), u0)
    assert float((g.cpu() - go).abs().max() / go.abs().max()) < TOL_D


# --- LogScalars: k and / or omega solved in log form ---------------------------------------------
LOG_CASES = [
    ("channel2d_sst", "limited", "limited", "lsq", 1e-5, False, True),
    ("channel3d_sst", "limited", "limited", "lsq", 1e-4, True, True),
    ("cavity_sst", "upwind", "upwind", "gg", 1e-2, True, False),
]


def _lids(c):
    return _ids(c[:5]) + ("-logk" if c[5] else "") + ("-logw" if c[6] else "")


def _log_setup(case, cuda, seed):
    name, fs, ss, gr, nu, kl, wl = case
    print(f"case {_lids(case)} seed {seed}", flush=True)
    mesh, rans, oset = RC.build(name, fs, ss, gr, nu=nu, device=cuda, k_log=kl, omega_log=wl)
    kind = "cavity" if name.startswith("cavity") else "channel"
    st = RC.to_solved(torch.tensor(RC.seeded_state(mesh, seed, kind)), mesh.dim, kl, wl)
    return mesh, rans, oset, st


@pytest.mark.parametrize("case", LOG_CASES, ids=_lids)
def test_rans_log_residual_jvp_vjp_match_oracle(case, cuda):
    mesh, rans, oset, st = _log_setup(case, cuda, seed=5)
    kl, wl = case[5], case[6]
    C, nb = mesh.n_cells, mesh.dim + 3
    mu = torch.full((C,), case[4], dtype=F64)
    rho = torch.ones(C, dtype=F64)

    def f(z):
        return RC.oracle_residual(oset, z, mu, rho, case[2], k_log=kl, omega_log=wl)

    ref = f(st)
    got = rans.residual(st.to(cuda))
    errs = RC.block_errors(got, ref, nb)
    print("log residual block errors", errs)
    assert max(errs) < TOL_R
    rng = np.random.default_rng(21)
    v = torch.tensor(rng.normal(size=st.numel()))
    _, ref_t = torch.func.jvp(f, (st,), (v,))
    R, dR = rans.residual_jvp(st.to(cuda), v.to(cuda))
    errs = RC.block_errors(dR, ref_t, nb)
    print("log jvp block errors", errs)
    assert max(errs) < TOL_D
    lam = torch.tensor(rng.normal(size=st.numel()))
    x = st.clone().requires_grad_(True)
    (gx,) = torch.autograd.grad(f(x), x, lam)
    bar = rans.residual_vjp(st.to(cuda), lam.to(cuda))
    errs = RC.block_errors(bar, gx, nb)
    print("log vjp block errors", errs)
    assert max(errs) < TOL_D
    R1, b1 = rans._raw.residual_vjp(rans._detached(st.to(cuda)), lam.to(cuda))
    assert torch.equal(b1[0], bar)
    assert max(RC.block_errors(R1, ref, nb)) < TOL_R


@pytest.mark.parametrize("case", LOG_CASES[:1], ids=_lids)
def test_rans_log_omega_keeps_transport_rows_and_linear_fixation(case, cuda):
    """Under log(omega) the transport rows equal the direct residual at the same physical state
    (test_coupled_rans.py:547-570) and the wall fixation rows have unit slope in the solved unknown
    (fixed_value.py:96-124)."""
    mesh, rans, oset, st = _log_setup(case, cuda, seed=6)
    d, C = mesh.dim, mesh.n_cells
    direct_mesh, direct, _ = RC.build(*case[:4], nu=case[4], device=cuda)
    phys = rans.physical_fields(st.to(cuda))
    st_direct = direct.state_from_physical(*phys)
    Rl = rans.residual(st.to(cuda)).reshape(d + 3, C)
    Rd = direct.residual(st_direct).reshape(d + 3, C)
    wall = oset[3].wall_cells
    interior = torch.ones(C, dtype=torch.bool)
    interior[wall] = False
    # (the physical states agree to an ulp of exp(log(.)), hence rounding-level tolerances)
    assert max(RC.block_errors(Rl[: d + 2], Rd[: d + 2], d + 2)) < TOL_R
    scale = float(Rd[d + 2][interior].abs().max())
    assert float((Rl[d + 2][interior] - Rd[d + 2][interior]).abs().max()) < TOL_R * scale
    # d R_w / d w_omega on a wall cell is exactly 1 (log-ratio row), checked through the JVP
    e = torch.zeros((d + 3) * C, dtype=F64, device=cuda)
    cw = int(wall[0])
    e[(d + 2) * C + cw] = 1.0
    _, Je = rans.residual_jvp(st.to(cuda), e)
    assert abs(float(Je[(d + 2) * C + cw]) - 1.0) < 1e-14


def test_rans_outlet_objective_vjp_matches_central_differences(cuda):
    """The 3D SST channel's objective J = mean over outlet faces of u^2 (outlet face velocity = owner
    value): with lambda = dJ/dx plus a seeded N(0, 0.1) part (on this short channel the outlet seed
    alone does not reach the inlet rows), the HIP VJP of <lambda, R(x; u_in, nu)> with respect to the
    state, the inlet velocity vector and the molecular viscosity matches central differences of the
    HIP residual to 1e-6 relative."""
    name, fs, ss, gr, nu = ("channel3d_sst", "limited", "limited", "lsq", 1e-4)
    mesh, rans, oset, st = _setup((name, fs, ss, gr, nu), cuda, seed=8)
    C, d = mesh.n_cells, mesh.dim
    out = torch.from_numpy(np.asarray(mesh.face_patches.indices("right"), dtype=np.int64))
    own = torch.from_numpy(mesh.owner.astype(np.int64))[out]
    x = st.to(cuda)
    lam = torch.zeros_like(x)
    lam[:C].index_add_(0, own.to(cuda), 2.0 * x[:C][own.to(cuda)] / out.numel())  # dJ/du at the owners
    lam += 0.1 * torch.tensor(np.random.default_rng(10).normal(size=x.numel()), device=cuda)
    left = torch.from_numpy(np.asarray(mesh.face_patches.indices("left"), dtype=np.int64))

    def phi(x_, uin, nu_):
        from aquaflux_amd.boundary import BoundaryConditions
        from aquaflux_amd.flow import VelocityInlet

        _, r_, _ = RC.build(name, fs, ss, gr, nu=nu_, device=cuda)
        conds = dict(r_.momentum.boundary.conditions)
        conds["left"] = VelocityInlet(tuple(float(u) for u in uin))
        mom = type(r_.momentum)(**r_.momentum.__dict__)
        mom.boundary = BoundaryConditions(conds).resolve(mesh.face_patches)
        return float(lam @ type(r_).build(mom, r_.turbulence).residual_raw(x_))

    bars = rans._raw.vjp(rans._detached(x), lam)
    uin0 = np.array([1.0, 0.0, 0.0])
    g_nu = float(bars[1].sum() + bars[3].sum())           # mu = rho nu with rho = 1, and the SST nu
    g_uin = bars[4][:, left.to(cuda)].sum(dim=1).cpu().numpy()
    rng = np.random.default_rng(9)
    v = torch.tensor(rng.normal(size=x.numel()), device=cuda) * x.abs().clamp_min(1e-3)
    checks = []
    h = 1e-6
    checks.append(("state", float(bars[0] @ v), (phi(x + h * v, uin0, nu) - phi(x - h * v, uin0, nu)) / (2 * h)))
    hn = 1e-4 * nu  # nu = 1e-4: a smaller step is dominated by rounding of <lambda, R>
    checks.append(("nu", g_nu, (phi(x, uin0, nu + hn) - phi(x, uin0, nu - hn)) / (2 * hn)))
    for i in range(d):
        e = np.zeros(d)
        e[i] = 1e-6
        checks.append((f"u_in[{i}]", float(g_uin[i]), (phi(x, uin0 + e, nu) - phi(x, uin0 - e, nu)) / 2e-6))
    scale = max(abs(c[2]) for c in checks)
    for label, ad, fd in checks:
        print(label, ad, fd, flush=True)
        assert abs(fd) > 1e-8 * scale, label  # every derivative is exercised
        assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-3 * scale), label


def test_rans_pseudo_time_diagonal(cuda):
    """The pseudo-transient shift basis of the coupled RANS (turbulence/coupled.py:607-640): a_P on
    the velocity rows, zero on continuity, a_P / rho times d(phi)/d(w) on the k and omega rows."""
    mesh, rans, _ = RC.build("channel2d_sst", "limited", "limited", "lsq", nu=1e-3, device=cuda, k_log=False,
                             omega_log=True)
    st = RC.to_solved(torch.tensor(RC.seeded_state(mesh, 3, "channel")), mesh.dim, False, True).to(cuda)
    d, C = mesh.dim, mesh.n_cells
    diag = rans.pseudo_time_diagonal(st).reshape(d + 3, C)
    aP = rans.closure_fields(st)["a_p"][0]
    assert torch.equal(diag[0], aP) and torch.equal(diag[1], aP)
    assert torch.equal(diag[d], torch.zeros_like(aP))
    assert torch.allclose(diag[d + 1], aP, rtol=1e-15, atol=0.0)  # rho = 1, direct k
    omega = torch.exp(st.reshape(d + 3, C)[d + 2])
    assert torch.allclose(diag[d + 2], aP * omega, rtol=1e-14, atol=0.0)


# --- node coordinates -----------------------------------------------------------------------------
GEO_CASES = [
    ("cavity_sst", "upwind", "upwind", "gg", 1e-2),
    ("channel2d_sst", "limited", "limited", "lsq", 1e-5),
    ("channel3d_sst", "limited", "upwind", "lsq", 1e-5),
    ("box3d_sst", "limited", "limited", "gg", 1e-3),
    ("brick3d_sst", "limited", "limited", "lsq", 1e-4),
    ("channel2d_sst", "limited", "limited", "cgg", 1e-5),
    ("channel3d_sst", "upwind", "upwind", "cgg6r0.9", 1e-4),
]


@pytest.mark.parametrize("case", GEO_CASES, ids=_ids)
def test_rans_node_coordinate_tangent_matches_oracle(case, cuda):
    """dR/dX . Xdot with the geometry, the wall distance (nearest wall-face centroid) and its gradient
    moving with the nodes -- dual-number geometry through every SST pass -- against torch
    forward-mode AD of the oracle built the way a differentiated build runs (traced wall distance)."""
    import torch.autograd.forward_ad as fwAD

    name, fs, ss, gr, nu = case
    print("case", _ids(case), flush=True)
    base = RC.CASES[name][0]
    mesh0, _ = RC.flow_cases.oracle_case(base, fs, gr)
    st = torch.tensor(RC.seeded_state(mesh0, 3, "cavity" if name.startswith(("cavity", "box")) else "channel"))
    mesh, f = RC.oracle_node_function(name, fs, ss, gr, st, nu=nu)
    dX = torch.tensor(np.random.default_rng(5).normal(size=mesh.node_coords.shape)) * 1e-3
    R_ref, dR_ref = torch.func.jvp(f, (torch.tensor(mesh.node_coords, dtype=F64),), (dX,))
    with fwAD.dual_level():
        Xd = fwAD.make_dual(torch.tensor(mesh.node_coords, dtype=F64, device=cuda), dX.to(cuda))
        _, rans, _ = RC.build(name, fs, ss, gr, nu=nu, device=cuda, node_coords=Xd)
        out = fwAD.unpack_dual(rans.residual(st.to(cuda)))
        R, dR = out.primal.clone(), out.tangent.clone()
    nb = mesh.dim + 3
    assert max(RC.block_errors(R, R_ref, nb)) < TOL_R
    errs = RC.block_errors(dR, dR_ref, nb)
    print("node tangent block errors", errs)
    assert max(errs) < TOL_D


@pytest.mark.parametrize("case", GEO_CASES, ids=_ids)
def test_rans_node_coordinate_adjoint_matches_oracle(case, cuda):
    """d <lambda, R(state; X)> / dX through the hand-written RANS adjoint with geometry, wall-distance
    and wall-distance-gradient cotangents (afx_rans_vjp_geometry, then the distance and gradient
    set-up and the geometry adjoints), against torch autograd of the oracle's differentiated build,
    and <J Xdot, lambda> against the forward-mode tangent of the same HIP path."""
    import torch.autograd.forward_ad as fwAD

    name, fs, ss, gr, nu = case
    print("case", _ids(case), flush=True)
    base = RC.CASES[name][0]
    mesh0, _ = RC.flow_cases.oracle_case(base, fs, gr)
    st = torch.tensor(RC.seeded_state(mesh0, 3, "cavity" if name.startswith(("cavity", "box")) else "channel"))
    mesh, f = RC.oracle_node_function(name, fs, ss, gr, st, nu=nu)
    rng = np.random.default_rng(8)
    nb = mesh.dim + 3
    scale = torch.tensor(np.repeat([1.0] * (nb - 2) + [1e3, 1e-2], mesh.n_cells))  # comparable k / omega rows
    lam = torch.tensor(rng.normal(size=st.numel())) * scale
    Xo = torch.tensor(mesh.node_coords, dtype=F64, requires_grad=True)
    (ref,) = torch.autograd.grad((f(Xo) * lam).sum(), Xo)
    X = torch.tensor(mesh.node_coords, dtype=F64, device=cuda, requires_grad=True)
    _, rans, _ = RC.build(name, fs, ss, gr, nu=nu, device=cuda, node_coords=X)
    (got,) = torch.autograd.grad((rans.residual(st.to(cuda)) * lam.to(cuda)).sum(), X)
    err = float((got.cpu() - ref).abs().max() / ref.abs().max())
    print("node adjoint err", err, float(ref.abs().max()))
    assert err < TOL_D
    dX = torch.tensor(rng.normal(size=mesh.node_coords.shape)) * 1e-3
    with fwAD.dual_level():
        Xd = fwAD.make_dual(torch.tensor(mesh.node_coords, dtype=F64, device=cuda), dX.to(cuda))
        _, rd, _ = RC.build(name, fs, ss, gr, nu=nu, device=cuda, node_coords=Xd)
        dR = fwAD.unpack_dual(rd.residual(st.to(cuda))).tangent.clone()
    a = float((dR.cpu() * lam).sum())
    b = float((got.cpu() * dX).sum())
    assert abs(a - b) <= 1e-11 * float((dR.cpu() * lam).abs().sum())


# --- the OpenMP CPU line ------------------------------------------------------------------------
HOST_CASES = [c for c in CASES if not c[3].startswith("cgg")]


@pytest.mark.parametrize("case", HOST_CASES, ids=_ids)
def test_cpu_line_matches_device(case, cuda):
    """afx_rans_residual_vjp_host (the kernels' cell bodies run by OpenMP threads on host copies, in
    the caller's face order) against the device path: R to 1e-12, every cotangent to 1e-11, per
    field block."""
    mesh, rans, oset, st = _setup(case, cuda, seed=9)
    rng = np.random.default_rng(13)
    lam = torch.tensor(rng.normal(size=st.numel()), device=cuda)
    prim = rans._detached(st.to(cuda))
    R, bars = rans._raw.residual_vjp(prim, lam)
    Rh, barsh = rans._raw.residual_vjp_host(prim, lam, threads=4)
    nb = mesh.dim + 3
    assert max(RC.block_errors(Rh, R.cpu(), nb)) < TOL_R
    assert max(RC.block_errors(barsh[0], bars[0].cpu(), nb)) < TOL_D
    for name, h, d in zip(("mu", "rho", "nu", "u_b", "p_b"), barsh[1:], bars[1:]):
        d = d.cpu()
        err = float((h - d).abs().max() / d.abs().max().clamp_min(1e-300))
        print(name, err)
        assert err < TOL_D, name


def test_cpu_line_log_scalars(cuda):
    """The CPU line with k and omega solved in log form."""
    mesh, rans, _ = RC.build("channel2d_sst", "limited", "limited", "lsq", nu=1e-4, device=cuda, k_log=True,
                             omega_log=True)
    st = torch.tensor(RC.seeded_state(mesh, 2, "channel"))
    st = RC.to_solved(st, mesh.dim, True, True)
    lam = torch.tensor(np.random.default_rng(4).normal(size=st.numel()), device=cuda)
    prim = rans._detached(st.to(cuda))
    R, bars = rans._raw.residual_vjp(prim, lam)
    Rh, barsh = rans._raw.residual_vjp_host(prim, lam, threads=2)
    nb = mesh.dim + 3
    assert max(RC.block_errors(Rh, R.cpu(), nb)) < TOL_R
    assert max(RC.block_errors(barsh[0], bars[0].cpu(), nb)) < TOL_D
