"""The HIP path against the committed golden fixtures (tests/golden/*.npz): residual to 1e-12 and
tangent / cotangents to 1e-11, each relative to its field block's largest entry, on the stored
meshes, inputs and directions."""

import numpy as np
import pytest
import torch

from tests import flow_cases as FC
from tests import rans_cases as RC
from tests import scalar_cases as SC
from tests.golden import make_golden as MG

pytestmark = pytest.mark.gpu
F64 = torch.float64
TOL_R = 1e-12
TOL_D = 1e-11


def _load(name):
    with np.load(MG.path(name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _check(label, got, ref, nb, tol):
    got = got.detach().double().cpu().numpy().reshape(nb, -1)
    ref = np.asarray(ref, dtype=np.float64).reshape(nb, -1)
    for i in range(nb):
        scale = float(np.abs(ref[i]).max())
        err = float(np.abs(got[i] - ref[i]).max())
        print(label, i, err / max(scale, 1e-300), flush=True)
        assert err <= tol * scale or scale == 0.0 and err < 1e-14, (label, i)


def _same_mesh(mesh, z):
    assert np.array_equal(np.asarray(mesh.owner, dtype=np.int32), z["owner"])
    assert np.array_equal(np.asarray(mesh.neighbour, dtype=np.int32), z["neighbour"])
    assert np.array_equal(np.asarray(mesh.node_coords, dtype=np.float64), z["node_coords"])


@pytest.mark.parametrize("name", sorted(MG.FLOW))
def test_flow_matches_golden(name, cuda):
    print("golden", name, flush=True)
    z = _load(name)
    base, scheme, gradient = MG.FLOW[name][:3]
    mesh, geom, mc = FC.build(base, scheme, gradient, device=cuda)
    _same_mesh(mesh, z)
    nb = int(z["n_fields"])
    T = lambda a: torch.tensor(a, dtype=F64, device=cuda)  # noqa: E731
    x, v, w = T(z["state"]), T(z["tangent"]), T(z["cotangent"])
    R, Jv = mc.residual_jvp(x, v)
    _check("residual", R, z["residual"], nb, TOL_R)
    _check("jvp", Jv, z["jvp"], nb, TOL_D)
    mu, rho, ub, pb, _ = mc._inputs()
    bars = mc._raw.vjp([x, mu.detach(), rho.detach(), ub.detach(), pb.detach()], w)
    _check("vjp state", bars[0], z["vjp_state"], nb, TOL_D)
    _check("vjp mu", bars[1], z["vjp_mu"], 1, TOL_D)
    _check("vjp rho", bars[2], z["vjp_rho"], 1, TOL_D)
    _check("vjp u_b", bars[3].T, z["vjp_u_b"], 1, TOL_D)
    _check("vjp p_b", bars[4], z["vjp_p_b"], 1, TOL_D)


@pytest.mark.parametrize("name", sorted(MG.RANS))
def test_rans_matches_golden(name, cuda):
    print("golden", name, flush=True)
    z = _load(name)
    case, fs, ss, gr, nu, k_log, w_log = MG.RANS[name]
    mesh, rans, _ = RC.build(case, fs, ss, gr, nu=nu, device=cuda, k_log=k_log, omega_log=w_log)
    _same_mesh(mesh, z)
    nb = int(z["n_fields"])
    T = lambda a: torch.tensor(a, dtype=F64, device=cuda)  # noqa: E731
    x, v, w = T(z["state"]), T(z["tangent"]), T(z["cotangent"])
    R, Jv = rans.residual_jvp(x, v)
    _check("residual", R, z["residual"], nb, TOL_R)
    _check("jvp", Jv, z["jvp"], nb, TOL_D)
    bars = rans._raw.vjp(rans._detached(x), w)
    _check("vjp state", bars[0], z["vjp_state"], nb, TOL_D)
    for i, key in enumerate(("vjp_mu", "vjp_rho", "vjp_nu"), start=1):
        _check(key, bars[i], z[key], 1, TOL_D)
    _check("vjp u_b", bars[4].T, z["vjp_u_b"], 1, TOL_D)
    _check("vjp p_b", bars[5], z["vjp_p_b"], 1, TOL_D)


@pytest.mark.parametrize("name", sorted(MG.SCALAR))
def test_scalar_matches_golden(name, cuda):
    import torch.autograd.forward_ad as fwAD

    print("golden", name, flush=True)
    z = _load(name)
    case, adv, grad, first = MG.SCALAR[name]
    mesh, u, bcs = SC.make(case)
    _same_mesh(mesh, z)
    transient = first is not None
    T = lambda a, **kw: torch.tensor(a, dtype=F64, device=cuda, **kw)  # noqa: E731
    X = T(z["node_coords"], requires_grad=True)
    geom, st = SC.build_gpu(mesh, bcs, adv, grad, z["gamma"], cuda, node_coords=X, transient=transient)
    hist = (T(z["phi_old"]), None if first else T(z["phi_older"]), float(z["dt"]), bool(first)) if transient else ()

    # residual and tangent along the stored direction (geometry held fixed)
    fn = st.residual(SC.gpu_flux(geom, u).detach())
    with fwAD.dual_level():
        ph = fwAD.make_dual(T(z["state"]), T(z["tangent"]))
        out = fwAD.unpack_dual(fn(ph, *hist))
        R, Jv = out.primal, out.tangent
    _check("residual", R, z["residual"], 1, TOL_R)
    _check("jvp", Jv, z["jvp"], 1, TOL_D)

    # cotangents: state and node coordinates through the residual and geometry adjoints
    x = T(z["state"], requires_grad=True)
    leaves = [x, X]
    if transient:
        po = T(z["phi_old"], requires_grad=True)
        poo = None if first else T(z["phi_older"], requires_grad=True)
        leaves += [po] + ([] if first else [poo])
        Rg = st.residual(SC.gpu_flux(geom, u))(x, po, poo, float(z["dt"]), bool(first))
    else:
        Rg = st.residual(SC.gpu_flux(geom, u))(x)
    got = list(torch.autograd.grad(Rg, leaves, T(z["cotangent"])))
    _check("vjp state", got.pop(0), z["vjp_state"], 1, TOL_D)
    _check("vjp node coords", got.pop(0), z["vjp_node_coords"], 1, TOL_D)
    if transient:
        _check("vjp phi_old", got.pop(0), z["vjp_phi_old"], 1, TOL_D)
        if not first:
            _check("vjp phi_older", got.pop(0), z["vjp_phi_older"], 1, TOL_D)
    # the diffusivity cotangent from the raw entry (flux held fixed)
    kw = dict(phi_old=hist[0], phi_older=hist[1], dt=hist[2], first_step=hist[3]) if transient else {}
    bars, _ = st.residual_vjp(SC.gpu_flux(geom, u).detach(), x.detach(), T(z["cotangent"]), **kw)
    _check("vjp gamma", bars[1], z["vjp_gamma"], 1, TOL_D)
