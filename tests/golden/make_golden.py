"""Writes the golden fixtures under tests/golden/: small seeded cases evaluated by the oracle.

Each ``<case>.npz`` holds the mesh the case's generator produces (owner, neighbour, node
coordinates: a later change of a generator shows up as a fixture mismatch), the seeded inputs
(state, tangent, cotangent, history levels), and the oracle's outputs: the residual, its
tangent along the stored direction and its cotangents with respect to the state and the
parameters.  The reference ships no numeric golden arrays of its own (its tests hold inline closed
forms, pinned in tests/test_oracle_pins.py), so these vectors come from the oracle, which those
pins fix.

    python tests/golden/make_golden.py            # rewrite every fixture
    python tests/golden/make_golden.py flow_cavity # one case
"""

from __future__ import annotations

import dataclasses
import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import fv  # noqa: E402
from tests import flow_cases as FC  # noqa: E402
from tests import rans_cases as RC  # noqa: E402
from tests import scalar_cases as SC  # noqa: E402

F64 = torch.float64
HERE = os.path.dirname(os.path.abspath(__file__))
DT = 0.037

# name -> (kind, spec); a flow spec may end with the state seed (default 20)
FLOW = {
    "flow_cavity_upwind_gg": ("cavity", "upwind", "gg", "cavity"),
    # the 64 x 64 lid-driven cavity at full size, state seed 0 (U(-1,1) u, v then N(0, 0.1) p)
    "flow_cavity64_upwind_gg": ("cavity64", "upwind", "gg", "cavity", 0),
    "flow_channel2d_limited_lsq": ("channel2d", "limited", "lsq", "channel"),
    "flow_channel3d_limited_cgg": ("channel3d", "limited", "cgg", "channel"),
}
RANS = {
    "rans_channel2d_limited_lsq": ("channel2d_sst", "limited", "limited", "lsq", 1e-5, False, False),
    "rans_box3d_limited_gg": ("box3d_sst", "limited", "limited", "gg", 1e-3, False, False),
    "rans_channel3d_limited_lsq_logkw": ("channel3d_sst", "limited", "limited", "lsq", 1e-4, True, True),
}
SCALAR = {
    "scalar_tet3_limited_gg": ("tet3", "limited", "gg", None),
    "scalar_quad_mixed_limited_cgg_bdf2": ("quad_mixed", "limited", "cgg", False),
    "scalar_tet3_upwind_lsq_bdf1": ("tet3", "upwind", "lsq", True),
}
CASES = {**{k: ("flow", v) for k, v in FLOW.items()}, **{k: ("rans", v) for k, v in RANS.items()},
         **{k: ("scalar", v) for k, v in SCALAR.items()}}


def mesh_arrays(mesh):
    return {
        "owner": np.asarray(mesh.owner, dtype=np.int32),
        "neighbour": np.asarray(mesh.neighbour, dtype=np.int32),
        "node_coords": np.asarray(mesh.node_coords, dtype=np.float64),
    }


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64)


# fixtures above this size are full-size cases, listed here with their cap
LARGE = {"flow_cavity64_upwind_gg": 1 << 20}


def flow_case(spec):
    name, scheme, gradient, kind = spec[:4]
    seed = spec[4] if len(spec) > 4 else 20
    mesh, (m, g, s) = FC.oracle_case(name, scheme, gradient)
    C, F, d = mesh.n_cells, mesh.n_faces, mesh.dim
    rng = np.random.default_rng(20)
    x = torch.tensor(FC.seeded_state(mesh, seed=seed, kind=kind))
    v = torch.tensor(rng.normal(size=x.numel()))
    w = torch.tensor(rng.normal(size=x.numel()))
    mu = torch.full((C,), 0.01, dtype=F64)
    rho = torch.ones(C, dtype=F64)

    def f(x_, mu_, rho_, ub_, pb_):
        return fv.flow_residual(m, g, dataclasses.replace(s, u_b=ub_, p_b=pb_), x_, mu_, rho_)

    R, Jv = torch.func.jvp(lambda z: f(z, mu, rho, s.u_b, s.p_b), (x,), (v,))
    leaves = [t.clone().requires_grad_(True) for t in (x, mu, rho, s.u_b, s.p_b)]
    bars = torch.autograd.grad(f(*leaves), leaves, w, allow_unused=True)
    bars = [torch.zeros_like(lf) if b is None else b for b, lf in zip(bars, leaves)]
    return dict(mesh_arrays(mesh), state=_np(x), tangent=_np(v), cotangent=_np(w), mu=_np(mu), rho=_np(rho),
                residual=_np(R), jvp=_np(Jv), vjp_state=_np(bars[0]), vjp_mu=_np(bars[1]), vjp_rho=_np(bars[2]),
                vjp_u_b=_np(bars[3]), vjp_p_b=_np(bars[4]), n_fields=np.int32(d + 1), n_faces=np.int32(F))


def rans_case(spec):
    name, fs, ss, gr, nu, k_log, w_log = spec
    mesh, oset = RC.oracle_case(name, fs, gr, nu=nu)
    m, g, s, t = oset
    C, d = mesh.n_cells, mesh.dim
    kind = "cavity" if name.startswith(("cavity", "box")) else "channel"
    rng = np.random.default_rng(21)
    x = RC.to_solved(torch.tensor(RC.seeded_state(mesh, 21, kind)), d, k_log, w_log)
    v = torch.tensor(rng.normal(size=x.numel())) * x.abs().clamp_min(1e-3)
    w = torch.tensor(rng.normal(size=x.numel()))
    mu = torch.full((C,), nu, dtype=F64)
    rho = torch.ones(C, dtype=F64)

    def f(x_, mu_, rho_, nu_, ub_, pb_):
        return RC.oracle_residual(oset, x_, mu_, rho_, ss, nu=nu_, u_b=ub_, p_b=pb_, k_log=k_log, omega_log=w_log)

    R, Jv = torch.func.jvp(lambda z: f(z, mu, rho, t.nu, s.u_b, s.p_b), (x,), (v,))
    leaves = [a.clone().requires_grad_(True) for a in (x, mu, rho, t.nu, s.u_b, s.p_b)]
    bars = torch.autograd.grad(f(*leaves), leaves, w, allow_unused=True)
    bars = [torch.zeros_like(lf) if b is None else b for b, lf in zip(bars, leaves)]
    return dict(mesh_arrays(mesh), state=_np(x), tangent=_np(v), cotangent=_np(w), mu=_np(mu), rho=_np(rho),
                residual=_np(R), jvp=_np(Jv), vjp_state=_np(bars[0]), vjp_mu=_np(bars[1]), vjp_rho=_np(bars[2]),
                vjp_nu=_np(bars[3]), vjp_u_b=_np(bars[4]), vjp_p_b=_np(bars[5]), n_fields=np.int32(d + 3))


def scalar_case(spec):
    name, adv, grad, first = spec
    mesh, u, bcs = SC.make(name)
    C = mesh.n_cells
    phi, gamma = SC.seeded(mesh, 22)
    old, older = SC.history(mesh, 22)
    rng = np.random.default_rng(22)
    v = torch.tensor(rng.normal(size=C))
    w = torch.tensor(rng.normal(size=C))
    s = SC.oracle_setup(mesh, bcs, adv, grad)
    transient = first is not None
    T = torch.tensor

    def f(ph, ga, X, po, poo):
        m = fv.OMesh.from_mesh(mesh, node_coords=X)
        g = fv.geometry(m)
        if transient:
            return fv.scalar_residual(m, g, s, ph, SC.oracle_flux(g, u), ga, po, poo, DT, first)
        return fv.scalar_residual(m, g, s, ph, SC.oracle_flux(g, u), ga)

    X = T(mesh.node_coords, dtype=F64)
    R, Jv = torch.func.jvp(lambda z: f(z, T(gamma), X, T(old), T(older)), (T(phi),), (v,))
    leaves = [a.clone().requires_grad_(True) for a in (T(phi), T(gamma), X, T(old), T(older))]
    bars = torch.autograd.grad(f(*leaves), leaves, w, allow_unused=True)
    bars = [torch.zeros_like(lf) if b is None else b for b, lf in zip(bars, leaves)]
    out = dict(mesh_arrays(mesh), state=np.asarray(phi), tangent=_np(v), cotangent=_np(w), gamma=np.asarray(gamma),
               velocity=np.asarray(u, dtype=np.float64), residual=_np(R), jvp=_np(Jv), vjp_state=_np(bars[0]),
               vjp_gamma=_np(bars[1]), vjp_node_coords=_np(bars[2]), n_fields=np.int32(1))
    if transient:
        out.update(phi_old=np.asarray(old), phi_older=np.asarray(older), dt=np.float64(DT),
                   first_step=np.bool_(first), vjp_phi_old=_np(bars[3]), vjp_phi_older=_np(bars[4]))
    return out


def evaluate(name):
    """The oracle's fixture contents for one case (also used by the tests to re-derive them)."""
    kind, spec = CASES[name]
    return {"flow": flow_case, "rans": rans_case, "scalar": scalar_case}[kind](spec)


def path(name):
    return os.path.join(HERE, name + ".npz")


def main(names):
    torch.set_num_threads(1)
    for name in names or CASES:
        data = evaluate(name)
        np.savez_compressed(path(name), **data)
        print(name, os.path.getsize(path(name)), "bytes")


if __name__ == "__main__":
    main(sys.argv[1:])
