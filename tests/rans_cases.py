"""Shared coupled-RANS (SST) test cases: the same problem built for the HIP path and for the oracle."""

from __future__ import annotations

import dataclasses

import numpy as np
import torch

from aquaflux_amd.boundary import BoundaryConditions, Dirichlet, ZeroGradient
from aquaflux_amd.discretization import FirstOrderUpwind, LimitedUpwind
from aquaflux_amd.schemes import VenkatakrishnanLimiter
from aquaflux_amd.turbulence import CoupledRANS, DirectScalars, LogScalars, SSTModel, SSTTurbulence
from oracle import fv
from oracle import sst as S

from tests import flow_cases

F64 = torch.float64

# name -> (flow mesh/boundary case, wall patches, k inlet value, omega inlet value)
CASES = {
    "cavity_sst": ("cavity", ("top", "bottom", "left", "right"), None, None),
    "channel2d_sst": ("channel2d", ("bottom", "top"), 1e-3, 1.0),
    "channel3d_sst": ("channel3d", ("bottom", "top", "back", "front"), 1e-3, 1.0),
    "box3d_sst": ("box3d", ("top", "bottom", "left", "right", "back", "front"), None, None),
    "brick2d_sst": ("brick2d", ("bottom", "top"), 1e-3, 1.0),
    "brick3d_sst": ("brick3d", ("bottom", "top", "back", "front"), 1e-3, 1.0),
}


def _scalar_bcs(mesh, walls, inlet_value, omega):
    bcs = {}
    for name in mesh.face_patches.names:
        if name in ("interior", "boundary"):
            continue
        if name in walls:
            bcs[name] = ZeroGradient() if omega else Dirichlet(0.0)
        elif name == "left" and inlet_value is not None:
            bcs[name] = Dirichlet(inlet_value)
        else:
            bcs[name] = ZeroGradient()
    return bcs


def _oracle_scalar(mesh, bcs):
    F = mesh.n_faces
    kind = torch.zeros(F, dtype=torch.long)
    value = torch.zeros(F, dtype=F64)
    for name, bc in bcs.items():
        idx = torch.from_numpy(np.asarray(mesh.face_patches.indices(name), dtype=np.int64))
        if isinstance(bc, Dirichlet):
            kind[idx] = fv.BC_DIRICHLET
            value[idx] = float(bc.value)
        else:
            kind[idx] = fv.BC_ZERO_GRADIENT
    return fv.ScalarSetup(kind=kind, value=value, t_inf=torch.zeros(F, dtype=F64))


def build(name, flow_scheme="upwind", scalar_scheme="upwind", gradient="gg", nu=1e-2, rho=1.0, device="cuda",
          k_log=False, omega_log=False, node_coords=None):
    base, walls, k_in, w_in = CASES[name]
    mesh, geom, mc = flow_cases.build(base, flow_scheme, gradient, mu=nu * rho, rho=rho, device=device,
                                      node_coords=node_coords)
    adv = {"upwind": FirstOrderUpwind(), "limited": LimitedUpwind(VenkatakrishnanLimiter(5.0))}[scalar_scheme]
    gs = flow_cases.gradient_scheme(gradient)
    kb = _scalar_bcs(mesh, walls, k_in, omega=False)
    wb = _scalar_bcs(mesh, walls, w_in, omega=True)
    C = mesh.n_cells
    nu_t = torch.full((C,), nu, dtype=F64)
    turb = SSTTurbulence.build(SSTModel(), mesh, geom, gs, adv, rho, nu_t, list(walls), BoundaryConditions(kb),
                               BoundaryConditions(wb))
    rans = CoupledRANS.build(mc, turb, LogScalars() if k_log else DirectScalars(),
                             LogScalars() if omega_log else DirectScalars())
    # the oracle's view
    m = fv.OMesh.from_mesh(mesh)
    g = fv.geometry(m)
    s = flow_cases.oracle_setup(mesh, mc, flow_scheme, gradient)
    return mesh, rans, (m, g, s, _oracle_sst(mesh, m, g, s, walls, kb, wb, nu, rho))


def _oracle_sst(mesh, m, g, s, walls, kb, wb, nu, rho):
    wall_faces = np.concatenate([np.asarray(mesh.face_patches.indices(w), dtype=np.int64) for w in walls])
    return S.build_sst_setup(m, g, S.gradient_spec(s), S.SSTModel(), torch.full((mesh.n_cells,), nu, dtype=F64),
                             rho, wall_faces, _oracle_scalar(mesh, kb), _oracle_scalar(mesh, wb))


def oracle_case(name, flow_scheme="upwind", gradient="gg", nu=1e-2, rho=1.0):
    """The oracle's view of a case without building the HIP operator (no GPU needed): the mesh and
    ``(OMesh, OGeometry, FlowSetup, SSTSetup)``."""
    base, walls, k_in, w_in = CASES[name]
    mesh, (m, g, s) = flow_cases.oracle_case(base, flow_scheme, gradient)
    kb = _scalar_bcs(mesh, walls, k_in, omega=False)
    wb = _scalar_bcs(mesh, walls, w_in, omega=True)
    return mesh, (m, g, s, _oracle_sst(mesh, m, g, s, walls, kb, wb, nu, rho))


def oracle_node_function(name, flow_scheme, scalar_scheme, gradient, state, nu=1e-2, rho=1.0):
    """X -> R(state; X) of the oracle with everything the node coordinates move rebuilt inside:
    geometry, and the SST set-up's wall distance (the reference's traced path) and its gradient."""
    base, walls, k_in, w_in = CASES[name]
    mesh, (m0, g0, s) = flow_cases.oracle_case(base, flow_scheme, gradient)
    kb = _oracle_scalar(mesh, _scalar_bcs(mesh, walls, k_in, omega=False))
    wb = _oracle_scalar(mesh, _scalar_bcs(mesh, walls, w_in, omega=True))
    wall_faces = np.concatenate([np.asarray(mesh.face_patches.indices(w), dtype=np.int64) for w in walls])
    C = mesh.n_cells
    nu_t, mu, rh = torch.full((C,), nu, dtype=F64), torch.full((C,), nu * rho, dtype=F64), torch.full((C,), rho,
                                                                                                       dtype=F64)

    def f(X):
        m = fv.OMesh.from_mesh(mesh, node_coords=X)
        g = fv.geometry(m)
        t = S.build_sst_setup(m, g, S.gradient_spec(s), S.SSTModel(), nu_t, rho, wall_faces, kb, wb, traced=True)
        return S.rans_residual(m, g, s, t, state, mu, rh, scalar_scheme)

    return mesh, f


def seeded_state(mesh, seed=0, kind="channel"):
    """Turbulent-like seeded state: u = 1 + 0.05 N(0,1), v(,w) = 0.05 N(0,1), p = N(0, 0.1),
    k ~ LogU(1e-5, 1e-2), omega ~ LogU(1, 1e3); the cavity takes U(-0.5, 0.5) velocities."""
    rng = np.random.default_rng(seed)
    C, d = mesh.n_cells, mesh.dim
    if kind == "cavity":
        vel = [rng.uniform(-0.5, 0.5, C) for _ in range(d)]
    else:
        vel = [1.0 + 0.05 * rng.normal(size=C)] + [0.05 * rng.normal(size=C) for _ in range(d - 1)]
    p = rng.normal(0.0, 0.1, C)
    k = np.exp(rng.uniform(np.log(1e-5), np.log(1e-2), C))
    w = np.exp(rng.uniform(np.log(1.0), np.log(1e3), C))
    return np.concatenate(vel + [p, k, w])


def oracle_residual(setup, state, mu, rho, scalar_scheme, nu=None, u_b=None, p_b=None, k_log=False,
                    omega_log=False):
    m, g, s, t = setup
    if nu is not None:
        t = dataclasses.replace(t, nu=nu)
    if u_b is not None or p_b is not None:
        s = dataclasses.replace(s, u_b=s.u_b if u_b is None else u_b, p_b=s.p_b if p_b is None else p_b)
    return S.rans_residual(m, g, s, t, state, mu, rho, scalar_scheme, k_log=k_log, omega_log=omega_log)


def to_solved(state, dim, k_log, omega_log):
    """Physical seeded state -> solved state (log of the transformed k / omega blocks)."""
    b = state.reshape(dim + 3, -1).clone()
    if k_log:
        b[dim + 1] = torch.log(b[dim + 1])
    if omega_log:
        b[dim + 2] = torch.log(b[dim + 2])
    return b.reshape(-1)


def block_errors(a, b, nblocks):
    """Per-block max|a-b| / max|b| (each field block normalised by its own scale)."""
    a = a.detach().double().cpu().reshape(nblocks, -1)
    b = b.detach().double().cpu().reshape(nblocks, -1)
    return [float((a[i] - b[i]).abs().max() / max(float(b[i].abs().max()), 1e-300)) for i in range(nblocks)]
