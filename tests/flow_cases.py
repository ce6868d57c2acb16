"""Shared flow test cases: the same problem built for the HIP path and for the CPU oracle."""

from __future__ import annotations

import numpy as np
import torch

from aquaflux_amd.boundary import BoundaryConditions
from aquaflux_amd.discretization import FirstOrderUpwind, LimitedUpwind
from aquaflux_amd.flow import MomentumContinuity, MovingWall, NoSlipWall, PressureOutlet, VelocityInlet
from aquaflux_amd.mesh import perturbed_grid_2d, perturbed_grid_3d, structured_grid_2d, structured_grid_3d
from aquaflux_amd.properties import Constant, PropertyModel
from aquaflux_amd.schemes import (
    CompactGreenGauss,
    CorrectedGreenGauss,
    LeastSquaresGradient,
    SweptGradientSolve,
    VenkatakrishnanLimiter,
)
from oracle import fv

F64 = torch.float64


def make_mesh(name):
    if name == "cavity":
        return structured_grid_2d(12, 10, named_boundaries=True)
    if name == "cavity64":
        return structured_grid_2d(64, 64, named_boundaries=True)
    if name == "channel2d":
        return perturbed_grid_2d(16, 8, lx=4.0, ly=1.0, perturb=0.2, seed=0, named_boundaries=True)
    if name == "channel3d":
        return perturbed_grid_3d(6, 4, 4, lx=2.0, perturb=0.15, seed=1, named_boundaries=True)
    if name == "channel3d_mg":  # large enough for a three-level Schur multigrid
        return perturbed_grid_3d(32, 16, 16, lx=2.0, perturb=0.15, seed=1, named_boundaries=True)
    if name == "box3d":
        return structured_grid_3d(5, 4, 3, named_boundaries=True)
    if name in ("brick2d", "brick3d"):
        # general polygons / polyhedra (hexagonal bricks, eight-faced prisms) read from a polyMesh
        from tests.polymesh_cases import brick_mesh

        return brick_mesh(2 if name == "brick2d" else 3)
    raise KeyError(name)


def flow_bcs(name, dim, inlet_velocity=None):
    if name.startswith("cavity") or name == "box3d":
        lid = (1.0, 0.0) if dim == 2 else (1.0, 0.0, 0.0)
        bcs = {"top": MovingWall(lid), "bottom": NoSlipWall(), "left": NoSlipWall(), "right": NoSlipWall()}
        if dim == 3:
            bcs.update({"back": NoSlipWall(), "front": NoSlipWall()})
        return bcs, 0
    uin = inlet_velocity if inlet_velocity is not None else ((1.0, 0.0) if dim == 2 else (1.0, 0.0, 0.0))
    bcs = {"left": VelocityInlet(uin), "right": PressureOutlet(0.0), "bottom": NoSlipWall(), "top": NoSlipWall()}
    if dim == 3:
        bcs.update({"back": NoSlipWall(), "front": NoSlipWall()})
    return bcs, None


def build(name, scheme="upwind", gradient="gg", mu=0.01, rho=1.0, device="cuda", inlet_velocity=None,
          node_coords=None):
    mesh = make_mesh(name)
    geom = mesh.geometry(node_coords=node_coords, device=device)
    bcs, pin = flow_bcs(name, mesh.dim, inlet_velocity)
    adv = {"none": None, "upwind": FirstOrderUpwind(), "limited": LimitedUpwind(VenkatakrishnanLimiter(5.0))}[scheme]
    gs = gradient_scheme(gradient)
    props = PropertyModel({"viscosity": Constant(mu), "density": Constant(rho)})
    mc = MomentumContinuity.build(mesh, geom, props, gs, BoundaryConditions(bcs), advection_scheme=adv, pressure_pin=pin)
    return mesh, geom, mc


def oracle_case(name, scheme="upwind", gradient="gg", inlet_velocity=None):
    """The oracle's view of a case without building the HIP operator (no GPU needed): the mesh and
    ``(OMesh, OGeometry, FlowSetup)``."""
    import types

    mesh = make_mesh(name)
    bcs, pin = flow_bcs(name, mesh.dim, inlet_velocity)
    view = types.SimpleNamespace(boundary=BoundaryConditions(bcs).resolve(mesh.face_patches), pressure_pin=pin)
    m = fv.OMesh.from_mesh(mesh)
    return mesh, (m, fv.geometry(m), oracle_setup(mesh, view, scheme, gradient))


def gradient_scheme(gradient):
    """'gg' compact Green-Gauss, 'lsq' least squares, 'cgg' / 'cgg<n>' / 'cgg<n>r<relax>' corrected."""
    if gradient == "gg":
        return CompactGreenGauss()
    if gradient == "lsq":
        return LeastSquaresGradient()
    sweeps, relax = gradient_sweeps(gradient)
    return CorrectedGreenGauss(SweptGradientSolve(sweeps=sweeps, relaxation=relax))


def gradient_sweeps(gradient):
    rest = gradient[3:]
    relax = 1.0
    if "r" in rest:
        rest, r = rest.split("r")
        relax = float(r)
    return (int(rest) if rest else 4), relax


def oracle_gradient(gradient):
    return {"gg": "green_gauss", "lsq": "least_squares"}.get(gradient, "corrected")


def oracle_setup(mesh, mc, scheme, gradient):
    """The oracle's view of the same boundary data."""
    d, F = mesh.dim, mesh.n_faces
    kind = torch.zeros(F, dtype=torch.long)
    ub = torch.zeros((F, d), dtype=F64)
    pb = torch.zeros(F, dtype=F64)
    for _, bc, faces in mc.boundary.items():
        idx = torch.from_numpy(faces)
        kind[idx] = bc.kind
        v = bc.prescribed_velocity(d)
        if v is not None:
            ub[idx] = v.detach().cpu()
        p = bc.prescribed_pressure()
        if p is not None:
            pb[idx] = float(p)
    return fv.FlowSetup(
        kind=kind,
        u_b=ub,
        p_b=pb,
        advection=scheme,
        gradient=oracle_gradient(gradient),
        pressure_pin=mc.pressure_pin,
        **(dict(zip(("gradient_sweeps", "gradient_relaxation"), gradient_sweeps(gradient)))
           if gradient.startswith("cgg") else {}),
    )


def seeded_state(mesh, seed=0, kind="cavity"):
    """Seeded states of the survey's recipes: U(-1,1) velocity + N(0,0.1) pressure for the cavity;
    u = 1 + 0.05 N(0,1), v(,w) = 0.05 N(0,1), p = N(0, 0.1) for channels."""
    rng = np.random.default_rng(seed)
    C, d = mesh.n_cells, mesh.dim
    if kind == "cavity":
        vel = [rng.uniform(-1, 1, C) for _ in range(d)]
    else:
        vel = [1.0 + 0.05 * rng.normal(size=C)] + [0.05 * rng.normal(size=C) for _ in range(d - 1)]
    p = rng.normal(0.0, 0.1, C)
    return np.concatenate(vel + [p])


def rel_err(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).abs().max() / max(float(b.abs().max()), 1e-300))
