"""Shared scalar-transport test cases: the same problem for the HIP path and for the oracle."""

from __future__ import annotations

import numpy as np
import torch

from aquaflux_amd.boundary import BoundaryConditions, Convective, Dirichlet, Neumann, ZeroGradient
from aquaflux_amd.discretization import FirstOrderUpwind, LimitedUpwind, TransientTerm
from aquaflux_amd.mesh import perturbed_grid_2d, tet_grid_3d
from aquaflux_amd.properties import FieldProperty
from aquaflux_amd.schemes import (
    CompactGreenGauss,
    CorrectedGreenGauss,
    LeastSquaresGradient,
    SweptGradientSolve,
    VenkatakrishnanLimiter,
)
from aquaflux_amd.transport import ScalarTransport, face_mass_flux
from oracle import fv

F64 = torch.float64

# name -> (mesh builder, velocity, {patch: closure})
CASES = {
    # the tetrahedral advection-diffusion case at test size (6 * 3^3 tets, jitter 0.15 h, Pe = 100)
    "tet3": (lambda: tet_grid_3d(3, perturb=0.15, seed=3), (1.0, 0.0, 0.0),
             {"left": Dirichlet(1.0), "right": ZeroGradient(), "bottom": ZeroGradient(), "top": ZeroGradient(),
              "back": ZeroGradient(), "front": ZeroGradient()}),
    # every closure kind on a perturbed quad mesh
    "quad_mixed": (lambda: perturbed_grid_2d(10, 6, lx=2.0, ly=1.0, perturb=0.2, seed=0, named_boundaries=True),
                   (1.0, 0.3),
                   {"left": Dirichlet(1.0), "right": ZeroGradient(), "bottom": Neumann(0.7),
                    "top": Convective(2.0, 0.25)}),
    # eight-faced prisms with hexagonal caps, read from an OpenFOAM polyMesh
    "brick3d": (lambda: _brick3d(), (1.0, 0.3, 0.1),
                {"left": Dirichlet(1.0), "right": ZeroGradient(), "bottom": Neumann(0.7), "top": Convective(2.0, 0.25),
                 "back": ZeroGradient(), "front": ZeroGradient()}),
}


def _brick3d():
    from tests.polymesh_cases import brick_mesh

    return brick_mesh(3)

ADV = {"none": None, "upwind": FirstOrderUpwind(), "limited": LimitedUpwind(VenkatakrishnanLimiter(5.0))}
GRAD = {"gg": CompactGreenGauss(), "lsq": LeastSquaresGradient(), "none": None,
        "cgg": CorrectedGreenGauss(SweptGradientSolve(sweeps=5, relaxation=0.9))}
ORACLE_GRAD = {"gg": "green_gauss", "lsq": "least_squares", "none": "none", "cgg": "corrected"}


def make(name):
    builder, u, bcs = CASES[name]
    return builder(), u, bcs


def oracle_setup(mesh, bcs, adv, grad):
    F = mesh.n_faces
    kind = torch.zeros(F, dtype=torch.long)
    value = torch.zeros(F, dtype=F64)
    t_inf = torch.zeros(F, dtype=F64)
    for name, bc in bcs.items():
        idx = torch.from_numpy(np.asarray(mesh.face_patches.indices(name), dtype=np.int64))
        kind[idx] = bc.kind
        v, t = bc.values()
        value[idx] = float(v)
        t_inf[idx] = float(t)
    return fv.ScalarSetup(kind=kind, value=value, t_inf=t_inf, advection=adv, gradient=ORACLE_GRAD[grad],
                          limiter=adv == "limited", limiter_k=5.0, gradient_sweeps=5, gradient_relaxation=0.9)


def oracle_flux(g, u):
    un = sum(float(u[j]) * g.normal[:, j] for j in range(g.normal.shape[1]))
    return un * g.area


def seeded(mesh, seed=0):
    rng = np.random.default_rng(seed)
    C = mesh.n_cells
    phi = rng.uniform(0.0, 1.0, C)
    gamma = 0.01 * (1.0 + 0.5 * rng.uniform(size=C))
    return phi, gamma


def history(mesh, seed=0):
    """Two earlier time levels near phi (the transient cases)."""
    rng = np.random.default_rng(100 + seed)
    C = mesh.n_cells
    return rng.uniform(0.0, 1.0, C), rng.uniform(0.0, 1.0, C)


def build_gpu(mesh, bcs, adv, grad, gamma, device, node_coords=None, transient=False):
    geom = mesh.geometry(node_coords=node_coords, device=device)
    st = ScalarTransport.build(mesh, geom, FieldProperty(torch.tensor(gamma, dtype=F64, device=device)),
                               BoundaryConditions(bcs), ADV[adv], gradient_scheme=GRAD[grad],
                               transient=TransientTerm() if transient else None)
    return geom, st


def gpu_flux(geom, u):
    return face_mass_flux(geom, u)
