"""The allocation-free scalar residual ``afx_scalar_residual_noalloc`` (workspace in a file-scope device
array) against ``afx_scalar_residual`` with a caller workspace, through the C ABI: bitwise equal on two
calls in a row, the second on a smaller mesh (the array keeps the first call's contents), and a mesh
beyond the array's size refused with status 2 before any kernel runs."""

import ctypes

import numpy as np
import pytest
import torch

from aquaflux_amd import _lib
from tests import scalar_cases as SC

pytestmark = pytest.mark.gpu


def _case(name, adv, grad, seed, cuda):
    mesh, u, bcs = SC.make(name)
    phi, gamma = SC.seeded(mesh, seed)
    geom, st = SC.build_gpu(mesh, bcs, adv, grad, gamma, cuda)
    flux = SC.gpu_flux(geom, u).detach().contiguous()
    raw = st._raw
    prim = (torch.tensor(phi, device=cuda), torch.tensor(gamma, device=cuda), flux, st._val.detach().contiguous(),
            st._tinf.detach().contiguous(), None, None)
    geo = tuple(t.detach() for t in st._geometry())
    return raw, prim, geo


def _noalloc(raw, prim, geo):
    R = torch.full_like(prim[0], float("nan"))
    gs = raw._gs(geo)
    ins = _lib.AfxScalarInputs(*[_lib.ptr(t) for t in prim])
    rc = _lib.lib().afx_scalar_residual_noalloc(raw.dm.handle, ctypes.byref(gs), ctypes.byref(raw._cfg(None)),
                                               ctypes.byref(ins), R.data_ptr(), _lib.stream_of(raw.dm.device))
    return rc, R


@pytest.mark.parametrize("first,second", [(("tet3", "limited", "gg"), ("quad_mixed", "upwind", "lsq")),
                                          (("tet3", "limited", "cgg"), ("brick3d", "limited", "lsq"))])
def test_noalloc_residual_equals_the_workspace_residual(cuda, first, second):
    print("case", first, "then", second, flush=True)
    runs = []
    for i, case in enumerate((first, second)):
        raw, prim, geo = _case(*case, seed=i, cuda=cuda)
        runs.append((raw, prim, geo))
    assert runs[1][1][0].numel() < runs[0][1][0].numel()
    for raw, prim, geo in runs:
        rc, R = _noalloc(raw, prim, geo)
        assert rc == 0, _lib.lib().afx_last_error()
        ref = raw.residual(prim, geo)
        torch.cuda.synchronize()
        assert torch.equal(R, ref)


def test_noalloc_refuses_a_mesh_beyond_the_array(cuda):
    from aquaflux_amd.mesh import structured_grid_3d

    cap = ctypes.c_int64()
    assert _lib.lib().afx_scalar_noalloc_max_bytes(ctypes.byref(cap)) == 0
    n = int(np.ceil((cap.value / 88 + 1) ** (1 / 3)))  # 88 bytes of workspace per 3D cell
    mesh = structured_grid_3d(n, n, n, named_boundaries=True)
    print("case structured", n, "^3 =", mesh.n_cells, "cells", flush=True)
    from aquaflux_amd.boundary import BoundaryConditions, Dirichlet, ZeroGradient
    from aquaflux_amd.discretization import FirstOrderUpwind
    from aquaflux_amd.properties import Constant
    from aquaflux_amd.schemes import CompactGreenGauss
    from aquaflux_amd.transport import ScalarTransport, face_mass_flux

    geom = mesh.geometry(device=cuda)
    bcs = {p: (Dirichlet(1.0) if p == "left" else ZeroGradient())
           for p in ("left", "right", "bottom", "top", "back", "front")}
    st = ScalarTransport.build(mesh, geom, Constant(0.01), BoundaryConditions(bcs), FirstOrderUpwind(),
                               gradient_scheme=CompactGreenGauss())
    raw = st._raw
    prim = (torch.zeros(mesh.n_cells, dtype=torch.float64, device=cuda),
            torch.full((mesh.n_cells,), 0.01, dtype=torch.float64, device=cuda),
            face_mass_flux(geom, (1.0, 0.0, 0.0)).detach().contiguous(), st._val.detach().contiguous(),
            st._tinf.detach().contiguous(), None, None)
    rc, R = _noalloc(raw, prim, tuple(t.detach() for t in st._geometry()))
    assert rc == 2
    assert b"more than" in _lib.lib().afx_last_error()
    assert torch.isnan(R).all()
