"""The cell kernels' block order changes scheduling only: residual, tangent and cotangents are
bitwise the same in memory order, in the slab-tile order set from the first geometry, and in an
arbitrary permutation of the blocks."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _orders(nb):
    rng = np.random.default_rng(0)
    return [("memory", None), ("random", rng.permutation(nb).astype(np.int32))]


def test_sst_residual_vjp_is_bitwise_independent_of_block_order(cuda):
    import bench

    print("case sst channel 32x16x16", flush=True)
    mesh, op, x, lam, _, _ = bench.build_workload(cuda, "sst", 32, 16, 16)
    dm = mesh.device_mesh(cuda)
    assert dm.block_order is not None  # set from the first geometry
    raw = op._raw
    prim = op._detached(x)
    v = torch.randn_like(x) * x.abs().clamp_min(1e-3)

    def run():
        R, bars = raw.residual_vjp(prim, lam)
        _, Jv = op.residual_jvp(x, v)
        return torch.cat([R, Jv] + [b.reshape(-1) for b in bars])

    ref = run()
    nb = (mesh.n_cells + 255) // 256
    for name, order in _orders(nb):
        dm.set_block_order(order)
        assert torch.equal(run(), ref), name


def test_tet_scalar_node_adjoint_is_bitwise_independent_of_block_order(cuda):
    from aquaflux_amd.boundary import BoundaryConditions, Dirichlet, ZeroGradient
    from aquaflux_amd.discretization import LimitedUpwind
    from aquaflux_amd.mesh import tet_grid_3d
    from aquaflux_amd.properties import Constant
    from aquaflux_amd.schemes import CompactGreenGauss, VenkatakrishnanLimiter
    from aquaflux_amd.transport import ScalarTransport, face_mass_flux

    print("case tet 6 x 12^3", flush=True)
    mesh = tet_grid_3d(12, perturb=0.15, seed=3)
    X = torch.tensor(mesh.node_coords, dtype=torch.float64, device=cuda, requires_grad=True)
    bcs = BoundaryConditions({p: (Dirichlet(1.0) if p == "left" else ZeroGradient())
                              for p in ("left", "right", "bottom", "top", "back", "front")})
    geom = mesh.geometry(node_coords=X, device=cuda)
    st = ScalarTransport.build(mesh, geom, Constant(0.01), bcs, LimitedUpwind(VenkatakrishnanLimiter(5.0)),
                               gradient_scheme=CompactGreenGauss())
    rng = np.random.default_rng(3)
    phi = torch.tensor(rng.uniform(0.0, 1.0, mesh.n_cells), device=cuda, requires_grad=True)
    lam = torch.tensor(rng.normal(size=mesh.n_cells), device=cuda)

    def run():
        g = mesh.geometry(node_coords=X, device=cuda)
        st.geometry = g
        R = st.residual(face_mass_flux(g, (1.0, 0.0, 0.0)))(phi)
        gX, gp = torch.autograd.grad((R * lam).sum(), (X, phi))
        return torch.cat([R.detach(), gX.reshape(-1), gp])

    dm = mesh.device_mesh(cuda)
    assert dm.block_order is not None
    ref = run()
    for name, order in _orders((mesh.n_cells + 255) // 256):
        dm.set_block_order(order)
        assert torch.equal(run(), ref), name


def test_block_order_rejects_a_non_permutation(cuda):
    from aquaflux_amd.mesh import structured_grid_3d

    mesh = structured_grid_3d(16, 16, 4, named_boundaries=True)
    mesh.geometry(device=cuda)
    dm = mesh.device_mesh(cuda)
    nb = (mesh.n_cells + 255) // 256
    with pytest.raises(ValueError, match="permutation"):
        dm.set_block_order(np.zeros(nb, dtype=np.int32))
    with pytest.raises(ValueError, match="n_blocks"):
        dm.set_block_order(np.arange(nb + 1, dtype=np.int32))
