"""Host-side GMRES logic (CPU): the status-check schedule of the device GMRES and the torch
restatement of the algorithm used for host tensors."""

import math

import torch

from aquaflux_amd.solve import _next_check_gap, gmres


def test_check_schedule():
    assert _next_check_gap([1.0, 0.5], 1e-3, 3) == 3
    assert _next_check_gap([1.0], 1e-3, None) == 1
    # rate 1/2 per step from 2^-3: 7 more steps reach 2^-10 ~ 1e-3 (floor of log(tol/r)/log(rate))
    resid = [2.0**-k for k in range(4)]
    need = math.floor(math.log(1e-3 / resid[-1]) / math.log(0.5))
    assert _next_check_gap(resid, 1e-3, None) == min(8, need)
    assert _next_check_gap([1.0, 1.0, 1.0], 1e-3, None) == 8  # stagnating: the cap
    assert _next_check_gap([1.0, 1e-2, 1e-4], 1e-4 * 0.99, None) == 1  # nearly there: look next step


def test_host_gmres_restarts_and_preconditioner():
    g = torch.Generator().manual_seed(0)
    n = 300
    A = torch.diag(torch.linspace(1.0, 4.0, n, dtype=torch.float64)) + 0.2 * torch.randn(n, n, generator=g,
                                                                                        dtype=torch.float64) / n**0.5
    b = torch.randn(n, generator=g, dtype=torch.float64)
    dinv = 1.0 / torch.diagonal(A)
    for kw in ({}, {"precond": lambda v: dinv * v}, {"precond": lambda v: dinv * v, "flexible": True},
               {"basis_dtype": torch.float32}):
        x, info = gmres(lambda v: A @ v, b, rtol=1e-10, restart=8, max_iter=400, **kw)
        assert info.converged and info.restarts >= 1
        assert float((b - A @ x).norm() / b.norm()) <= 1e-10
