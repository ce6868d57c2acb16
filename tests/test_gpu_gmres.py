"""Device-resident GMRES (afx_gmres_*): the Arnoldi recurrence, Givens rotations, stopping test and
update run as kernels; the host reads the cycle status only when it checks.

Checks: the true residual meets the tolerance the reference's solve_linear uses
(``||b - A x|| <= atol + rtol ||b||``, aquaflux/solve/linear.py:199-298); the iterate agrees with
the same algorithm run on the host in torch; results are bitwise reproducible and independent of
when the host looks; the stored basis is orthonormal; restarts, an initial guess, right and
flexible preconditioning, FP32 bases and vectors, breakdown and a zero right-hand side behave as
the host algorithm does.
"""

import ctypes

import numpy as np
import pytest
import torch

from aquaflux_amd import _lib
from aquaflux_amd.solve import _gmres_host, gmres

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _system(n, seed, device, dtype=F64):
    """Nonsymmetric, well-conditioned A = diag(1..3) + 0.3 N(0,1)/sqrt(n), and b ~ N(0,1)."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    A = torch.diag(torch.linspace(1.0, 3.0, n, dtype=F64)) + 0.3 * torch.randn(n, n, generator=g, dtype=F64) / n**0.5
    b = torch.randn(n, generator=g, dtype=F64)
    return A.to(device=device, dtype=dtype), b.to(device=device, dtype=dtype)


def _true_rel(A, x, b):
    return float(torch.linalg.vector_norm((b - A @ x).double()) / torch.linalg.vector_norm(b.double()))


@pytest.mark.parametrize("basis", [torch.float64, torch.float32])
@pytest.mark.parametrize("restart", [40, 7])
def test_gmres_matches_the_host_algorithm(cuda, basis, restart):
    n = 3000
    print(f"case gmres n={n} basis={basis} restart={restart}", flush=True)
    A, b = _system(n, 1, cuda)
    x, info = gmres(lambda v: A @ v, b, rtol=1e-10, restart=restart, max_iter=400, basis_dtype=basis)
    Ah, bh = A.cpu(), b.cpu()
    xh, ih = _gmres_host(lambda v: Ah @ v, bh, x0=None, rtol=1e-10, atol=0.0, restart=restart, max_iter=400,
                         precond=None, flexible=False, reorthogonalize="dgks", basis_dtype=basis)
    assert info.converged and ih.converged
    assert _true_rel(A, x, b) <= 1e-10
    assert abs(info.iterations - ih.iterations) <= 1
    assert float((x.cpu() - xh).abs().max() / xh.abs().max()) <= 1e-8


def test_gmres_is_bitwise_reproducible_and_independent_of_checks(cuda):
    A, b = _system(2000, 2, cuda)
    runs = [gmres(lambda v: A @ v, b, rtol=1e-11, restart=30, max_iter=300, check_every=c) for c in (1, 1, 5, None)]
    x0, i0 = runs[0]
    for x, i in runs[1:]:
        assert torch.equal(x, x0)
        assert i.iterations == i0.iterations and i.restarts == i0.restarts
        assert i.residual_norm == i0.residual_norm


@pytest.mark.parametrize("n,ld,m", [(5000, 5000, 12), (4097, 4160, 12), (3001, 3001, 12), (1500, 1536, 70)])
def test_gmres_basis_is_orthonormal(cuda, n, ld, m):
    """m Arnoldi steps through the C-ABI: V^T V = I to 1e-14 (FP64 basis) and 1e-6 (FP32), and the
    residual estimate matches the true residual of the updated iterate.  Sizes cover 16-byte loads
    with a ragged tail (4097), rows that are not 16-byte aligned (ld 3001: one entry per load) and a
    basis longer than the 64 vectors the fused sweep holds in registers (m = 70)."""
    L = _lib.lib()
    print(f"case arnoldi n={n} ld={ld} m={m}", flush=True)
    A, b = _system(n, 3, cuda)
    st_len, wk_len = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(L.afx_gmres_sizes(n, m, ctypes.byref(st_len), ctypes.byref(wk_len)), "afx_gmres_sizes")
    s = _lib.stream_of(cuda)
    for dt, prec, tol in ((F64, 1, 1e-14), (torch.float32, 0, 1e-6)):
        state = torch.zeros(st_len.value, dtype=F64, device=cuda)
        work = torch.empty(wk_len.value, dtype=F64, device=cuda)
        Vs = torch.zeros((m + 1, ld), dtype=dt, device=cuda)
        V = Vs[:, :n]
        _lib.check(L.afx_gmres_start(n, m, prec, 1, b.data_ptr(), 0.0, 0.0, 1, V.data_ptr(), ld, state.data_ptr(),
                                     work.data_ptr(), s), "afx_gmres_start")
        for j in range(m):
            w = (A @ V[j].double()).contiguous()
            _lib.check(L.afx_gmres_arnoldi(n, m, j, prec, 1, w.data_ptr(), V.data_ptr(), ld, 0, state.data_ptr(),
                                           work.data_ptr(), s), "afx_gmres_arnoldi")
        assert float(Vs[:, n:].abs().max()) == 0.0 if ld > n else True  # nothing written past the rows
        G = V.double() @ V.double().T
        assert float((G - torch.eye(m + 1, dtype=F64, device=cuda)).abs().max()) <= tol
        assert int(state[0]) == m and state[1].item() == 0.0
        x = torch.zeros(n, dtype=F64, device=cuda)
        _lib.check(L.afx_gmres_update(n, m, prec, V.data_ptr(), ld, state.data_ptr(), 1, x.data_ptr(), s),
                   "afx_gmres_update")
        true = float(torch.linalg.vector_norm(b - A @ x))
        est = float(state[6])
        assert abs(true - est) <= (1e-10 if prec else 1e-5) * float(torch.linalg.vector_norm(b))


def test_gmres_initial_guess_and_preconditioners(cuda):
    n = 2500
    A, b = _system(n, 4, cuda)
    dinv = 1.0 / torch.diagonal(A)
    x0 = torch.randn(n, dtype=F64, device=cuda)
    x, info = gmres(lambda v: A @ v, b, x0=x0, rtol=1e-10, restart=25, max_iter=300)
    assert info.converged and _true_rel(A, x, b) <= 1e-10
    x, info = gmres(lambda v: A @ v, b, rtol=1e-10, restart=25, max_iter=300, precond=lambda v: dinv * v)
    assert info.converged and _true_rel(A, x, b) <= 1e-10
    # flexible with an FP32 basis and a preconditioner that changes between calls
    calls = [0]

    def varying(v):
        calls[0] += 1
        return (dinv * (1.0 + 0.05 * (calls[0] % 3))).to(v.dtype) * v

    x, info = gmres(lambda v: A @ v, b, rtol=1e-10, restart=25, max_iter=300, precond=varying, flexible=True,
                    basis_dtype=torch.float32)
    assert info.converged and _true_rel(A, x, b) <= 1e-10


def test_gmres_fp32_vectors(cuda):
    """An FP32 system (the inner solves' precision) with FP32 products and update."""
    n = 2000
    A, b = _system(n, 5, cuda, dtype=torch.float32)
    x, info = gmres(lambda v: A @ v, b, rtol=1e-5, restart=30, max_iter=30, check_every=None)
    assert x.dtype == torch.float32 and info.converged
    assert _true_rel(A.double(), x.double(), b.double()) <= 2e-5


def test_gmres_breakdown_and_zero_rhs(cuda):
    n = 1000
    b = torch.randn(n, dtype=F64, device=cuda)
    x, info = gmres(lambda v: 2.0 * v, b, rtol=1e-12, restart=10)
    assert info.converged and info.iterations == 1
    assert float((x - 0.5 * b).abs().max()) <= 1e-15
    x, info = gmres(lambda v: 2.0 * v, torch.zeros(n, dtype=F64, device=cuda), rtol=1e-12, restart=10)
    assert info.converged and info.iterations == 0 and float(x.abs().max()) == 0.0
