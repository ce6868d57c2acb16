"""Newton-Krylov drivers on the GPU (``aquaflux/solve/newton.py``, ``linear.py``, ``implicit.py``).

* :func:`gmres` -- restarted, right-preconditioned GMRES; ``flexible=True`` gives FGMRES (the
  preconditioner may change between iterations, e.g. an inner Krylov solve).  Orthogonalisation is
  classical Gram-Schmidt applied twice (CGS2), two matrix-vector products against the basis per
  step, which keeps the basis orthogonal to working precision while running as BLAS-2 on the GPU.
  Stops when ``||r|| <= atol + rtol ||b||`` (``solve/linear.py:199-298``).
* :class:`DampedNewtonStep` -- Newton correction from GMRES on the exact FP64 tangent (the HIP JVP),
  then a backtracking ladder ``alpha in {1, 1/2, ..., 2**-steps}`` accepting the first strict
  descent (``solve/implicit.py:270-455``).  The Krylov solve is preconditioned by the compact block
  Jacobian: point-block Jacobi, or an inner GMRES(30) on that Jacobian, in FP32 or FP64 -- the FP32
  inner solve is an iterative refinement under the FP64 outer FGMRES.
* :class:`ImplicitNewtonSolver` -- drives Newton to ``||R|| <= atol + rtol ||R0||``
  (``solve/implicit.py:568-635``, ``forward_step.py:352-354``) and differentiates the root by the
  implicit-function theorem: the backward pass solves ``J^T lambda = g`` with the HIP adjoint and
  returns ``-(dR/dtheta)^T lambda`` for the parameter tensors (``solve/implicit.py:733-761``).
"""

from __future__ import annotations

import ctypes
import dataclasses
import math
import time
import weakref

import numpy as np
import torch

from . import _lib


@dataclasses.dataclass
class KrylovInfo:
    iterations: int = 0
    restarts: int = 0
    residual_norm: float = math.inf
    converged: bool = False


# a new Arnoldi vector is orthogonalised a second time only when the first classical Gram-Schmidt
# pass removed more than 1 - 1/sqrt(2) of its norm (the Daniel-Gragg-Kaufman-Stewart criterion,
# "twice is enough"): each pass reads the whole Krylov basis, so this halves the basis traffic when
# the basis is well conditioned, which is the usual case
DGKS_ETA = 1.0 / math.sqrt(2.0)

# status doubles of the device GMRES state (afx_gmres_*)
_ST_K, _ST_DONE, _ST_BREAK, _ST_RES, _ST_HEAD = 0, 1, 2, 6, 8


def gmres(matvec, b, *, x0=None, rtol=1e-10, atol=0.0, restart=30, max_iter=1000, precond=None, flexible=False,
          reorthogonalize: str = "dgks", basis_dtype=None, check_every=1):
    """Solve ``A x = b``; returns ``(x, KrylovInfo)``.  ``precond`` is applied on the right.
    ``reorthogonalize``: ``"dgks"`` (a second Gram-Schmidt pass when the first lost more than
    1 - 1/sqrt(2) of the vector's norm) or ``"always"`` (two passes every step).  ``basis_dtype``
    stores the Krylov (and flexible) basis and runs Gram-Schmidt in another precision (FP32 halves
    the orthogonalisation traffic of a solve that needs only a few digits).  Residuals and products
    stay in ``b``'s precision; dot products, norms and the update ``x + Z y`` widen each stored basis
    entry and sum in FP64.

    For CUDA tensors the Arnoldi recurrence runs on the device (``afx_gmres_*``): Gram-Schmidt, the
    Hessenberg column, the Givens rotations, the stopping test and the update are kernels, and the
    host synchronises only to read the cycle's status.  ``check_every``: an int reads it after every
    that many steps; ``None`` predicts from the residual estimates' recent rate how many steps
    remain and reads it then (at most every 8 steps).  Steps enqueued after the device has stopped
    the cycle do nothing, so the result does not depend on when the host looks; only the products
    and preconditioner calls of those steps are wasted.  Host (CPU) tensors run the same algorithm
    in torch, for tests of host-side logic."""
    if reorthogonalize not in ("dgks", "always"):
        raise ValueError(f"unknown reorthogonalize {reorthogonalize!r}")
    if not b.is_cuda:
        return _gmres_host(matvec, b, x0=x0, rtol=rtol, atol=atol, restart=restart, max_iter=max_iter,
                           precond=precond, flexible=flexible, reorthogonalize=reorthogonalize,
                           basis_dtype=basis_dtype)
    if b.dtype not in (torch.float32, torch.float64):
        raise ValueError(f"gmres: b must be float32 or float64, got {b.dtype}")
    dtype, dev, shape = b.dtype, b.device, b.shape
    b = b.contiguous().reshape(-1)
    n = b.numel()
    m = int(restart)
    bdt = dtype if basis_dtype is None else basis_dtype
    if bdt not in (torch.float32, torch.float64):
        raise ValueError(f"gmres: basis_dtype must be float32 or float64, got {bdt}")
    L = _lib.lib()
    st_len, wk_len = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(L.afx_gmres_sizes(n, m, ctypes.byref(st_len), ctypes.byref(wk_len)), "afx_gmres_sizes")
    state = torch.zeros(st_len.value, dtype=torch.float64, device=dev)
    work = torch.empty(wk_len.value, dtype=torch.float64, device=dev)
    ld = -(-n // 64) * 64  # rows 256-byte aligned, so that the sweeps load 16 bytes at a time
    V = torch.empty((m + 1, ld), dtype=bdt, device=dev)[:, :n]
    Z = torch.empty((m, ld), dtype=bdt, device=dev)[:, :n] if (flexible and precond is not None) else None
    prec = 0 if bdt == torch.float32 else 1
    vprec = 0 if dtype == torch.float32 else 1
    reo = 1 if reorthogonalize == "always" else 0
    stream = _lib.stream_of(dev)
    x = torch.zeros_like(b) if x0 is None else x0.contiguous().reshape(-1).to(dtype).clone()
    info = KrylovInfo()
    if x0 is not None:  # the tolerance is relative to ||b||, not to the first residual
        rtol_dev, atol_dev = 0.0, atol + rtol * float(torch.linalg.vector_norm(b))
    else:
        rtol_dev, atol_dev = rtol, atol
    while True:
        r = b if (x0 is None and info.restarts == 0) else b - matvec(x)
        r = r.contiguous()
        _lib.check(L.afx_gmres_start(n, m, prec, vprec, r.data_ptr(), rtol_dev, atol_dev, int(info.restarts == 0),
                                     V.data_ptr(), V.stride(0), state.data_ptr(), work.data_ptr(), stream),
                   "afx_gmres_start")
        limit = min(m, max_iter - info.iterations)
        if info.restarts > 0 or limit <= 0:
            st = state[:_ST_HEAD].cpu().numpy()
            info.residual_norm = float(st[_ST_RES])
            if st[_ST_DONE]:
                info.converged = True
                return x.view(shape), info
            if limit <= 0:
                return x.view(shape), info
        launched = 0
        next_check = 1 if check_every is None else max(1, int(check_every))
        st = None
        while launched < limit:
            j = launched
            z = V[j] if precond is None else precond(V[j])
            if Z is not None:
                Z[j] = z
                z = Z[j]  # the stored (possibly rounded) vector, so that A Z = V H holds as stored
            w = matvec(z if z.dtype == dtype else z.to(dtype)).contiguous()
            if w.dtype != dtype or w.numel() != n:
                raise ValueError("gmres: matvec must return a vector of b's size and dtype")
            _lib.check(L.afx_gmres_arnoldi(n, m, j, prec, vprec, w.data_ptr(), V.data_ptr(), V.stride(0), reo,
                                           state.data_ptr(), work.data_ptr(), stream), "afx_gmres_arnoldi")
            launched += 1
            if launched >= next_check or launched == limit:
                st = state[: _ST_HEAD + launched].cpu().numpy()
                if st[_ST_DONE]:
                    break
                next_check = launched + _next_check_gap(st[_ST_HEAD:_ST_HEAD + launched], st[4], check_every)
        if st is None or not st[_ST_DONE]:
            st = state[: _ST_HEAD + launched].cpu().numpy()
        k = int(st[_ST_K])
        if Z is not None:
            _lib.check(L.afx_gmres_update(n, m, prec, Z.data_ptr(), Z.stride(0), state.data_ptr(), vprec, x.data_ptr(),
                                          stream), "afx_gmres_update")
        elif precond is None:
            _lib.check(L.afx_gmres_update(n, m, prec, V.data_ptr(), V.stride(0), state.data_ptr(), vprec, x.data_ptr(),
                                          stream), "afx_gmres_update")
        else:
            upd = torch.zeros(n, dtype=bdt, device=dev)
            _lib.check(L.afx_gmres_update(n, m, prec, V.data_ptr(), V.stride(0), state.data_ptr(), prec,
                                          upd.data_ptr(), stream), "afx_gmres_update")
            x = x + precond(upd).to(dtype)
        info.iterations += k
        info.restarts += 1
        info.residual_norm = float(st[_ST_RES])
        if k == 0:  # stopped at the start (||r|| <= tol): converged; or no step possible
            info.converged = bool(st[_ST_DONE])
            return x.view(shape), info


def _next_check_gap(resid, tol, check_every):
    """Steps until the host next reads the GMRES status: ``check_every`` when given, else the steps
    the residual estimates' geometric rate over the last (up to) four steps needs to reach ``tol``,
    between 1 and 8."""
    if check_every is not None:
        return max(1, int(check_every))
    cap = 8
    q = len(resid)
    if q < 2 or not (resid[-1] > 0.0):
        return 1 if q < 2 else cap
    p = max(0, q - 5)
    rate = (resid[-1] / resid[p]) ** (1.0 / (q - 1 - p)) if resid[p] > 0.0 else 1.0
    if not (0.0 < rate < 1.0) or not (tol > 0.0):
        return cap
    need = math.log(tol / resid[-1]) / math.log(rate)
    return int(min(cap, max(1.0, math.floor(need))))


def _gmres_host(matvec, b, *, x0, rtol, atol, restart, max_iter, precond, flexible, reorthogonalize, basis_dtype):
    """The same restarted GMRES on host tensors, in torch (tests of host-side logic)."""
    dtype, dev = b.dtype, b.device
    n = b.numel()
    bnorm = float(torch.linalg.vector_norm(b))
    tol = atol + rtol * bnorm
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    info = KrylovInfo()
    if bnorm == 0.0:
        info.residual_norm, info.converged = 0.0, True
        return x, info
    m = int(restart)
    bdt = dtype if basis_dtype is None else basis_dtype
    V = torch.empty((m + 1, n), dtype=bdt, device=dev)
    Z = torch.empty((m, n), dtype=bdt, device=dev) if (flexible and precond is not None) else None
    while True:
        r = b - matvec(x) if (x0 is not None or info.restarts > 0) else b.clone()
        beta = float(torch.linalg.vector_norm(r))
        info.residual_norm = beta
        if beta <= tol:
            info.converged = True
            return x, info
        if info.iterations >= max_iter:
            return x, info
        V[0] = (r / beta).to(bdt)
        H = np.zeros((m + 1, m))
        cs = np.zeros(m)
        sn = np.zeros(m)
        g = np.zeros(m + 1)
        g[0] = beta
        k = 0
        for j in range(m):
            z = V[j] if precond is None else precond(V[j])
            if Z is not None:
                Z[j] = z
                z = Z[j]
            w = matvec(z.to(dtype)).to(torch.float64)
            Vj = V[: j + 1].to(torch.float64)
            w = w.to(bdt).to(torch.float64)
            w0 = float(torch.linalg.vector_norm(w))
            h = Vj @ w
            w = (w - Vj.T @ h).to(bdt).to(torch.float64)
            hn = float(torch.linalg.vector_norm(w))
            if reorthogonalize == "always" or not hn > DGKS_ETA * w0:
                h2 = Vj @ w
                w = (w - Vj.T @ h2).to(bdt).to(torch.float64)
                h = h + h2
                hn = float(torch.linalg.vector_norm(w))
            H[: j + 1, j] = h.numpy()
            H[j + 1, j] = hn
            if hn > 0:
                V[j + 1] = (w / hn).to(bdt)
            for i in range(j):  # apply previous rotations
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            den = math.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = (1.0, 0.0) if den == 0.0 else (H[j, j] / den, H[j + 1, j] / den)
            H[j, j] = den
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            k = j + 1
            info.iterations += 1
            if abs(g[j + 1]) <= tol or info.iterations >= max_iter or hn == 0.0:
                break
        y = np.linalg.solve(np.triu(H[:k, :k]) + np.diag(np.where(np.diag(H[:k, :k]) == 0, 1.0, 0.0)), g[:k])
        yt = torch.tensor(y, dtype=torch.float64)
        if Z is not None:
            x = x + (Z[:k].to(torch.float64).T @ yt).to(dtype)
        elif precond is None:
            x = x + (V[:k].to(torch.float64).T @ yt).to(dtype)
        else:
            upd = (V[:k].to(torch.float64).T @ yt).to(bdt)
            x = x + precond(upd).to(dtype)
        info.restarts += 1
        info.residual_norm = abs(g[k])


class BlockJacobian:
    """Compact block Jacobian of a flow residual at a state (block-ELL, FP32 or FP64)."""

    def __init__(self, op, state, precision: str = "fp32"):
        self.op = op
        raw = op._raw
        dm = raw.dm
        self.dm = dm
        self.C = op.mesh.n_cells
        self.B = op.mesh.dim + 1
        self.prec = 0 if precision == "fp32" else 1
        self.dtype = torch.float32 if self.prec == 0 else torch.float64
        ns = ctypes.c_int32()
        _lib.check(_lib.lib().afx_bell_slots(dm.handle, ctypes.byref(ns)), "afx_bell_slots")
        self.nslot = int(ns.value)
        dev = dm.device
        self.vals = torch.empty(self.nslot * self.B * self.B * self.C, dtype=self.dtype, device=dev)
        self.cols = torch.empty(self.nslot * self.C, dtype=torch.int32, device=dev)
        self.tslot = torch.empty(self.nslot * self.C, dtype=torch.int32, device=dev)
        fields = torch.empty(((self.B * (self.B - 1)) + self.B) * self.C, dtype=torch.float64, device=dev)
        mu, rho, ub, pb, gb = op._inputs()
        ins = raw.inputs(state.contiguous(), mu.detach(), rho.detach(), ub.detach(), pb.detach(),
                         None if gb is None else gb.detach())
        st = _lib.stream_of(dev)
        _lib.check(
            _lib.lib().afx_flow_jacobian(dm.handle, ctypes.byref(raw.gs), ctypes.byref(raw.cfg), ctypes.byref(ins),
                                         self.prec, self.vals.data_ptr(), self.cols.data_ptr(),
                                         self.tslot.data_ptr(), fields.data_ptr(), st),
            "afx_flow_jacobian",
        )
        self.inv = torch.empty(self.B * self.B * self.C, dtype=self.dtype, device=dev)
        flag = torch.zeros(1, dtype=torch.int32, device=dev)
        _lib.check(
            _lib.lib().afx_pbj_setup(self.C, self.B, self.prec, self.vals.data_ptr(), self.inv.data_ptr(),
                                     flag.data_ptr(), st),
            "afx_pbj_setup",
        )
        self.singular = bool(flag.item())

    # vectors: field-major FP64 (residual layout) <-> cell-major in the Jacobian's precision
    def to_cell_major(self, v):
        return v.reshape(self.B, self.C).T.contiguous().reshape(-1).to(self.dtype)

    def from_cell_major(self, v):
        return v.reshape(self.C, self.B).T.contiguous().reshape(-1).to(torch.float64)

    def matvec(self, x):
        y = torch.empty_like(x)
        _lib.check(
            _lib.lib().afx_bell_spmv(self.C, self.B, self.nslot, self.prec, self.vals.data_ptr(), self.cols.data_ptr(),
                                     x.data_ptr(), y.data_ptr(), _lib.stream_of(x.device)),
            "afx_bell_spmv",
        )
        return y

    def rmatvec(self, x):
        """``y = A^T x``."""
        y = torch.empty_like(x)
        _lib.check(
            _lib.lib().afx_bell_spmv_t(self.C, self.B, self.nslot, self.prec, self.vals.data_ptr(),
                                       self.cols.data_ptr(), self.tslot.data_ptr(), x.data_ptr(), y.data_ptr(),
                                       _lib.stream_of(x.device)),
            "afx_bell_spmv_t",
        )
        return y

    def pbj(self, x):
        y = torch.empty_like(x)
        _lib.check(
            _lib.lib().afx_pbj_apply(self.C, self.B, self.prec, self.inv.data_ptr(), x.data_ptr(), y.data_ptr(),
                                     _lib.stream_of(x.device)),
            "afx_pbj_apply",
        )
        return y

    def add_diagonal(self, shift):
        """Adds a per-row shift (field-major FP64 vector) to the diagonal blocks and refreshes the
        point-block Jacobi inverses."""
        B, C = self.B, self.C
        sc = shift.reshape(B, C).to(self.dtype)
        diag = self.vals[: B * B * C].view(B, B, C)
        for r in range(B):
            diag[r, r] += sc[r]
        flag = torch.zeros(1, dtype=torch.int32, device=self.vals.device)
        _lib.check(
            _lib.lib().afx_pbj_setup(C, B, self.prec, self.vals.data_ptr(), self.inv.data_ptr(), flag.data_ptr(),
                                     _lib.stream_of(self.vals.device)),
            "afx_pbj_setup",
        )

    def pbj_transpose(self, x):
        """``y = D^-T x`` (the adjoint solve's preconditioner)."""
        inv = self.inv.reshape(self.B, self.B, self.C)  # inv[r, k, c]
        return torch.einsum("rkc,cr->ck", inv, x.reshape(self.C, self.B)).reshape(-1).contiguous()


@dataclasses.dataclass
class DampedNewtonStep:
    """Newton correction + backtracking line search (``solve/implicit.py:458-565``).

    The linear solve stops at a global relative residual ``||J delta + R|| <= rtol ||R||`` -- the
    reference's robust forward stop (``relative_residual_gmres``, ``solve/linear.py:139-200``), not
    lineax's componentwise test whose absolute floor would freeze Newton once ``||R||`` falls below
    it.

    ``preconditioner``: ``None`` (plain GMRES on the JVP), ``"pbj"`` (point-block Jacobi of the
    block Jacobian), ``"inner"`` (FGMRES whose preconditioner is GMRES(``inner_restart``) on the
    block Jacobian with point-block Jacobi) or ``"block"`` (FGMRES with the SIMPLE-family
    ``block_preconditioner``, a :class:`~aquaflux_amd.block_preconditioner.BlockPreconditioner` of
    the flow assembler, built with its defaults when None, its saddle blocks read off the block
    Jacobian and its Schur built at the pseudo-time-shifted ``a_P``); ``precision`` is the block
    Jacobian's arithmetic.
    ``jacobian``: ``"compact"`` (the flow residual's reach-1 dual-number assembly with gradients,
    a_P and limiters frozen) or ``"probed"`` (the exact Jacobian of any residual by coloured probing
    of its tangent, ``sparse_jacobian.py:722-848``, with ``column_reach`` per state field; the plan
    is built once per mesh and kept).
    """

    preconditioner: str | None = "inner"
    precision: str = "fp32"
    line_search: int = 10
    rtol: float = 1e-3
    atol: float = 0.0
    restart: int = 30
    max_iter: int = 300
    # a 50-vector inner basis halves the outer iterations of a 2M-cell Newton solve against 30
    # (the inner SpMV is cheap beside the outer dual-number JVP); 60 costs more than it saves
    inner_restart: int = 50
    inner_iters: int = 50
    inner_rtol: float = 1e-2
    # pseudo-transient continuation (switched evolution relaxation): None = plain Newton
    ptc_cfl0: float | None = None
    ptc_cfl_max: float = 1e12
    jacobian: str = "compact"
    column_reach: tuple | None = None
    block_preconditioner: object = None
    outer_basis: str = "fp64"  # precision of the outer Krylov basis ("fp64" or "fp32")
    # probe plans per mesh (held weakly: a plan dies with its mesh) and column reach
    _plans: weakref.WeakKeyDictionary = dataclasses.field(default_factory=weakref.WeakKeyDictionary, repr=False)

    def block_jacobian(self, op, phi, precision: str | None = None):
        """The block Jacobian of ``op`` at ``phi`` in ``precision`` (default: the step's own)."""
        precision = self.precision if precision is None else precision
        if self.jacobian == "compact":
            from .flow import MomentumContinuity

            if not isinstance(op, MomentumContinuity):
                raise ValueError(f"the compact Jacobian is the coupled-flow residual's own; a {type(op).__name__} "
                                 "needs jacobian='probed' with column_reach")
            return BlockJacobian(op, phi, precision)
        if self.jacobian != "probed":
            raise ValueError(f"unknown jacobian {self.jacobian!r}")
        from .jacobian import materialize_block_jacobian, mesh_probe_plan

        if self.column_reach is None:
            raise ValueError("the probed Jacobian needs column_reach (the stencil reach of each state field)")
        if getattr(op, "mesh", None) is None:
            raise ValueError("the probed Jacobian needs the operator's mesh (pass mesh= with a residual function)")
        per_mesh = self._plans.setdefault(op.mesh, {})
        key = tuple(self.column_reach)
        plan = per_mesh.get(key)
        if plan is None:
            plan = mesh_probe_plan(op.mesh, self.column_reach)
            per_mesh[key] = plan
        return materialize_block_jacobian(lambda v: op.residual_jvp(phi, v)[1], plan, device=phi.device,
                                          precision=precision)

    def pseudo_time_shift(self, op, phi, cfl):
        """Diagonal ``a_P / CFL`` on the momentum rows (zero on continuity): a local pseudo-time step
        scaled by the momentum diagonal, as the reference's pseudo-transient shift basis is.  An
        operator with its own basis (``pseudo_time_diagonal``, e.g. the coupled RANS) supplies it."""
        if hasattr(op, "pseudo_time_diagonal"):
            return op.pseudo_time_diagonal(phi) / cfl
        _, aP, _ = op.fields(phi)
        d = op.mesh.dim
        shift = torch.zeros_like(phi).reshape(d + 1, -1)
        shift[:d] = (aP / cfl)[None, :]
        return shift.reshape(-1)

    def correction(self, op, phi, r, cfl=None):
        J = None
        precond = None
        flexible = False
        shift = None if cfl is None or cfl >= self.ptc_cfl_max else self.pseudo_time_shift(op, phi, cfl)
        if self.preconditioner is not None:
            J = self.block_jacobian(op, phi)
            if shift is not None:
                J.add_diagonal(shift)
            if self.preconditioner == "pbj":
                precond = lambda v: J.from_cell_major(J.pbj(J.to_cell_major(v)))  # noqa: E731
            elif self.preconditioner == "inner":
                flexible = True

                def precond(v):
                    vc = J.to_cell_major(v)
                    z, _ = gmres(J.matvec, vc, rtol=self.inner_rtol, restart=self.inner_restart,
                                 max_iter=self.inner_iters, precond=J.pbj, check_every=None)
                    return J.from_cell_major(z)
            elif self.preconditioner == "block":
                flexible = True
                bp = self.block_preconditioner
                if bp is None:
                    from .block_preconditioner import BlockPreconditioner

                    bp = BlockPreconditioner.build(op)
                    self.block_preconditioner = bp
                elif bp.assembler is not op:
                    raise ValueError("block_preconditioner was built for another assembler")
                a_p = bp.frozen_momentum_diagonal(phi)
                precond = bp.apply_at(phi, J, a_p if shift is None else a_p * (1.0 + 1.0 / cfl))
            else:
                raise ValueError(f"unknown preconditioner {self.preconditioner!r}")

        def matvec(v):
            jv = op.residual_jvp(phi, v)[1]
            return jv if shift is None else jv + shift * v

        if self.outer_basis not in ("fp64", "fp32"):
            raise ValueError(f"outer_basis must be 'fp64' or 'fp32', got {self.outer_basis!r}")
        delta, info = gmres(matvec, -r, rtol=self.rtol, atol=self.atol, restart=self.restart, max_iter=self.max_iter,
                            precond=precond, flexible=flexible,
                            basis_dtype=torch.float32 if self.outer_basis == "fp32" else None)
        return delta, info


def backtracking_line_search(residual, phi, delta, reference_norm, steps, norm=None):
    """First (largest) ``alpha`` in ``{1, 1/2, ..., 2**-steps}`` with ``||R(phi + alpha delta)||`` below
    ``reference_norm``; else the smallest rung with a finite residual (``solve/implicit.py:270-455``)."""
    norm = norm or (lambda v: float(torch.linalg.vector_norm(v)))
    fallback = None
    for i in range(steps + 1):
        alpha = 0.5**i
        cand = phi + alpha * delta
        r = residual(cand)
        rn = norm(r)
        if math.isfinite(rn):
            fallback = (cand, r, rn, alpha)
            if rn < reference_norm:
                return cand, r, rn, alpha
    if fallback is None:
        raise FloatingPointError("line search: every rung produced a non-finite residual")
    return fallback


@dataclasses.dataclass
class NewtonReport:
    steps: int = 0
    residual_norms: list = dataclasses.field(default_factory=list)
    krylov_iterations: list = dataclasses.field(default_factory=list)
    alphas: list = dataclasses.field(default_factory=list)
    seconds: float = 0.0
    converged: bool = False


@dataclasses.dataclass(frozen=True)
class GMRES:
    """Settings of a restarted GMRES solve, in the role of the reference's ``lineax.GMRES(rtol, atol,
    restart, max_steps)`` linear-solver objects (``solve/linear.py:40-136``): passed as
    ``ImplicitNewtonSolver(solver=..., adjoint_solver=...)``."""

    rtol: float = 1e-10
    atol: float = 0.0
    restart: int = 50
    max_steps: int = 2000


class ResidualFunction:
    """A residual callable ``residual_fn(phi, theta)`` as the operator the Newton drivers take: the
    residual with no autograd, its tangent in ``phi`` by forward-mode AD and its adjoint in ``phi`` by
    reverse-mode AD (through the HIP operators' own tangent and adjoint), and the residual with
    ``theta`` live for the implicit-function gradient.  ``mesh`` (optional) enables block-Jacobian
    preconditioners by coloured probing (``DampedNewtonStep(jacobian="probed", column_reach=...)``)."""

    def __init__(self, residual_fn, theta, mesh=None):
        self.fn = residual_fn
        self.theta = theta
        self.mesh = mesh
        self._theta_detached = _tree_map(lambda t: t.detach() if isinstance(t, torch.Tensor) else t, theta)

    def residual_raw(self, phi):
        with torch.no_grad():
            return self.fn(phi, self._theta_detached)

    def residual(self, phi):
        return self.fn(phi, self.theta)

    def residual_jvp(self, phi, v):
        import torch.autograd.forward_ad as fwAD

        with torch.no_grad(), fwAD.dual_level():
            out = self.fn(fwAD.make_dual(phi.detach(), v), self._theta_detached)
            prim, tan = fwAD.unpack_dual(out)
            return prim, (tan if tan is not None else torch.zeros_like(prim))

    def residual_vjp(self, phi, w):
        with torch.enable_grad():
            x = phi.detach().requires_grad_(True)
            (g,) = torch.autograd.grad(self.fn(x, self._theta_detached), (x,), w, allow_unused=True)
        return g if g is not None else torch.zeros_like(phi)


def _tree_map(f, tree):
    if isinstance(tree, dict):
        return {k: _tree_map(f, v) for k, v in tree.items()}
    if isinstance(tree, (tuple, list)):
        return type(tree)(_tree_map(f, v) for v in tree)
    return f(tree)


def _tree_leaves(tree):
    if isinstance(tree, dict):
        return [x for v in tree.values() for x in _tree_leaves(v)]
    if isinstance(tree, (tuple, list)):
        return [x for v in tree for x in _tree_leaves(v)]
    return [tree]


@dataclasses.dataclass
class ImplicitNewtonSolver:
    """Newton to ``||R|| <= atol + rtol ||R0||`` with an IFT adjoint (``solve/implicit.py:767-864``).

    ``solver`` / ``adjoint_solver`` (:class:`GMRES` settings, or None) override the forward step's
    Krylov tolerances and the adjoint solve's (``adjoint_rtol`` / ``adjoint_restart`` /
    ``adjoint_max_iter``), as the reference's ``solver`` / ``adjoint_solver`` fields do."""

    rtol: float = 1e-10
    atol: float = 1e-12
    max_steps: int = 50
    forward_step: DampedNewtonStep = dataclasses.field(default_factory=DampedNewtonStep)
    adjoint_rtol: float = 1e-10
    adjoint_restart: int = 50
    adjoint_max_iter: int = 2000
    progress: object = None  # optional callable(step, ||R||, krylov_iterations, alpha, seconds)
    solver: GMRES | None = None
    adjoint_solver: GMRES | None = None

    def __post_init__(self):
        if self.solver is not None:
            self.forward_step = dataclasses.replace(self.forward_step, rtol=self.solver.rtol, atol=self.solver.atol,
                                                    restart=self.solver.restart, max_iter=self.solver.max_steps)
        if self.adjoint_solver is not None:
            self.adjoint_rtol = self.adjoint_solver.rtol
            self.adjoint_restart = self.adjoint_solver.restart
            self.adjoint_max_iter = self.adjoint_solver.max_steps

    def forward(self, op, phi0):
        """Runs Newton from ``phi0`` (no autograd); returns ``(phi, NewtonReport)``."""
        rep = NewtonReport()
        t0 = time.perf_counter()
        with torch.no_grad():
            phi = phi0.detach().clone().contiguous()
            r = op.residual_raw(phi)
            r0 = float(torch.linalg.vector_norm(r))
            rn = r0
            rep.residual_norms.append(rn)
            target = self.atol + self.rtol * r0
            while rn > target:
                if rep.steps >= self.max_steps:
                    torch.cuda.synchronize()
                    rep.seconds = time.perf_counter() - t0
                    raise RuntimeError(
                        f"Newton did not converge in {self.max_steps} steps: ||R|| = {rn:.3e} > {target:.3e}"
                    )
                fs = self.forward_step
                if fs.ptc_cfl0 is not None:
                    # switched evolution relaxation: CFL grows as the residual falls
                    cfl = min(fs.ptc_cfl_max, fs.ptc_cfl0 * r0 / rn)
                    delta, kinfo = fs.correction(op, phi, r, cfl)
                    phi, r, rn, alpha = backtracking_line_search(op.residual_raw, phi, delta, 4.0 * rn, fs.line_search)
                else:
                    delta, kinfo = fs.correction(op, phi, r)
                    phi, r, rn, alpha = backtracking_line_search(op.residual_raw, phi, delta, rn, fs.line_search)
                if not math.isfinite(rn):
                    raise FloatingPointError("non-finite residual in Newton")
                rep.steps += 1
                rep.residual_norms.append(rn)
                rep.krylov_iterations.append(kinfo.iterations)
                rep.alphas.append(alpha)
                if self.progress is not None:
                    self.progress(rep.steps, rn / r0, kinfo.iterations, alpha, time.perf_counter() - t0)
            torch.cuda.synchronize()
        rep.seconds = time.perf_counter() - t0
        rep.converged = True
        return phi, rep

    def solve(self, residual_fn, phi0, theta=(), *, mesh=None):
        """Root ``phi*`` of the residual, differentiable w.r.t. the tensors in ``theta`` by the
        implicit-function theorem (``solve/implicit.py:813-864``).

        ``residual_fn`` is either the reference's ``residual_fn(phi, theta)`` callable (any
        composition of this package's residual operators and torch operations; ``theta`` a tensor or
        a tuple / list / dict of them; ``mesh`` enables probed block-Jacobian preconditioners), or an
        operator with ``residual_raw`` / ``residual_jvp`` / ``residual_vjp`` / ``residual`` (a
        ``MomentumContinuity``, ``CoupledRANS``, ...), in which case ``theta`` lists the tensors that
        operator depends on."""
        if hasattr(residual_fn, "residual_raw"):
            op, params = residual_fn, tuple(_tree_leaves(theta))
        else:
            op = ResidualFunction(residual_fn, theta, mesh)
            params = tuple(_tree_leaves(theta))
        params = tuple(p for p in params if isinstance(p, torch.Tensor))
        phi, rep = self.forward(op, phi0)
        self.last_report = rep
        if not any(isinstance(p, torch.Tensor) and p.requires_grad for p in params):
            return phi
        return _ImplicitRoot.apply(self, op, phi, *params)

    def adjoint(self, op, phi, cotangent):
        """``lambda`` solving ``J(phi)^T lambda = cotangent``: FGMRES on the HIP adjoint, preconditioned
        by GMRES on the transposed block Jacobian of the forward step (compact or probed, here in
        FP64) with transposed point-block Jacobi (the forward preconditioner, transposed, as
        ``solve/implicit.py:709-730`` does)."""
        J = self.forward_step.block_jacobian(op, phi, precision="fp64")
        step = self.forward_step

        def precond(v):
            vc = J.to_cell_major(v)
            z, _ = gmres(J.rmatvec, vc, rtol=step.inner_rtol, restart=step.inner_restart,
                         max_iter=step.inner_iters, precond=J.pbj_transpose, check_every=None)
            return J.from_cell_major(z)

        lam, info = gmres(lambda w: op.residual_vjp(phi, w), cotangent, rtol=self.adjoint_rtol,
                          restart=self.adjoint_restart, max_iter=self.adjoint_max_iter, precond=precond, flexible=True)
        if not info.converged:
            raise RuntimeError(f"adjoint GMRES did not converge: ||r|| = {info.residual_norm:.3e}")
        return lam


class _ImplicitRoot(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, op, phi, *params):
        ctx.solver, ctx.op = solver, op
        ctx.save_for_backward(phi, *params)
        return phi.clone()

    @staticmethod
    def backward(ctx, g):
        phi, *params = ctx.saved_tensors
        lam = ctx.solver.adjoint(ctx.op, phi, g.contiguous())
        with torch.enable_grad():
            leaves = [p for p in params if p.requires_grad]
            R = ctx.op.residual(phi.detach())
            grads = torch.autograd.grad(R, leaves, -lam, allow_unused=True)
        it = iter(grads)
        out = [next(it) if p.requires_grad else None for p in params]
        return (None, None, None, *out)
