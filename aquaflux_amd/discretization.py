"""Advection-scheme strategy objects (``aquaflux/discretization/advection.py``)."""

from __future__ import annotations

import dataclasses

import torch

from .schemes import Limiter, VenkatakrishnanLimiter


class AdvectionScheme:
    code: int = -1


class FirstOrderUpwind(AdvectionScheme):
    """Upwind cell value; boundary inflow takes the weak value (``advection.py:96-110``)."""

    code = 1


@dataclasses.dataclass(frozen=True)
class LimitedUpwind(AdvectionScheme):
    """``phi_C + psi_C grad phi_C . (x_f - x_C)`` from the upwind cell (``advection.py:113-158``).

    Only the Venkatakrishnan limiter is implemented on the GPU; ``limiter=None`` is not supported
    (the reference's unlimited variant) and raises at build time.
    """

    limiter: Limiter | None = dataclasses.field(default_factory=VenkatakrishnanLimiter)
    code = 2

    def limiter_k(self) -> float:
        if not isinstance(self.limiter, VenkatakrishnanLimiter):
            raise ValueError("LimitedUpwind on the GPU supports only VenkatakrishnanLimiter")
        return float(self.limiter.k)


class TransientTerm:
    """Backward-difference accumulation ``V_P dphi/dt`` (``discretization/transient.py:23-62``):
    BDF1 ``V (phi - phi_old) / dt`` on the first step, BDF2 ``V (1.5 phi - 2 phi_old + 0.5 phi_older)
    / dt`` afterwards.

    Inside a residual this is a strategy marker: the scalar residual kernel adds the term to each
    cell's flux sum.  :meth:`residual` evaluates the stencil on its own (elementwise tensor
    arithmetic, no mesh).
    """

    def residual(self, phi, phi_old, phi_older, dt, first_step, volume):
        if first_step:
            return volume * (phi - phi_old) / dt
        return volume * (1.5 * phi - 2.0 * phi_old + 0.5 * phi_older) / dt


class DifferenceRow:
    """Fixation row ``phi - target`` (``discretization/fixed_value.py:80-93``)."""

    code = 0

    def row(self, phi, target):
        return phi - target

    def jacobian_scale(self, phi, chain):
        return chain


class LogRatioRow:
    """Fixation row ``log(phi / target)``, linear in ``w`` for ``phi = e^w``
    (``discretization/fixed_value.py:96-124``)."""

    code = 1

    def row(self, phi, target):
        return torch.log(phi / target)

    def jacobian_scale(self, phi, chain):
        return chain / phi
