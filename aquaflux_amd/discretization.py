"""Advection-scheme strategy objects (``aquaflux/discretization/advection.py``)."""

from __future__ import annotations

import dataclasses

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
