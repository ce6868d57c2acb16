"""Numerical strategy objects: gradient reconstruction and slope limiting (``aquaflux/schemes``).

These are configuration objects: the arithmetic lives in the HIP kernels, which dispatch on them.
"""

from __future__ import annotations

import dataclasses


class GradientScheme:
    """Strategy interface (``aquaflux/schemes/gradient.py:94-126``)."""

    code: int = -1


class CompactGreenGauss(GradientScheme):
    """One-shot Green-Gauss with linearly interpolated face values (``gradient.py:129-153``)."""

    code = 0


@dataclasses.dataclass(frozen=True)
class SweptGradientSolve:
    """Fixed preconditioned-Richardson sweeps ``x <- x + r V^-1 (b - A_g x)`` from the peeled
    ``x0 = r V^-1 b`` (``aquaflux/schemes/gradient.py:546-705``).  ``warn_tol`` is accepted for API
    compatibility; the GPU path does not emit the under-resolution warning."""

    sweeps: int = 4
    relaxation: float = 1.0
    warn_tol: float | None = None

    def __post_init__(self):
        if self.relaxation <= 0.0:
            raise ValueError("SweptGradientSolve: relaxation must be positive")


@dataclasses.dataclass(frozen=True)
class CorrectedGreenGauss(GradientScheme):
    """Green-Gauss with the non-orthogonal skewness correction, ``A_g G = B phi`` solved by swept
    Richardson sweeps preconditioned by ``1/V`` (``aquaflux/schemes/gradient.py:1122-1337``).  Only
    :class:`SweptGradientSolve` and the inverse-volume preconditioner are provided on the GPU."""

    solver: SweptGradientSolve = dataclasses.field(default_factory=SweptGradientSolve)
    code = 3

    @property
    def sweeps(self) -> int:
        return int(self.solver.sweeps)

    @property
    def relaxation(self) -> float:
        return float(self.solver.relaxation)


class LeastSquaresGradient(GradientScheme):
    """Inverse-distance-squared weighted least squares over the face neighbours.

    Not present in the reference (its schemes are Green-Gauss variants); exact for linear fields on
    any mesh.  Boundary faces contribute the face-centroid offset and the boundary value.
    """

    code = 1


class Limiter:
    pass


@dataclasses.dataclass(frozen=True)
class VenkatakrishnanLimiter(Limiter):
    """Venkatakrishnan (1993) limiter, ``eps^2 = V K^3`` (``aquaflux/schemes/limiter.py:66-117``)."""

    k: float = 5.0
