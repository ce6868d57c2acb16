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
