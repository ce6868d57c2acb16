"""Named boundary-condition collections and scalar closures (``aquaflux/boundary``)."""

from __future__ import annotations

import numpy as np


class BoundaryConditions:
    """``{patch: closure}`` bound to a mesh by :meth:`resolve` (``boundary/collection.py:30-137``)."""

    def __init__(self, conditions: dict, _faces: dict | None = None):
        self.conditions = dict(conditions)
        self.faces = _faces

    def resolve(self, face_patches) -> "BoundaryConditions":
        if self.faces is not None:
            return self
        faces = {name: np.asarray(face_patches.indices(name), dtype=np.int64) for name in self.conditions}
        return BoundaryConditions(self.conditions, _faces=faces)

    def items(self):
        if self.faces is None:
            raise ValueError(
                "BoundaryConditions must be bound to a mesh via resolve(face_patches) before apply()"
            )
        for name, bc in self.conditions.items():
            yield name, bc, self.faces[name]


# --- scalar closures (boundary/conditions.py:46-201) --------------------------------------------

BC_NONE, BC_DIRICHLET, BC_ZERO_GRADIENT, BC_NEUMANN, BC_CONVECTIVE = 0, 1, 2, 3, 4


class BoundaryCondition:
    kind = BC_NONE

    def values(self):
        """(value, t_inf) carried per face."""
        return 0.0, 0.0


class Dirichlet(BoundaryCondition):
    """``phi_ip = value``."""

    kind = BC_DIRICHLET

    def __init__(self, value):
        self.value = value

    def values(self):
        return self.value, 0.0


class ZeroGradient(BoundaryCondition):
    """``phi_ip = phi_P + grad_P . (d - (d.n) n)``."""

    kind = BC_ZERO_GRADIENT


class Neumann(BoundaryCondition):
    """``phi_ip = phi_P + corr - (q / Gamma_P)(d.n)``."""

    kind = BC_NEUMANN

    def __init__(self, flux):
        self.flux = flux

    def values(self):
        return self.flux, 0.0


class Convective(BoundaryCondition):
    """Robin exchange ``Gamma_P dphi/dn = h (Tinf - phi_ip)``."""

    kind = BC_CONVECTIVE

    def __init__(self, h, t_inf):
        self.h = h
        self.t_inf = t_inf

    def values(self):
        return self.h, self.t_inf
