"""Named boundary-condition collections and scalar closures (``aquaflux/boundary``)."""

from __future__ import annotations

import numpy as np
import torch


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
# ``face_value`` keeps the reference's per-face closure seam (``BoundaryCondition.face_value``,
# ``boundary/conditions.py:69-108``) as torch arithmetic on the caller's arrays; the residual
# kernels evaluate the same closures inside their face loops.

BC_NONE, BC_DIRICHLET, BC_ZERO_GRADIENT, BC_NEUMANN, BC_CONVECTIVE = 0, 1, 2, 3, 4


def _dot(a, b):
    """Left-to-right component sum of ``a * b`` over the last axis (``vectors.dot``)."""
    p = a * b
    out = p[..., 0]
    for i in range(1, p.shape[-1]):
        out = out + p[..., i]
    return out


def tangential_correction(grad_owner, d, normal):
    """``grad_P . (d - (d.n) n)`` (``boundary/conditions.py:14-36``)."""
    return _dot(grad_owner, d - _dot(d, normal)[..., None] * normal)


class BoundaryCondition:
    kind = BC_NONE

    def values(self):
        """(value, t_inf) carried per face."""
        return 0.0, 0.0

    def face_values(self, face_centroid):
        """Per-face ``(value, t_inf)`` on faces with centroids ``(n, d)`` (constants broadcast)."""
        n, dev = face_centroid.shape[0], face_centroid.device
        out = []
        for x in self.values():
            xt = x if isinstance(x, torch.Tensor) else torch.tensor(float(x), dtype=torch.float64)
            out.append(xt.to(device=dev, dtype=torch.float64).reshape(()).expand(n))
        return tuple(out)

    def face_value(self, phi_owner, grad_owner, d, normal, gamma_owner, face_centroid):
        """Boundary-face values ``phi_ip`` ``(n,)`` (``BoundaryCondition.face_value``)."""
        raise NotImplementedError


class Dirichlet(BoundaryCondition):
    """``phi_ip = value`` (``boundary/conditions.py:111-123``)."""

    kind = BC_DIRICHLET

    def __init__(self, value):
        self.value = value

    def values(self):
        return self.value, 0.0

    def face_value(self, phi_owner, grad_owner, d, normal, gamma_owner, face_centroid):
        v = self.value if isinstance(self.value, torch.Tensor) else torch.tensor(float(self.value))
        return v.to(device=phi_owner.device, dtype=phi_owner.dtype).expand(phi_owner.shape)


class DirichletField(BoundaryCondition):
    """Prescribed value ``phi_ip = field_fn(x_ip)`` at the face centroid (``boundary/conditions.py:
    126-142``): ``field_fn`` maps centroids ``(n, d)`` to values ``(n,)``.  The residual kernels treat
    it as a Dirichlet face whose value is ``field_fn`` evaluated once at build, at the geometry's face
    centroids."""

    kind = BC_DIRICHLET

    def __init__(self, field_fn):
        self.field_fn = field_fn

    def values(self):
        raise ValueError("DirichletField has no single value; its values depend on the face centroids")

    def face_values(self, face_centroid):
        v = self.field_fn(face_centroid)
        v = v if isinstance(v, torch.Tensor) else torch.as_tensor(np.asarray(v), dtype=torch.float64)
        v = v.to(device=face_centroid.device, dtype=torch.float64)
        if v.shape != (face_centroid.shape[0],):
            raise ValueError(f"DirichletField: field_fn must return shape ({face_centroid.shape[0]},), "
                             f"got {tuple(v.shape)}")
        return v, torch.zeros_like(v)

    def face_value(self, phi_owner, grad_owner, d, normal, gamma_owner, face_centroid):
        return self.face_values(face_centroid)[0]


class ZeroGradient(BoundaryCondition):
    """``phi_ip = phi_P + grad_P . (d - (d.n) n)`` (``boundary/conditions.py:145-153``)."""

    kind = BC_ZERO_GRADIENT

    def face_value(self, phi_owner, grad_owner, d, normal, gamma_owner, face_centroid):
        return phi_owner + tangential_correction(grad_owner, d, normal)


class Neumann(BoundaryCondition):
    """``phi_ip = phi_P + corr - (q / Gamma_P)(d.n)`` (``boundary/conditions.py:156-173``)."""

    kind = BC_NEUMANN

    def __init__(self, flux):
        self.flux = flux

    def values(self):
        return self.flux, 0.0

    def face_value(self, phi_owner, grad_owner, d, normal, gamma_owner, face_centroid):
        corr = tangential_correction(grad_owner, d, normal)
        return phi_owner + corr - self.flux / gamma_owner * _dot(d, normal)


class Convective(BoundaryCondition):
    """Robin exchange ``Gamma_P dphi/dn = h (Tinf - phi_ip)`` (``boundary/conditions.py:176-201``):
    ``phi_ip = (phi_P + corr + beta T_inf) / (1 + beta)``, ``beta = (h / Gamma_P)(d.n)``."""

    kind = BC_CONVECTIVE

    def __init__(self, h, t_inf):
        self.h = h
        self.t_inf = t_inf

    def values(self):
        return self.h, self.t_inf

    def face_value(self, phi_owner, grad_owner, d, normal, gamma_owner, face_centroid):
        corr = tangential_correction(grad_owner, d, normal)
        beta = self.h / gamma_owner * _dot(d, normal)
        return (phi_owner + corr + beta * self.t_inf) / (1.0 + beta)
