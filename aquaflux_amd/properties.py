"""Per-cell material properties (``aquaflux/properties/model.py``, ``property.py``).

A property value may be a Python float or a torch tensor; a tensor that requires grad makes the
residual differentiable with respect to it (the reference's differentiable ``Constant.value``).
"""

from __future__ import annotations

import torch


class Property:
    def evaluate(self, n_cells: int, device) -> torch.Tensor:  # pragma: no cover - interface
        raise NotImplementedError

    def scaled(self, factor) -> "Property":  # pragma: no cover - interface
        raise NotImplementedError


class Constant(Property):
    """A uniform value everywhere (``aquaflux/properties/property.py:73-89``)."""

    def __init__(self, value):
        self.value = value

    def evaluate(self, n_cells, device):
        v = self.value
        if isinstance(v, torch.Tensor):
            return v.to(device=device, dtype=torch.float64).reshape(()).expand(n_cells).contiguous()
        return torch.full((n_cells,), float(v), dtype=torch.float64, device=device)

    def scaled(self, factor):
        return Constant(self.value * factor)


class FieldProperty(Property):
    """An explicit per-cell field (``aquaflux/properties/property.py:140-182``)."""

    def __init__(self, values):
        self.values = values

    def evaluate(self, n_cells, device):
        v = torch.as_tensor(self.values, dtype=torch.float64).to(device)
        if v.shape != (n_cells,):
            raise ValueError(f"FieldProperty must have shape ({n_cells},), got {tuple(v.shape)}")
        return v.contiguous()

    def scaled(self, factor):
        return FieldProperty(self.values * factor)


class PropertyModel:
    """Named properties -> per-cell arrays (``aquaflux/properties/model.py:17-78``)."""

    def __init__(self, properties: dict):
        self.properties = dict(properties)

    def require(self, *names):
        missing = [n for n in names if n not in self.properties]
        if missing:
            raise ValueError(f"PropertyModel is missing required properties {missing}")

    def evaluate(self, n_cells: int, device) -> dict:
        return {k: p.evaluate(n_cells, device) for k, p in self.properties.items()}

    def with_scaled(self, name, factor) -> "PropertyModel":
        out = dict(self.properties)
        out[name] = out[name].scaled(factor)
        return PropertyModel(out)

    def __getitem__(self, name):
        return self.properties[name]
