"""The committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py) on the
CPU: the mesh generators still produce the stored meshes bit for bit, the oracle still produces
the stored residuals, tangents and cotangents, and every stored tangent / cotangent pair satisfies
the adjoint identity <w, J v> = <J^T w, v>."""

import numpy as np
import pytest
import torch

from tests.golden import make_golden as MG

NAMES = sorted(MG.CASES)


def _load(name):
    with np.load(MG.path(name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _blocks(a, nb):
    return np.asarray(a, dtype=np.float64).reshape(nb, -1)


@pytest.mark.parametrize("name", NAMES)
def test_fixture_exists_and_is_small(name):
    import os

    assert os.path.getsize(MG.path(name)) < MG.LARGE.get(name, 64 * 1024)


@pytest.mark.parametrize("name", NAMES)
def test_adjoint_identity_holds_in_the_fixture(name):
    z = _load(name)
    lhs = float(np.dot(z["cotangent"], z["jvp"]))
    rhs = float(np.dot(z["vjp_state"], z["tangent"]))
    scale = float(np.abs(z["cotangent"]).dot(np.abs(z["jvp"])))
    assert abs(lhs - rhs) <= 1e-12 * scale


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_the_fixture(name):
    torch.set_num_threads(1)
    z = _load(name)
    fresh = MG.evaluate(name)
    assert set(fresh) == set(z)
    for key in ("owner", "neighbour", "node_coords"):
        assert np.array_equal(fresh[key], z[key]), key
    nb = int(z["n_fields"])
    for key in ("residual", "jvp", "vjp_state"):
        a, b = _blocks(fresh[key], nb), _blocks(z[key], nb)
        for i in range(nb):
            scale = max(float(np.abs(b[i]).max()), 1e-300)
            assert float(np.abs(a[i] - b[i]).max()) <= 1e-13 * scale, (key, i)
    for key in z:
        if key.startswith("vjp_") and key != "vjp_state":
            scale = max(float(np.abs(z[key]).max()), 1e-300)
            assert float(np.abs(fresh[key] - z[key]).max()) <= 1e-13 * scale, key
