"""The slab-tile block order of the cell kernels (host side): a permutation that keeps each XCD's
contiguous share of the blocks and walks it tile by tile along the share's thinnest axis."""

import numpy as np

from aquaflux_amd.device import slab_tile_block_order
from aquaflux_amd.mesh import structured_grid_3d


def _centroids(mesh):
    X = mesh.node_coords
    off, idx = mesh.face_node_offsets, mesh.face_node_indices
    fid = np.repeat(np.arange(mesh.n_faces), np.diff(off))
    fc = np.stack([np.bincount(fid, weights=X[idx, j]) / np.diff(off) for j in range(3)], axis=1)
    inter = mesh.neighbour >= 0
    num = np.stack([np.bincount(mesh.owner, weights=fc[:, j], minlength=mesh.n_cells)
                    + np.bincount(mesh.neighbour[inter], weights=fc[inter, j], minlength=mesh.n_cells)
                    for j in range(3)], axis=1)
    den = np.bincount(mesh.owner, minlength=mesh.n_cells) + np.bincount(mesh.neighbour[inter], minlength=mesh.n_cells)
    return num / den[:, None]


def test_order_is_a_permutation_keeping_each_xcd_share():
    mesh = structured_grid_3d(128, 64, 64, lx=2.0, named_boundaries=True)
    order = slab_tile_block_order(_centroids(mesh))
    nb = mesh.n_cells // 256
    assert np.array_equal(np.sort(order), np.arange(nb))
    share = nb // 8
    for x in range(8):
        assert set(order[x * share:(x + 1) * share] // share) == {x}


def test_structured_channel_walks_tiles_of_rows_through_the_slab():
    """128 x 64 x 64 cells, x fastest: a block is two x-rows at one z; an XCD's share is 8 z-planes.
    With 64 blocks per tile the order is: 4 tiles of 8 row pairs, each walked z-plane by z-plane."""
    mesh = structured_grid_3d(128, 64, 64, lx=2.0, named_boundaries=True)
    order = slab_tile_block_order(_centroids(mesh), per_tile=64)
    expect = [z * 32 + yp for x in range(8) for yt in range(0, 32, 8) for z in range(8 * x, 8 * x + 8)
              for yp in range(yt, yt + 8)]
    assert np.array_equal(order, np.array(expect))


def test_small_and_planar_inputs():
    assert np.array_equal(slab_tile_block_order(np.random.default_rng(0).random((300, 2))), [0, 1])
    o = slab_tile_block_order(np.random.default_rng(1).random((5000, 3)))
    assert np.array_equal(np.sort(o), np.arange(20))
