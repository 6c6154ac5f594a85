"""Grid lattices and n-th nearest-neighbour bond lists (oracle).

Restates ``quantax/sites/lattice.py:52-93`` (index/xyz layout, row-major with
the last axis fastest), ``quantax/sites/lattice.py:139-179`` (shortest-image
distance under periodic boundaries) and ``quantax/sites/sites.py:224-283``
(``get_neighbor`` / ``_compute_neighbor``: the n-th smallest distance, pairs
from ``np.argwhere`` with i < j, i.e. sorted by i then j).
"""
from __future__ import annotations

import numpy as np


def grid_xyz(extent):
    """xyz_from_index for a one-site-per-cell grid (``lattice.py:73-79``)."""
    extent = tuple(int(e) for e in extent)
    n = int(np.prod(extent))
    idx = np.arange(n)
    xyz = []
    for a in range(len(extent)):
        later = int(np.prod(extent[a + 1:])) if a + 1 < len(extent) else 1
        xyz.append(idx // later % extent[a])
    return np.stack(xyz, axis=1)


def grid_dist(extent, boundary=1):
    """Distance matrix with shortest periodic image per axis (``lattice.py:139-179``)."""
    extent = np.asarray(extent, dtype=np.int64)
    bc = np.broadcast_to(np.asarray(boundary), extent.shape)
    xyz = grid_xyz(extent)
    d = np.abs(xyz[:, None, :] - xyz[None, :, :]).astype(np.float64)
    for a in range(len(extent)):
        if bc[a] != 0:
            d[..., a] = np.minimum(d[..., a], extent[a] - d[..., a])
    return np.linalg.norm(d, axis=-1)


def get_neighbor(extent, n_neighbor=1, boundary=1):
    """n-th nearest-neighbour pairs (``sites.py:224-283``).

    Returns an (nbonds, 2) int array when ``n_neighbor`` is an int, else a list.
    """
    dist = grid_dist(extent, boundary)
    tol = 1e-6
    ns = [n_neighbor] if isinstance(n_neighbor, (int, np.integer)) else list(n_neighbor)
    out = []
    min_dist = tol
    for _ in range(1, max(ns) + 1):
        min_dist = np.min(dist[dist > min_dist])
        nb = np.argwhere(np.abs((dist - min_dist) / min_dist) < tol)
        nb = nb[nb[:, 0] < nb[:, 1]]
        out.append(nb)
        min_dist *= 1 + tol
    res = [out[n - 1] for n in ns]
    return res[0] if isinstance(n_neighbor, (int, np.integer)) else res


def site_neighbor_table(extent, n_neighbor=1, boundary=1):
    """Per-site neighbour table padded with -1 (``common_samplers.py:36-55``).

    Row i lists, in increasing order, the sites bonded to i; width is the
    maximum neighbour count.
    """
    ns = [n_neighbor] if isinstance(n_neighbor, (int, np.integer)) else list(n_neighbor)
    nbs = get_neighbor(extent, ns, boundary)
    nbs = np.concatenate(nbs, axis=0)
    N = int(np.prod(extent))
    mat = np.zeros((N, N), dtype=bool)
    mat[nbs[:, 0], nbs[:, 1]] = True
    mat = mat | mat.T
    width = int(mat.sum(axis=1).max())
    table = -np.ones((N, width), dtype=np.int32)
    for i in range(N):
        j = np.flatnonzero(mat[i])
        table[i, : j.size] = j
    return table
