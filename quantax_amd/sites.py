"""Grid lattices with periodic/open boundaries (spins only).

Mirrors ``quantax/sites/lattice.py`` and ``sites/common_lattices.py:7-46``:
site index = row-major over the extent (last axis fastest); the n-th nearest
neighbours are the pairs (i < j, sorted by i then j) at the n-th smallest
non-zero shortest-image distance (``sites/sites.py:224-283``).  For grids the
squared distance is an integer, so shells are found by exact comparison.
"""
from __future__ import annotations

from typing import Optional, Sequence, Tuple, Union

import numpy as np

from . import global_defs


class Lattice:
    def __init__(self, extent: Sequence[int], boundary: Union[int, Sequence[int]] = 1,
                 Nparticles: Union[None, Tuple[int, int]] = None):
        self._extent = tuple(int(e) for e in extent)
        nd = len(self._extent)
        self._boundary = np.broadcast_to(np.asarray(boundary, dtype=int), (nd,)).copy()
        if np.any(self._boundary == -1):
            raise ValueError("Spin system can't have anti-periodic boundary conditions.")
        self._N = int(np.prod(self._extent))
        if Nparticles is not None:
            Nparticles = tuple(int(v) for v in Nparticles)
            if sum(Nparticles) != self._N:
                raise ValueError("Nup + Ndown must equal the number of sites for spins.")
        self._Nparticles = Nparticles
        strides = [int(np.prod(self._extent[a + 1:])) for a in range(nd)]
        idx = np.arange(self._N)
        self._xyz = np.stack([(idx // st) % e for st, e in zip(strides, self._extent)], axis=1)
        self._shells = None
        global_defs.set_sites(self)

    # -- descriptors
    @property
    def Nsites(self) -> int:
        return self._N

    @property
    def Nmodes(self) -> int:
        return self._N

    @property
    def N(self) -> int:
        return self._N

    @property
    def shape(self) -> Tuple[int, ...]:
        return (1,) + self._extent

    @property
    def extent(self) -> Tuple[int, ...]:
        return self._extent

    @property
    def ndim(self) -> int:
        return len(self._extent)

    @property
    def boundary(self) -> np.ndarray:
        return self._boundary.copy()

    @property
    def Nparticles(self):
        return self._Nparticles if self._Nparticles is not None else self._N

    @property
    def xyz_from_index(self) -> np.ndarray:
        return self._xyz.copy()

    # -- neighbours
    def _dist2(self) -> np.ndarray:
        d = np.abs(self._xyz[:, None, :] - self._xyz[None, :, :])
        for a, (e, bc) in enumerate(zip(self._extent, self._boundary)):
            if bc != 0:
                d[..., a] = np.minimum(d[..., a], e - d[..., a])
        return (d * d).sum(-1)

    def get_neighbor(self, n_neighbor: Union[int, Sequence[int]] = 1):
        d2 = self._dist2()
        if self._shells is None:
            self._shells = np.unique(d2[d2 > 0])
        ns = [n_neighbor] if np.isscalar(n_neighbor) else list(n_neighbor)
        out = []
        iu, ju = np.triu_indices(self._N, k=1)
        for n in ns:
            if n < 1 or n > self._shells.size:
                raise ValueError(f"no {n}-th neighbour shell on this lattice")
            sel = d2[iu, ju] == self._shells[n - 1]
            out.append(np.stack([iu[sel], ju[sel]], axis=1))
        return out[0] if np.isscalar(n_neighbor) else out

    def neighbor_table(self, n_neighbor: Union[int, Sequence[int]] = 1) -> np.ndarray:
        """Per-site neighbour rows padded with -1 (``common_samplers.py:36-55``)."""
        ns = [n_neighbor] if np.isscalar(n_neighbor) else list(n_neighbor)
        pairs = np.concatenate(self.get_neighbor(ns), axis=0)
        adj = [set() for _ in range(self._N)]
        for i, j in pairs:
            adj[i].add(int(j))
            adj[j].add(int(i))
        width = max(len(a) for a in adj)
        tab = -np.ones((self._N, width), dtype=np.int32)
        for i, a in enumerate(adj):
            row = sorted(a)
            tab[i, : len(row)] = row
        return tab


def Grid(extent: Sequence[int], boundary=1, Nparticles=None) -> Lattice:
    return Lattice(extent, boundary, Nparticles)


def Chain(L: int, boundary=1, Nparticles=None) -> Lattice:
    return Lattice([L], boundary, Nparticles)


def Square(L: int, boundary=1, Nparticles=None) -> Lattice:
    return Lattice([L, L], boundary, Nparticles)
