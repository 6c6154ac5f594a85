"""Symmetries of spin systems and the projection of a wavefunction onto a
symmetry sector.

Mirrors ``quantax/symmetry/symmetry.py:85-432`` (``Symmetry``: group elements
generated from permutation generators and sectors, characters, spin
inversion, ``get_symm_spins``, ``symmetrize``, ``@``),
``common_symmetries.py:13-278`` (Identity, Z2Inversion, SpinInverse,
LinearTransform, Flip, Rotation, C4v, D6) and ``translation.py:7-99``
(Translation, TransND) for real wavefunctions on grid lattices:

    psi_sym(s) = sum_g c_g psi(T_g s),   c_g = chi_g chi_0 / |G|,

with the images T_g s = s[perm_g] followed, under spin inversion, by
-s[perm_g].  The group elements are listed in the reference's order (each
generator's powers g^0, g^1, ... nested after the previous generators'), so
characters and images line up with a quantax run.  Characters are real:
a sector must give chi = +-1, as the reference requires when its default
dtype is real.  On the device the images come from ``qtx_symm_images`` and
the projection from ``qtx_symmetrize``.
"""
from __future__ import annotations

from typing import Optional, Sequence, Union

import numpy as np
import torch

from . import _lib
from .global_defs import get_sites


def _group(generators: np.ndarray, sectors: Sequence[int], N: int):
    """Group elements and characters from generators (``symmetry.py:11-53``)."""
    s0 = np.arange(N)
    perm = s0[None, :]
    character = np.ones(1)
    for g, sec in zip(generators, sectors):
        powers = [s0]
        p = g.copy()
        while not np.array_equal(p, s0):
            powers.append(p)
            p = p[g]
        powers = np.stack(powers)
        n = powers.shape[0]
        if not 0 <= sec < n:
            raise ValueError(f"Sector {sec} out of range.")
        if (2 * sec) % n != 0:
            raise ValueError("Wavefunctions here are real, but the sector gives complex characters.")
        chi = -1.0 if sec else 1.0
        perm = perm[:, powers].reshape(-1, N)
        character = np.outer(character, chi ** np.arange(n)).reshape(-1)
    return perm, character


class Symmetry:
    """A symmetry group with a one- (or higher-, through ``character``)
    dimensional representation; ``symmetry.py:85-432``."""

    def __init__(self, generator: Optional[np.ndarray] = None, sector: Union[int, Sequence[int]] = 0,
                 Z2_inversion: int = 0, perm: Optional[np.ndarray] = None, character=None):
        sites = get_sites()
        N = sites.Nmodes
        self._N = N
        if generator is None:
            generator = np.arange(N)[None, :]
        generator = np.atleast_2d(np.asarray(generator, dtype=np.int64))
        if generator.shape[1] != N:
            raise ValueError(f"Got a generator with size {generator.shape[1]}, incompatible with the system size {N}.")
        self._generator = generator
        self._sector = [int(sector)] * generator.shape[0] if np.isscalar(sector) else [int(v) for v in sector]
        if Z2_inversion not in (-1, 0, 1):
            raise ValueError("Z2_inversion must be -1, 0 or 1")
        self._Z2 = int(Z2_inversion)
        if perm is None or character is None:
            p, c = _group(generator, self._sector, N)
            perm = p if perm is None else perm
            character = c if character is None else character
        # C order: products of groups come out of fancy indexing with other strides
        self._perm = np.ascontiguousarray(np.asarray(perm, dtype=np.int64).reshape(-1, N))
        self._character = np.asarray(character, dtype=np.float64).reshape(-1)
        if self._character.size != self._perm.shape[0]:
            raise ValueError("one character per group element")
        self._dev = None

    # -- descriptors
    @property
    def Nmodes(self) -> int:
        return self._N

    @property
    def perm(self) -> np.ndarray:
        return self._perm.copy()

    @property
    def character(self) -> np.ndarray:
        """Characters of the permutation elements (``symmetry.py:220-223``)."""
        return self._character.copy()

    @property
    def Z2_inversion(self) -> int:
        return self._Z2

    @property
    def nsymm(self) -> int:
        n = self._character.size
        return n if self._Z2 == 0 else 2 * n

    @property
    def is_identity(self) -> bool:
        return self.nsymm == 1

    @property
    def coefficients(self) -> np.ndarray:
        """c_g = chi_g chi_0 / |G| over all images (``symmetry.py:386-391``)."""
        chi = self._character
        if self._Z2 != 0:
            chi = np.concatenate([chi, self._Z2 * chi])
        return chi * chi[0] / chi.size

    # -- host forms of the reference functions
    def get_symm_spins(self, spins) -> np.ndarray:
        """All images of one configuration (``symmetry.py:324-341``)."""
        spins = np.asarray(spins)
        if spins.ndim > 1:
            raise ValueError(f"Input spins should be 1D, got dimension {spins.ndim}")
        out = spins[self._perm]
        if self._Z2 != 0:
            out = np.concatenate([out, -out], axis=0)
        return out

    def symmetrize(self, psi) -> float:
        """psi_sym from the images' psi values (``symmetry.py:343-392``)."""
        psi = np.asarray(psi)
        if psi.ndim > 1:
            raise ValueError(f"Input wavefunction should be 1D, got dimension {psi.ndim}")
        return float(np.sum(psi * self.coefficients))

    # -- device
    def _device(self):
        if self._dev is None:
            dev = _lib.device()
            self._dev = (torch.from_numpy(np.ascontiguousarray(self._perm, dtype=np.int32)).to(dev).contiguous(),
                         torch.from_numpy(np.ascontiguousarray(self.coefficients)).to(dev).contiguous())
        return self._dev

    def images(self, s: torch.Tensor) -> torch.Tensor:
        """(ns, N) int8 -> (ns * nsymm, N) int8, images of each sample contiguous."""
        s = s.contiguous()
        perm, _ = self._device()
        ns = s.shape[0]
        out = torch.empty((ns * self.nsymm, self._N), dtype=torch.int8, device=s.device)
        _lib.call("qtx_symm_images", _lib.stream(), s.data_ptr(), ns, self._N, perm.data_ptr(), perm.shape[0],
                  1 if self._Z2 else 0, out.data_ptr())
        return out

    def project(self, logabs: torch.Tensor, weights: bool = False):
        """(ns * nsymm,) image log|psi| -> (sign, logabs) of psi_sym per sample,
        and the Jacobian weights of the images when asked."""
        _, coef = self._device()
        nimg = self.nsymm
        ns = logabs.numel() // nimg
        la = logabs.contiguous()
        sg = torch.empty(ns, dtype=torch.float64, device=la.device)
        lo = torch.empty(ns, dtype=torch.float64, device=la.device)
        w = torch.empty((ns, nimg), dtype=torch.float64, device=la.device) if weights else None
        _lib.call("qtx_symmetrize", _lib.stream(), la.data_ptr(), None, coef.data_ptr(), ns, nimg, sg.data_ptr(),
                  lo.data_ptr(), _lib.ptr(w))
        return (sg, lo, w) if weights else (sg, lo)

    def __matmul__(self, other: "Symmetry") -> "Symmetry":
        """Product group (``symmetry.py:394-432``)."""
        generator = np.concatenate([self._generator, other._generator], axis=0)
        sector = [*self._sector, *other._sector]
        perm = self._perm[:, other._perm].reshape(-1, self._N)
        character = np.outer(self._character, other._character).reshape(-1)
        if self._Z2 == 0:
            z2 = other._Z2
        elif other._Z2 == 0 or other._Z2 == self._Z2:
            z2 = self._Z2
        else:
            raise ValueError("Symmetry with different Z2_inversion can't be added")
        return Symmetry(generator, sector, z2, perm, character)


_identity = {}


def Identity() -> Symmetry:
    """No symmetry (``common_symmetries.py:13-18``)."""
    key = get_sites().Nmodes
    if key not in _identity:
        _identity[key] = Symmetry()
    return _identity[key]


def Z2Inversion(eigval: int = 1) -> Symmetry:
    """Global spin flip with eigenvalue +-1 (``common_symmetries.py:21-41``)."""
    if eigval not in (1, -1):
        raise ValueError("'eigval' of Z2Inversion should be 1 or -1.")
    return Symmetry(Z2_inversion=eigval)


def SpinInverse(eigval: int = 1) -> Symmetry:
    """Spin inversion of a spin system; 0 means no inversion
    (``common_symmetries.py:44-71``)."""
    if eigval == 0:
        return Identity()
    return Z2Inversion(eigval)


def _standardize(coord: np.ndarray, lat) -> np.ndarray:
    """Wrap coordinates into the periodic box (``common_symmetries.py:89-100``,
    unit basis vectors)."""
    coord = coord.copy()
    ext = np.asarray(lat.extent, dtype=np.float64)
    per = lat.boundary != 0
    c = coord[:, per] % ext[per]
    c[np.isclose(c, ext[per])] = 0.0
    coord[:, per] = c
    return coord


def LinearTransform(matrix: np.ndarray, center: Optional[np.ndarray] = None, sector: int = 0,
                    character=None) -> Symmetry:
    """Symmetry generated by a linear map of the lattice about ``center``
    (default: the mean site position), ``common_symmetries.py:103-143``."""
    lat = get_sites()
    xyz = lat.xyz_from_index.astype(np.float64)
    if center is None:
        center = xyz.mean(axis=0)
    coord = _standardize(xyz, lat)
    new = np.einsum("ij,nj->ni", np.asarray(matrix, dtype=np.float64), xyz - center) + center
    new = _standardize(new, lat)
    match = np.all(np.isclose(coord[:, None, :], new[None, :, :]), axis=-1)
    if not np.all(match.sum(axis=1) == 1):
        raise ValueError("The transformation does not map the lattice to itself.")
    generator = np.argmax(match, axis=1)
    if character is None:
        return Symmetry(generator, sector)
    perm, _ = _group(generator[None, :], [0], lat.Nmodes)
    return Symmetry(generator, sector, perm=perm, character=np.asarray(character, dtype=np.float64))


def Flip(axis: Union[int, Sequence[int]] = 0, center: Optional[np.ndarray] = None, sector: int = 0) -> Symmetry:
    """Reflection along ``axis`` (``common_symmetries.py:146-162``)."""
    m = np.ones(get_sites().ndim)
    m[np.asarray(axis)] = -1
    return LinearTransform(np.diag(m), center, sector)


def Rotation(angle: float, axes: Sequence[int] = (0, 1), center: Optional[np.ndarray] = None, sector: int = 0,
             character=None) -> Symmetry:
    """Rotation in the plane of ``axes`` (``common_symmetries.py:165-202``)."""
    nd = get_sites().ndim
    if max(axes) >= nd:
        raise ValueError(f"The rotated axis {max(axes)} is out-of-bound for a {nd}-D system")
    c, s = np.cos(angle), np.sin(angle)
    m = np.eye(nd)
    x, y = axes
    m[x, x], m[x, y], m[y, x], m[y, y] = c, -s, s, c
    return LinearTransform(m, center, sector, character)


def C4v(center: Optional[np.ndarray] = None, repr: str = "A1") -> Symmetry:
    """C4v point group of the square lattice (``common_symmetries.py:205-240``)."""
    if repr in ("A1", "A2", "B1", "B2"):
        rsec = 0 if repr in ("A1", "A2") else 2
        fsec = 0 if repr in ("A1", "B1") else 1
        return Rotation(np.pi / 2, center=center, sector=rsec) @ Flip(center=center, sector=fsec)
    if repr == "E":
        return Rotation(np.pi, center=center, character=[2.0, -2.0])
    raise ValueError("'repr' should be one of the following: 'A1', 'A2', 'B1', 'B2' or 'E'")


def D6(center: Optional[np.ndarray] = None, repr: str = "A1") -> Symmetry:
    """D6 point group (``common_symmetries.py:243-278``); needs a lattice that
    a 60-degree rotation maps to itself, which no grid lattice is."""
    if repr in ("A1", "A2", "B1", "B2"):
        rsec = 0 if repr in ("A1", "A2") else 3
        fsec = 0 if repr in ("A1", "B1") else 1
        return Rotation(np.pi / 3, center=center, sector=rsec) @ Flip(center=center, sector=fsec)
    if repr == "E1":
        return Rotation(np.pi / 3, character=[2.0, 1.0, -1.0, -2.0, -1.0, 1.0])
    if repr == "E2":
        return Rotation(np.pi / 3, character=[2.0, -1.0, -1.0, 2.0, -1.0, -1.0])
    raise ValueError("'repr' should be one of the following: 'A1', 'A2', 'B1', 'B2', 'E1' or 'E2'")


def Translation(vectors: Sequence, sector: Union[int, Sequence[int]] = 0) -> Symmetry:
    """Translations by ``vectors`` in lattice units (``translation.py:7-56``):
    site i of the image holds the spin at x_i + v."""
    lat = get_sites()
    vectors = np.asarray(vectors, dtype=np.int64).reshape(-1, lat.ndim)
    if np.any((vectors != 0) & (lat.boundary[None, :] == 0)):
        raise ValueError("Translation symmetry can't be imposed on open boundary.")
    ext = np.asarray(lat.extent, dtype=np.int64)
    strides = np.array([int(np.prod(ext[a + 1:])) for a in range(lat.ndim)], dtype=np.int64)
    xyz = lat.xyz_from_index.astype(np.int64)
    gens = []
    for vec in vectors:
        for a, v in enumerate(vec):
            if v != 0 and ext[a] % v != 0:
                raise ValueError(f"Translation vector must be compatible with lattice shape, got lattice shape "
                                 f"{tuple(ext)} and vector {tuple(vec)}.")
        gens.append(((xyz + vec[None, :]) % ext) @ strides)
    return Symmetry(np.stack(gens), sector)


def TransND(sector: Union[int, Sequence[int]] = 0) -> Symmetry:
    """All translations by unit lattice vectors (``translation.py:93-98``)."""
    return Translation(np.identity(get_sites().ndim, dtype=np.int64), sector)
