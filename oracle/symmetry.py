"""Symmetry groups and the projected wavefunction (oracle, test infrastructure).

Restates, for spin systems with real characters on grid lattices,
``quantax/symmetry/symmetry.py:11-53`` (``_get_perm``: group elements as
nested powers of the generators, characters exp(-2 pi i q m / M) -> +-1),
``:324-341`` (``get_symm_spins``), ``:343-392`` (``symmetrize``), ``:394-432``
(``@``), ``common_symmetries.py:89-240`` (LinearTransform, Flip, Rotation,
C4v) and ``translation.py:7-99`` (Translation, TransND).  A group here is the
pair (perm, character) plus the spin-inversion eigenvalue z2.
"""
from __future__ import annotations

import numpy as np


def group(generators, sectors, N):
    """``_get_perm`` (symmetry.py:11-53) for real characters."""
    perm = np.arange(N)[None]
    chars = np.array([1.0])
    for g, q in zip(np.atleast_2d(generators), sectors):
        elems = [np.arange(N)]
        while True:
            nxt = elems[-1][g]
            if np.array_equal(nxt, elems[0]):
                break
            elems.append(nxt)
        M = len(elems)
        assert (2 * q) % M == 0, "complex character"
        chi = np.cos(2 * np.pi * q / M) ** np.arange(M)  # +-1 for q in {0, M/2}
        perm = np.stack([p[e] for p in perm for e in elems])
        chars = np.array([a * b for a in chars for b in chi])
    return perm, np.rint(chars)


def product(a, b):
    """``Symmetry.__matmul__`` (symmetry.py:413-432): (perm_a, chi_a, z2a) @ (perm_b, chi_b, z2b)."""
    (pa, ca, za), (pb, cb, zb) = a, b
    perm = np.stack([p[q] for p in pa for q in pb])
    chars = np.array([x * y for x in ca for y in cb])
    z2 = za if zb == 0 else zb
    return perm, chars, z2


def images(s, perm, z2):
    """``get_symm_spins`` (symmetry.py:324-341)."""
    out = s[perm]
    return np.concatenate([out, -out]) if z2 else out


def coefficients(chars, z2):
    """symmetry.py:386-391: c = chi chi_0 / |G| over [chi, z2 chi]."""
    chi = np.concatenate([chars, z2 * chars]) if z2 else chars
    return chi * chi[0] / chi.size


def psi_sym(logpsi_fn, s, perm, chars, z2):
    """psi_sym(s) = sum_g c_g psi(T_g s) for a batch s (Ns, N); logpsi_fn gives
    log|psi| of positive wavefunctions (RBM, ResConv)."""
    c = coefficients(chars, z2)
    out = np.empty(s.shape[0])
    for i, x in enumerate(s):
        la = logpsi_fn(images(x, perm, z2))
        m = la.max()
        out[i] = np.exp(m) * np.sum(c * np.exp(la - m))
    return out


def linear_transform(xyz, extent, matrix, center=None):
    """``LinearTransform`` generator (common_symmetries.py:103-143) on a grid
    with unit basis vectors and periodic boundaries."""
    xyz = np.asarray(xyz, np.float64)
    ext = np.asarray(extent, np.float64)
    center = xyz.mean(0) if center is None else center
    new = (xyz - center) @ np.asarray(matrix, np.float64).T + center
    wrap = lambda c: np.where(np.isclose(c % ext, ext), 0.0, c % ext)
    a, b = wrap(xyz), wrap(new)
    match = np.all(np.isclose(a[:, None, :], b[None, :, :]), -1)
    assert np.all(match.sum(1) == 1)
    return np.argmax(match, 1)


def rotation(xyz, extent, angle):
    c, s = np.cos(angle), np.sin(angle)
    return linear_transform(xyz, extent, [[c, -s], [s, c]])


def flip(xyz, extent, axis=0):
    m = np.ones(len(extent))
    m[axis] = -1
    return linear_transform(xyz, extent, np.diag(m))


def translation(xyz, extent, vec):
    """``Translation`` generator (translation.py:26-41): site i -> x_i + v."""
    ext = np.asarray(extent)
    strides = np.array([int(np.prod(ext[a + 1:])) for a in range(len(ext))])
    return ((np.asarray(xyz) + np.asarray(vec)[None]) % ext) @ strides


def logpsi_sym_fn(logabs_fn, perm, chars, z2):
    """s (Ns, N) -> (sign, log|psi_sym|) for a positive wavefunction given by
    its log-amplitude function (the form ``operator.oloc`` takes)."""
    c = coefficients(chars, z2)

    def fn(s):
        s = np.asarray(s)
        sg, la = np.empty(s.shape[0]), np.empty(s.shape[0])
        for i, x in enumerate(s):
            l = logabs_fn(images(x, perm, z2))
            m = l.max()
            v = np.sum(c * np.exp(l - m))
            sg[i] = np.sign(v)
            la[i] = m + np.log(abs(v)) if v != 0 else -np.inf
        return sg, la

    return fn
