"""Symmetry groups on the host: the package's group construction against the
oracle's, and both against what group theory fixes — closure, the group
orders, multiplicative 1D characters, and the projection property
psi_sym(T_h s) = chi_h psi_sym(s), psi_sym(-s) = z2 psi_sym(s) for any psi."""
import numpy as np
import pytest

import quantax_amd as qtx
from oracle import models as om
from oracle import symmetry as osym


def _oracle(kind, lat, **kw):
    xyz, ext = lat.xyz_from_index, lat.extent
    N = lat.Nsites
    if kind == "c4v":
        rsec, fsec = kw["rsec"], kw["fsec"]
        r = osym.group(osym.rotation(xyz, ext, np.pi / 2)[None], [rsec], N)
        f = osym.group(osym.flip(xyz, ext)[None], [fsec], N)
        return osym.product((*r, 0), (*f, 0))
    if kind == "trans":
        gens = np.stack([osym.translation(xyz, ext, v) for v in np.identity(len(ext), dtype=int)])
        p, c = osym.group(gens, kw["sec"], N)
        return p, c, 0


@pytest.mark.parametrize("rep,rsec,fsec", [("A1", 0, 0), ("A2", 0, 1), ("B1", 2, 0), ("B2", 2, 1)])
def test_c4v_matches_oracle_and_is_a_group(rep, rsec, fsec):
    lat = qtx.sites.Square(4)
    d = qtx.symmetry.C4v(repr=rep)
    p, c, _ = _oracle("c4v", lat, rsec=rsec, fsec=fsec)
    assert np.array_equal(d.perm, p) and np.array_equal(d.character, c)
    assert d.nsymm == 8 and np.array_equal(p[0], np.arange(16))
    elems = {tuple(x): i for i, x in enumerate(p)}
    assert len(elems) == 8
    for a in range(8):
        for b in range(8):
            k = elems[tuple(p[a][p[b]])]  # closure
            assert c[k] == c[a] * c[b]  # one-dimensional representation


def test_space_group_orders():
    lat = qtx.sites.Square(4)
    t = qtx.symmetry.TransND()
    assert t.nsymm == 16
    full = t @ qtx.symmetry.C4v() @ qtx.symmetry.SpinInverse(-1)
    assert full.nsymm == 256 and len({tuple(x) for x in full.perm}) == 128
    lat6 = qtx.sites.Square(6)
    assert (qtx.symmetry.TransND() @ qtx.symmetry.C4v()).nsymm == 288
    p, c, _ = _oracle("trans", lat6, sec=[0, 3])
    d = qtx.symmetry.TransND(sector=[0, 3])
    assert np.array_equal(d.perm, p) and np.array_equal(d.character, c)


@pytest.mark.parametrize("rep,z2", [("A1", 1), ("B2", -1), ("A2", 0), ("B1", 1)])
def test_projection_property(rep, z2, rng):
    lat = qtx.sites.Square(4)
    symm = qtx.symmetry.C4v(repr=rep)
    if z2:
        symm = symm @ qtx.symmetry.SpinInverse(z2)
    W = rng.standard_normal((24, 16)) * 0.3
    b = rng.standard_normal(24) * 0.1
    fn = osym.logpsi_sym_fn(lambda x: om.rbm_logpsi(W, b, x)[1], symm.perm, symm.character, symm.Z2_inversion)
    s = (2 * rng.integers(0, 2, (6, 16)) - 1).astype(np.int8)
    sg, la = fn(s)
    ref = sg * np.exp(la)
    for h, chi in zip(symm.perm, symm.character):
        sg2, la2 = fn(s[:, h])
        assert np.allclose(sg2 * np.exp(la2), chi * ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    if z2:
        sg2, la2 = fn(-s)
        assert np.allclose(sg2 * np.exp(la2), z2 * ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # the package's host symmetrize agrees with the oracle on the same psi values
    x = s[0]
    psi_img = np.exp(om.rbm_logpsi(W, b, symm.get_symm_spins(x))[1])
    assert np.isclose(symm.symmetrize(psi_img), ref[0], rtol=1e-12)


def test_invalid_symmetries():
    qtx.sites.Square(4)
    with pytest.raises(ValueError):
        qtx.symmetry.Rotation(np.pi / 2, sector=1)  # complex character
    with pytest.raises(ValueError):
        qtx.symmetry.D6()  # a grid is not mapped to itself by 60 degrees
    e = qtx.symmetry.C4v(repr="E")
    assert np.array_equal(e.coefficients, [2.0, -2.0])


def test_group_tables_are_c_ordered():
    """The device copy of the element table is taken from raw memory: products
    of groups (fancy indexing) must still be stored row by row."""
    qtx.sites.Square(4)
    for symm in (qtx.symmetry.C4v() @ qtx.symmetry.SpinInverse(1),
                 qtx.symmetry.TransND() @ qtx.symmetry.C4v(repr="B1")):
        assert symm._perm.flags["C_CONTIGUOUS"]
