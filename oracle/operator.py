"""Spin operators in QuSpin op_list form and local estimators (oracle).

Restates:
* ``quantax/operator/site_operator.py:39-132`` — sigma_x/z carry strength 2,
  sigma_+/- strength 1 (Pauli convention: sigma_z sigma_z = s_i s_j).
* ``quantax/operator/operator.py:291-301`` (``@`` concatenates op strings and
  multiplies strengths), ``:402-418`` (``+`` merges equal op strings, keeping
  first-seen order), ``:442-455`` (scalar ``*``).
* ``quantax/operator/common_operators.py:23-71`` — Heisenberg and Ising.
* ``quantax/operator/operator.py:30-78`` (``_apply_site_operator``),
  ``:81-93`` (``_apply_diag``), ``:96-119`` (``_apply_off_diag``),
  ``:144-165`` (``_get_conn``), ``:168-184`` (``_get_Olocx``),
  ``:510-562`` (``Oloc``), ``:564-595`` (``expectation``).
"""
from __future__ import annotations

import numpy as np

from .lattice import get_neighbor

DIAG_OPS = ("I", "n", "z")


class OpList:
    """A QuSpin-format op_list ``[[opstr, [[J, i, j, ...], ...]], ...]``."""

    def __init__(self, op_list):
        self.op_list = [[o, [list(t) for t in terms]] for o, terms in op_list]

    @staticmethod
    def site(opstr, strength, i):
        return OpList([[opstr, [[strength, i]]]])

    def __matmul__(self, other):  # operator.py:291-301
        out = []
        for o1, t1 in self.op_list:
            for o2, t2 in other.op_list:
                terms = []
                for J1, *i1 in t1:
                    for J2, *i2 in t2:
                        terms.append([J1 * J2, *i1, *i2])
                out.append([o1 + o2, terms])
        return OpList(out)

    def __add__(self, other):  # operator.py:402-418
        if not isinstance(other, OpList):
            if other == 0:
                return self
            return NotImplemented
        out = [[o, [list(t) for t in terms]] for o, terms in self.op_list]
        names = tuple(o for o, _ in out)
        for o2, t2 in other.op_list:
            if o2 in names:
                out[names.index(o2)][1] += [list(t) for t in t2]
            else:
                out.append([o2, [list(t) for t in t2]])
        return OpList(out)

    __radd__ = __add__

    def __mul__(self, c):  # operator.py:442-455
        return OpList([[o, [[J * c, *idx] for J, *idx in terms]] for o, terms in self.op_list])

    __rmul__ = __mul__

    def __neg__(self):
        return self * (-1)


def sigma_x(i):
    return OpList.site("x", 2.0, i)


def sigma_z(i):
    return OpList.site("z", 2.0, i)


def sigma_p(i):
    return OpList.site("+", 1.0, i)


def sigma_m(i):
    return OpList.site("-", 1.0, i)


def _sum(ops):
    acc = 0
    for o in ops:
        acc = acc + o
    return acc


def heisenberg(extent, J=1.0, n_neighbor=1, msr=False, boundary=1):
    """``common_operators.py:23-53``."""
    Js = [J] if np.isscalar(J) else list(J)
    ns = [n_neighbor] if np.isscalar(n_neighbor) else list(n_neighbor)
    assert len(Js) == len(ns)
    nbs = get_neighbor(extent, ns, boundary)

    def hij(i, j, sign):
        hx = 2 * sign * (sigma_p(i) @ sigma_m(j) + sigma_m(i) @ sigma_p(j))
        hz = sigma_z(i) @ sigma_z(j)
        return hx + hz

    H = 0
    for k, nb in enumerate(nbs):
        sign = -1 if msr and ns[k] == 1 else 1
        H = H + Js[k] * _sum(hij(int(i), int(j), sign) for i, j in nb)
    return H


def ising(extent, h=0.0, J=1.0, boundary=1):
    """``common_operators.py:56-71``."""
    N = int(np.prod(extent))
    H = -h * _sum(sigma_x(i) for i in range(N))
    nb = get_neighbor(extent, 1, boundary)
    H = H + (-J) * _sum(sigma_z(int(i)) @ sigma_z(int(j)) for i, j in nb)
    return H


# --------------------------------------------------------------------------
# application to basis states (operator.py:30-184)
# --------------------------------------------------------------------------

def _apply_site(x, op, J, idx):
    """Vectorised ``_apply_site_operator`` for spins (operator.py:30-78).

    x: (B, N) int8, J: (B,) float64, idx: (B,) int.  Returns new (x, J).
    """
    rows = np.arange(x.shape[0])
    xi = x[rows, idx]
    if op == "I":
        return x, J
    if op == "n":
        return x, np.where(xi > 0, J, 0.0)
    if op == "z":
        return x, J * xi / 2
    if op in ("x", "y"):
        J = J / 2
    if op == "+":
        J = np.where(xi < 0, J, np.nan)
        x = x.copy()
        x[rows, idx] = 1
    elif op == "-":
        J = np.where(xi > 0, J, np.nan)
        x = x.copy()
        x[rows, idx] = -1
    elif op == "x":
        x = x.copy()
        x[rows, idx] *= -1
    elif op == "y":
        raise NotImplementedError("sigma_y needs complex dtypes")
    return x, J


def apply_diag(s, op_list):
    """``_apply_diag`` (operator.py:81-93): (Ns, N) int8 -> (Ns,) float64."""
    s = np.asarray(s)
    Hz = np.zeros(s.shape[0])
    for opstr, terms in op_list.op_list:
        if all(o in DIAG_OPS for o in opstr):
            for J, *index in terms:
                Jv = np.full(s.shape[0], float(J))
                for o, i in zip(opstr, index):
                    _, Jv = _apply_site(s, o, Jv, np.full(s.shape[0], i))
                Hz = Hz + Jv
    return Hz


def apply_off_diag(s, op_list):
    """``_apply_off_diag`` (operator.py:96-119).

    Returns ``{nflips: (s_conn (Ns, nconn, N) int8, H (Ns, nconn) f64)}`` in the
    first-seen order of nflips, candidates in op_list order with NaN = invalid.
    """
    s = np.asarray(s)
    Ns = s.shape[0]
    out = {}
    for opstr, terms in op_list.op_list:
        nflips = sum(1 for o in opstr if o not in DIAG_OPS)
        if nflips == 0:
            continue
        sc_list, H_list = [], []
        for J, *index in terms:
            x = s.copy()
            Jv = np.full(Ns, float(J))
            for o, i in zip(reversed(opstr), reversed(index)):
                x, Jv = _apply_site(x, o, Jv, np.full(Ns, int(i)))
            sc_list.append(x)
            H_list.append(Jv)
        sc = np.stack(sc_list, axis=1)
        H = np.stack(H_list, axis=1)
        if nflips in out:
            out[nflips][0].append(sc)
            out[nflips][1].append(H)
        else:
            out[nflips] = [[sc], [H]]
    return {k: (np.concatenate(v[0], axis=1), np.concatenate(v[1], axis=1)) for k, v in out.items()}


def get_conn(s_conn, H_conn):
    """``_get_conn`` (operator.py:144-165) without padding: valid = not NaN and
    not isclose(H, 0); order = np.nonzero over (sample, candidate)."""
    valid = ~(np.isnan(H_conn) | np.isclose(H_conn, 0.0))
    seg, ci = np.nonzero(valid)
    return seg, s_conn[seg, ci], H_conn[seg, ci]


def oloc(s, logpsi_fn, op_list, logpsi_s=None):
    """``Operator.Oloc`` (operator.py:510-562) for real positive-sign states.

    ``logpsi_fn(spins) -> (sign, logabs)`` evaluates psi on any batch.
    E_loc(s) = diag(s) + sum_{s'} H_{s s'} psi(s')/psi(s)  (``_get_Olocx``).
    """
    s = np.asarray(s)
    if logpsi_s is None:
        sg, la = logpsi_fn(s)
    else:
        sg, la = logpsi_s
    E = apply_diag(s, op_list)
    for nflips, (sc, H) in apply_off_diag(s, op_list).items():
        seg, sconn, Hc = get_conn(sc, H)
        if seg.size == 0:
            continue
        sg2, la2 = logpsi_fn(sconn)
        ratio = (sg2 / sg[seg]) * np.exp(la2 - la[seg])
        E = E + np.bincount(seg, weights=ratio * Hc, minlength=s.shape[0])
    return E


def expectation(E, reweight=None, return_var=False):
    """``Operator.expectation`` (operator.py:564-595)."""
    r = 1.0 if reweight is None else reweight
    m = np.mean(E * r)
    if return_var:
        return m, np.mean(np.abs(E) ** 2 * r) - np.abs(m) ** 2
    return m


# --------------------------------------------------------------------------
# compiled candidate table: the form the device kernels consume
# --------------------------------------------------------------------------

class CompiledHam:
    """Flat tables equivalent to an op_list of spin operators.

    diag terms  : coefficient c, sites (a, b) (b = -1 for one-site terms) and
                  kind: 0 -> c*s_a*s_b, 1 -> c*s_a, 2 -> c (identity).
    off-diagonal: candidate k flips sites a_k (and b_k if >= 0); it is valid
                  iff s[a] == req_a (0 = any) and s[b] == req_b; amplitude c_k.
    Candidate order equals the concatenation order of ``apply_off_diag``
    (groups by first-seen nflips, op_list order inside a group).
    """

    def __init__(self, N, diag_c, diag_a, diag_b, diag_kind, off_a, off_b, off_ra, off_rb, off_c, nflips):
        self.N = N
        self.diag_c = np.asarray(diag_c, np.float64)
        self.diag_a = np.asarray(diag_a, np.int32)
        self.diag_b = np.asarray(diag_b, np.int32)
        self.diag_kind = np.asarray(diag_kind, np.int32)
        self.off_a = np.asarray(off_a, np.int32)
        self.off_b = np.asarray(off_b, np.int32)
        self.off_ra = np.asarray(off_ra, np.int8)
        self.off_rb = np.asarray(off_rb, np.int8)
        self.off_c = np.asarray(off_c, np.float64)
        self.nflips = np.asarray(nflips, np.int32)


def compile_op_list(op_list, N):
    """Compile an op_list into ``CompiledHam`` (supported: op strings of length
    <= 2 over {I, z, +, -, x}; flips on distinct sites; diagonal factors only
    in purely diagonal strings)."""
    diag_c, diag_a, diag_b, diag_kind = [], [], [], []
    groups = {}
    order = []
    for opstr, terms in op_list.op_list:
        if len(opstr) > 2:
            raise NotImplementedError(f"op string {opstr!r} longer than 2")
        nfl = sum(1 for o in opstr if o not in DIAG_OPS)
        if nfl == 0:
            for J, *idx in terms:
                c = float(J)
                sites = []
                for o, i in zip(opstr, idx):
                    if o == "z":
                        c /= 2
                        sites.append(int(i))
                    elif o == "n":
                        raise NotImplementedError("'n' operator on spins")
                if len(sites) == 2:
                    diag_c.append(c); diag_a.append(sites[0]); diag_b.append(sites[1]); diag_kind.append(0)
                elif len(sites) == 1:
                    diag_c.append(c); diag_a.append(sites[0]); diag_b.append(-1); diag_kind.append(1)
                else:
                    diag_c.append(c); diag_a.append(0); diag_b.append(-1); diag_kind.append(2)
            continue
        if any(o in ("z", "n") for o in opstr):
            raise NotImplementedError(f"mixed diagonal/off-diagonal op string {opstr!r}")
        if nfl not in groups:
            groups[nfl] = []
            order.append(nfl)
        for J, *idx in terms:
            c = float(J)
            fl = []
            for o, i in zip(opstr, idx):
                if o == "I":
                    continue
                if o == "x":
                    c /= 2
                    fl.append((int(i), 0))
                elif o == "+":
                    fl.append((int(i), -1))
                elif o == "-":
                    fl.append((int(i), 1))
                else:
                    raise NotImplementedError(o)
            if len(fl) == 2 and fl[0][0] == fl[1][0]:
                raise NotImplementedError("two flips on one site")
            a, ra = fl[0]
            b, rb = fl[1] if len(fl) == 2 else (-1, 0)
            groups[nfl].append((a, b, ra, rb, c, nfl))
    off = [t for g in order for t in groups[g]]
    # candidates with |c| <= 1e-8 are never valid (isclose(H, 0), operator.py:155)
    off = [t for t in off if not np.isclose(t[4], 0.0)]
    if off:
        a, b, ra, rb, c, nf = map(list, zip(*off))
    else:
        a = b = ra = rb = c = nf = []
    return CompiledHam(N, diag_c, diag_a, diag_b, diag_kind, a, b, ra, rb, c, nf)


def compiled_diag(s, ch):
    s = np.asarray(s, np.float64)
    E = np.zeros(s.shape[0])
    for c, a, b, k in zip(ch.diag_c, ch.diag_a, ch.diag_b, ch.diag_kind):
        if k == 0:
            E += c * s[:, a] * s[:, b]
        elif k == 1:
            E += c * s[:, a]
        else:
            E += c
    return E


def compiled_conn(s, ch):
    """Connected configurations from the compiled table, sample-major.

    Returns (segment, s_conn, H, cand_index)."""
    s = np.asarray(s)
    Ns = s.shape[0]
    nc = ch.off_a.size
    if nc == 0:
        return (np.zeros(0, np.int64), np.zeros((0, s.shape[1]), np.int8), np.zeros(0), np.zeros(0, np.int64))
    sa = s[:, ch.off_a]
    okb = ch.off_b >= 0
    sb = s[:, np.where(okb, ch.off_b, 0)]
    valid = (ch.off_ra[None, :] == 0) | (sa == ch.off_ra[None, :])
    valid &= (~okb[None, :]) | (ch.off_rb[None, :] == 0) | (sb == ch.off_rb[None, :])
    seg, k = np.nonzero(valid)
    sconn = s[seg].copy()
    sconn[np.arange(seg.size), ch.off_a[k]] *= -1
    hb = okb[k]
    sconn[np.arange(seg.size)[hb], ch.off_b[k][hb]] *= -1
    return seg, sconn, ch.off_c[k], k


def compiled_oloc(s, logpsi_fn, ch, logpsi_s=None):
    s = np.asarray(s)
    sg, la = logpsi_fn(s) if logpsi_s is None else logpsi_s
    E = compiled_diag(s, ch)
    seg, sconn, Hc, _ = compiled_conn(s, ch)
    if seg.size:
        sg2, la2 = logpsi_fn(sconn)
        ratio = (sg2 / sg[seg]) * np.exp(la2 - la[seg])
        E = E + np.bincount(seg, weights=ratio * Hc, minlength=s.shape[0])
    return E
