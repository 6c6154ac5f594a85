"""Spin operators and local estimators.

``Operator`` keeps the reference's QuSpin-format ``op_list``
(``quantax/operator/operator.py:190-212``) with the same algebra — ``+``
merges equal op strings in first-seen order (``:402-418``), ``*`` scales
strengths (``:442-455``), ``@`` concatenates op strings (``:291-301``) — and
the Pauli-convention site operators (``operator/site_operator.py:39-132``:
sigma_x/z carry strength 2, sigma_+/- strength 1).

For the device it is compiled once into flat tables (``qtx_ham``): diagonal
terms and off-diagonal candidates in the order ``_apply_off_diag`` produces
them (``operator.py:96-119``: groups by first-seen flip count, op_list order
inside a group; |H| <= 1e-8 candidates dropped as ``_get_conn`` does,
``:155``).  ``Oloc`` then runs the model's fused local-energy kernel.
"""
from __future__ import annotations

from numbers import Number
from typing import Optional

import numpy as np
import torch

from . import _lib
from .global_defs import get_sites

_DIAG = ("I", "n", "z")


class Operator:
    def __init__(self, op_list: list):
        self._op_list = [[o, [list(t) for t in terms]] for o, terms in op_list]
        self._dev = None
        self._connectivity = None

    @property
    def op_list(self) -> list:
        return self._op_list

    # ------------------------------------------------------------ algebra
    def __add__(self, other):
        if isinstance(other, Number):
            if not np.isclose(other, 0.0):
                raise ValueError("Constant shift is not implemented for Operator.")
            return self
        if not isinstance(other, Operator):
            return NotImplemented
        out = [[o, [list(t) for t in terms]] for o, terms in self._op_list]
        names = [o for o, _ in out]
        for o2, t2 in other._op_list:
            if o2 in names:
                out[names.index(o2)][1] += [list(t) for t in t2]
            else:
                out.append([o2, [list(t) for t in t2]])
                names.append(o2)
        return Operator(out)

    __radd__ = __add__

    def __sub__(self, other):
        return self + (-other)

    def __mul__(self, c):
        if isinstance(c, torch.Tensor) or isinstance(c, np.ndarray):
            c = c.item()
        if not isinstance(c, Number):
            return NotImplemented
        return Operator([[o, [[J * c, *idx] for J, *idx in terms]] for o, terms in self._op_list])

    __rmul__ = __mul__

    def __truediv__(self, c):
        return self * (1.0 / c)

    def __neg__(self):
        return self * (-1)

    def __matmul__(self, other):
        if not isinstance(other, Operator):
            return NotImplemented
        out = []
        for o1, t1 in self._op_list:
            for o2, t2 in other._op_list:
                out.append([o1 + o2, [[J1 * J2, *i1, *i2] for J1, *i1 in t1 for J2, *i2 in t2]])
        return Operator(out)

    def __repr__(self):
        return f"Operator({self._op_list!r})"

    # ------------------------------------------------------------ compilation
    def tables(self):
        """Host tables (numpy) of the compiled operator."""
        N = get_sites().Nsites
        diag, groups, order = [], {}, []
        for opstr, terms in self._op_list:
            nfl = sum(1 for o in opstr if o not in _DIAG)
            if len(opstr) > 2:
                raise NotImplementedError(f"op string {opstr!r}: at most two site operators are supported")
            if nfl == 0:
                for J, *idx in terms:
                    c, sites = float(J), []
                    for o, i in zip(opstr, idx):
                        if o == "z":
                            c /= 2
                            sites.append(int(i))
                        elif o == "n":
                            raise NotImplementedError("'n' on spins is not supported")
                    kind = {2: 0, 1: 1, 0: 2}[len(sites)]
                    a = sites[0] if sites else 0
                    b = sites[1] if len(sites) == 2 else -1
                    diag.append((c, a, b, kind))
                continue
            if any(o in ("z", "n") for o in opstr):
                raise NotImplementedError(f"op string {opstr!r} mixes diagonal and flip operators")
            if nfl not in groups:
                groups[nfl] = []
                order.append(nfl)
            for J, *idx in terms:
                c, flips = float(J), []
                for o, i in zip(opstr, idx):
                    if o == "I":
                        continue
                    if o == "x":
                        c /= 2
                        flips.append((int(i), 0))
                    elif o == "+":
                        flips.append((int(i), -1))  # sigma+ acts on a down spin
                    elif o == "-":
                        flips.append((int(i), 1))
                    else:
                        raise NotImplementedError(f"operator {o!r}")
                if len(flips) == 2 and flips[0][0] == flips[1][0]:
                    raise NotImplementedError("two flip operators on one site")
                a, ra = flips[0]
                b, rb = flips[1] if len(flips) == 2 else (-1, 0)
                if not np.isclose(c, 0.0):
                    groups[nfl].append((a, b, ra, rb, c))
        cands = [t for g in order for t in groups[g]]
        for c, a, b, _ in diag:
            if not (0 <= a < N and b < N):
                raise ValueError("site index out of range")
        for a, b, *_ in cands:
            if not (0 <= a < N and b < N):
                raise ValueError("site index out of range")
        d = np.array(diag, dtype=object).reshape(-1, 4)
        o = np.array(cands, dtype=object).reshape(-1, 5)
        return dict(
            N=N,
            diag_c=d[:, 0].astype(np.float64), diag_a=d[:, 1].astype(np.int32),
            diag_b=d[:, 2].astype(np.int32), diag_kind=d[:, 3].astype(np.int32),
            off_a=o[:, 0].astype(np.int32), off_b=o[:, 1].astype(np.int32),
            off_ra=o[:, 2].astype(np.int8), off_rb=o[:, 3].astype(np.int8), off_c=o[:, 4].astype(np.float64),
        )

    def device_tables(self):
        """Device copies of the tables plus the ctypes ``qtx_ham`` struct (cached)."""
        if self._dev is None:
            t = self.tables()
            dev = _lib.device()
            keep = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in t.items() if k != "N"}
            h = _lib.QtxHam()
            h.N = t["N"]
            h.ndiag = int(t["diag_c"].size)
            h.noff = int(t["off_c"].size)
            for k, v in keep.items():
                setattr(h, k, v.data_ptr() if v.numel() else None)
            self._dev = (keep, h)
        return self._dev[1]

    # ------------------------------------------------------------ estimators
    def apply_diag(self, s: torch.Tensor) -> torch.Tensor:
        s = s.contiguous()
        out = torch.empty(s.shape[0], dtype=torch.float64, device=s.device)
        _lib.call("qtx_ham_diag", _lib.stream(), self.device_tables(), s.data_ptr(), s.shape[0], out.data_ptr())
        return out

    def connected(self, s: torch.Tensor, with_offsets: bool = False):
        """Compacted connected configurations (segment, candidate, s', H), the
        counterpart of ``_apply_off_diag`` + ``_get_conn`` (operator.py:96-165);
        with_offsets: also the (ns + 1) per-sample offsets into them."""
        s = s.contiguous()
        ns, N = s.shape
        H = self.device_tables()
        cnt = torch.empty(ns, dtype=torch.int32, device=s.device)
        _lib.call("qtx_conn_count", _lib.stream(), H, s.data_ptr(), ns, cnt.data_ptr())
        off = torch.zeros(ns + 1, dtype=torch.int64, device=s.device)
        off[1:] = torch.cumsum(cnt.to(torch.int64), 0)
        K = int(off[-1].item())
        seg = torch.empty(K, dtype=torch.int32, device=s.device)
        cand = torch.empty(K, dtype=torch.int32, device=s.device)
        sc = torch.empty((K, N), dtype=torch.int8, device=s.device)
        hc = torch.empty(K, dtype=torch.float64, device=s.device)
        if K:
            _lib.call("qtx_conn_list", _lib.stream(), H, s.data_ptr(), ns, off.data_ptr(), seg.data_ptr(),
                      cand.data_ptr(), sc.data_ptr(), hc.data_ptr())
        if with_offsets:
            return seg, cand, sc, hc, off
        return seg, cand, sc, hc

    def Oloc(self, state, samples) -> torch.Tensor:
        """E_loc(s) = sum_s' <s|O|s'> psi(s')/psi(s) (``operator.py:510-562``)."""
        from .sampler import Samples

        if isinstance(samples, Samples):
            s, psi = samples.spins, samples.psi
        else:
            s = torch.as_tensor(samples, device=_lib.device()).to(torch.int8)
            psi = state(s)
        if getattr(state, "is_symmetric", False):
            eloc, nconn = self._oloc_generic(state, s, psi)
        else:
            eloc, nconn = state.model.eloc(self, s, psi.logabs)
        self._connectivity = nconn
        return eloc

    def _oloc_generic(self, state, s, psi):
        """Materialised connected configurations, psi of each (the state's
        own forward, e.g. symmetrised), then the per-sample sum
        (``operator.py:168-184``).  Connected configurations are evaluated in
        chunks whose images stay near 1 GB."""
        s = s.contiguous()
        ns = s.shape[0]
        chunk = max(1024, (1 << 30) // (state.symm.nsymm * s.shape[1]))
        diag = self.apply_diag(s)
        _, _, sc, hc, off = self.connected(s, with_offsets=True)
        K = sc.shape[0]
        sg_c = torch.empty(K, dtype=torch.float64, device=s.device)
        la_c = torch.empty(K, dtype=torch.float64, device=s.device)
        for k0 in range(0, K, chunk):
            p = state(sc[k0: k0 + chunk])
            sg_c[k0: k0 + chunk] = p.sign
            la_c[k0: k0 + chunk] = p.logabs
        sg = psi.sign.to(torch.float64).contiguous()
        la = psi.logabs.contiguous()
        eloc = torch.empty(ns, dtype=torch.float64, device=s.device)
        _lib.call("qtx_conn_combine", _lib.stream(), diag.data_ptr(), off.data_ptr(), _lib.ptr(hc if K else None),
                  sg_c.data_ptr() if K else None, la_c.data_ptr() if K else None, sg.data_ptr(), la.data_ptr(), ns,
                  eloc.data_ptr())
        return eloc, (off[1:] - off[:-1]).to(torch.int32)

    def expectation(self, state, samples, return_var: bool = False):
        """``operator.py:564-595``."""
        from .sampler import Samples

        r = samples.reweight_factor if isinstance(samples, Samples) else None
        E = self.Oloc(state, samples)
        r = torch.ones_like(E) if r is None else r
        m = torch.mean(E * r)
        if return_var:
            v = torch.mean(torch.abs(E) ** 2 * r) - torch.abs(m) ** 2
            return m.item(), v.item()
        return m.item()


def _site(opstr, strength, i):
    return Operator([[opstr, [[strength, int(i)]]]])


def sigma_x(i):
    return _site("x", 2.0, i)


def sigma_z(i):
    return _site("z", 2.0, i)


def sigma_p(i):
    return _site("+", 1.0, i)


def sigma_m(i):
    return _site("-", 1.0, i)


def _total(ops):
    acc = 0
    for o in ops:
        acc = acc + o
    return acc


def Heisenberg(J=1.0, n_neighbor=1, msr: bool = False) -> Operator:
    """H = sum_n J_n sum_<ij>_n sigma_i . sigma_j (``common_operators.py:23-53``)."""
    sites = get_sites()
    Js = [J] if isinstance(J, Number) else list(J)
    ns = [n_neighbor] if isinstance(n_neighbor, Number) else list(n_neighbor)
    if len(Js) != len(ns):
        raise ValueError("'J' and 'n_neighbor' should have the same length.")
    neighbors = sites.get_neighbor(ns)

    def hij(i, j, sign):
        return 2 * sign * (sigma_p(i) @ sigma_m(j) + sigma_m(i) @ sigma_p(j)) + sigma_z(i) @ sigma_z(j)

    H = 0
    for k, nb in enumerate(neighbors):
        sign = -1 if msr and ns[k] == 1 else 1
        H = H + Js[k] * _total(hij(int(i), int(j), sign) for i, j in nb)
    return H


def Ising(h=0.0, J=1.0) -> Operator:
    """H = -J sum_<ij> s^z_i s^z_j - h sum_i s^x_i (``common_operators.py:56-71``)."""
    sites = get_sites()
    H = -h * _total(sigma_x(i) for i in range(sites.Nsites))
    H = H + (-J) * _total(sigma_z(int(i)) @ sigma_z(int(j)) for i, j in sites.get_neighbor())
    return H
