"""Exact diagonalisation from an op_list (oracle; stands in for QuSpin).

``Operator.diagonalize`` (quantax/operator/operator.py:331-359) calls QuSpin's
``hamiltonian.eigsh``; QuSpin is absent here, so the sparse matrix is built by
applying ``apply_diag`` / ``apply_off_diag`` of this oracle to every basis
state of the (optionally Sz-restricted) space.
"""
from __future__ import annotations

import itertools

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .operator import apply_diag, apply_off_diag


def basis_states(N, nup=None):
    """All +-1 configurations (int8), optionally with exactly ``nup`` up spins."""
    if nup is None:
        ints = np.arange(2 ** N, dtype=np.int64)
    else:
        ints = np.array(
            [sum(1 << (N - 1 - k) for k in c) for c in itertools.combinations(range(N), nup)],
            dtype=np.int64,
        )
        ints.sort()
    bits = (ints[:, None] >> np.arange(N - 1, -1, -1)[None, :]) & 1
    return (2 * bits - 1).astype(np.int8), ints


def _to_int(s):
    N = s.shape[-1]
    w = (1 << np.arange(N - 1, -1, -1)).astype(np.int64)
    return ((s > 0).astype(np.int64) * w).sum(-1)


def sparse_hamiltonian(op_list, N, nup=None):
    s, ints = basis_states(N, nup)
    pos = {int(v): k for k, v in enumerate(ints)}
    dim = s.shape[0]
    rows, cols, vals = [np.arange(dim)], [np.arange(dim)], [apply_diag(s, op_list)]
    for _, (sc, H) in apply_off_diag(s, op_list).items():
        valid = ~np.isnan(H)
        r, c = np.nonzero(valid)
        tgt = _to_int(sc[r, c])
        if nup is None:
            col = tgt
        else:
            col = np.array([pos.get(int(t), -1) for t in tgt])
            keep = col >= 0
            r, c, col = r[keep], c[keep], col[keep]
        rows.append(r)
        cols.append(col)
        vals.append(H[r, c])
    # <s|H|s'> with s' reached from s: matrix element row s, column s'
    M = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(dim, dim))
    return M.tocsr(), s


def ground_energy(op_list, N, nup=None, k=1):
    H, _ = sparse_hamiltonian(op_list, N, nup)
    if H.shape[0] <= 64:
        w = np.linalg.eigvalsh(H.toarray())
        return w[:k]
    w = spla.eigsh(H, k=k, which="SA", return_eigenvectors=False, tol=1e-12)
    return np.sort(w)
