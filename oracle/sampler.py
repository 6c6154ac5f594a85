"""Metropolis sampling on supplied streams (oracle).

Restates ``quantax/sampler/metropolis.py:171-322`` (sweep loop, acceptance
|psi'/psi|^reweight > 1 - u or |psi| == 0, only when s' != s) and the
proposal kernels ``sampler/common_samplers.py:14-33`` (LocalFlip: one random
site flipped) and ``:36-162`` (SpinExchange: a random site holding the
minority "hopping" species, ``:138-142``, swapped with a random entry of its
-1-padded neighbour row; -1 means no move).

The random choices come from Philox words (``oracle.philox``) instead of JAX
keys: LocalFlip site = (w0 * N) >> 32; SpinExchange: the k-th hopping site in
increasing index order with k = (w0 * n_hop) >> 32, neighbour slot
(w1 * width) >> 32; u from (w2, w3).  The log-domain test
reweight * (logpsi' - logpsi) > log1p(-u) is the monotone image of the
reference's test; decisions within ``tie_tol`` of the threshold are counted.
"""
from __future__ import annotations

import numpy as np

from .philox import bounded, sweep_words, words_to_uniform

LOCALFLIP = 0
EXCHANGE = 1


def propose(kind, s, w, nbr=None, hopping=1):
    """Apply one proposal per chain.  s (Ns, N) int8, w (Ns, 4) uint32."""
    Ns, N = s.shape
    rows = np.arange(Ns)
    s_new = s.copy()
    if kind == LOCALFLIP:
        pos = bounded(w[:, 0], N)
        s_new[rows, pos] *= -1
        return s_new
    is_hop = s == hopping
    nhop = is_hop.sum(axis=1)
    k = bounded(w[:, 0], np.maximum(nhop, 1)) if np.all(nhop == nhop[0]) else np.array(
        [bounded(w[c, 0], max(int(nhop[c]), 1)) for c in range(Ns)])
    csum = np.cumsum(is_hop, axis=1)
    # k-th (0-based) hopping site: first index where the running count reaches k+1
    site = np.argmax(csum > k[:, None], axis=1)
    width = nbr.shape[1]
    slot = bounded(w[:, 1], width)
    nb = nbr[site, slot]
    nb = np.where(nb < 0, site, nb)
    valid = nhop > 0
    a = s[rows, site]
    b = s[rows, nb]
    s_new[rows[valid], site[valid]] = b[valid]
    s_new[rows[valid], nb[valid]] = a[valid]
    return s_new


def metropolis(logpsi_fn, s0, words, kind, nbr=None, hopping=1, reweight=2.0, tie_tol=1e-9, logpsi0=None):
    """Run ``words.shape[0]`` Metropolis steps on all chains.

    Returns (s, logabs, n_accept per chain, near-threshold count per chain).
    """
    s = np.array(s0, dtype=np.int8, copy=True)
    Ns = s.shape[0]
    la = logpsi_fn(s)[1] if logpsi0 is None else np.array(logpsi0, np.float64)
    nacc = np.zeros(Ns, np.int64)
    near = np.zeros(Ns, np.int64)
    for t in range(words.shape[0]):
        w = words[t]
        s_new = propose(kind, s, w, nbr, hopping)
        updated = np.any(s_new != s, axis=1)
        la_new = np.full(Ns, -np.inf)
        if updated.any():
            la_new[updated] = logpsi_fn(s_new[updated])[1]
        u = words_to_uniform(w[:, 2], w[:, 3])
        thr = np.log1p(-u)
        with np.errstate(invalid="ignore"):
            d = reweight * (la_new - la)
        acc = (d > thr) | (la == -np.inf)
        acc &= updated
        near += (updated & (np.abs(d - thr) < tie_tol)).astype(np.int64)
        s[acc] = s_new[acc]
        la[acc] = la_new[acc]
        nacc += acc
    return s, la, nacc, near


def sweep(logpsi_fn, s0, seed, nsteps, kind, nbr=None, hopping=1, reweight=2.0, step0=0, tie_tol=1e-9):
    words = sweep_words(seed, nsteps, s0.shape[0], step0=step0)
    return metropolis(logpsi_fn, s0, words, kind, nbr, hopping, reweight, tie_tol)


def reweight_factor(logabs, reweight=2.0):
    """``Sampler._get_reweight_factor`` (sampler/sampler.py:66-69)."""
    r = np.exp((2.0 - reweight) * (logabs - np.max(logabs)))
    return r / r.mean()
