"""Philox4x32-10 counter-based generator (oracle).

The reference draws proposals and acceptance uniforms from JAX threefry keys
split on the host (quantax/global_defs.py:66-91, sampler/common_samplers.py:31,
65-73, sampler/metropolis.py:303).  That stream cannot be reproduced without
JAX, so the device path consumes *supplied* streams instead; this module is
the published Philox4x32-10 algorithm (Salmon, Moraes, Dror, Shaw, "Parallel
random numbers: as easy as 1, 2, 3", SC'11) used to fill them, restated in
numpy so the device generator can be checked bit for bit.

Stream layout used by the Metropolis kernels: one 4x32-bit block per
(step, chain), counter = (chain, step, tag, 0), key = (seed lo, seed hi).
Words 0/1 drive the proposal, words 2/3 form the uniform
u = ((w2 << 32 | w3) >> 11) * 2^-53 in [0, 1).
"""
from __future__ import annotations

import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)

TAG_SWEEP = 0x51A7
TAG_INIT = 0x1D17


def philox4x32_10(ctr, key):
    """ctr: (..., 4) uint32, key: (..., 2) uint32 -> (..., 4) uint32."""
    c = [np.asarray(ctr[..., i], dtype=np.uint64) for i in range(4)]
    k0 = np.asarray(key[..., 0], dtype=np.uint64)
    k1 = np.asarray(key[..., 1], dtype=np.uint64)
    for r in range(10):
        if r > 0:
            k0 = (k0 + np.uint64(W0)) & MASK
            k1 = (k1 + np.uint64(W1)) & MASK
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c = [
            hi1 ^ c[1] ^ k0,
            lo1,
            hi0 ^ c[3] ^ k1,
            lo0,
        ]
    return np.stack([x.astype(np.uint32) for x in c], axis=-1)


def seed_key(seed):
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return np.array([seed & 0xFFFFFFFF, seed >> 32], dtype=np.uint32)


def sweep_words(seed, nsteps, nchains, step0=0, chain0=0):
    """Stream words, shape (nsteps, nchains, 4) uint32."""
    st = np.arange(step0, step0 + nsteps, dtype=np.uint64)
    ch = np.arange(chain0, chain0 + nchains, dtype=np.uint64)
    ctr = np.zeros((nsteps, nchains, 4), dtype=np.uint32)
    ctr[..., 0] = ch[None, :].astype(np.uint32)
    ctr[..., 1] = st[:, None].astype(np.uint32)
    ctr[..., 2] = TAG_SWEEP
    key = np.broadcast_to(seed_key(seed), (nsteps, nchains, 2))
    return philox4x32_10(ctr, key)


def words_to_uniform(w2, w3):
    """u = ((w2 << 32 | w3) >> 11) * 2^-53, exactly representable in float64."""
    v = (np.asarray(w2, np.uint64) << np.uint64(32)) | np.asarray(w3, np.uint64)
    return (v >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def bounded(w, n):
    """Multiply-shift map of a 32-bit word to [0, n): (w * n) >> 32."""
    return ((np.asarray(w, np.uint64) * np.uint64(n)) >> np.uint64(32)).astype(np.int64)


def init_spins(seed, nchains, N, nup=None):
    """Initial chains (``utils/basis.py:121-133, 160-219``, restated).

    nup None: each site +-1 from one bit of Philox output (block per
    (chain, site//128), bit = site % 128).  nup given: the first ``nup`` sites
    set +1 and a Fisher-Yates shuffle from the last index down, draw for index
    i from block (chain, i) word 0 mapped to [0, i] by multiply-shift.
    """
    key = seed_key(seed)
    if nup is None:
        nblk = (N + 127) // 128
        ctr = np.zeros((nchains, nblk, 4), np.uint32)
        ctr[..., 0] = np.arange(nchains, dtype=np.uint32)[:, None]
        ctr[..., 1] = np.arange(nblk, dtype=np.uint32)[None, :]
        ctr[..., 2] = TAG_INIT
        ctr[..., 3] = 1
        w = philox4x32_10(ctr, np.broadcast_to(key, (nchains, nblk, 2)))
        site = np.arange(N)
        blk, bit = site // 128, site % 128
        word = w[:, blk, bit // 32]
        b = (word >> (bit % 32).astype(np.uint32)) & 1
        return (2 * b.astype(np.int8) - 1).astype(np.int8)
    s = -np.ones((nchains, N), np.int8)
    s[:, :nup] = 1
    ctr = np.zeros((nchains, N, 4), np.uint32)
    ctr[..., 0] = np.arange(nchains, dtype=np.uint32)[:, None]
    ctr[..., 1] = np.arange(N, dtype=np.uint32)[None, :]
    ctr[..., 2] = TAG_INIT
    w = philox4x32_10(ctr, np.broadcast_to(key, (nchains, N, 2)))
    rows = np.arange(nchains)
    for i in range(N - 1, 0, -1):
        j = bounded(w[:, i, 0], i + 1)
        a = s[rows, i].copy()
        s[rows, i] = s[rows, j]
        s[rows, j] = a
    return s
