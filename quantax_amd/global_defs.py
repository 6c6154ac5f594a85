"""Process-wide defaults mirroring ``quantax/global_defs.py``.

The reference keeps the lattice (``get_sites``, ``global_defs.py:115-126``)
and a JAX PRNG key (``:51-91``) as process globals.  Here the lattice global
is kept for API compatibility, while random streams are counter-based
(Philox4x32-10): ``get_subkeys`` hands out 64-bit seeds derived from the
global seed with splitmix64, and every kernel takes its seed explicitly.
"""
from __future__ import annotations

import numpy as np

_SEED = 42
_COUNTER = 0
_SITES = None
DTYPE = np.float64  # the reference's default computation dtype (global_defs.py:13)


def _splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


def set_random_seed(seed: int) -> None:
    """Reset the global seed and the subkey counter (``global_defs.py:54-63``)."""
    global _SEED, _COUNTER
    _SEED = int(seed) & 0xFFFFFFFFFFFFFFFF
    _COUNTER = 0


def get_subkeys(num=None):
    """Fresh 64-bit seeds (one, or a list of ``num``), replacing ``get_subkeys``
    (``global_defs.py:75-91``)."""
    global _COUNTER
    n = 1 if num is None else int(num)
    out = []
    for _ in range(n):
        _COUNTER += 1
        out.append(_splitmix64(_SEED ^ _splitmix64(_COUNTER)))
    return out[0] if num is None else out


def set_default_dtype(dtype) -> None:
    global DTYPE
    DTYPE = np.dtype(dtype).type


def get_default_dtype():
    return DTYPE


def set_sites(sites) -> None:
    global _SITES
    _SITES = sites


def get_sites():
    if _SITES is None:
        raise RuntimeError("No lattice defined: create one with quantax_amd.sites.Square/Chain/Grid first.")
    return _SITES


def get_lattice():
    return get_sites()
