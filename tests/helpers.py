"""Shared test helpers: build the same problem for the device and the oracle."""
import numpy as np

from oracle import lattice as olat
from oracle import models as om
from oracle import operator as oop


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def max_rel(a, b, floor=1.0):
    """max |a - b| / max(|b|, floor): relative error with an absolute floor."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def heis_oracle(extent, J=1.0, n_neighbor=1, msr=False):
    return oop.heisenberg(extent, J, n_neighbor, msr)


def rand_spins(rng, ns, N, nup=None):
    if nup is None:
        return (2 * rng.integers(0, 2, size=(ns, N)) - 1).astype(np.int8)
    s = -np.ones((ns, N), np.int8)
    s[:, :nup] = 1
    for r in s:
        rng.shuffle(r)
    return s


def rbm_oracle_fn(W, b):
    return lambda s: om.rbm_logpsi(W, b, s)


def resconv_oracle_fn(spec, theta, approximate=True):
    return lambda s: om.resconv_forward(spec, theta, s, approximate)
