"""The oracle's momentum SR variants (SPRING, MARCH, AdamSR), pinned by what
their algebra fixes independently of the update formulas:

* with a full-rank Obar (Ns < Np) and the exact pseudo-inverse, every step of
  every variant solves Obar step = Ebar;
* the part of the step the solve adds lies in the row space of Obar scaled by
  the variant's metric: step - mu last (SPRING), V^2 (step - mu last) (MARCH),
  V^2 (step - m) (AdamSR) are in the row space of Obar;
* SPRING with mu = 0 is SR; MARCH's first step (V = 0 -> 1, last = 0) is SR;
  AdamSR's first step is SR to the 1e-8 of its V offset.
"""
import numpy as np
import pytest

from oracle import sr as osr


def _pinv(A, b):
    return osr.auto_pinv_eig(A, b, rtol=1e-12)


def _seq(rng, steps=4, ns=40, np_=150):
    out = []
    for _ in range(steps):
        A = rng.standard_normal((ns, np_)) / np.sqrt(ns)
        b = rng.standard_normal(ns) / np.sqrt(ns)
        out.append((A, b))
    return out


def _in_rowspace(A, x):
    Q, _ = np.linalg.qr(A.T)
    return np.linalg.norm(x - Q @ (Q.T @ x)) / np.linalg.norm(x)


@pytest.mark.parametrize("kind", ["spring", "march", "adam"])
def test_momentum_steps_solve_the_sr_equation(kind, rng):
    """Every step solves Obar step = Ebar; SPRING's and AdamSR's added parts
    lie in the (scaled) row space (MARCH: the test below)."""
    seq = _seq(rng)
    np_ = seq[0][0].shape[1]
    opt = {"spring": osr.Spring(np_, _pinv, mu=0.9), "march": osr.March(np_, _pinv),
           "adam": osr.AdamSR(np_, _pinv)}[kind]
    for A, b in seq:
        prev = (opt.last.copy() if kind != "adam" else None)
        step = opt.solve(A, b)
        assert np.linalg.norm(A @ step - b) / np.linalg.norm(b) < 1e-10
        if kind == "spring":
            assert _in_rowspace(A, step - opt.mu * prev) < 1e-10
        elif kind == "adam":
            m = opt.m / (1 - opt.mu ** opt.t)
            V = (opt.v / (1 - opt.beta ** opt.t)) ** 0.25 + 1e-8
            x = V ** 2 * (step - m)
            # on the first step Obar m = Ebar already and the added part is round-off
            if np.linalg.norm(x) > 1e-8 * np.linalg.norm(V ** 2 * step):
                assert _in_rowspace(A, x) < 1e-9


def test_march_metric_row_space(rng):
    seq = _seq(rng)
    np_ = seq[0][0].shape[1]
    opt = osr.March(np_, _pinv)
    for A, b in seq:
        last, Vs = opt.last.copy(), opt.V.copy()
        V = np.ones(np_) if np.allclose(Vs, 0) else (Vs / (1 - opt.beta ** (opt.t + 1))) ** 0.25 + 1e-8
        step = opt.solve(A, b)
        assert _in_rowspace(A, V ** 2 * (step - opt.mu * last)) < 1e-9


def test_first_steps_reduce_to_sr(rng):
    (A, b), = _seq(rng, steps=1)
    np_ = A.shape[1]
    sr = _pinv(A, b)
    assert np.allclose(osr.Spring(np_, _pinv, mu=0.0).solve(A, b), sr, rtol=0, atol=1e-12 * np.abs(sr).max())
    assert np.linalg.norm(osr.March(np_, _pinv).solve(A, b) - sr) / np.linalg.norm(sr) < 1e-12
    assert np.linalg.norm(osr.AdamSR(np_, _pinv).solve(A, b) - sr) / np.linalg.norm(sr) < 1e-6
    s = osr.Spring(np_, _pinv, mu=0.0)
    for A2, b2 in _seq(rng, steps=3):
        assert np.linalg.norm(s.solve(A2, b2) - _pinv(A2, b2)) < 1e-12 * np.linalg.norm(_pinv(A2, b2))
