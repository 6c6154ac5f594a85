"""SR / MinSR step in float64 (oracle).

Restates ``quantax/optimizer/sr.py:74-195`` (Obar, Ebar, energy, VarE, solve)
and ``quantax/optimizer/solver.py:12-201, 262-294`` (shifted Cholesky and soft
pseudo-inverse solvers with the SNR truncation, automatic SR/MinSR choice,
the pseudo-inverse of a given T), plus the parameter update
``quantax/state/variational.py:558-579`` (skip on non-finite, theta -= step).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla


def get_Ebar(eloc, r=None):
    """``SR.get_Ebar`` (sr.py:180-195): returns (Ebar, energy, VarE)."""
    eloc = np.asarray(eloc, np.float64)
    Ns = eloc.size
    r = np.ones(Ns) if r is None else np.asarray(r, np.float64)
    Emean = np.mean(eloc * r)
    VarE = np.mean(np.abs(eloc - Emean) ** 2 * r)
    Ebar = (eloc - np.mean(eloc)) * np.sqrt(r / Ns)
    return Ebar, Emean, VarE


def get_Obar(O, r=None):
    """``QNGD.get_Obar`` / ``_Omat_to_Obar`` (sr.py:74-88): returns (Obar, Omean)."""
    O = np.asarray(O, np.float64)
    Ns = O.shape[0]
    r = np.ones(Ns) if r is None else np.asarray(r, np.float64)
    Omean = np.mean(O * r[:, None], axis=0)
    Obar = (O - np.mean(O, axis=0, keepdims=True)) * np.sqrt(r / Ns)[:, None]
    return Obar, Omean


def _rtol(rtol):
    return 1e-12 if rtol is None else rtol  # _get_rtol for float64 (solver.py:12-21)


def minnorm_shift_eig(A, b, rshift=None, ashift=1e-4):
    """solver.py:50-62: T = A A^T + (rshift tr T + ashift) I; x = A^T T^-1 b."""
    T = A @ A.T
    shift = _rtol(rshift) * np.trace(T) + ashift
    T = T + shift * np.eye(T.shape[0])
    y = sla.cho_solve(sla.cho_factor(T, lower=True), b)
    return A.T @ y


def lstsq_shift_eig(A, b, rshift=None, ashift=1e-4):
    """solver.py:65-77: S = A^T A + (rshift tr S + ashift) I; x = S^-1 A^T b."""
    S = A.T @ A
    F = A.T @ b
    shift = _rtol(rshift) * np.trace(S) + ashift
    S = S + shift * np.eye(S.shape[0])
    return sla.cho_solve(sla.cho_factor(S, lower=True), F)


def auto_shift_eig(A, b, rshift=None, ashift=1e-4):
    if A.shape[0] < A.shape[1]:
        return minnorm_shift_eig(A, b, rshift, ashift)
    return lstsq_shift_eig(A, b, rshift, ashift)


def eigs_inv(vals, rtol=None, atol=0.0):
    """``_get_eigs_inv`` (solver.py:94-101)."""
    va = np.abs(vals)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = 1 + ((_rtol(rtol) * np.max(va) + atol) / va) ** 6
        inv = 1.0 / (vals * f)
    return np.where(va > 0.0, inv, 0.0)


def sum_without_noise(c, tol_snr):
    """``_sum_without_noise`` (solver.py:114-125): the column sums of c (m, n)
    divided by 1 + (tol_snr / snr)^6, snr = |mean| / sqrt(mean |c - mean|^2 / m),
    applied only when tol_snr > 1e-6.  A zero sum stays zero (the reference's
    expression is 0 / (1 + (tol / (0/0))^6) = NaN for an all-zero column)."""
    c = np.asarray(c, np.float64)
    m = c.shape[0]
    x = np.sum(c, axis=0)
    if not tol_snr > 1e-6:
        return x
    mean = x / m
    sd = np.sqrt(np.mean(np.abs(c - mean[None, :]) ** 2, axis=0) / m)
    with np.errstate(divide="ignore", invalid="ignore"):
        snr = np.abs(mean) / sd
        out = x / (1 + (tol_snr / snr) ** 6)
    return np.where(x == 0.0, 0.0, out)


def minnorm_pinv_eig(A, b, rtol=None, atol=0.0, tol_snr=0.0):
    """solver.py:128-149: x = A^T U diag(eig_inv) rho, rho = SNR-truncated
    sums of c_ts = U_ts b_t."""
    T = A @ A.T
    vals, U = np.linalg.eigh(T)
    inv = eigs_inv(vals, rtol, atol)
    rho = sum_without_noise(U * b[:, None], tol_snr)
    return A.T @ (U @ (inv * rho))


def lstsq_pinv_eig(A, b, rtol=None, atol=0.0, tol_snr=0.0):
    """solver.py:152-164: x = V diag(eig_inv) rho, rho = SNR-truncated sums of
    c_sk = (A V)_sk b_s."""
    S = A.T @ A
    vals, V = np.linalg.eigh(S)
    inv = eigs_inv(vals, rtol, atol)
    rho = sum_without_noise((A @ V) * b[:, None], tol_snr)
    return V @ (inv * rho)


def auto_pinv_eig(A, b, rtol=None, atol=0.0, tol_snr=0.0):
    """solver.py:167-201."""
    if A.shape[0] < A.shape[1]:
        return minnorm_pinv_eig(A, b, rtol, atol, tol_snr)
    return lstsq_pinv_eig(A, b, rtol, atol, tol_snr)


def minsr_pinv_eig(T, b, rtol=None, atol=0.0, tol_snr=0.0):
    """solver.py:262-294: x = U diag(eig_inv) rho for a given symmetric T."""
    vals, U = np.linalg.eigh(np.asarray(T, np.float64))
    inv = eigs_inv(vals, rtol, atol)
    rho = sum_without_noise(U * np.asarray(b, np.float64)[:, None], tol_snr)
    return U @ (inv * rho)


def sr_step(O, eloc, solver, r=None):
    """``QNGD.get_step`` (sr.py:115-123) for real parameters, imag_time=True."""
    Ebar, E, VarE = get_Ebar(eloc, r)
    Obar, _ = get_Obar(O, r)
    return solver(Obar, Ebar), E, VarE


def update(theta, step, dtype=None):
    """``Variational.update`` (variational.py:558-579): None when skipped."""
    step = np.asarray(step)
    if not np.all(np.isfinite(step)):
        return None
    dtype = np.asarray(theta).dtype if dtype is None else dtype
    return (np.asarray(theta) + (-step).astype(dtype)).astype(dtype)


# ---------------------------------------------------------------- momentum SR variants
# Each takes the SR solve as a callable solver(Obar, Ebar) -> step (sr.py:90-113
# with the real, imaginary-time branch: step = solver(Obar, Ebar)).

class Spring:
    """``SPRING.solve`` (sr.py:247-255): Ebar -= mu Obar last; step =
    solve(Obar, Ebar) + mu last; last = step."""

    def __init__(self, nparams, solver, mu=0.9):
        self.solver, self.mu = solver, mu
        self.last = np.zeros(nparams)

    def solve(self, Obar, Ebar):
        Ebar = Ebar - self.mu * (Obar @ self.last)
        step = self.solver(Obar, Ebar) + self.mu * self.last
        self.last = step
        return step


class March:
    """``MARCH.solve`` (sr.py:321-341): second-moment scaling V of the
    parameter changes, bias-corrected, V**0.25 + 1e-8 (1 while V is all ~0)."""

    def __init__(self, nparams, solver, mu=0.95, beta=0.995):
        self.solver, self.mu, self.beta = solver, mu, beta
        self.last = np.zeros(nparams)
        self.V = np.zeros(nparams)
        self.t = 0

    def solve(self, Obar, Ebar):
        self.t += 1
        Ebar = Ebar - self.mu * (Obar @ self.last)
        if np.allclose(self.V, 0):
            V = np.ones_like(self.V)
        else:
            V = (self.V / (1 - self.beta ** self.t)) ** 0.25 + 1e-8
        step = self.solver(Obar / V[None, :], Ebar)
        step = step / V + self.mu * self.last
        self.V = self.beta * self.V + (1 - self.beta) * np.abs(step - self.last) ** 2
        self.last = step
        return step


class AdamSR:
    """``AdamSR.solve`` (sr.py:405-421): a plain SR step g feeds Adam moments
    m, v (bias-corrected); the step solves for the residual Ebar - Obar m in
    the metric scaled by V = v**0.25 + 1e-8."""

    def __init__(self, nparams, solver, mu=0.95, beta=0.995):
        self.solver, self.mu, self.beta = solver, mu, beta
        self.m = np.zeros(nparams)
        self.v = np.zeros(nparams)
        self.t = 0

    def solve(self, Obar, Ebar):
        self.t += 1
        g = self.solver(Obar, Ebar)
        self.m = self.mu * self.m + (1 - self.mu) * g
        self.v = self.beta * self.v + (1 - self.beta) * np.abs(g) ** 2
        m = self.m / (1 - self.mu ** self.t)
        v = self.v / (1 - self.beta ** self.t)
        V = v ** 0.25 + 1e-8
        Ebar = Ebar - Obar @ m
        step = self.solver(Obar / V[None, :], Ebar)
        return step / V + m
