"""SR / MinSR optimizer and the Euler driver.

Mirrors ``quantax/optimizer/sr.py:17-195`` (QNGD/SR: Ebar, Obar, solve,
get_step, energy, VarE), the solver factories of
``quantax/optimizer/solver.py:50-201`` and ``Euler`` of
``quantax/optimizer/driver.py:30-39``.  A solver here is a small descriptor
(mode, shifts, tolerances, SNR tolerance, Gram precision, branch) executed by
``qtx_minsr_solve`` / ``qtx_minsr_solve_snr``: the ``auto_*`` factories take
MinSR (T = Obar Obar^T, Ns x Ns) when Ns < Np, else SR (S = Obar^T Obar); the
``minnorm_*`` / ``lstsq_*`` factories take their branch whatever the shape.
``minsr_pinv_eig`` solves T x = b for a given T (``solver.py:262-294``).
Obar is formed on the fly on the device; nothing of size Ns x Np is
materialised beyond O itself.  The
momentum variants SPRING, MARCH and AdamSR (``sr.py:198-429``) reuse the same
solve, with Obar v products and the column scaling of Obar done by
``qtx_obar_matvec`` and ``qtx_scale_columns``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib
from .utils import DataTracer

_GRAM = {"fp64": _lib.QTX_GRAM_FP64, "fp32": _lib.QTX_GRAM_FP32, "bf16x3": _lib.QTX_GRAM_BF16X3}


@dataclass
class Solver:
    mode: int
    rshift: float = 0.0
    ashift: float = 0.0
    rtol: float = 1e-12
    atol: float = 0.0
    gram: str = "fp64"
    branch: str = "auto"  # "auto" | "minnorm" | "lstsq": the reference factory this came from
    tol_snr: float = 0.0  # SNR truncation of the eigen-components (pinv solvers), <= 1e-6: none
    # rocBLAS / rocSOLVER scratch in the solve's workspace (torch's allocator)
    # rather than allocated by rocBLAS per stream handle
    caller_scratch: bool = True

    @property
    def snr(self) -> bool:
        return self.tol_snr > 1e-6  # solver.py:124

    @property
    def cmode(self) -> int:
        """The mode word of the C-ABI: the base mode, the branch flag of a
        minnorm_* / lstsq_* factory, the SNR workspace flag and the
        caller-scratch flag."""
        m = self.mode
        if self.branch == "minnorm":
            m |= _lib.QTX_SOLVE_MINNORM
        elif self.branch == "lstsq":
            m |= _lib.QTX_SOLVE_LSTSQ
        if self.snr:
            m |= _lib.QTX_SOLVE_SNR
        if self.caller_scratch:
            m |= _lib.QTX_SOLVE_BLAS_WS
        return m

    def ws_bytes(self, ns: int, np_: int) -> int:
        """Workspace bytes of a solve of this shape: the library's own part,
        plus (caller_scratch) the rocBLAS / rocSOLVER scratch by a size query
        on the current stream's handle, cached per shape and mode."""
        own = int(_lib.lib().qtx_minsr_ws_bytes(ns, np_, self.cmode, _GRAM[self.gram]))
        if not self.caller_scratch:
            return own
        return own + _blas_bytes(ns, np_, self.cmode, _GRAM[self.gram], own)


_BLAS_BYTES = {}


def _blas_bytes(ns: int, np_: int, cmode: int, gram: int, own: int) -> int:
    key = (ns, np_, cmode, gram)
    if key not in _BLAS_BYTES:
        tmp = torch.empty(max(own, 1), dtype=torch.uint8, device=_lib.device())
        out = C.c_size_t(0)
        _lib.call("qtx_minsr_blas_ws_bytes", _lib.stream(), ns, np_, cmode, gram, tmp.data_ptr(), tmp.numel(),
                  C.byref(out))
        _BLAS_BYTES[key] = int(out.value)
    return _BLAS_BYTES[key]


def _rt(v):
    return 1e-12 if v is None else float(v)  # _get_rtol for float64 (solver.py:12-21)


def _snr(v) -> float:
    v = float(v)
    if not v >= 0.0:
        raise ValueError("tol_snr must be >= 0")
    return v


def minnorm_shift_eig(rshift: Optional[float] = None, ashift: float = 1e-4, gram: str = "fp64") -> Solver:
    """x = Obar^T (Obar Obar^T + shift)^{-1} Ebar whatever the shape (solver.py:50-62)."""
    return Solver(_lib.QTX_SOLVE_CHOL_SHIFT, rshift=_rt(rshift), ashift=ashift, gram=gram, branch="minnorm")


def lstsq_shift_eig(rshift: Optional[float] = None, ashift: float = 1e-4, gram: str = "fp64") -> Solver:
    """x = (Obar^T Obar + shift)^{-1} Obar^T Ebar whatever the shape (solver.py:65-77)."""
    return Solver(_lib.QTX_SOLVE_CHOL_SHIFT, rshift=_rt(rshift), ashift=ashift, gram=gram, branch="lstsq")


def auto_shift_eig(rshift: Optional[float] = None, ashift: float = 1e-4, gram: str = "fp64") -> Solver:
    return Solver(_lib.QTX_SOLVE_CHOL_SHIFT, rshift=_rt(rshift), ashift=ashift, gram=gram)


def minnorm_pinv_eig(rtol: Optional[float] = None, atol: float = 0.0, tol_snr: float = 0.0,
                     gram: str = "fp64") -> Solver:
    """MinSR by the soft pseudo-inverse of T = Obar Obar^T (solver.py:128-149)."""
    return Solver(_lib.QTX_SOLVE_EIGH_PINV, rtol=_rt(rtol), atol=atol, gram=gram, branch="minnorm",
                  tol_snr=_snr(tol_snr))


def lstsq_pinv_eig(rtol: Optional[float] = None, atol: float = 0.0, tol_snr: float = 0.0,
                   gram: str = "fp64") -> Solver:
    """SR by the soft pseudo-inverse of S = Obar^T Obar (solver.py:152-164)."""
    return Solver(_lib.QTX_SOLVE_EIGH_PINV, rtol=_rt(rtol), atol=atol, gram=gram, branch="lstsq",
                  tol_snr=_snr(tol_snr))


def auto_pinv_eig(rtol: Optional[float] = None, atol: float = 0.0, tol_snr: float = 0.0,
                  gram: str = "fp64") -> Solver:
    """MinSR when Ns < Np, else SR (solver.py:167-201)."""
    return Solver(_lib.QTX_SOLVE_EIGH_PINV, rtol=_rt(rtol), atol=atol, gram=gram, tol_snr=_snr(tol_snr))


def minsr_pinv_eig(rtol: Optional[float] = None, atol: float = 0.0, tol_snr: float = 0.0):
    """The solver of ``T x = b`` for a given symmetric T (solver.py:262-294):
    x = U g(D) rho, T = U D U^T by rocSOLVER dsyevd, rho = U^T b with the SNR
    truncation when tol_snr > 1e-6.  Returns ``solve(T, b) -> x`` on FP64
    device tensors."""
    rt, at, ts = _rt(rtol), float(atol), _snr(tol_snr)

    def solve(T: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
        n = b.shape[0]
        if T.shape != (n, n):
            raise ValueError(f"T must be {n} x {n}, got {tuple(T.shape)}")
        if T.dtype != torch.float64 or b.dtype != torch.float64:
            raise TypeError("minsr_pinv_eig works in float64")
        T, b = T.contiguous(), b.contiguous()
        need = int(_lib.lib().qtx_pinv_eig_ws_bytes(n))
        ws = torch.empty(need, dtype=torch.uint8, device=T.device)
        x = torch.empty(n, dtype=torch.float64, device=T.device)
        _lib.call("qtx_pinv_eig", _lib.stream(), T.data_ptr(), n, b.data_ptr(), rt, at, ts, x.data_ptr(),
                  ws.data_ptr(), ws.numel())
        return x

    return solve


def solve_step(O: torch.Tensor, Omean: torch.Tensor, ebar: torch.Tensor, solver: Solver,
               r: Optional[torch.Tensor] = None, ws: Optional[torch.Tensor] = None) -> torch.Tensor:
    """step = solver(Obar, Ebar) with Obar = (O - Omean) sqrt(r / Ns) formed on the fly."""
    ns, np_ = O.shape
    od = _lib.QTX_F32 if O.dtype == torch.float32 else _lib.QTX_F64
    need = solver.ws_bytes(ns, np_)
    if ws is None or ws.numel() < need:
        ws = torch.empty(need, dtype=torch.uint8, device=O.device)
    step = torch.empty(np_, dtype=torch.float64, device=O.device)
    if solver.snr:
        _lib.call("qtx_minsr_solve_snr", _lib.stream(), od, O.data_ptr(), Omean.data_ptr(), _lib.ptr(r),
                  ebar.data_ptr(), ns, np_, solver.cmode, _GRAM[solver.gram], solver.rtol, solver.atol,
                  solver.tol_snr, step.data_ptr(), ws.data_ptr(), ws.numel())
    else:
        _lib.call("qtx_minsr_solve", _lib.stream(), od, O.data_ptr(), Omean.data_ptr(), _lib.ptr(r),
                  ebar.data_ptr(), ns, np_, solver.cmode, _GRAM[solver.gram], solver.rshift, solver.ashift,
                  solver.rtol, solver.atol, step.data_ptr(), ws.data_ptr(), ws.numel())
    return step


def last_solve_info() -> dict:
    """What the last solve on this thread did: ``path`` is "chol_shift",
    "partial_fractions", "syevd", "shifted_cholesky" or "chebyshev"; Lanczos
    steps and lambda_max when run;
    ``fallback``: why the partial-fraction or Chebyshev route was left, if it was."""
    path, steps, lmax, fb = C.c_int(), C.c_int(), C.c_double(), C.c_int()
    _lib.call("qtx_minsr_last_path", C.byref(path), C.byref(steps), C.byref(lmax), C.byref(fb))
    names = {0: "chol_shift", 1: "partial_fractions", 2: "syevd", 3: "shifted_cholesky", 4: "chebyshev"}
    reasons = {0: None, 1: "small_cutoff", 2: "nonfinite", 3: "nonpositive_lambda_max", 4: "cholesky_failed",
               5: "snr_truncation", 6: "shifted_cholesky_failed", 7: "shifted_cholesky_not_converged",
               8: "chebyshev_interval"}
    return {"path": names.get(path.value, "none"), "lanczos_steps": steps.value, "lambda_max": lmax.value,
            "fallback": reasons.get(fb.value)}


class QNGD:
    def __init__(self, state, imag_time: bool = True, solver: Optional[Solver] = None):
        if not imag_time:
            raise NotImplementedError("real-time evolution is not supported")
        self._state = state
        self._solver = auto_pinv_eig() if solver is None else solver
        self._Omean = None
        self._ws = None

    @property
    def state(self):
        return self._state

    @property
    def solver(self) -> Solver:
        return self._solver

    def get_Ebar(self, samples):
        raise NotImplementedError

    def get_Obar(self, samples):
        """Returns (O, Omean): Obar = (O - mean O) sqrt(r / Ns) is applied
        inside the solve (``sr.py:74-88``)."""
        O = self._state.jacobian(samples.spins)
        mean = torch.empty(O.shape[1], dtype=torch.float64, device=O.device)
        _lib.call("qtx_colmean", _lib.stream(), _lib.QTX_F32 if O.dtype == torch.float32 else _lib.QTX_F64,
                  O.data_ptr(), O.shape[0], O.shape[1], mean.data_ptr())
        self._Omean = mean
        return O, mean

    def get_step(self, samples) -> torch.Tensor:
        ebar = self.get_Ebar(samples)
        O, mean = self.get_Obar(samples)
        r = None if samples.reweight_factor is None or self._trivial_r(samples) else samples.reweight_factor
        ns, np_ = O.shape
        need = self._solver.ws_bytes(ns, np_)
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device=O.device)
        return solve_step(O, mean, ebar, self._solver, r, self._ws)

    @staticmethod
    def _trivial_r(samples) -> bool:
        return bool(torch.all(samples.reweight_factor == 1.0).item())


class SR(QNGD):
    """Stochastic reconfiguration, MinSR when Ns < Np (``sr.py:131-195``)."""

    def __init__(self, state, hamiltonian, imag_time: bool = True, solver: Optional[Solver] = None):
        super().__init__(state, imag_time, solver)
        self._hamiltonian = hamiltonian
        self._energy = None
        self._VarE = None
        self._eloc = None

    @property
    def hamiltonian(self):
        return self._hamiltonian

    @property
    def energy(self) -> Optional[float]:
        return None if self._energy is None else float(self._energy)

    @property
    def VarE(self) -> Optional[float]:
        return None if self._VarE is None else float(self._VarE)

    def get_Ebar(self, samples) -> torch.Tensor:
        eloc = self._hamiltonian.Oloc(self._state, samples)
        self._eloc = eloc
        ns = eloc.shape[0]
        r = samples.reweight_factor
        r = None if r is None or self._trivial_r(samples) else r.contiguous()
        ebar = torch.empty(ns, dtype=torch.float64, device=eloc.device)
        out2 = torch.empty(2, dtype=torch.float64, device=eloc.device)
        _lib.call("qtx_ebar", _lib.stream(), eloc.data_ptr(), _lib.ptr(r), ns, ebar.data_ptr(), out2.data_ptr())
        self._energy, self._VarE = out2[0], out2[1]
        return ebar


class _MomentumSR(SR):
    """Shared plumbing of the momentum variants: ``get_step`` forms E_loc and
    O as SR does and hands (O, mean, Ebar) to ``solve``, which may scale O's
    columns in place (O is recomputed every step)."""

    _saved = ()

    def _vec(self):
        return torch.zeros(self._state.nparams, dtype=torch.float64, device="cuda")

    def get_step(self, samples) -> torch.Tensor:
        ebar = self.get_Ebar(samples)
        O, mean = self.get_Obar(samples)
        r = None if samples.reweight_factor is None or self._trivial_r(samples) else samples.reweight_factor
        return self.solve(O, mean, ebar, r)

    def _sr(self, O, mean, ebar, r):
        ns, np_ = O.shape
        need = self._solver.ws_bytes(ns, np_)
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device=O.device)
        return solve_step(O, mean, ebar, self._solver, r, self._ws)

    @staticmethod
    def _odt(O):
        return _lib.QTX_F32 if O.dtype == torch.float32 else _lib.QTX_F64

    def _obar_v(self, O, mean, r, alpha, v, ebar):
        """ebar + alpha Obar v (a new vector)."""
        y = ebar.clone()
        _lib.call("qtx_obar_matvec", _lib.stream(), self._odt(O), O.data_ptr(), mean.data_ptr(), _lib.ptr(r),
                  O.shape[0], O.shape[1], float(alpha), v.data_ptr(), 1.0, y.data_ptr())
        return y

    def _scale_cols(self, O, mean, d):
        _lib.call("qtx_scale_columns", _lib.stream(), self._odt(O), O.data_ptr(), mean.data_ptr(), O.shape[0],
                  O.shape[1], d.data_ptr())

    def save(self, file) -> None:
        """The optimizer's internal quantities (``sr.py:257-262`` and alike), as
        an .npz of this package (not the equinox leaf format)."""
        np.savez(file, **{k: (v.cpu().numpy() if torch.is_tensor(v) else np.asarray(v))
                          for k, v in ((k, getattr(self, k)) for k in self._saved)})

    def _load(self, file) -> None:
        with np.load(file, allow_pickle=False) as z:
            for k in self._saved:
                v = getattr(self, k)
                setattr(self, k, torch.from_numpy(z[k]).to(v.device) if torch.is_tensor(v) else type(v)(z[k]))


class SPRING(_MomentumSR):
    """SR with momentum (``sr.py:198-262``): Ebar -= mu Obar last, step =
    solve(Obar, Ebar) + mu last.  With the default mu = 0.9 the learning rate
    should be about 1/5 of SR's."""

    _saved = ("_mu", "_last_step")

    def __init__(self, state, hamiltonian, imag_time: bool = True, solver: Optional[Solver] = None,
                 mu: float = 0.9, file=None):
        super().__init__(state, hamiltonian, imag_time, solver)
        self._mu = float(mu)
        self._last_step = self._vec()
        if file is not None:
            self._load(file)

    def solve(self, O, mean, ebar, r=None) -> torch.Tensor:
        ebar = self._obar_v(O, mean, r, -self._mu, self._last_step, ebar)
        step = self._sr(O, mean, ebar, r) + self._mu * self._last_step
        self._last_step = step.clone()
        return step


class MARCH(_MomentumSR):
    """SR with first and second momenta of the parameter changes
    (``sr.py:266-349``).  Defaults mu = 0.95, beta = 0.995; learning rate
    about 1/5 of SR's."""

    _saved = ("_mu", "_beta", "_last_step", "_V", "_t")

    def __init__(self, state, hamiltonian, imag_time: bool = True, solver: Optional[Solver] = None,
                 mu: float = 0.95, beta: float = 0.995, file=None):
        super().__init__(state, hamiltonian, imag_time, solver)
        self._mu, self._beta = float(mu), float(beta)
        self._last_step = self._vec()
        self._V = self._vec()
        self._t = 0
        if file is not None:
            self._load(file)

    def solve(self, O, mean, ebar, r=None) -> torch.Tensor:
        self._t += 1
        ebar = self._obar_v(O, mean, r, -self._mu, self._last_step, ebar)
        if torch.allclose(self._V, torch.zeros_like(self._V)):
            V = torch.ones_like(self._V)
        else:
            V = (self._V / (1 - self._beta ** self._t)) ** 0.25 + 1e-8
        self._scale_cols(O, mean, 1.0 / V)
        step = self._sr(O, mean, ebar, r) / V + self._mu * self._last_step
        self._V = self._beta * self._V + (1 - self._beta) * (step - self._last_step) ** 2
        self._last_step = step.clone()
        return step


class AdamSR(_MomentumSR):
    """Adam moments of the SR step (``sr.py:352-429``): two solves per step."""

    _saved = ("_mu", "_beta", "_m", "_v", "_t")

    def __init__(self, state, hamiltonian, imag_time: bool = True, solver: Optional[Solver] = None,
                 mu: float = 0.95, beta: float = 0.995, file=None):
        super().__init__(state, hamiltonian, imag_time, solver)
        self._mu, self._beta = float(mu), float(beta)
        self._m = self._vec()
        self._v = self._vec()
        self._t = 0
        if file is not None:
            self._load(file)

    def solve(self, O, mean, ebar, r=None) -> torch.Tensor:
        self._t += 1
        g = self._sr(O, mean, ebar, r)
        self._m = self._mu * self._m + (1 - self._mu) * g
        self._v = self._beta * self._v + (1 - self._beta) * g ** 2
        m = self._m / (1 - self._mu ** self._t)
        v = self._v / (1 - self._beta ** self._t)
        V = v ** 0.25 + 1e-8
        ebar = self._obar_v(O, mean, r, -1.0, m, ebar)
        self._scale_cols(O, mean, 1.0 / V)
        return self._sr(O, mean, ebar, r) / V + m


class Euler:
    """One VMC step per call: sweep -> get_step -> update (``driver.py:30-39``)."""

    def __init__(self, state, sampler, optimizer, step_length: float):
        self._state = state
        self._sampler = sampler
        self._optimizer = optimizer
        self._step_length = step_length
        self._time = 0.0
        self.energy = DataTracer()
        self.VarE = DataTracer()

    def step(self) -> None:
        samples = self._sampler.sweep()
        step = self._optimizer.get_step(samples)
        self._state.update(self._step_length * step)
        self._time += self._step_length
        self.energy.append(self._optimizer.energy, self._time)
        self.VarE.append(self._optimizer.VarE, self._time)
