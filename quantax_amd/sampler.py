"""Metropolis samplers: ``quantax/sampler/metropolis.py`` with the proposal
kernels of ``sampler/common_samplers.py`` (LocalFlip, SpinExchange).

Semantics kept from the reference: ``nsamples`` chains advance in lock-step,
one sample per chain per ``sweep()``; default ``sweep_steps = 2 N`` and
``thermal_steps = 20 N`` (``metropolis.py:125-132``); acceptance
|psi'/psi|^reweight > 1 - u or |psi| == 0, only for s' != s
(``:291-322``); after a local-update sweep psi is re-evaluated by a full
forward and mismatches are counted (``:201-213``).  Random numbers come from
Philox streams keyed by (seed, chain, step) and generated on the device.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence, Union
from warnings import warn

import numpy as np
import torch

from . import _lib
from .global_defs import get_sites, get_subkeys
from .utils import LogArray


@dataclass
class Samples:
    """``quantax/sampler/samples.py:9-74``."""
    spins: torch.Tensor
    psi: LogArray
    state_internal: object = None
    reweight_factor: Optional[torch.Tensor] = None

    @property
    def nsamples(self) -> int:
        return self.spins.shape[0]


class Metropolis:
    kind = None
    nflips = None

    def __init__(self, state, nsamples: int, reweight: float = 2.0, thermal_steps: Optional[int] = None,
                 sweep_steps: Optional[int] = None, initial_spins=None, seed: Optional[int] = None,
                 tie_tol: float = 1e-9):
        self._state = state
        self._nsamples = int(nsamples)
        self._reweight = float(reweight)
        N = state.Nmodes
        self._thermal_steps = 20 * N if thermal_steps is None else int(thermal_steps)
        self._sweep_steps = 2 * N if sweep_steps is None else int(sweep_steps)
        self._seed = get_subkeys() if seed is None else int(seed)
        self._step = 0  # Philox step counter: consecutive sweeps draw disjoint streams
        self._tie_tol = tie_tol
        dev = _lib.device()
        self.n_accept = torch.zeros(self._nsamples, dtype=torch.int32, device=dev)
        self.n_near = torch.zeros(self._nsamples, dtype=torch.int32, device=dev)
        self.n_proposed = 0
        self.last_drift = 0
        self.kernel_events = []  # (start, end) HIP events around each Metropolis kernel launch
        if initial_spins is not None:
            s = torch.as_tensor(initial_spins).to(torch.int8)
            if s.ndim == 1:
                s = s[None, :].repeat(self._nsamples, 1)
            initial_spins = s.reshape(self._nsamples, N).to(dev)
        self._initial_spins = initial_spins
        self._nbr = None
        self._width = 0
        self._hopping = 1
        self.reset()

    @property
    def state(self):
        return self._state

    @property
    def nsamples(self) -> int:
        return self._nsamples

    @property
    def reweight(self) -> float:
        return self._reweight

    @property
    def Nmodes(self) -> int:
        return self._state.Nmodes

    def _initial(self) -> torch.Tensor:
        s = torch.empty((self._nsamples, self.Nmodes), dtype=torch.int8, device=_lib.device())
        _lib.call("qtx_init_spins", _lib.stream(), self._seed ^ 0x5DEECE66D, self._nsamples, self.Nmodes,
                  self._nup(), s.data_ptr())
        return s

    def _nup(self) -> int:
        return -1

    def reset(self) -> None:
        """Reset the chains to ``initial_spins`` (or random ones) and thermalise."""
        self._spins = self._initial() if self._initial_spins is None else self._initial_spins.clone()
        psi = self._state(self._spins)
        self._sign, self._logabs = psi.sign.contiguous(), psi.logabs.contiguous()
        if self._thermal_steps > 0:
            self.sweep(self._thermal_steps)

    @property
    def _generic(self) -> bool:
        """Symmetrised states (and the tests of this path) step through
        propose -> images -> forward -> project -> accept."""
        return self._state.is_symmetric or getattr(self, "force_generic", False)

    def _generic_steps(self, words: torch.Tensor, nsteps: int) -> None:
        st = self._state
        ns, N = self._spins.shape
        s_prop = torch.empty_like(self._spins)
        upd = torch.empty(ns, dtype=torch.int32, device=self._spins.device)
        for t in range(nsteps):
            w = words[t]
            _lib.call("qtx_propose", _lib.stream(), self.kind, _lib.ptr(self._nbr), self._width, self._hopping,
                      w.data_ptr(), self._spins.data_ptr(), ns, N, s_prop.data_ptr(), upd.data_ptr())
            psi = st(s_prop)
            sg, la = psi.sign.contiguous(), psi.logabs.contiguous()
            _lib.call("qtx_accept", _lib.stream(), w.data_ptr(), s_prop.data_ptr(), la.data_ptr(), sg.data_ptr(),
                      upd.data_ptr(), ns, N, self._reweight, self._tie_tol, self._spins.data_ptr(),
                      self._logabs.data_ptr(), self._sign.data_ptr(), self.n_accept.data_ptr(),
                      self.n_near.data_ptr())

    def sweep(self, nsweeps: Optional[int] = None) -> Samples:
        nsteps = self._sweep_steps if nsweeps is None else int(nsweeps)
        model = self._state.model
        ns = self._nsamples
        words = torch.empty((max(nsteps, 1), ns, 4), dtype=torch.int32, device=self._spins.device)
        if nsteps > 0:
            _lib.call("qtx_philox_fill", _lib.stream(), self._seed, _lib.TAG_SWEEP, self._step, nsteps, 0, ns,
                      words.data_ptr())
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record()
            if self._generic:
                self._generic_steps(words, nsteps)
            else:
                model.metropolis(self.kind, self._nbr, self._width, self._hopping, words, nsteps, self._spins,
                                 self._logabs, self.n_accept, self.n_near, self._reweight, self._tie_tol)
            ev1.record()
            self.kernel_events.append((ev0, ev1))
        self._step += nsteps
        self.n_proposed += nsteps
        if self._generic:
            la = self._logabs.clone()
            return Samples(self._spins.clone(), LogArray(self._sign.clone(), la), None, self._reweight_factor(la))
        uses_ref = self._state.use_ref and hasattr(model, "M")
        if uses_ref:
            # local-update drift check and replacement by the full forward (metropolis.py:201-213)
            full = model.logpsi(self._spins)
            close = (torch.abs(torch.exp(self._logabs) - torch.exp(full)) < 1e-8) | (
                torch.abs(torch.exp(self._logabs - full) - 1) < 1e-3)
            ndiff = int((~close).sum().item())
            self.last_drift = ndiff
            if ndiff > 0:
                warn(f"{ndiff} out of {ns} wavefunctions are not close in direct forward pass and local updates.")
            self._logabs = full
        la = self._logabs.clone()
        psi = LogArray(torch.ones_like(la), la)
        return Samples(self._spins.clone(), psi, None, self._reweight_factor(la))

    def _reweight_factor(self, logabs: torch.Tensor) -> torch.Tensor:
        """``Sampler._get_reweight_factor`` (sampler/sampler.py:66-69)."""
        if self._reweight == 2.0:
            return torch.ones_like(logabs)
        r = torch.exp((2.0 - self._reweight) * (logabs - logabs.max()))
        return r / r.mean()

    @property
    def acceptance(self) -> float:
        if self.n_proposed == 0:
            return 0.0
        return float(self.n_accept.sum().item()) / (self.n_proposed * self._nsamples)


class LocalFlip(Metropolis):
    """Flip one random spin per proposal (``common_samplers.py:14-33``)."""
    kind = _lib.QTX_LOCALFLIP
    nflips = 1


class SpinExchange(Metropolis):
    """Swap a random minority-species spin with a random neighbour
    (``common_samplers.py:85-162``)."""
    kind = _lib.QTX_EXCHANGE
    nflips = 2

    def __init__(self, state, nsamples: int, reweight: float = 2.0, thermal_steps: Optional[int] = None,
                 sweep_steps: Optional[int] = None, initial_spins=None, n_neighbor: Union[int, Sequence[int]] = 1,
                 seed: Optional[int] = None, tie_tol: float = 1e-9):
        sites = get_sites()
        if not isinstance(sites.Nparticles, tuple):
            raise ValueError("The number spin-up and spin-down particles should be specified in sites for "
                             "`SpinExchange` sampler.")
        self._nup_val = sites.Nparticles[0]
        tab = sites.neighbor_table(n_neighbor)
        self._tab = tab
        super_init = super().__init__
        self._pending = (tab, 1 if 2 * self._nup_val <= state.Nmodes else -1)
        super_init(state, nsamples, reweight, thermal_steps, sweep_steps, initial_spins, seed, tie_tol)

    def reset(self) -> None:
        if self._nbr is None:
            tab, hop = self._pending
            self._nbr = torch.from_numpy(tab).to(_lib.device())
            self._width = tab.shape[1]
            self._hopping = hop
        super().reset()

    def _nup(self) -> int:
        return self._nup_val
