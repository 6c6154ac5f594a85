"""Wavefunction models backed by the HIP kernels.

``RBM_Dense`` mirrors ``quantax/model/shallow_nets.py:35-126`` (psi = prod
cosh(W s + b), a RefModel with local updates) and ``ResConv`` mirrors
``quantax/model/conv_nets.py:26-183`` (residual 3x3 circular convolutions with
GELU, exp head averaged over channels and translations).  Parameters are one
flat device vector in ravel_pytree order; the kernels consume a packed copy
rebuilt whenever the vector changes.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import _lib
from .global_defs import get_sites, get_subkeys

_TORCH = {np.float32: torch.float32, np.float64: torch.float64}


def _qdtype(tdtype):
    return _lib.QTX_F32 if tdtype == torch.float32 else _lib.QTX_F64


class _Model:
    nparams: int
    theta: torch.Tensor

    def set_params(self, theta) -> None:
        theta = torch.as_tensor(theta).to(device=_lib.device(), dtype=self.theta.dtype).reshape(-1)
        if theta.numel() != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters, got {theta.numel()}")
        self.theta.copy_(theta)
        self.repack()


class RBM_Dense(_Model):
    """psi(s) = prod_i cosh((W s + b)_i), W: (features, N), b: (features,)."""

    def __init__(self, features: int, use_bias: bool = True, dtype=np.float32, init_std: Optional[float] = None,
                 seed: Optional[int] = None):
        self.N = get_sites().Nsites
        self.M = int(features)
        self.dtype = np.dtype(dtype).type
        if self.dtype not in _TORCH:
            raise ValueError("RBM_Dense supports float32 and float64 parameters")
        if not use_bias:
            raise NotImplementedError("RBM_Dense without bias is not supported")
        self.use_bias = use_bias
        self.nparams = self.M * self.N + self.M
        rng = np.random.default_rng(get_subkeys() if seed is None else seed)
        if init_std is None:
            # lecun-normal weights rescaled so the std of log|psi| over random
            # inputs is 0.1 sqrt(N), zero bias (shallow_nets.py:17-32, 70-74)
            W = rng.standard_normal((self.M, self.N)) / np.sqrt(self.N)
            W *= self._scale(rng)
            b = np.zeros(self.M)
        else:
            W = rng.standard_normal((self.M, self.N)) * init_std
            b = rng.standard_normal(self.M) * init_std if use_bias else np.zeros(self.M)
        theta = np.concatenate([W.ravel(), b]).astype(self.dtype)
        self.theta = torch.from_numpy(theta).to(_lib.device())
        self.packed = torch.empty_like(self.theta)
        self.repack()

    def _scale(self, rng) -> float:
        x = rng.standard_normal((1000, self.M))
        target = 0.1 * np.sqrt(self.N)
        grid = np.arange(0, 1, 0.01)
        err = [(np.std(np.sum(np.log(np.abs(np.cosh(x * a))), axis=1)) - target) ** 2 for a in grid]
        return float(grid[int(np.nanargmin(err))])

    @property
    def qdtype(self):
        return _qdtype(self.theta.dtype)

    def repack(self) -> None:
        _lib.call("qtx_rbm_pack", _lib.stream(), self.qdtype, self.theta.data_ptr(), self.M, self.N,
                  self.packed.data_ptr())

    def logpsi(self, s: torch.Tensor, h_out: bool = False):
        s = s.contiguous()
        ns = s.shape[0]
        la = torch.empty(ns, dtype=torch.float64, device=s.device)
        h = torch.empty((ns, self.M), dtype=self.theta.dtype, device=s.device) if h_out else None
        _lib.call("qtx_rbm_logpsi", _lib.stream(), self.qdtype, self.packed.data_ptr(), self.M, self.N,
                  s.data_ptr(), ns, la.data_ptr(), _lib.ptr(h))
        return (la, h) if h_out else la

    def metropolis(self, kind, nbr, width, hopping, words, nsteps, s, logabs, n_accept, n_near, reweight, tie_tol):
        _lib.call("qtx_rbm_metropolis", _lib.stream(), self.qdtype, self.packed.data_ptr(), self.M, self.N, kind,
                  _lib.ptr(nbr), width, hopping, words.data_ptr(), nsteps, s.shape[0], reweight, tie_tol,
                  s.data_ptr(), logabs.data_ptr(), n_accept.data_ptr(), n_near.data_ptr())

    def eloc(self, H, s: torch.Tensor, logabs: torch.Tensor):
        s = s.contiguous()
        ns = s.shape[0]
        e = torch.empty(ns, dtype=torch.float64, device=s.device)
        nc = torch.empty(ns, dtype=torch.int32, device=s.device)
        _lib.call("qtx_rbm_eloc", _lib.stream(), self.qdtype, self.packed.data_ptr(), self.M, self.N,
                  H.device_tables(), s.data_ptr(), ns, e.data_ptr(), nc.data_ptr())
        return e, nc

    def jacobian(self, s: torch.Tensor, out_dtype=None):
        s = s.contiguous()
        ns = s.shape[0]
        od = self.theta.dtype if out_dtype is None else out_dtype
        O = torch.empty((ns, self.nparams), dtype=od, device=s.device)
        _lib.call("qtx_rbm_jacobian", _lib.stream(), self.qdtype, self.packed.data_ptr(), self.M, self.N,
                  s.data_ptr(), ns, _qdtype(od), O.data_ptr())
        return O


class ResConv(_Model):
    """ResConv(nblocks, channels, 3) on the current square/rectangular grid.

    ``gelu``: "tanh" (the reference's ``jax.nn.gelu`` default,
    ``model/conv_nets.py:85, 87``), "erf", or "relu" -- an opt-in ReLU in
    place of GELU, which the reference does not have."""

    def __init__(self, nblocks: int, channels: int, kernel_size: int = 3, dtype=np.float32,
                 precision: str = "fp32", gelu: str = "tanh", seed: Optional[int] = None):
        lat = get_sites()
        if kernel_size != 3:
            raise NotImplementedError("kernel_size must be 3")
        if lat.ndim != 2 or not np.all(lat.boundary != 0):
            raise NotImplementedError("ResConv needs a periodic 2D grid")
        if np.dtype(dtype).type != np.float32:
            raise NotImplementedError("ResConv parameters are float32")
        self.nblocks, self.channels = int(nblocks), int(channels)
        self.L0, self.L1 = lat.extent
        self.N = lat.Nsites
        self.desc = _lib.QtxResConvDesc()
        self.desc.nblocks, self.desc.channels = self.nblocks, self.channels
        self.desc.L0, self.desc.L1 = self.L0, self.L1
        acts = {"tanh": _lib.QTX_ACT_GELU_TANH, "erf": _lib.QTX_ACT_GELU_ERF, "relu": _lib.QTX_ACT_RELU}
        if gelu not in acts:
            raise ValueError(f"gelu must be one of {sorted(acts)}")
        self.activation = gelu
        self.desc.activation = acts[gelu]
        self.set_precision(precision, repack=False)
        self.nparams = int(_lib.lib().qtx_resconv_nparams(self.desc))
        rng = np.random.default_rng(get_subkeys() if seed is None else seed)
        theta = np.zeros(self.nparams, np.float64)
        off = 0
        for i in range(self.nblocks):
            for cin, hb in ((1 if i == 0 else self.channels, True), (self.channels, i != self.nblocks - 1)):
                n = self.channels * cin * 9
                theta[off: off + n] = rng.standard_normal(n) * np.sqrt(2.0 / (cin * 9))  # he_normal, zero bias
                off += n + (self.channels if hb else 0)
        assert off == self.nparams
        self.theta = torch.from_numpy(theta.astype(np.float32)).to(_lib.device())
        self.packed = None
        self.repack()

    def set_precision(self, precision: str, repack: bool = True) -> None:
        self.precision = precision
        self.desc.precision = {"fp32": _lib.QTX_CONV_FP32, "bf16x3": _lib.QTX_CONV_BF16X3}[precision]
        if repack:
            self.repack()

    def repack(self) -> None:
        nbytes = int(_lib.lib().qtx_resconv_packed_bytes(self.desc))
        if nbytes == 0:
            raise _lib.QtxError("unsupported ResConv shape: " + _lib.lib().qtx_last_error().decode())
        if self.packed is None or self.packed.numel() != nbytes:
            self.packed = torch.zeros(nbytes, dtype=torch.uint8, device=self.theta.device)
        _lib.call("qtx_resconv_pack", _lib.stream(), self.desc, self.theta.data_ptr(), self.packed.data_ptr())

    def logpsi(self, s: torch.Tensor):
        s = s.contiguous()
        la = torch.empty(s.shape[0], dtype=torch.float64, device=s.device)
        _lib.call("qtx_resconv_logpsi", _lib.stream(), self.desc, self.packed.data_ptr(), s.data_ptr(), s.shape[0],
                  la.data_ptr())
        return la

    def metropolis(self, kind, nbr, width, hopping, words, nsteps, s, logabs, n_accept, n_near, reweight, tie_tol):
        _lib.call("qtx_resconv_metropolis", _lib.stream(), self.desc, self.packed.data_ptr(), kind, _lib.ptr(nbr),
                  width, hopping, words.data_ptr(), nsteps, s.shape[0], reweight, tie_tol, s.data_ptr(),
                  logabs.data_ptr(), n_accept.data_ptr(), n_near.data_ptr())

    def eloc(self, H, s: torch.Tensor, logabs: torch.Tensor):
        s = s.contiguous()
        ns = s.shape[0]
        e = torch.empty(ns, dtype=torch.float64, device=s.device)
        nc = torch.empty(ns, dtype=torch.int32, device=s.device)
        _lib.call("qtx_resconv_eloc", _lib.stream(), self.desc, self.packed.data_ptr(), H.device_tables(),
                  s.data_ptr(), logabs.contiguous().data_ptr(), ns, e.data_ptr(), nc.data_ptr())
        return e, nc

    def jacobian(self, s: torch.Tensor, out_dtype=None):
        if out_dtype not in (None, torch.float32):
            raise NotImplementedError("ResConv Jacobian is written in float32")
        s = s.contiguous()
        ns = s.shape[0]
        O = torch.empty((ns, self.nparams), dtype=torch.float32, device=s.device)
        wsb = int(_lib.lib().qtx_resconv_jacobian_ws_bytes(self.desc, ns))
        ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=s.device)
        _lib.call("qtx_resconv_jacobian", _lib.stream(), self.desc, self.packed.data_ptr(), s.data_ptr(), ns,
                  O.data_ptr(), ws.data_ptr(), wsb)
        return O
