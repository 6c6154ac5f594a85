"""Variational state: ``quantax/state/variational.py`` for real parameters.

``__call__`` returns log psi as a ``LogArray`` (the reference returns a
PsiArray cast to float64, ``variational.py:261-347``); ``jacobian`` returns
O = d log|psi| / d theta per sample in ravel_pytree order
(``variational.py:424-511``); ``update`` applies theta <- theta - step and
skips non-finite steps with a warning (``variational.py:558-579``).
``max_parallel`` is accepted for signature compatibility: the kernels tile
samples internally, so results never depend on it.  With a non-trivial
``symm`` (``variational.py:261-323``, ``436-491``) the forward evaluates the
model on every image and projects (``Symmetry.project``), and the Jacobian
is the images' Jacobians weighted by c_g psi_g / psi_sym.
"""
from __future__ import annotations

from typing import Optional
from warnings import warn

import numpy as np
import torch

from . import _lib
from .utils import LogArray


class Variational:
    def __init__(self, model, param_file: Optional[str] = None, symm=None, max_parallel=None, use_ref: bool = True):
        self._symm = None if symm is None or symm.is_identity else symm
        self._model = model
        self._max_parallel = max_parallel
        self._use_ref = use_ref
        if param_file is not None:
            self.load(param_file)

    @property
    def model(self):
        return self._model

    @property
    def nparams(self) -> int:
        return self._model.nparams

    @property
    def Nsites(self) -> int:
        return self._model.N

    @property
    def Nmodes(self) -> int:
        return self._model.N

    @property
    def dtype(self):
        return self._model.theta.dtype

    @property
    def use_ref(self) -> bool:
        return self._use_ref

    @property
    def symm(self):
        from .symmetry import Identity

        return Identity() if self._symm is None else self._symm

    @property
    def is_symmetric(self) -> bool:
        return self._symm is not None

    def __call__(self, s) -> LogArray:
        s = torch.as_tensor(s, device=_lib.device()).to(torch.int8).reshape(-1, self.Nmodes)
        if self._symm is None:
            la = self._model.logpsi(s)
            return LogArray(torch.ones_like(la), la)
        sg, la = self._symm.project(self._model.logpsi(self._symm.images(s)))
        return LogArray(sg, la)

    def jacobian(self, fock_states) -> torch.Tensor:
        s = torch.as_tensor(fock_states, device=_lib.device()).to(torch.int8).reshape(-1, self.Nmodes)
        if self._symm is None:
            return self._model.jacobian(s)
        symm = self._symm
        ns, nimg, np_ = s.shape[0], symm.nsymm, self.nparams
        O = None
        # images' Jacobians of at most ~4 GB at a time
        chunk = max(1, min(ns, (4 << 30) // max(1, nimg * np_ * 8)))
        for s0 in range(0, ns, chunk):
            sc = s[s0: s0 + chunk]
            img = symm.images(sc)
            _, _, w = symm.project(self._model.logpsi(img), weights=True)
            Oi = self._model.jacobian(img)
            if O is None:
                O = torch.empty((ns, np_), dtype=Oi.dtype, device=Oi.device)
            _lib.call("qtx_rows_combine", _lib.stream(), _lib.QTX_F32 if Oi.dtype == torch.float32 else _lib.QTX_F64,
                      Oi.data_ptr(), w.data_ptr(), sc.shape[0], nimg, np_, O[s0:].data_ptr())
        return O

    def get_params_flatten(self) -> torch.Tensor:
        return self._model.theta.clone()

    def set_params(self, theta) -> None:
        self._model.set_params(theta)

    def update(self, step: torch.Tensor) -> bool:
        """theta <- theta - step; returns False (and warns) when skipped."""
        step = torch.as_tensor(step, device=_lib.device()).to(torch.float64).contiguous()
        if step.numel() != self.nparams:
            raise ValueError("step size does not match the number of parameters")
        flag = torch.empty(1, dtype=torch.int32, device=step.device)
        m = self._model
        _lib.call("qtx_apply_update", _lib.stream(), _lib.QTX_F32 if m.theta.dtype == torch.float32 else _lib.QTX_F64,
                  m.theta.data_ptr(), step.data_ptr(), 1.0, self.nparams, flag.data_ptr())
        ok = bool(flag.item())
        if not ok:
            warn("Got invalid update step. The update is interrupted.")
            return False
        m.repack()
        return True

    def save(self, file) -> None:
        """The flat parameter vector as one .npy array (ravel_pytree order)."""
        np.save(file, self._model.theta.cpu().numpy(), allow_pickle=False)

    def load(self, file) -> None:
        """Parameters from ``save`` above, or from a quantax checkpoint written by
        ``Variational.save`` = ``equinox.tree_serialise_leaves`` (variational.py:
        581-587, read at :162-163): see ``read_leaf_stream``."""
        self._model.set_params(params_from_file(file, self.nparams))


def read_leaf_stream(file) -> list:
    """The arrays of a leaf stream as ``equinox.tree_serialise_leaves`` writes
    it: one ``.npy`` record per array-like leaf (arrays, and bool/int/float
    scalars), in pytree-flatten order, back to back; other leaves write
    nothing.  Read with ``allow_pickle=False``."""
    out = []
    with open(file, "rb") if isinstance(file, (str, bytes)) or hasattr(file, "__fspath__") else _nullctx(file) as f:
        while True:
            head = f.read(6)
            if len(head) < 6:
                break
            if head != b"\x93NUMPY":
                raise ValueError("not a stream of .npy records")
            f.seek(-6, 1)
            out.append(np.lib.format.read_array(f, allow_pickle=False))
    return out


class _nullctx:
    def __init__(self, f):
        self.f = f

    def __enter__(self):
        return self.f

    def __exit__(self, *a):
        return False


def params_from_file(file, nparams: int) -> np.ndarray:
    """Flat parameters from a single-array file or an equinox leaf stream: the
    floating-point leaves of rank >= 1 are the model's parameters, in ravel_pytree
    order (integer and boolean scalars such as ``nblocks`` or ``holomorphic``
    are configuration); their total size must equal nparams."""
    leaves = read_leaf_stream(file)
    if len(leaves) == 1 and leaves[0].size == nparams:
        return leaves[0].reshape(-1)
    params = [a for a in leaves if np.issubdtype(a.dtype, np.floating) and a.ndim >= 1]
    total = sum(a.size for a in params)
    if total != nparams:
        raise ValueError(f"the file holds {total} floating-point parameters in {len(params)} arrays; "
                         f"the model has {nparams}")
    return np.concatenate([a.reshape(-1) for a in params])
