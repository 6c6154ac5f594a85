"""quantax_amd — the quantax VMC step on MI355X (gfx950) HIP kernels.

Drop-in mirror of the reference's hot path (ChenAo-Phys/quantax):
``sites.Square/Chain``, ``operator.Heisenberg/Ising`` with ``Oloc`` and
``expectation``, ``model.RBM_Dense/ResConv``, ``state.Variational``,
``sampler.LocalFlip/SpinExchange.sweep()``, ``optimizer.SR(...).get_step()``
(and SPRING, MARCH, AdamSR), ``optimizer.Euler`` and the symmetry
projections of ``symmetry`` (C4v, Flip, Rotation, SpinInverse, TransND, ...).  All device work goes through the C ABI of
``_lib/libqtxhip.so`` (include/qtx_hip.h); there is no CPU fallback.
"""
from . import global_defs
from .global_defs import get_default_dtype, get_sites, get_subkeys, set_default_dtype, set_random_seed

__all__ = [
    "global_defs", "sites", "operator", "model", "state", "sampler", "optimizer", "utils", "symmetry",
    "set_random_seed", "get_subkeys", "get_sites", "set_default_dtype", "get_default_dtype",
]


def __getattr__(name):
    # submodules that need torch / the HIP library are imported lazily so the
    # host-only parts (sites, op_list algebra) import without a GPU
    import importlib

    if name in ("sites", "operator", "model", "state", "sampler", "optimizer", "utils", "symmetry", "_lib", "build"):
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
