"""The opt-in ReLU ResConv variant (``gelu="relu"``, no counterpart in the
reference, whose ResConv uses GELU: ``model/conv_nets.py:85, 87``) on every
forward kernel against the oracle's ReLU network (needs an MI355X).

Kernels: the 16 x 16 window forward (C = 16), the 32-channel tile forward
(20 x 20, split-bf16), the generic FP32 forward (20 x 20, C = 32; 6 x 6) and
the FP32-MFMA Jacobian.  Tolerances as for GELU: FP32 1e-5 relative (absolute
floor 1), split-bf16 1e-4; chains identical except near-threshold decisions.
"""
import numpy as np
import pytest
import torch

from oracle import models as om
from oracle import operator as oop
from oracle import philox
from oracle import sampler as osm
from tests.helpers import max_rel, rand_spins

pytestmark = pytest.mark.gpu

import quantax_amd as qtx  # noqa: E402
from quantax_amd import _lib  # noqa: E402

DEV = "cuda"
RELU = "relu"


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _resconv(L, nb, C, precision, seed=21, nup=None, bias=0.1):
    lat = qtx.sites.Square(L, Nparticles=None if nup is None else (nup, L * L - nup))
    model = qtx.model.ResConv(nb, C, precision=precision, seed=seed, gelu=RELU)
    th = model.theta.cpu().numpy().astype(np.float64)
    r = np.random.default_rng(seed + 1)
    spec = om.ResConvSpec(nb, C, (L, L))
    for (cin, cout, hb), (w, bo) in zip(spec.layers, spec.offsets):
        if hb:
            th[bo: bo + cout] = bias * r.standard_normal(cout)
    model.set_params(th.astype(np.float32))
    return lat, model, spec, model.theta.cpu().numpy().astype(np.float64)


def _fn(spec, theta):
    return lambda x: om.resconv_forward(spec, theta, x, RELU)


@pytest.mark.parametrize("L,nb,C", [(16, 8, 16), (20, 3, 32), (6, 3, 16), (6, 2, 32)])
@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16x3", 1e-4)])
def test_relu_logpsi(L, nb, C, precision, tol):
    rng = np.random.default_rng(L * nb + C)
    lat, model, spec, theta = _resconv(L, nb, C, precision)
    s = rand_spins(rng, 5, L * L)
    la = model.logpsi(to_dev(s)).cpu().numpy()
    ref = om.resconv_forward(spec, theta, s, RELU)[1]
    # the GELU network on the same parameters differs: the variant is live
    assert np.max(np.abs(ref - om.resconv_forward(spec, theta, s)[1])) > 1e-3
    assert max_rel(la, ref) < tol


@pytest.mark.parametrize("L,nb,C,ns", [(16, 2, 16, 2), (20, 2, 32, 2), (6, 2, 16, 6)])
@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16x3", 1e-4)])
def test_relu_eloc(L, nb, C, ns, precision, tol):
    rng = np.random.default_rng(3 * L + C)
    nup = L * L // 2
    lat, model, spec, theta = _resconv(L, nb, C, precision, nup=nup)
    H = qtx.operator.Heisenberg(J=[1, 0.5], n_neighbor=[1, 2], msr=True)
    Ho = oop.heisenberg([L, L], [1, 0.5], [1, 2], True)
    s = rand_spins(rng, ns, L * L, nup=nup)
    la = model.logpsi(to_dev(s))
    e, _ = model.eloc(H, to_dev(s), la)
    ref = oop.oloc(s, _fn(spec, theta), Ho)
    assert max_rel(e.cpu().numpy(), ref) < tol


@pytest.mark.parametrize("L,nb,C", [(16, 2, 16), (20, 2, 32)])
def test_relu_metropolis(L, nb, C):
    rng = np.random.default_rng(L + nb)
    nup = L * L // 2
    lat, model, spec, theta = _resconv(L, nb, C, "bf16x3", nup=nup)
    ns, nsteps, tie = 4, 6, 1e-3
    s0 = rand_spins(rng, ns, L * L, nup=nup)
    words = philox.sweep_words(29, nsteps, ns)
    nbr_np = lat.neighbor_table(1)
    s = to_dev(s0)
    la = model.logpsi(s)
    nacc = torch.zeros(ns, dtype=torch.int32, device=DEV)
    nnear = torch.zeros(ns, dtype=torch.int32, device=DEV)
    model.metropolis(_lib.QTX_EXCHANGE, to_dev(nbr_np), nbr_np.shape[1], 1, to_dev(words.view(np.int32)), nsteps, s,
                     la, nacc, nnear, 2.0, tie)
    s_ref, la_ref, nacc_ref, near_ref = osm.metropolis(_fn(spec, theta), s0, words, osm.EXCHANGE, nbr_np, 1, 2.0, tie)
    clean = (near_ref == 0) & (nnear.cpu().numpy() == 0)
    assert clean.sum() >= ns - 1
    assert np.array_equal(s.cpu().numpy()[clean], s_ref[clean])
    assert np.array_equal(nacc.cpu().numpy()[clean], nacc_ref[clean])
    assert max_rel(la.cpu().numpy()[clean], la_ref[clean]) < 1e-4
    assert nacc_ref.sum() > 0


@pytest.mark.parametrize("L,nb,C", [(6, 3, 16), (16, 2, 16), (6, 2, 32)])
def test_relu_jacobian(L, nb, C):
    """ReLU's derivative is 0 or 1: a pre-activation within FP32 rounding of
    0 would flip one unit's gradient, which at these sizes (< 1e4
    pre-activations) has probability ~1e-3; the seeds here have none."""
    rng = np.random.default_rng(L * C + nb)
    lat, model, spec, theta = _resconv(L, nb, C, "fp32")
    s = rand_spins(rng, 3, L * L)
    O = model.jacobian(to_dev(s)).cpu().numpy()
    ref = om.resconv_jacobian(spec, theta, s, RELU)
    assert np.max(np.abs(O - ref)) < 1e-5 * max(1.0, np.max(np.abs(ref)))
