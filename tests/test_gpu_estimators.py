"""Estimators and sampler rules on the device against the oracle (needs an
MI355X): sample weights for reweight != 2 (``sampler/sampler.py:66-69``) in
the sampler, E_loc statistics and the SR/MinSR step; ``Operator.expectation``
over a whole basis against the dense <H> and <H^2> - <H>^2; the drift check
after a local-update sweep (``metropolis.py:201-213``); the |psi| = 0
acceptance rule (``metropolis.py:304``); results across the internal chunk
boundaries (2048-sample ResConv Jacobian launches, 256 MiB FP64 Obar row
blocks of the SR branch)."""
import numpy as np
import pytest
import torch

from oracle import ed as oed
from oracle import models as om
from oracle import operator as oop
from oracle import philox
from oracle import sampler as osm
from oracle import sr as osr
from tests.helpers import max_rel, rand_spins

pytestmark = pytest.mark.gpu

import quantax_amd as qtx  # noqa: E402
from quantax_amd import _lib  # noqa: E402

DEV = "cuda"


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _rbm_state(L=4, M=32, nup=8, seed=2, std=0.1, chain=False):
    if chain:
        qtx.sites.Chain(L)
        N = L
    else:
        qtx.sites.Square(L, Nparticles=None if nup is None else (nup, L * L - nup))
        N = L * L
    model = qtx.model.RBM_Dense(M, dtype=np.float64, init_std=std, seed=seed)
    state = qtx.state.Variational(model)
    W, b = om.rbm_unpack(state.get_params_flatten().cpu().numpy(), M, N)
    return model, state, W, b


@pytest.mark.parametrize("reweight", [0.0, 1.0, 1.5, 2.0])
def test_reweight_factor(reweight):
    model, state, W, b = _rbm_state(std=0.3)
    smp = qtx.sampler.SpinExchange(state, 256, reweight=reweight, thermal_steps=64, seed=9).sweep()
    r = smp.reweight_factor.cpu().numpy()
    ref = osm.reweight_factor(smp.psi.logabs.cpu().numpy(), reweight)
    assert max_rel(r, ref, floor=1e-300) < 1e-13
    assert abs(r.mean() - 1.0) < 1e-13


@pytest.mark.parametrize("kind", ["flip", "exchange"])
def test_metropolis_reweight_1_bit_identical(kind, rng):
    """|psi|^1 sampling: the acceptance exponent enters the device's
    log-domain test; float64 chains equal the oracle's on the same words."""
    nup = 18 if kind == "exchange" else None
    qtx.sites.Square(6, Nparticles=None if nup is None else (nup, 36 - nup))
    model = qtx.model.RBM_Dense(40, dtype=np.float64, init_std=0.15, seed=3)
    W, b = om.rbm_unpack(model.theta.cpu().numpy(), 40, 36)
    lat = qtx.get_sites()
    ns, nsteps = 96, 108
    s0 = rand_spins(rng, ns, 36, nup)
    words = philox.sweep_words(78, nsteps, ns)
    if kind == "flip":
        kd, nbr, width, nbr_np = _lib.QTX_LOCALFLIP, None, 0, None
    else:
        kd, nbr_np = _lib.QTX_EXCHANGE, lat.neighbor_table(1)
        nbr, width = to_dev(nbr_np), nbr_np.shape[1]
    s = to_dev(s0)
    la = model.logpsi(s)
    nacc = torch.zeros(ns, dtype=torch.int32, device=DEV)
    nnear = torch.zeros(ns, dtype=torch.int32, device=DEV)
    model.metropolis(kd, nbr, width, 1, to_dev(words.view(np.int32)), nsteps, s, la, nacc, nnear, 1.0, 1e-9)
    s_ref, la_ref, nacc_ref, near_ref = osm.metropolis(
        lambda x: om.rbm_logpsi(W, b, x), s0, words, osm.LOCALFLIP if kind == "flip" else osm.EXCHANGE, nbr_np, 1,
        1.0, 1e-9)
    assert near_ref.sum() == 0
    assert np.array_equal(s.cpu().numpy(), s_ref)
    assert np.array_equal(nacc.cpu().numpy(), nacc_ref)
    assert max_rel(la.cpu().numpy(), la_ref) < 1e-10


@pytest.mark.parametrize("ns", [256, 1024])
@pytest.mark.parametrize("solver_kind", ["shift", "pinv"])
def test_sr_step_with_sample_weights(ns, solver_kind):
    """reweight = 1: energy, VarE, Ebar and the update vector with the
    weights r = |psi| / mean |psi| (MinSR at 256 samples, SR at 1024 > 544
    parameters) against the oracle on the same spins."""
    qtx.set_random_seed(5)
    model, state, W, b = _rbm_state(std=0.2)
    sampler = qtx.sampler.SpinExchange(state, ns, reweight=1.0, thermal_steps=96, seed=6)
    samples = sampler.sweep()
    r = samples.reweight_factor.cpu().numpy()
    assert r.std() > 1e-3  # the weights are not trivial
    if solver_kind == "shift":
        solver = qtx.optimizer.auto_shift_eig(rshift=0.0, ashift=1e-3)
        ref_solver = lambda A, v: osr.auto_shift_eig(A, v, rshift=0.0, ashift=1e-3)
    else:
        solver = qtx.optimizer.auto_pinv_eig(rtol=1e-8)
        ref_solver = lambda A, v: osr.auto_pinv_eig(A, v, rtol=1e-8)
    H = qtx.operator.Heisenberg(J=[1, 0.5], n_neighbor=[1, 2], msr=True)
    opt = qtx.optimizer.SR(state, H, solver=solver)
    step = opt.get_step(samples).cpu().numpy()
    s = samples.spins.cpu().numpy()
    Ho = oop.heisenberg([4, 4], [1, 0.5], [1, 2], True)
    eloc = oop.oloc(s, lambda x: om.rbm_logpsi(W, b, x), Ho)
    assert max_rel(opt._eloc.cpu().numpy(), eloc) < 1e-10
    O = om.rbm_jacobian(W, b, s)
    ref, E, VarE = osr.sr_step(O, eloc, ref_solver, r)
    assert abs(opt.energy - E) < 1e-10 * abs(E)
    assert abs(opt.VarE - VarE) < 1e-8 * max(1.0, abs(VarE))
    assert np.linalg.norm(step - ref) / np.linalg.norm(ref) < 1e-8


def test_ebar_with_weights(rng):
    ns = 777
    E = rng.standard_normal(ns) * 3 - 10
    r = rng.uniform(0.1, 2.0, ns)
    r /= r.mean()
    eb = torch.empty(ns, dtype=torch.float64, device=DEV)
    out2 = torch.empty(2, dtype=torch.float64, device=DEV)
    Ed, rd = to_dev(E), to_dev(r)
    _lib.call("qtx_ebar", _lib.stream(), Ed.data_ptr(), rd.data_ptr(), ns, eb.data_ptr(), out2.data_ptr())
    ref, Em, Ev = osr.get_Ebar(E, r)
    assert max_rel(eb.cpu().numpy(), ref, floor=1e-12) < 1e-12
    assert abs(out2[0].item() - Em) < 1e-13 * abs(Em)
    assert abs(out2[1].item() - Ev) < 1e-12 * abs(Ev)


def test_expectation_over_whole_basis():
    """reweight = 0 samples the basis uniformly; over all 2^12 configurations
    the weights are |psi|^2 / mean |psi|^2, so expectation() equals
    <psi|H|psi> / <psi|psi> and its variance <H^2> - <H>^2 of the dense H."""
    model, state, W, b = _rbm_state(L=12, M=24, nup=None, std=0.3, chain=True)
    H = qtx.operator.Heisenberg(J=[1.0, 0.4], n_neighbor=[1, 2])
    Ho = oop.heisenberg([12], [1.0, 0.4], [1, 2])
    Hs, basis = oed.sparse_hamiltonian(Ho, 12)
    sd = to_dev(basis)
    sampler = qtx.sampler.LocalFlip(state, 16, reweight=0.0, thermal_steps=0)
    psi = state(sd)
    r = sampler._reweight_factor(psi.logabs.contiguous())
    samples = qtx.sampler.Samples(sd, psi, None, r)
    m, v = H.expectation(state, samples, return_var=True)
    x = np.exp(om.rbm_logpsi(W, b, basis)[1] - om.rbm_logpsi(W, b, basis)[1].max())
    hx = Hs @ x
    e_ref = float(x @ hx / (x @ x))
    v_ref = float(hx @ hx / (x @ x)) - e_ref ** 2
    assert abs(m - e_ref) < 1e-10 * abs(e_ref)
    assert abs(v - v_ref) < 1e-8 * max(1.0, abs(v_ref))
    assert abs(H.expectation(state, samples) - m) < 1e-14 * abs(m)


def test_drift_check_replaces_local_update_psi(monkeypatch):
    """A local-update log psi that drifts from the full forward is counted
    with a warning and replaced (metropolis.py:201-213); a deviation below
    the 1e-3 relative threshold is not counted."""
    model, state, W, b = _rbm_state(L=16, M=16, nup=None, std=0.1, chain=True)
    sampler = qtx.sampler.LocalFlip(state, 32, thermal_steps=0, sweep_steps=16, seed=3)
    orig = model.metropolis
    drift = {3: 0.01, 7: -0.05, 20: 0.3, 11: 1e-5}

    def drifting(*args):
        orig(*args)
        la = args[7]
        for k, d in drift.items():
            la[k] += d

    monkeypatch.setattr(model, "metropolis", drifting)
    with pytest.warns(UserWarning, match="3 out of 32 wavefunctions"):
        smp = sampler.sweep()
    assert sampler.last_drift == 3
    assert torch.equal(smp.psi.logabs, model.logpsi(smp.spins))
    la_ref = om.rbm_logpsi(W, b, smp.spins.cpu().numpy())[1]
    assert max_rel(smp.psi.logabs.cpu().numpy(), la_ref) < 1e-12


def test_accept_zero_amplitude_always_moves(rng):
    """qtx_accept: a chain whose current |psi| is 0 (log psi = -inf) takes
    every proposal that changes it, whatever the uniform; the others follow
    the usual test.  Proposals are made unfavourable (log psi + 30 on the
    current state), so only the zero-amplitude chains move."""
    model, state, W, b = _rbm_state(L=16, M=16, nup=None, std=0.2, chain=True)
    ns, N = 64, 16
    s0 = rand_spins(rng, ns, N)
    words = philox.sweep_words(17, 1, ns)
    sd = to_dev(s0)
    wd = to_dev(words[0].view(np.int32))
    s_prop = torch.empty_like(sd)
    upd = torch.empty(ns, dtype=torch.int32, device=DEV)
    _lib.call("qtx_propose", _lib.stream(), _lib.QTX_LOCALFLIP, None, 0, 1, wd.data_ptr(), sd.data_ptr(), ns, N,
              s_prop.data_ptr(), upd.data_ptr())
    lp_prop = model.logpsi(s_prop).contiguous()
    la0 = om.rbm_logpsi(W, b, s0)[1] + 30.0
    zero = np.zeros(ns, bool)
    zero[::3] = True
    la0[zero] = -np.inf
    # half of the zero-amplitude chains propose a zero-amplitude state too:
    # their log ratio is NaN and only the |psi| = 0 clause accepts them
    both = np.zeros(ns, bool)
    both[::6] = True
    lp_prop[to_dev(both)] = -np.inf
    lp = to_dev(la0)
    sg = torch.ones(ns, dtype=torch.float64, device=DEV)
    sg_prop = torch.ones(ns, dtype=torch.float64, device=DEV)
    nacc = torch.zeros(ns, dtype=torch.int32, device=DEV)
    nnear = torch.zeros(ns, dtype=torch.int32, device=DEV)
    _lib.call("qtx_accept", _lib.stream(), wd.data_ptr(), s_prop.data_ptr(), lp_prop.data_ptr(), sg_prop.data_ptr(),
              upd.data_ptr(), ns, N, 2.0, 1e-9, sd.data_ptr(), lp.data_ptr(), sg.data_ptr(), nacc.data_ptr(),
              nnear.data_ptr())
    table = {x.tobytes(): v for x, v in zip(s_prop.cpu().numpy(), lp_prop.cpu().numpy())}
    assert len(table) == ns  # distinct proposals, so the lookup below is per chain
    fn = lambda x: (np.ones(len(x)), np.array([table[y.tobytes()] for y in x]))
    s_ref, la_ref, nacc_ref, _ = osm.metropolis(fn, s0, words, osm.LOCALFLIP, logpsi0=la0)
    assert np.array_equal(nacc_ref.astype(bool), zero)  # every zero-amplitude chain moved, no other did
    assert np.array_equal(sd.cpu().numpy(), s_ref)
    assert np.array_equal(lp.cpu().numpy(), la_ref)
    assert np.array_equal(nacc.cpu().numpy(), nacc_ref)


def test_resconv_jacobian_across_launch_chunks():
    """The ResConv Jacobian runs in launches of 2048 samples: 2085 samples
    (a partial second launch) give the same rows as sub-batches of 97, and
    rows from both launches match the oracle."""
    rng = np.random.default_rng(12)
    qtx.sites.Square(4)
    model = qtx.model.ResConv(2, 16, precision="fp32", seed=3)
    spec = om.ResConvSpec(2, 16, (4, 4))
    theta = model.theta.cpu().numpy().astype(np.float64)
    s = rand_spins(rng, 2085, 16)
    sd = to_dev(s)
    O = model.jacobian(sd)
    parts = torch.cat([model.jacobian(sd[i: i + 97]) for i in range(0, 2085, 97)])
    assert torch.equal(O, parts)
    rows = np.array([0, 1, 2046, 2047, 2048, 2049, 2084])
    ref = om.resconv_jacobian(spec, theta, s[rows])
    assert np.max(np.abs(O.cpu().numpy()[rows] - ref)) < 1e-5 * max(1.0, np.abs(ref).max())


def test_sr_branch_across_obar_row_blocks(rng):
    """ns >= np: S = Obar^T Obar is accumulated from FP64 Obar row blocks of
    at most 256 MiB; 140000 x 300 (320 MB in FP64) spans two blocks."""
    ns, np_ = 140000, 300
    O = (rng.standard_normal((ns, np_)) * 0.4 + 0.2).astype(np.float32)
    E = rng.standard_normal(ns) - 2.0
    assert ns * np_ * 8 > (256 << 20)
    Od = to_dev(O)
    mean = torch.empty(np_, dtype=torch.float64, device=DEV)
    _lib.call("qtx_colmean", _lib.stream(), _lib.QTX_F32, Od.data_ptr(), ns, np_, mean.data_ptr())
    eb = torch.empty(ns, dtype=torch.float64, device=DEV)
    out2 = torch.empty(2, dtype=torch.float64, device=DEV)
    Ed = to_dev(E)
    _lib.call("qtx_ebar", _lib.stream(), Ed.data_ptr(), None, ns, eb.data_ptr(), out2.data_ptr())
    step = qtx.optimizer.solve_step(Od, mean, eb, qtx.optimizer.auto_shift_eig(rshift=0.0, ashift=1e-3)).cpu().numpy()
    Obar, _ = osr.get_Obar(O.astype(np.float64))
    Ebar, _, _ = osr.get_Ebar(E)
    ref = osr.auto_shift_eig(Obar, Ebar, rshift=0.0, ashift=1e-3)
    assert np.linalg.norm(step - ref) / np.linalg.norm(ref) < 1e-10
