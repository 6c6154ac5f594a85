"""Device results against the committed golden vectors (needs an MI355X)."""
import numpy as np
import pytest
import torch

from tests.golden.make_golden import OUT
from tests.helpers import max_rel

pytestmark = pytest.mark.gpu

import quantax_amd as qtx  # noqa: E402
from quantax_amd import _lib  # noqa: E402

G = np.load(OUT, allow_pickle=False)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _j1j2():
    qtx.sites.Square(4, Nparticles=(8, 8))
    return qtx.operator.Heisenberg(J=[1.0, 0.5], n_neighbor=[1, 2], msr=True)


def test_golden_connected_configurations():
    H = _j1j2()
    s = dev(G["j1j2_spins"])
    assert np.allclose(H.apply_diag(s).cpu().numpy(), G["j1j2_diag"], rtol=0, atol=1e-12)
    seg, _, sc, hc = (x.cpu().numpy() for x in H.connected(s))
    assert np.array_equal(seg, G["j1j2_conn_segment"])
    assert np.array_equal(sc, G["j1j2_conn_spins"])
    assert np.array_equal(hc, G["j1j2_conn_h"])


def test_golden_rbm_float64():
    H = _j1j2()
    model = qtx.model.RBM_Dense(24, dtype=np.float64, init_std=0.1, seed=0)
    model.set_params(G["rbm_theta"])
    s = dev(G["j1j2_spins"])
    assert max_rel(model.logpsi(s).cpu().numpy(), G["rbm_logabs"]) < 1e-12
    assert max_rel(model.eloc(H, s, None)[0].cpu().numpy(), G["rbm_eloc"]) < 1e-10
    assert np.max(np.abs(model.jacobian(s).cpu().numpy() - G["rbm_O"])) < 1e-13


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16x3", 1e-4)])
def test_golden_resconv(precision, tol):
    H = _j1j2()
    model = qtx.model.ResConv(2, 16, precision=precision, seed=0)
    model.set_params(G["resconv_theta"])
    s = dev(G["j1j2_spins"])
    la = model.logpsi(s)
    assert max_rel(la.cpu().numpy(), G["resconv_logabs"]) < tol
    assert max_rel(model.eloc(H, s, la)[0].cpu().numpy(), G["resconv_eloc"]) < tol
    O = model.jacobian(s[:4]).cpu().numpy()
    assert np.max(np.abs(O - G["resconv_O"])) < 1e-5 * max(1.0, np.abs(G["resconv_O"]).max())


@pytest.mark.parametrize("kind", ["flip", "exchange"])
def test_golden_metropolis_chains(kind):
    pre = "mc" if kind == "flip" else "mcx"
    qtx.sites.Square(6, Nparticles=None if kind == "flip" else (18, 18))
    model = qtx.model.RBM_Dense(20, dtype=np.float64, init_std=0.1, seed=0)
    model.set_params(G["mc_theta"])
    s = dev(G[f"{pre}_s0"])
    la = model.logpsi(s)
    words = torch.empty((40, 32, 4), dtype=torch.int32, device="cuda")
    _lib.call("qtx_philox_fill", _lib.stream(), int(G[f"{pre}_seed"]), _lib.TAG_SWEEP, 0, 40, 0, 32,
              words.data_ptr())
    nacc = torch.zeros(32, dtype=torch.int32, device="cuda")
    nnear = torch.zeros(32, dtype=torch.int32, device="cuda")
    if kind == "flip":
        model.metropolis(_lib.QTX_LOCALFLIP, None, 0, 1, words, 40, s, la, nacc, nnear, 2.0, 1e-12)
    else:
        tab = qtx.get_sites().neighbor_table(1)
        model.metropolis(_lib.QTX_EXCHANGE, dev(tab), tab.shape[1], 1, words, 40, s, la, nacc, nnear, 2.0, 1e-12)
    assert np.array_equal(s.cpu().numpy(), G[f"{pre}_s"])
    assert np.array_equal(nacc.cpu().numpy(), G[f"{pre}_nacc"])
    assert max_rel(la.cpu().numpy(), G[f"{pre}_logabs"]) < 1e-10


@pytest.mark.parametrize("mode", ["chol", "eigh"])
def test_golden_sr_step(mode):
    O = dev(G["rbm_O"])
    E = dev(G["rbm_eloc"])
    ns, np_ = O.shape
    mean = torch.empty(np_, dtype=torch.float64, device="cuda")
    _lib.call("qtx_colmean", _lib.stream(), _lib.QTX_F64, O.data_ptr(), ns, np_, mean.data_ptr())
    eb = torch.empty(ns, dtype=torch.float64, device="cuda")
    out2 = torch.empty(2, dtype=torch.float64, device="cuda")
    _lib.call("qtx_ebar", _lib.stream(), E.data_ptr(), None, ns, eb.data_ptr(), out2.data_ptr())
    if mode == "chol":
        solver = qtx.optimizer.auto_shift_eig(rshift=0.0, ashift=1e-3)
        ref = G["sr_chol"]
    else:
        solver = qtx.optimizer.auto_pinv_eig(rtol=1e-10)
        ref = G["sr_eigh"]
    step = qtx.optimizer.solve_step(O, mean, eb, solver).cpu().numpy()
    assert np.linalg.norm(step - ref) / np.linalg.norm(ref) < 1e-9
