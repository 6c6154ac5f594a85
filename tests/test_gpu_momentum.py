"""Momentum SR variants on the device (SPRING, MARCH, AdamSR) against the
oracle on the same (O, E) sequences, and the two kernels they add: Obar v and
the column scaling of O (needs an MI355X)."""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle import sr as osr

pytestmark = pytest.mark.gpu

import quantax_amd as qtx  # noqa: E402
from quantax_amd import _lib  # noqa: E402

DEV = "cuda"


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _mean(Od):
    mean = torch.empty(Od.shape[1], dtype=torch.float64, device=DEV)
    dt = _lib.QTX_F32 if Od.dtype == torch.float32 else _lib.QTX_F64
    _lib.call("qtx_colmean", _lib.stream(), dt, Od.data_ptr(), Od.shape[0], Od.shape[1], mean.data_ptr())
    return mean


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("with_r", [False, True])
def test_obar_matvec(dtype, with_r, rng):
    ns, np_ = 300, 1037
    O = (rng.standard_normal((ns, np_)) * 0.3 + 0.1).astype(dtype)
    r = rng.uniform(0.5, 1.5, ns) if with_r else None
    v = rng.standard_normal(np_)
    y0 = rng.standard_normal(ns)
    Od = _dev(O)
    mean = _mean(Od)
    y = _dev(y0.copy())
    rd = None if r is None else _dev(r)  # device tensors stay referenced until the call is done
    vd = _dev(v)
    dt = _lib.QTX_F32 if dtype == np.float32 else _lib.QTX_F64
    _lib.call("qtx_obar_matvec", _lib.stream(), dt, Od.data_ptr(), mean.data_ptr(), _lib.ptr(rd), ns, np_, -0.7,
              vd.data_ptr(), 2.0, y.data_ptr())
    O64 = O.astype(np.float64)
    Obar = (O64 - O64.mean(0)) * np.sqrt((np.ones(ns) if r is None else r) / ns)[:, None]
    ref = 2.0 * y0 - 0.7 * (Obar @ v)
    assert np.max(np.abs(y.cpu().numpy() - ref)) < 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_scale_columns(dtype, rng):
    ns, np_ = 70, 2001
    O = rng.standard_normal((ns, np_)).astype(dtype)
    d = rng.uniform(0.2, 3.0, np_)
    Od = _dev(O)
    mean = _mean(Od)
    m0 = mean.cpu().numpy()
    dt = _lib.QTX_F32 if dtype == np.float32 else _lib.QTX_F64
    dd = _dev(d)
    _lib.call("qtx_scale_columns", _lib.stream(), dt, Od.data_ptr(), mean.data_ptr(), ns, np_, dd.data_ptr())
    ref = (O.astype(np.float64) * d).astype(dtype)
    assert np.array_equal(Od.cpu().numpy(), ref)
    assert np.max(np.abs(mean.cpu().numpy() - m0 * d)) <= 1e-15 * np.max(np.abs(m0 * d))


@pytest.mark.parametrize("kind", ["spring", "march", "adam"])
@pytest.mark.parametrize("rtol", [1e-3, 1e-8])
def test_momentum_optimizer_matches_oracle(kind, rtol, rng):
    """Four steps on fresh (O, E) each (O is scaled in place by MARCH and
    AdamSR); rtol 1e-3 takes the partial-fraction route, 1e-8 dsyevd."""
    ns, np_ = 96, 400
    state = SimpleNamespace(nparams=np_)
    solver = qtx.optimizer.auto_pinv_eig(rtol=rtol, gram="fp64")
    dev = {"spring": qtx.optimizer.SPRING, "march": qtx.optimizer.MARCH, "adam": qtx.optimizer.AdamSR}[kind](
        state, None, solver=solver)
    ref_solver = lambda A, b: osr.auto_pinv_eig(A, b, rtol=rtol)
    ref = {"spring": osr.Spring(np_, ref_solver), "march": osr.March(np_, ref_solver),
           "adam": osr.AdamSR(np_, ref_solver)}[kind]
    for _ in range(4):
        O = rng.standard_normal((ns, np_)) * 0.2 + 0.05
        O[:, : np_ // 3] *= 0.1
        E = rng.standard_normal(ns) - 2.0
        Obar, _ = osr.get_Obar(O)
        Ebar, _, _ = osr.get_Ebar(E)
        want = ref.solve(Obar, Ebar)
        Od = _dev(O)
        mean = _mean(Od)
        eb = torch.empty(ns, dtype=torch.float64, device=DEV)
        o2 = torch.empty(2, dtype=torch.float64, device=DEV)
        Ed = _dev(E)
        _lib.call("qtx_ebar", _lib.stream(), Ed.data_ptr(), None, ns, eb.data_ptr(), o2.data_ptr())
        got = dev.solve(Od, mean, eb).cpu().numpy()
        assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-9


def test_spring_save_load_roundtrip(tmp_path, rng):
    np_ = 50
    state = SimpleNamespace(nparams=np_)
    opt = qtx.optimizer.MARCH(state, None, solver=qtx.optimizer.auto_pinv_eig(rtol=1e-3, gram="fp64"))
    opt._last_step = _dev(rng.standard_normal(np_))
    opt._V = _dev(rng.uniform(0, 1, np_))
    opt._t = 7
    f = tmp_path / "march.npz"
    opt.save(str(f))
    new = qtx.optimizer.MARCH(state, None, file=str(f))
    assert new._t == 7 and torch.equal(new._V, opt._V) and torch.equal(new._last_step, opt._last_step)


def _tfim_run(opt_cls, lr, steps, **kw):
    qtx.set_random_seed(0)
    qtx.sites.Chain(16)
    H = qtx.operator.Ising(h=1.0)
    model = qtx.model.RBM_Dense(16, dtype=np.float64, init_std=0.01, seed=0)
    state = qtx.state.Variational(model)
    sampler = qtx.sampler.LocalFlip(state, 1024, seed=1)
    opt = opt_cls(state, H, solver=qtx.optimizer.lstsq_shift_eig(rshift=0.0, ashift=1e-3), **kw)
    drv = qtx.optimizer.Euler(state, sampler, opt, lr)
    for _ in range(steps):
        drv.step()
    return drv.energy.mean(20)


def test_tfim16_spring_beats_sr_at_equal_step():
    """TFIM L=16, h=1, RBM(16) float64: at the same small learning rate the
    momentum of SPRING (mu 0.9) gets closer to the exact energy than SR."""
    e0 = -2.0 / np.sin(np.pi / 32)
    e_sr = _tfim_run(qtx.optimizer.SR, 2e-3, 150)
    e_sp = _tfim_run(qtx.optimizer.SPRING, 2e-3, 150, mu=0.9)
    assert e_sp < e_sr, (e_sp, e_sr)
    assert abs(e_sp - e0) / abs(e0) < 0.05, e_sp
