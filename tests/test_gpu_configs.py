"""The kernels each benchmark configuration runs, at that configuration's
shapes, against the CPU oracle (needs an MI355X).

* 10 x 10 Heisenberg with Marshall sign, Sz = 0, RBM_Dense M = 400 (register
  class 8, the 40.4k-parameter model) and M = 100 (class 2): log psi, E_loc,
  O, SpinExchange chains on the same Philox words, and the 4096 x 40400
  CHOL_SHIFT MinSR solve (diag shift 1e-4) against scipy's Cholesky solve.
* 20 x 20 J1-J2 ResConv with 32 channels: log psi (2 and 16 blocks), E_loc
  (J1-J2 and transverse-field Ising) and Metropolis, in FP32 and split-bf16;
  E_loc at 6 x 6 with 32 channels.
* 16 x 16 J1-J2 ResConv with 8 blocks and 16 channels: E_loc and O.
* Served RBM jobs in register classes 1, 2, 8 and 16 through the grouped
  launches: chains, E_loc and O against the oracle, and each job's result
  the same alone and in the mix.

Tolerances as in test_gpu_parity: FP32 network paths 1e-5 relative (E_loc and
log psi with an absolute floor of 1), split-bf16 paths 1e-4, FP64 solves 1e-9
(the shifted T has condition number ~7e4 here, so summation-order differences
of ~1e-16 grow to ~1e-11), chains identical except near-threshold decisions.
"""
import numpy as np
import pytest
import torch

from oracle import models as om
from oracle import operator as oop
from oracle import philox
from oracle import sampler as osm
from oracle import sr as osr
from tests.helpers import max_rel, rand_spins

pytestmark = pytest.mark.gpu

import quantax_amd as qtx  # noqa: E402
from quantax_amd import _lib  # noqa: E402
from quantax_amd import serve  # noqa: E402

DEV = "cuda"


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


# ------------------------------------------------------------------ 10 x 10 RBM (c2)
def _rbm_10x10(M, std=0.05, seed=None):
    lat = qtx.sites.Square(10, Nparticles=(50, 50))
    model = qtx.model.RBM_Dense(M, dtype=np.float32, init_std=std, seed=M if seed is None else seed)
    W, b = om.rbm_unpack(model.theta.cpu().numpy().astype(np.float64), M, 100)
    return lat, model, W, b


@pytest.mark.parametrize("M", [400, 100])
def test_c2_rbm_logpsi_eloc_jacobian(M, rng):
    lat, model, W, b = _rbm_10x10(M)
    s = rand_spins(rng, 64, 100, nup=50)
    sd = to_dev(s)
    fn = lambda x: om.rbm_logpsi(W, b, x)
    assert max_rel(model.logpsi(sd).cpu().numpy(), fn(s)[1]) < 1e-5
    H = qtx.operator.Heisenberg(msr=True)
    Ho = oop.heisenberg([10, 10], 1.0, 1, True)
    e, nconn = model.eloc(H, sd, None)
    assert max_rel(e.cpu().numpy(), oop.oloc(s, fn, Ho)) < 1e-5
    seg, _, _ = oop.get_conn(*list(oop.apply_off_diag(s, Ho).values())[0])
    assert np.array_equal(nconn.cpu().numpy(), np.bincount(seg, minlength=64))
    O = model.jacobian(sd).cpu().numpy()
    assert O.shape == (64, M * 100 + M)
    assert np.max(np.abs(O - om.rbm_jacobian(W, b, s))) < 2e-6


@pytest.mark.parametrize("M", [400, 100])
def test_c2_exchange_chains(M, rng):
    """A full 2N-step sweep (200 SpinExchange proposals) of 64 chains."""
    lat, model, W, b = _rbm_10x10(M)
    ns, nsteps, tie = 64, 200, 1e-4
    s0 = rand_spins(rng, ns, 100, nup=50)
    words = philox.sweep_words(31, nsteps, ns)
    nbr_np = lat.neighbor_table(1)
    s = to_dev(s0)
    la = model.logpsi(s)
    nacc = torch.zeros(ns, dtype=torch.int32, device=DEV)
    nnear = torch.zeros(ns, dtype=torch.int32, device=DEV)
    model.metropolis(_lib.QTX_EXCHANGE, to_dev(nbr_np), nbr_np.shape[1], 1, to_dev(words.view(np.int32)), nsteps, s,
                     la, nacc, nnear, 2.0, tie)
    s_ref, la_ref, nacc_ref, near_ref = osm.metropolis(lambda x: om.rbm_logpsi(W, b, x), s0, words, osm.EXCHANGE,
                                                       nbr_np, 1, 2.0, tie)
    clean = (near_ref == 0) & (nnear.cpu().numpy() == 0)
    assert clean.sum() >= ns - 4
    assert np.array_equal(s.cpu().numpy()[clean], s_ref[clean])
    assert np.array_equal(nacc.cpu().numpy()[clean], nacc_ref[clean])
    assert max_rel(la.cpu().numpy()[clean], la_ref[clean]) < 1e-5
    assert 0 < nacc_ref.sum() < ns * nsteps


@pytest.mark.parametrize("gram,tol", [("fp64", 1e-9), ("bf16x3", 1e-4)])
def test_c2_minsr_chol_shift_4096x40400(gram, tol, rng):
    """MinSR with T = Obar Obar^T of 4096 x 4096 (Np = 40400), diag shift
    1e-4, on the device's own O and E_loc of 4096 Sz = 0 samples: FP64 Gram
    (the plain path) against scipy cho_solve to 1e-9, split-bf16 Gram to the
    1e-4 update-vector bound of the mixed path."""
    lat, model, W, b = _rbm_10x10(400, std=0.01, seed=0)
    ns = 4096
    s = rand_spins(rng, ns, 100, nup=50)
    sd = to_dev(s)
    H = qtx.operator.Heisenberg(msr=True)
    e, _ = model.eloc(H, sd, None)
    O = model.jacobian(sd)
    mean = torch.empty(O.shape[1], dtype=torch.float64, device=DEV)
    _lib.call("qtx_colmean", _lib.stream(), _lib.QTX_F32, O.data_ptr(), ns, O.shape[1], mean.data_ptr())
    eb = torch.empty(ns, dtype=torch.float64, device=DEV)
    out2 = torch.empty(2, dtype=torch.float64, device=DEV)
    _lib.call("qtx_ebar", _lib.stream(), e.data_ptr(), None, ns, eb.data_ptr(), out2.data_ptr())
    solver = qtx.optimizer.minnorm_shift_eig(rshift=0.0, ashift=1e-4, gram=gram)
    step = qtx.optimizer.solve_step(O, mean, eb, solver).cpu().numpy()
    assert qtx.optimizer.last_solve_info()["path"] == "chol_shift"
    On = O.cpu().numpy().astype(np.float64)
    en = e.cpu().numpy()
    Obar, Omean = osr.get_Obar(On)
    Ebar, E, VarE = osr.get_Ebar(en)
    assert np.max(np.abs(mean.cpu().numpy() - Omean)) < 1e-12
    assert abs(out2[0].item() - E) < 1e-12 * abs(E)
    ref = osr.minnorm_shift_eig(Obar, Ebar, rshift=0.0, ashift=1e-4)
    err = np.linalg.norm(step - ref) / np.linalg.norm(ref)
    assert err < tol, err


# ------------------------------------------------------------------ ResConv at the c3 / c4 shapes
_ORACLE = {}


def _resconv(L, nb, C, precision, seed=11, nup=None, bias=0.1):
    lat = qtx.sites.Square(L, Nparticles=None if nup is None else (nup, L * L - nup))
    model = qtx.model.ResConv(nb, C, precision=precision, seed=seed)
    th = model.theta.cpu().numpy().astype(np.float64)
    r = np.random.default_rng(seed + 1)
    spec = om.ResConvSpec(nb, C, (L, L))
    for (cin, cout, hb), (w, bo) in zip(spec.layers, spec.offsets):
        if hb:
            th[bo: bo + cout] = bias * r.standard_normal(cout)
    model.set_params(th.astype(np.float32))
    return lat, model, spec, model.theta.cpu().numpy().astype(np.float64)


def _ham(L, kind):
    if kind == "j1j2":
        return (qtx.operator.Heisenberg(J=[1, 0.5], n_neighbor=[1, 2], msr=True),
                oop.heisenberg([L, L], [1, 0.5], [1, 2], True))
    return qtx.operator.Ising(h=0.7, J=1.3), oop.ising([L, L], h=0.7, J=1.3)


def _oracle_eloc(key, s, spec, theta, Ho):
    """The oracle's E_loc, computed once per (shape, seed, operator) and
    shared by the FP32 and split-bf16 runs (same parameters, same spins)."""
    if key not in _ORACLE:
        _ORACLE[key] = oop.oloc(s, lambda x: om.resconv_forward(spec, theta, x), Ho)
    return _ORACLE[key]


@pytest.mark.parametrize("nb", [2, 16])
@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16x3", 1e-4)])
def test_c4_resconv_logpsi_20x20(nb, precision, tol):
    rng = np.random.default_rng(100 + nb)
    lat, model, spec, theta = _resconv(20, nb, 32, precision)
    s = rand_spins(rng, 6, 400)
    la = model.logpsi(to_dev(s)).cpu().numpy()
    assert max_rel(la, om.resconv_forward(spec, theta, s)[1]) < tol


@pytest.mark.parametrize("L,nb,C,ns", [(20, 2, 32, 2), (6, 2, 32, 12), (16, 8, 16, 2)])
@pytest.mark.parametrize("ham", ["j1j2", "ising"])
@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16x3", 1e-4)])
def test_resconv_eloc_config_shapes(L, nb, C, ns, ham, precision, tol):
    """E_loc of the 20 x 20, 32-channel kernels (c4), of the 32-channel
    kernels on a lattice of three 16-site tiles, and of the full 8-block
    16 x 16 window kernel (c3)."""
    if (L, nb) == (16, 8) and ham == "ising":
        pytest.skip("the 16 x 16 window kernel's single-flip path is covered in test_gpu_parity")
    rng = np.random.default_rng(7 * L + nb)
    nup = L * L // 2 if ham == "j1j2" else None
    lat, model, spec, theta = _resconv(L, nb, C, precision, nup=nup)
    H, Ho = _ham(L, ham)
    s = rand_spins(rng, ns, L * L, nup=nup)
    la = model.logpsi(to_dev(s))
    e, nconn = model.eloc(H, to_dev(s), la)
    ref = _oracle_eloc((L, nb, C, ham), s, spec, theta, Ho)
    assert max_rel(e.cpu().numpy(), ref) < tol
    seg = np.concatenate([oop.get_conn(sc, Hc)[0] for sc, Hc in oop.apply_off_diag(s, Ho).values()])
    assert np.array_equal(nconn.cpu().numpy(), np.bincount(seg, minlength=ns))


@pytest.mark.parametrize("precision,tie,lptol", [("fp32", 1e-4, 1e-5), ("bf16x3", 1e-3, 1e-4)])
def test_c4_resconv_metropolis_20x20(precision, tie, lptol):
    rng = np.random.default_rng(5)
    lat, model, spec, theta = _resconv(20, 2, 32, precision, nup=200)
    ns, nsteps = 4, 6
    s0 = rand_spins(rng, ns, 400, nup=200)
    words = philox.sweep_words(23, nsteps, ns)
    nbr_np = lat.neighbor_table(1)
    s = to_dev(s0)
    la = model.logpsi(s)
    nacc = torch.zeros(ns, dtype=torch.int32, device=DEV)
    nnear = torch.zeros(ns, dtype=torch.int32, device=DEV)
    model.metropolis(_lib.QTX_EXCHANGE, to_dev(nbr_np), nbr_np.shape[1], 1, to_dev(words.view(np.int32)), nsteps, s,
                     la, nacc, nnear, 2.0, tie)
    s_ref, la_ref, nacc_ref, near_ref = osm.metropolis(lambda x: om.resconv_forward(spec, theta, x), s0, words,
                                                       osm.EXCHANGE, nbr_np, 1, 2.0, tie)
    clean = (near_ref == 0) & (nnear.cpu().numpy() == 0)
    assert clean.sum() >= ns - 1
    assert np.array_equal(s.cpu().numpy()[clean], s_ref[clean])
    assert np.array_equal(nacc.cpu().numpy()[clean], nacc_ref[clean])
    assert max_rel(la.cpu().numpy()[clean], la_ref[clean]) < lptol
    assert nacc_ref.sum() > 0


def test_c3_resconv_jacobian_8_blocks():
    rng = np.random.default_rng(8)
    lat, model, spec, theta = _resconv(16, 8, 16, "fp32")
    s = rand_spins(rng, 2, 256)
    O = model.jacobian(to_dev(s)).cpu().numpy()
    ref = om.resconv_jacobian(spec, theta, s)
    assert np.max(np.abs(O - ref)) < 1e-5 * max(1.0, np.max(np.abs(ref)))


# ------------------------------------------------------------------ served jobs (c5)
def _specs_big():
    # register classes R = ceil(M / 64) rounded to a power of two:
    # L=10 a=4 (M=400) -> 8, L=12 a=4 (M=576) -> 16, L=8 a=2 (M=128) -> 2, L=6 a=1 (M=36) -> 1
    return [
        serve.JobSpec(job_id=0, L=10, J2=0.5, alpha=4, ns=64, nsteps=2, seed=111, init_std=0.05),
        serve.JobSpec(job_id=1, L=12, J2=0.2, alpha=4, ns=48, nsteps=2, seed=222, init_std=0.05),
        serve.JobSpec(job_id=2, L=8, J2=0.4, alpha=2, ns=80, nsteps=2, seed=333, init_std=0.05),
        serve.JobSpec(job_id=3, L=6, J2=0.0, alpha=1, ns=96, nsteps=2, seed=444, init_std=0.05),
    ]


def test_c5_grouped_launch_against_oracle():
    """One grouped sweep (2N SpinExchange steps per job, words generated in
    the kernel from each job's seed), then grouped E_loc and Jacobian, for
    jobs of four register classes at once; each job against the oracle."""
    jobs = [serve.JobRun(sp, gram="fp64") for sp in _specs_big()]
    s0 = [j.s.cpu().numpy() for j in jobs]
    nst = [2 * j.spec.N for j in jobs]
    tie = 1e-4  # above the FP32 error of the local-update log ratios
    arr = (_lib.QtxRbmJob * len(jobs))()
    for i, (j, n) in enumerate(zip(jobs, nst)):
        arr[i] = j.job_struct(n)
        arr[i].tie_tol = tie
    nws = int(_lib.lib().qtx_rbm_group_ws_bytes(len(jobs)))
    ws = torch.empty(nws, dtype=torch.uint8, device=DEV)
    for name in ("qtx_rbm_group_metropolis", "qtx_rbm_group_eloc", "qtx_rbm_group_jacobian"):
        _lib.call(name, _lib.stream(), arr, len(jobs), ws.data_ptr(), nws)
    torch.cuda.synchronize()
    for j, s_init, n in zip(jobs, s0, nst):
        sp = j.spec
        W, b = om.rbm_unpack(j.model.theta.cpu().numpy().astype(np.float64), sp.M, sp.N)
        fn = lambda x, W=W, b=b: om.rbm_logpsi(W, b, x)
        lat = qtx.sites.Square(sp.L, Nparticles=(sp.N // 2, sp.N - sp.N // 2))
        words = philox.sweep_words(sp.seed, n, sp.ns, step0=0)
        s_ref, la_ref, nacc_ref, near_ref = osm.metropolis(fn, s_init, words, osm.EXCHANGE, lat.neighbor_table(1),
                                                           1, 2.0, tie)
        s_dev = j.s.cpu().numpy()
        clean = (near_ref == 0) & (j.n_near.cpu().numpy() == 0)
        assert clean.sum() >= sp.ns - 2, sp
        assert np.array_equal(s_dev[clean], s_ref[clean]), sp
        assert np.array_equal(j.n_accept.cpu().numpy()[clean], nacc_ref[clean]), sp
        assert max_rel(j.logabs.cpu().numpy()[clean], la_ref[clean]) < 1e-5, sp
        Ho = oop.heisenberg([sp.L, sp.L], [1.0, sp.J2], [1, 2], True)
        assert max_rel(j.eloc.cpu().numpy(), oop.oloc(s_dev, fn, Ho)) < 1e-5, sp
        assert np.max(np.abs(j.O.cpu().numpy() - om.rbm_jacobian(W, b, s_dev))) < 2e-6, sp


def test_c5_alone_equals_mix_large_classes():
    """Jobs of register classes 1, 2, 8 and 16 (the 16 x 16 alpha = 4 job is
    class 16 as well): each job's energies and parameters are the same alone
    and in the mix, one arriving while the others run."""
    specs = _specs_big()
    specs.append(serve.JobSpec(job_id=4, L=16, J2=0.6, alpha=4, ns=32, nsteps=2, seed=555, arrival=0.2))
    srv = serve.Server(keep_results=True)
    try:
        mix = srv.run(specs)
    finally:
        srv.close()
    assert mix.total_steps == sum(s.nsteps for s in specs)
    for sp in specs:
        one = serve.Server(keep_results=True)
        try:
            alone = one.run([serve.JobSpec(**{**sp.__dict__, "arrival": 0.0})])
        finally:
            one.close()
        e1, e2 = alone.results[sp.job_id]["energy"], mix.results[sp.job_id]["energy"]
        assert np.all(np.isfinite(e1))
        assert np.allclose(e1, e2, rtol=1e-10, atol=0), (sp, e1, e2)
        assert np.allclose(alone.results[sp.job_id]["theta"], mix.results[sp.job_id]["theta"], rtol=1e-5, atol=1e-7)
