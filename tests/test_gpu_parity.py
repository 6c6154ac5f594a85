"""HIP path vs the CPU oracle on the same seeded inputs (needs an MI355X).

Tolerances: float32 network paths 1e-5 relative, float64 paths
1e-10 relative; integer/byte work (streams, initial chains, connected
configurations, chains of the float64 RBM) bit-exact.  Relative errors of
log-amplitudes and local energies use an absolute floor of 1 so that values
crossing zero are not over-weighted.
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

DEV = "cuda"


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


# ------------------------------------------------------------------ streams
def test_philox_fill_bit_exact():
    words = torch.empty((7, 33, 4), dtype=torch.int32, device=DEV)
    _lib.call("qtx_philox_fill", _lib.stream(), 123456789123, _lib.TAG_SWEEP, 5, 7, 0, 33, words.data_ptr())
    got = words.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, philox.sweep_words(123456789123, 7, 33, step0=5))


@pytest.mark.parametrize("N,nup", [(16, None), (100, 50), (37, 11), (300, None)])
def test_init_spins_bit_exact(N, nup):
    s = torch.empty((65, N), dtype=torch.int8, device=DEV)
    _lib.call("qtx_init_spins", _lib.stream(), 987654321, 65, N, -1 if nup is None else nup, s.data_ptr())
    got = s.cpu().numpy()
    ref = philox.init_spins(987654321, 65, N, nup)
    assert np.array_equal(got, ref)
    if nup is not None:
        assert np.all(got.sum(1) == 2 * nup - N)


# ------------------------------------------------------------------ operator
@pytest.mark.parametrize("case", ["j1j2_4x4", "heis_6x6", "ising_chain16", "heis_chain10"])
def test_connected_configs_bit_exact(case, rng):
    if case == "j1j2_4x4":
        qtx.sites.Square(4)
        H = qtx.operator.Heisenberg(J=[1, 0.5], n_neighbor=[1, 2], msr=True)
        Ho = oop.heisenberg([4, 4], [1, 0.5], [1, 2], True)
    elif case == "heis_6x6":
        qtx.sites.Square(6)
        H = qtx.operator.Heisenberg()
        Ho = oop.heisenberg([6, 6])
    elif case == "ising_chain16":
        qtx.sites.Chain(16)
        H = qtx.operator.Ising(h=0.7, J=1.3)
        Ho = oop.ising([16], h=0.7, J=1.3)
    else:
        qtx.sites.Chain(10)
        H = qtx.operator.Heisenberg(msr=True)
        Ho = oop.heisenberg([10], msr=True)
    N = qtx.get_sites().Nsites
    s = rand_spins(rng, 50, N)
    diag = H.apply_diag(to_dev(s)).cpu().numpy()
    # FP64 sums of the diagonal terms: exact for dyadic couplings, otherwise
    # equal up to the summation order
    assert np.allclose(diag, oop.apply_diag(s, Ho), rtol=0, atol=1e-12)
    seg, cand, sc, hc = (x.cpu().numpy() for x in H.connected(to_dev(s)))
    offd = oop.apply_off_diag(s, Ho)
    assert len(offd) == 1
    (sc_ref, H_ref), = offd.values()
    seg_ref, sconn_ref, h_ref = oop.get_conn(sc_ref, H_ref)
    assert np.array_equal(seg, seg_ref)
    assert np.array_equal(sc, sconn_ref)
    assert np.array_equal(hc, h_ref)


# ------------------------------------------------------------------ RBM
def _rbm(L, M, dtype, seed=3, nup=None, chain=False):
    lat = qtx.sites.Chain(L) if chain else qtx.sites.Square(L, Nparticles=None if nup is None else (nup, L * L - nup))
    model = qtx.model.RBM_Dense(M, dtype=dtype, init_std=0.15, seed=seed)
    theta = model.theta.cpu().numpy().astype(np.float64)
    W, b = om.rbm_unpack(theta, M, lat.Nsites)
    return lat, model, W, b


@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-5), (np.float64, 1e-10)])
@pytest.mark.parametrize("M", [12, 64, 200, 700])
def test_rbm_logpsi(dtype, tol, M, rng):
    lat, model, W, b = _rbm(6, M, dtype)
    s = rand_spins(rng, 40, lat.Nsites)
    la = model.logpsi(to_dev(s)).cpu().numpy()
    assert max_rel(la, om.rbm_logpsi(W, b, s)[1]) < tol


@pytest.mark.parametrize("dtype,tol", [(np.float32, 2e-6), (np.float64, 1e-12)])
def test_rbm_jacobian(dtype, tol, rng):
    lat, model, W, b = _rbm(4, 24, dtype)
    s = rand_spins(rng, 30, lat.Nsites)
    O = model.jacobian(to_dev(s)).cpu().numpy()
    assert O.shape == (30, 24 * 16 + 24)
    assert np.max(np.abs(O - om.rbm_jacobian(W, b, s))) < tol


@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-5), (np.float64, 1e-10)])
def test_rbm_eloc_j1j2(dtype, tol, rng):
    lat, model, W, b = _rbm(4, 32, dtype)
    H = qtx.operator.Heisenberg(J=[1, 0.5], n_neighbor=[1, 2], msr=True)
    Ho = oop.heisenberg([4, 4], [1, 0.5], [1, 2], True)
    s = rand_spins(rng, 64, 16, nup=8)
    e, nconn = model.eloc(H, to_dev(s), None)
    ref = oop.oloc(s, lambda x: om.rbm_logpsi(W, b, x), Ho)
    assert max_rel(e.cpu().numpy(), ref) < tol
    seg, _, _ = oop.get_conn(*list(oop.apply_off_diag(s, Ho).values())[0])
    assert np.array_equal(nconn.cpu().numpy(), np.bincount(seg, minlength=64))


def test_rbm_eloc_ising_f64(rng):
    lat, model, W, b = _rbm(16, 16, np.float64, chain=True)
    H = qtx.operator.Ising(h=1.0)
    Ho = oop.ising([16], h=1.0)
    s = rand_spins(rng, 128, 16)
    e, _ = model.eloc(H, to_dev(s), None)
    ref = oop.oloc(s, lambda x: om.rbm_logpsi(W, b, x), Ho)
    assert max_rel(e.cpu().numpy(), ref) < 1e-10


@pytest.mark.parametrize("kind", ["flip", "exchange"])
def test_rbm_metropolis_chains_f64_bit_identical(kind, rng):
    nup = 18 if kind == "exchange" else None
    lat, model, W, b = _rbm(6, 40, np.float64, nup=nup)
    N = 36
    ns, nsteps = 96, 3 * N
    s0 = rand_spins(rng, ns, N, nup)
    words = philox.sweep_words(77, nsteps, ns)
    if kind == "flip":
        kd, nbr, width, hop = _lib.QTX_LOCALFLIP, None, 0, 1
        nbr_np = None
    else:
        kd = _lib.QTX_EXCHANGE
        nbr_np = lat.neighbor_table(1)
        nbr, width, hop = to_dev(nbr_np), nbr_np.shape[1], 1
    s = to_dev(s0)
    la = model.logpsi(s)
    nacc = torch.zeros(ns, dtype=torch.int32, device=DEV)
    nnear = torch.zeros(ns, dtype=torch.int32, device=DEV)
    model.metropolis(kd, nbr, width, hop, to_dev(words.view(np.int32)), nsteps, s, la, nacc, nnear, 2.0, 1e-9)
    s_ref, la_ref, nacc_ref, near_ref = osm.metropolis(
        lambda x: om.rbm_logpsi(W, b, x), s0, words, osm.LOCALFLIP if kind == "flip" else osm.EXCHANGE,
        nbr_np, 1, 2.0, 1e-9)
    assert near_ref.sum() == 0
    assert np.array_equal(s.cpu().numpy(), s_ref)
    assert np.array_equal(nacc.cpu().numpy(), nacc_ref)
    assert max_rel(la.cpu().numpy(), la_ref) < 1e-10
    assert 0 < nacc_ref.sum() < ns * nsteps


def test_rbm_metropolis_f32_identical_outside_ties(rng):
    lat, model, W, b = _rbm(6, 40, np.float32)
    N, ns, nsteps = 36, 128, 72
    s0 = rand_spins(rng, ns, N)
    words = philox.sweep_words(5, nsteps, ns)
    s = to_dev(s0)
    la = model.logpsi(s)
    nacc = torch.zeros(ns, dtype=torch.int32, device=DEV)
    nnear = torch.zeros(ns, dtype=torch.int32, device=DEV)
    tie = 1e-4
    model.metropolis(_lib.QTX_LOCALFLIP, None, 0, 1, to_dev(words.view(np.int32)), nsteps, s, la, nacc, nnear, 2.0,
                     tie)
    s_ref, la_ref, _, near_ref = osm.metropolis(lambda x: om.rbm_logpsi(W, b, x), s0, words, osm.LOCALFLIP,
                                                tie_tol=tie)
    clean = (near_ref == 0) & (nnear.cpu().numpy() == 0)
    assert clean.sum() >= ns - 4
    assert np.array_equal(s.cpu().numpy()[clean], s_ref[clean])
    assert max_rel(la.cpu().numpy()[clean], la_ref[clean]) < 1e-5


# ------------------------------------------------------------------ ResConv
def _resconv(L, nb, C, precision="fp32", seed=11, gelu="tanh", nup=None):
    lat = qtx.sites.Square(L, Nparticles=None if nup is None else (nup, L * L - nup))
    model = qtx.model.ResConv(nb, C, precision=precision, gelu=gelu, seed=seed)
    # random non-zero biases so every parameter block is exercised
    th = model.theta.cpu().numpy().astype(np.float64)
    r = np.random.default_rng(seed + 1)
    spec = om.ResConvSpec(nb, C, (L, L))
    for (cin, cout, hb), (w, bo) in zip(spec.layers, spec.offsets):
        if hb:
            th[bo: bo + cout] = 0.1 * r.standard_normal(cout)
    model.set_params(th.astype(np.float32))
    theta = model.theta.cpu().numpy().astype(np.float64)
    return lat, model, spec, theta


@pytest.mark.parametrize("L,nb,C", [(4, 2, 16), (6, 3, 16), (16, 8, 16), (10, 2, 16), (20, 2, 16), (6, 2, 32),
                                    (20, 3, 32)])
def test_resconv_logpsi_fp32(L, nb, C, rng):
    lat, model, spec, theta = _resconv(L, nb, C)
    s = rand_spins(rng, 9, L * L)
    la = model.logpsi(to_dev(s)).cpu().numpy()
    assert max_rel(la, om.resconv_forward(spec, theta, s)[1]) < 1e-5


def test_resconv_logpsi_gelu_erf(rng):
    lat, model, spec, theta = _resconv(6, 2, 16, gelu="erf")
    s = rand_spins(rng, 9, 36)
    la = model.logpsi(to_dev(s)).cpu().numpy()
    assert max_rel(la, om.resconv_forward(spec, theta, s, approximate=False)[1]) < 1e-5


@pytest.mark.parametrize("L,nb,C", [(4, 2, 16), (16, 8, 16), (6, 2, 32)])
def test_resconv_logpsi_bf16x3(L, nb, C, rng):
    lat, model, spec, theta = _resconv(L, nb, C, precision="bf16x3")
    s = rand_spins(rng, 9, L * L)
    la = model.logpsi(to_dev(s)).cpu().numpy()
    assert max_rel(la, om.resconv_forward(spec, theta, s)[1]) < 1e-4


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16x3", 1e-4)])
def test_resconv_eloc_j1j2(precision, tol, rng):
    lat, model, spec, theta = _resconv(4, 2, 16, precision=precision, nup=8)
    H = qtx.operator.Heisenberg(J=[1, 0.5], n_neighbor=[1, 2], msr=True)
    Ho = oop.heisenberg([4, 4], [1, 0.5], [1, 2], True)
    s = rand_spins(rng, 24, 16, nup=8)
    la = model.logpsi(to_dev(s))
    e, _ = model.eloc(H, to_dev(s), la)
    ref = oop.oloc(s, lambda x: om.resconv_forward(spec, theta, x), Ho)
    assert max_rel(e.cpu().numpy(), ref) < tol


@pytest.mark.parametrize("L,nb,C,gelu", [(6, 2, 16, "tanh"), (16, 3, 16, "tanh"), (5, 2, 32, "tanh"),
                                         (20, 2, 32, "tanh"), (6, 2, 16, "erf"), (3, 2, 16, "tanh")])
def test_resconv_jacobian(L, nb, C, gelu, rng):
    """Per-sample gradients (FP32 MFMA GEMMs) against the float64 oracle;
    5 x 5 and 3 x 3 have partial 16-site tiles and a site count that is not a
    multiple of 4."""
    lat, model, spec, theta = _resconv(L, nb, C, gelu=gelu)
    s = rand_spins(rng, 5, L * L)
    O = model.jacobian(to_dev(s)).cpu().numpy()
    ref = om.resconv_jacobian(spec, theta, s, approximate=(gelu == "tanh"))
    assert np.max(np.abs(O - ref)) < 1e-5 * max(1.0, np.max(np.abs(ref)))


def test_resconv_metropolis_identical_outside_ties(rng):
    lat, model, spec, theta = _resconv(4, 2, 16, nup=8)
    N, ns, nsteps = 16, 16, 24
    s0 = rand_spins(rng, ns, N, nup=8)
    words = philox.sweep_words(9, nsteps, ns)
    nbr_np = lat.neighbor_table(1)
    s = to_dev(s0)
    la = model.logpsi(s)
    nacc = torch.zeros(ns, dtype=torch.int32, device=DEV)
    nnear = torch.zeros(ns, dtype=torch.int32, device=DEV)
    tie = 1e-4
    model.metropolis(_lib.QTX_EXCHANGE, to_dev(nbr_np), nbr_np.shape[1], 1, to_dev(words.view(np.int32)), nsteps, s,
                     la, nacc, nnear, 2.0, tie)
    s_ref, la_ref, nacc_ref, near_ref = osm.metropolis(lambda x: om.resconv_forward(spec, theta, x), s0, words,
                                                       osm.EXCHANGE, nbr_np, 1, 2.0, tie)
    clean = (near_ref == 0) & (nnear.cpu().numpy() == 0)
    assert clean.sum() >= ns - 2
    assert np.array_equal(s.cpu().numpy()[clean], s_ref[clean])
    assert max_rel(la.cpu().numpy()[clean], la_ref[clean]) < 1e-5


# ------------------------------------------------------------------ SR / MinSR
def _solver_case(ns, np_, dtype, rng):
    O = rng.standard_normal((ns, np_)) * 0.5 + 0.3
    O = O.astype(dtype)
    E = rng.standard_normal(ns) - 3.0
    return O, E


@pytest.mark.parametrize("gram,tol", [("fp64", 1e-10), ("fp32", 2e-4), ("bf16x3", 2e-4)])
@pytest.mark.parametrize("mode", ["chol", "eigh"])
@pytest.mark.parametrize("shape", [(96, 300), (200, 1000), (300, 40)])
def test_minsr_solve(gram, tol, mode, shape, rng):
    ns, np_ = shape
    O, E = _solver_case(ns, np_, np.float32, rng)
    Od = to_dev(O)
    mean = torch.empty(np_, dtype=torch.float64, device=DEV)
    _lib.call("qtx_colmean", _lib.stream(), _lib.QTX_F32, Od.data_ptr(), ns, np_, mean.data_ptr())
    Ebar, Em, Ev = osr.get_Ebar(E)
    eb = torch.empty(ns, dtype=torch.float64, device=DEV)
    out2 = torch.empty(2, dtype=torch.float64, device=DEV)
    Ed = to_dev(E)  # referenced until the call is done
    _lib.call("qtx_ebar", _lib.stream(), Ed.data_ptr(), None, ns, eb.data_ptr(), out2.data_ptr())
    assert max_rel(eb.cpu().numpy(), Ebar) < 1e-12
    assert abs(out2[0].item() - Em) < 1e-12 * abs(Em) and abs(out2[1].item() - Ev) < 1e-12 * abs(Ev)
    Obar, Omean = osr.get_Obar(O.astype(np.float64))
    assert np.max(np.abs(mean.cpu().numpy() - Omean)) < 1e-12
    if mode == "chol":
        solver = qtx.optimizer.auto_shift_eig(rshift=0.0, ashift=1e-3, gram=gram)
        ref = osr.auto_shift_eig(Obar, Ebar, rshift=0.0, ashift=1e-3)
    else:
        solver = qtx.optimizer.auto_pinv_eig(rtol=1e-8, gram=gram)
        ref = osr.auto_pinv_eig(Obar, Ebar, rtol=1e-8)
    if ns >= np_:
        solver.gram = "fp64"  # the SR branch always forms S in FP64
    step = qtx.optimizer.solve_step(Od, mean, eb, solver).cpu().numpy()
    err = np.linalg.norm(step - ref) / np.linalg.norm(ref)
    assert err < (1e-10 if ns >= np_ else tol), err


def _solve_inputs(O, E):
    ns, np_ = O.shape
    Od = to_dev(O)
    mean = torch.empty(np_, dtype=torch.float64, device=DEV)
    _lib.call("qtx_colmean", _lib.stream(), _lib.QTX_F32, Od.data_ptr(), ns, np_, mean.data_ptr())
    eb = torch.empty(ns, dtype=torch.float64, device=DEV)
    out2 = torch.empty(2, dtype=torch.float64, device=DEV)
    Ed = to_dev(E)  # referenced until the call is done
    _lib.call("qtx_ebar", _lib.stream(), Ed.data_ptr(), None, ns, eb.data_ptr(), out2.data_ptr())
    return Od, mean, eb


@pytest.mark.parametrize("rtol", [1e-2, 1e-4, 1e-5])
@pytest.mark.parametrize("shape", [(200, 1000), (512, 3000), (300, 40)])
def test_pinv_partial_fractions(rtol, shape, rng):
    """The eigenvector-free soft inverse against the oracle's eigh form, and
    the Lanczos lambda_max against numpy's eigvalsh."""
    ns, np_ = shape
    O, E = _solver_case(ns, np_, np.float32, rng)
    O[:, : np_ // 2] *= 0.05  # spread the spectrum
    Od, mean, eb = _solve_inputs(O, E)
    Obar, _ = osr.get_Obar(O.astype(np.float64))
    Ebar, _, _ = osr.get_Ebar(E)
    ref = osr.auto_pinv_eig(Obar, Ebar, rtol=rtol)
    solver = qtx.optimizer.auto_pinv_eig(rtol=rtol, gram="fp64")
    step = qtx.optimizer.solve_step(Od, mean, eb, solver).cpu().numpy()
    info = qtx.optimizer.last_solve_info()
    assert info["path"] == "partial_fractions", info
    Tm = Obar @ Obar.T if ns < np_ else Obar.T @ Obar
    lmax = np.linalg.eigvalsh(Tm).max()
    assert abs(info["lambda_max"] - lmax) < 1e-12 * lmax, (info, lmax)
    err = np.linalg.norm(step - ref) / np.linalg.norm(ref)
    assert err < 1e-10, err
    # forced dsyevd on the same inputs agrees as well
    solver.mode = _lib.QTX_SOLVE_EIGH_PINV_DIRECT
    step2 = qtx.optimizer.solve_step(Od, mean, eb, solver).cpu().numpy()
    assert qtx.optimizer.last_solve_info()["path"] == "syevd"
    assert np.linalg.norm(step2 - ref) / np.linalg.norm(ref) < 1e-10


@pytest.mark.parametrize("shape,rtol", [((2000, 4000), 1e-4), ((1100, 1500), 1e-2)])
def test_pinv_partial_fractions_several_blocks(shape, rtol, rng):
    """ns above the 1024-wide diagonal blocks of the batched Cholesky (the
    last block partial): panel solves, strided-batched trsm and syrk (below
    n = 2048, where EIGH_PINV keeps the partial fractions)."""
    ns, np_ = shape
    O, E = _solver_case(ns, np_, np.float32, rng)
    O[:, : np_ // 2] *= 0.05
    Od, mean, eb = _solve_inputs(O, E)
    Obar, _ = osr.get_Obar(O.astype(np.float64))
    Ebar, _, _ = osr.get_Ebar(E)
    ref = osr.auto_pinv_eig(Obar, Ebar, rtol=rtol)
    step = qtx.optimizer.solve_step(Od, mean, eb, qtx.optimizer.auto_pinv_eig(rtol=rtol, gram="fp64")).cpu().numpy()
    assert qtx.optimizer.last_solve_info()["path"] == "partial_fractions"
    assert np.linalg.norm(step - ref) / np.linalg.norm(ref) < 1e-10


def test_chol_shift_several_blocks(rng):
    ns, np_ = 2100, 3000
    O, E = _solver_case(ns, np_, np.float32, rng)
    Od, mean, eb = _solve_inputs(O, E)
    Obar, _ = osr.get_Obar(O.astype(np.float64))
    Ebar, _, _ = osr.get_Ebar(E)
    ref = osr.auto_shift_eig(Obar, Ebar, rshift=0.0, ashift=1e-3)
    step = qtx.optimizer.solve_step(Od, mean, eb, qtx.optimizer.auto_shift_eig(rshift=0.0, ashift=1e-3)).cpu().numpy()
    assert qtx.optimizer.last_solve_info()["path"] == "chol_shift"
    assert np.linalg.norm(step - ref) / np.linalg.norm(ref) < 1e-10


def test_chol_shift_indefinite_reports_block(rng):
    """A negative shift larger than the smallest eigenvalue makes T + shift
    indefinite: the factorisation reports the diagonal block of the first
    non-positive pivot and the solve fails loudly."""
    O, E = _solver_case(1500, 3000, np.float32, rng)
    Od, mean, eb = _solve_inputs(O, E)
    with pytest.raises(_lib.QtxError, match="potrf failed in diagonal block"):
        qtx.optimizer.solve_step(Od, mean, eb, qtx.optimizer.auto_shift_eig(rshift=0.0, ashift=-1e3))


def test_pinv_small_cutoff_routes(rng):
    """Cutoffs below 1e-5 lambda_max (the reference default rtol = 1e-12):
    a well-conditioned T takes the shifted Cholesky route and equals the
    eigh soft inverse; duplicate samples (a second null direction of T, an
    eigenvalue below 150 c) make the shifted factorisation fail and the solve
    takes dsyevd; both against the oracle."""
    for dup in (False, True):
        O, E = _solver_case(128, 500, np.float32, rng)
        if dup:
            O[7] = O[3]
            E[7] = E[3]
        Od, mean, eb = _solve_inputs(O, E)
        for rtol in (1e-8, 1e-12):
            step = qtx.optimizer.solve_step(Od, mean, eb, qtx.optimizer.auto_pinv_eig(rtol=rtol)).cpu().numpy()
            info = qtx.optimizer.last_solve_info()
            Obar, _ = osr.get_Obar(O.astype(np.float64))
            Ebar, _, _ = osr.get_Ebar(E)
            ref = osr.auto_pinv_eig(Obar, Ebar, rtol=rtol)
            err = np.linalg.norm(step - ref) / np.linalg.norm(ref)
            print(f"duplicates {dup} rtol {rtol}: {info['path']} / {info['fallback']}, err {err:.2e}")
            if dup:
                assert info["path"] == "syevd" and info["fallback"] == "shifted_cholesky_failed", info
            else:
                assert info["path"] == "shifted_cholesky", info
            assert err < 1e-9, err


def test_pinv_zero_gram(rng):
    """Identical samples: Obar = 0, every eigenvalue is 0 and the step is 0."""
    O = np.tile(rng.standard_normal((1, 700)).astype(np.float32), (64, 1))
    E = rng.standard_normal(64)
    Od, mean, eb = _solve_inputs(O, E)
    step = qtx.optimizer.solve_step(Od, mean, eb, qtx.optimizer.auto_pinv_eig(rtol=1e-3)).cpu().numpy()
    assert np.all(step == 0.0)


def test_apply_update_skips_nonfinite():
    qtx.sites.Chain(8)
    model = qtx.model.RBM_Dense(4, dtype=np.float32, init_std=0.1, seed=1)
    state = qtx.state.Variational(model)
    th0 = state.get_params_flatten().cpu().numpy()
    step = np.linspace(-1, 1, state.nparams)
    assert state.update(torch.from_numpy(step))
    th1 = state.get_params_flatten().cpu().numpy()
    assert np.array_equal(th1, osr.update(th0, step))
    bad = step.copy()
    bad[3] = np.nan
    with pytest.warns(UserWarning):
        assert not state.update(torch.from_numpy(bad))
    assert np.array_equal(state.get_params_flatten().cpu().numpy(), th1)


# ------------------------------------------------------------------ row-tile kernels (16-wide lattices)
def _resconv_rect(L0, L1, nb, C, precision, seed=5, nup=None, gelu="tanh"):
    lat = qtx.sites.Grid([L0, L1], Nparticles=None if nup is None else (nup, L0 * L1 - nup))
    model = qtx.model.ResConv(nb, C, precision=precision, gelu=gelu, seed=seed)
    th = model.theta.cpu().numpy().astype(np.float64)
    r = np.random.default_rng(seed + 1)
    spec = om.ResConvSpec(nb, C, (L0, L1))
    for (cin, cout, hb), (w, bo) in zip(spec.layers, spec.offsets):
        if hb:
            th[bo: bo + cout] = 0.1 * r.standard_normal(cout)
    model.set_params(th.astype(np.float32))
    return lat, model, spec, model.theta.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16x3", 1e-4)])
@pytest.mark.parametrize("L0", [16, 12, 5])
def test_resconv_rows_logpsi(precision, tol, L0, rng):
    lat, model, spec, theta = _resconv_rect(L0, 16, 3, 16, precision)
    s = rand_spins(rng, 7, L0 * 16)
    la = model.logpsi(to_dev(s)).cpu().numpy()
    assert max_rel(la, om.resconv_forward(spec, theta, s)[1]) < tol


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16x3", 1e-4)])
def test_resconv_window_gelu_erf_16x16(precision, tol, rng):
    """The 16 x 16 window kernel with the exact (erf) GELU, both precisions."""
    lat, model, spec, theta = _resconv_rect(16, 16, 3, 16, precision, gelu="erf")
    s = rand_spins(rng, 7, 256)
    la = model.logpsi(to_dev(s)).cpu().numpy()
    assert max_rel(la, om.resconv_forward(spec, theta, s, approximate=False)[1]) < tol


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16x3", 1e-4)])
def test_resconv_rows_eloc_j1j2_16x16(precision, tol, rng):
    lat, model, spec, theta = _resconv_rect(16, 16, 2, 16, precision, nup=128)
    H = qtx.operator.Heisenberg(J=[1, 0.5], n_neighbor=[1, 2], msr=True)
    Ho = oop.heisenberg([16, 16], [1, 0.5], [1, 2], True)
    s = rand_spins(rng, 2, 256, nup=128)
    la = model.logpsi(to_dev(s))
    e, nconn = model.eloc(H, to_dev(s), la)
    ref = oop.oloc(s, lambda x: om.resconv_forward(spec, theta, x), Ho)
    assert max_rel(e.cpu().numpy(), ref) < tol


@pytest.mark.parametrize("L,precision,tol", [(4, "fp32", 1e-5), (4, "bf16x3", 1e-4), (16, "fp32", 1e-5),
                                              (16, "bf16x3", 1e-4)])
def test_resconv_eloc_ising_single_flips(L, precision, tol, rng):
    """Transverse-field Ising: every off-diagonal term flips one site, the
    case where the E_loc kernels apply and undo a single flip in place (4 x 4:
    tile kernel, 16 x 16: window kernel); every site is a connected
    configuration."""
    lat, model, spec, theta = _resconv_rect(L, L, 2, 16, precision)
    H = qtx.operator.Ising(h=0.7, J=1.3)
    Ho = oop.ising([L, L], h=0.7, J=1.3)
    s = rand_spins(rng, 3 if L == 16 else 16, L * L)
    la = model.logpsi(to_dev(s))
    e, nconn = model.eloc(H, to_dev(s), la)
    ref = oop.oloc(s, lambda x: om.resconv_forward(spec, theta, x), Ho)
    assert max_rel(e.cpu().numpy(), ref) < tol
    assert (nconn.cpu().numpy() == L * L).all()


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16x3", 1e-4)])
@pytest.mark.parametrize("L0,ham", [(12, "j1j2"), (5, "ising")])
def test_resconv_rows_eloc(L0, ham, precision, tol, rng):
    """E_loc on the 16-wide row kernel, two configurations per workgroup: each
    of the two proposal copies gets its own candidate, and a sample with an
    odd number of valid candidates leaves the second copy idle in its last
    batch."""
    lat, model, spec, theta = _resconv_rect(L0, 16, 2, 16, precision,
                                            nup=(L0 * 8) if ham == "j1j2" else None)
    if ham == "j1j2":
        H = qtx.operator.Heisenberg(J=[1, 0.5], n_neighbor=[1, 2], msr=True)
        Ho = oop.heisenberg([L0, 16], [1, 0.5], [1, 2], True)
        s = rand_spins(rng, 4, L0 * 16, nup=L0 * 8)
    else:
        H = qtx.operator.Ising(h=0.7, J=1.3)
        Ho = oop.ising([L0, 16], h=0.7, J=1.3)
        s = rand_spins(rng, 4, L0 * 16)
    la = model.logpsi(to_dev(s))
    e, nconn = model.eloc(H, to_dev(s), la)
    ref = oop.oloc(s, lambda x: om.resconv_forward(spec, theta, x), Ho)
    assert max_rel(e.cpu().numpy(), ref) < tol
    seg = np.concatenate([oop.get_conn(sc, Hc)[0] for sc, Hc in oop.apply_off_diag(s, Ho).values()])
    assert np.array_equal(nconn.cpu().numpy(), np.bincount(seg, minlength=4))


def test_resconv_eloc_large_log_ratios(rng):
    """Connected configurations whose |log psi(s') - log psi(s)| passes 80
    take the FP64 exp in the E_loc kernel (FP32 overflows near 88).  The last
    convolution's weights are scaled until the oracle sees such ratios (and
    none beyond FP64's range)."""
    lat, model, spec, theta0 = _resconv_rect(4, 4, 2, 16, "fp32")
    H = qtx.operator.Heisenberg(J=[1, 0.5], n_neighbor=[1, 2], msr=True)
    Ho = oop.heisenberg([4, 4], [1, 0.5], [1, 2], True)
    s = rand_spins(rng, 16, 16, nup=8)
    cin, cout, _ = spec.layers[-1]
    w0, _ = spec.offsets[-1]
    conn = [oop.get_conn(sc, Hc)[:2] for sc, Hc in oop.apply_off_diag(s, Ho).values()]
    for scale in (50.0, 100.0, 200.0, 400.0, 800.0):
        th = theta0.copy()
        th[w0: w0 + cout * cin * 9] *= scale
        th = th.astype(np.float32).astype(np.float64)
        f = lambda x, th=th: om.resconv_forward(spec, th, x)
        la0 = f(s)[1]
        d = np.concatenate([np.abs(f(sconn)[1] - la0[seg]) for seg, sconn in conn])
        if d.max() > 80.0:
            break
    assert 80.0 < d.max() < 600.0, d.max()
    model.set_params(th.astype(np.float32))
    la = model.logpsi(to_dev(s))
    e, _ = model.eloc(H, to_dev(s), la)
    ref = oop.oloc(s, f, Ho)
    assert np.all(np.isfinite(e.cpu().numpy()))
    # log psi runs to ~200 here and the scaled layer's pre-activations are as
    # large: FP32 rounding through the network leaves a few tens of ulps of
    # log psi in each log ratio, i.e. ratio errors of ~1e-4
    tol = 1e-5 + 32 * 2.0 ** -24 * float(np.max(np.abs(la0)))
    assert max_rel(e.cpu().numpy(), ref) < tol, tol


@pytest.mark.parametrize("precision,tie,lptol", [("fp32", 1e-4, 1e-5), ("bf16x3", 1e-3, 1e-4)])
def test_resconv_rows_metropolis_16x16(precision, tie, lptol, rng):
    lat, model, spec, theta = _resconv_rect(16, 16, 2, 16, precision, nup=128)
    ns, nsteps = 6, 8
    s0 = rand_spins(rng, ns, 256, nup=128)
    words = philox.sweep_words(21, nsteps, ns)
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
    assert max_rel(la.cpu().numpy()[clean], la_ref[clean]) < lptol


@pytest.mark.parametrize("gram,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 1e-5)])
@pytest.mark.parametrize("shape", [(300, 1000), (256, 20000), (1100, 5000)])
def test_gram_entry_point(gram, tol, shape, rng):
    ns, np_ = shape
    O = (rng.standard_normal((ns, np_)) * 0.2 + 0.05).astype(np.float32)
    Od = to_dev(O)
    mean = torch.empty(np_, dtype=torch.float64, device=DEV)
    _lib.call("qtx_colmean", _lib.stream(), _lib.QTX_F32, Od.data_ptr(), ns, np_, mean.data_ptr())
    T = torch.empty((ns, ns), dtype=torch.float64, device=DEV)
    g = {"fp64": 0, "fp32": 1, "bf16x3": 2}[gram]
    nws = int(_lib.lib().qtx_gram_ws_bytes(ns, np_, g))
    assert (nws > 0) == (gram == "bf16x3")
    ws = torch.empty(max(nws, 1), dtype=torch.uint8, device=DEV)
    _lib.call("qtx_gram", _lib.stream(), _lib.QTX_F32, Od.data_ptr(), mean.data_ptr(), None, ns, np_, g, T.data_ptr(),
              ws.data_ptr(), nws)
    Obar, _ = osr.get_Obar(O.astype(np.float64))
    ref = Obar @ Obar.T
    Tn = T.cpu().numpy()
    assert np.array_equal(Tn, Tn.T)
    assert np.max(np.abs(Tn - ref)) / np.max(np.abs(ref)) < tol


def test_gram_split_beyond_2pow32_elements():
    """ns * np above 2^32 (the split pass once launched rows * kpad work-items,
    which wraps a launch's 32-bit grid size).  O = a (x) u with a_s in
    {-2..2} is exact in FP32, so T_st = (a_s - abar)(a_t - abar) sum(u^2) / ns
    in closed form; the workspace starts as NaN, so any element the split pass
    skips poisons T."""
    ns, np_ = 256, (1 << 24) + 4096
    assert ns * np_ > (1 << 32)
    gen = torch.Generator(device=DEV).manual_seed(7)
    a = torch.randint(-2, 3, (ns,), generator=gen, device=DEV).to(torch.float32)
    u = torch.randn(np_, generator=gen, device=DEV, dtype=torch.float32)
    O = a[:, None] * u[None, :]
    mean = torch.empty(np_, dtype=torch.float64, device=DEV)
    _lib.call("qtx_colmean", _lib.stream(), _lib.QTX_F32, O.data_ptr(), ns, np_, mean.data_ptr())
    nws = int(_lib.lib().qtx_gram_ws_bytes(ns, np_, _lib.QTX_GRAM_BF16X3))
    ws = torch.full((nws,), 0xFF, dtype=torch.uint8, device=DEV)
    T = torch.empty((ns, ns), dtype=torch.float64, device=DEV)
    _lib.call("qtx_gram", _lib.stream(), _lib.QTX_F32, O.data_ptr(), mean.data_ptr(), None, ns, np_,
              _lib.QTX_GRAM_BF16X3, T.data_ptr(), ws.data_ptr(), nws)
    del O, ws
    ad = a.double() - a.double().mean()
    ref = torch.outer(ad, ad) * (u.double() ** 2).sum() / ns
    assert torch.isfinite(T).all()
    assert ((T - ref).abs().max() / ref.abs().max()).item() < 1e-5
