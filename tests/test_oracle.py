"""Pins of the CPU oracle (no GPU): known answers, closed forms, brute force.

Each check is chosen so that a plausible mistake in the restatement (a
dropped term, a sign, a transposed index) fails it — none of them retypes the
oracle's own formula.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import ed
from oracle import lattice as olat
from oracle import models as om
from oracle import operator as oop
from oracle import philox
from oracle import sampler as osm
from oracle import sr as osr
from tests.helpers import rand_spins


# ------------------------------------------------------------------ Philox
def test_philox_random123_known_answers():
    """Known-answer vectors of Philox4x32-10 from the Random123 distribution."""
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for c, k, out in kat:
        got = philox.philox4x32_10(np.array(c, np.uint32), np.array(k, np.uint32))
        assert [int(v) for v in got] == out


def test_uniform_and_bounded_ranges():
    w = philox.sweep_words(1, 50, 200).reshape(-1, 4)
    u = philox.words_to_uniform(w[:, 2], w[:, 3])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01
    assert philox.words_to_uniform(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53
    b = philox.bounded(w[:, 0], 7)
    assert b.min() == 0 and b.max() == 6
    assert np.all(np.bincount(b) > 1200)


def test_init_spins_invariants():
    s = philox.init_spins(5, 300, 36, nup=18)
    assert np.all(s.sum(1) == 0) and set(np.unique(s)) == {-1, 1}
    # Fisher-Yates is uniform over positions: every site is up about half the time
    assert np.all(np.abs((s > 0).mean(0) - 0.5) < 0.12)
    u = philox.init_spins(5, 400, 20)
    assert abs((u > 0).mean() - 0.5) < 0.03


# ------------------------------------------------------------------ lattice
@pytest.mark.parametrize("L", [3, 4, 5, 6, 8, 10, 16])
def test_square_bond_counts(L):
    nn, nnn = olat.get_neighbor([L, L], [1, 2])
    assert len(nn) == 2 * L * L and len(nnn) == 2 * L * L
    # every site has 4 nearest and 4 next-nearest neighbours
    for nb in (nn, nnn):
        deg = np.bincount(nb.ravel(), minlength=L * L)
        assert np.all(deg == 4)
    assert np.all(nn[:, 0] < nn[:, 1])
    assert np.all(np.diff(nn[:, 0] * L * L + nn[:, 1]) > 0)  # sorted by (i, j)


def test_chain_bond_order_matches_argwhere_convention():
    nb = olat.get_neighbor([16], 1)
    assert nb.tolist()[:3] == [[0, 1], [0, 15], [1, 2]]
    assert len(nb) == 16
    assert olat.get_neighbor([6], 1, boundary=0).tolist() == [[i, i + 1] for i in range(5)]


# ------------------------------------------------------------------ ED known answers
def test_ed_tfim_chain_closed_form():
    for L in (8, 12):
        k = np.pi * (2 * np.arange(L) + 1) / L
        exact = -np.sum(np.sqrt(2 - 2 * np.cos(k)))
        assert abs(ed.ground_energy(oop.ising([L], h=1.0), L)[0] - exact) < 1e-9
    assert abs(-np.sum(np.sqrt(2 - 2 * np.cos(np.pi * (2 * np.arange(16) + 1) / 16)))
               - (-2 / np.sin(np.pi / 32))) < 1e-12


def test_ed_heisenberg_4x4_reference_values():
    """tutorials/exact_diag.ipynb:191-192 (Pauli convention, Sz = 0)."""
    w = ed.ground_energy(oop.heisenberg([4, 4]), 16, nup=8, k=2)
    assert abs(w[0] - (-44.913932833715506)) < 1e-8
    assert abs(w[1] - (-42.599539490653896)) < 1e-8


def test_ed_j1j2_4x4_reference_value():
    """tutorials/J1J2.ipynb:374 (J2 = 0.5, Marshall sign rule)."""
    w = ed.ground_energy(oop.heisenberg([4, 4], [1.0, 0.5], [1, 2], msr=True), 16, nup=8)
    assert abs(w[0] - (-33.831693405579394)) < 1e-8


# ------------------------------------------------------------------ local energy
def _dense_expectation(op, psi_all, N):
    Hs, basis = ed.sparse_hamiltonian(op, N)
    v = psi_all
    return float(v @ (Hs @ v) / (v @ v)), basis


@pytest.mark.parametrize("which", ["ising", "heis", "j1j2"])
def test_eloc_weighted_mean_equals_rayleigh_quotient(which, rng):
    """sum |psi|^2 E_loc / sum |psi|^2 == <psi|H|psi>/<psi|psi> over the full basis."""
    if which == "ising":
        N, op = 8, oop.ising([8], h=0.7, J=1.1)
    elif which == "heis":
        N, op = 8, oop.heisenberg([8], msr=True)
    else:
        N, op = 9, oop.heisenberg([3, 3], [1.0, 0.5], [1, 2], msr=True)
    W = rng.standard_normal((6, N)) * 0.4
    b = rng.standard_normal(6) * 0.2
    basis, _ = ed.basis_states(N)
    la = om.rbm_logpsi(W, b, basis)[1]
    psi = np.exp(la - la.max())
    E = oop.oloc(basis, lambda x: om.rbm_logpsi(W, b, x), op)
    lhs = np.sum(psi ** 2 * E) / np.sum(psi ** 2)
    rhs, _ = _dense_expectation(op, psi, N)
    assert abs(lhs - rhs) < 1e-10 * max(1, abs(rhs))


@pytest.mark.parametrize("reweight", [0.0, 0.5, 1.0, 2.0, 3.0])
def test_reweight_factor_closed_form(reweight, rng):
    """r = |psi|^(2 - n) / mean |psi|^(2 - n) (sampler/sampler.py:66-69) from
    psi itself, without the oracle's max shift: identical, mean 1, r = 1 at
    n = 2, and unchanged when every log psi moves by the same constant."""
    la = rng.standard_normal(300) * 2.0
    psi = np.exp(la)
    direct = np.abs(psi) ** (2.0 - reweight)
    direct /= direct.mean()
    r = osm.reweight_factor(la, reweight)
    assert np.max(np.abs(r - direct) / direct) < 1e-13
    assert abs(r.mean() - 1.0) < 1e-14
    assert np.array_equal(osm.reweight_factor(la + 700.0, reweight) > 0, np.ones(300, bool))
    assert np.max(np.abs(osm.reweight_factor(la + 700.0, reweight) - r) / r) < 1e-12
    if reweight == 2.0:
        assert np.all(r == 1.0)


def test_expectation_with_weights_equals_dense_moments(rng):
    """Uniform samples of the whole basis (reweight 0) with the weights
    |psi|^2 / mean |psi|^2: expectation() is <psi|H|psi>/<psi|psi> and its
    variance <psi|H^2|psi>/<psi|psi> - <H>^2 of the dense H
    (operator.py:564-595); the unweighted mean is not."""
    N, op = 10, oop.heisenberg([10], [1.0, 0.3], [1, 2])
    W = rng.standard_normal((8, N)) * 0.5
    b = rng.standard_normal(8) * 0.3
    basis, _ = ed.basis_states(N)
    la = om.rbm_logpsi(W, b, basis)[1]
    E = oop.oloc(basis, lambda x: om.rbm_logpsi(W, b, x), op)
    m, v = oop.expectation(E, osm.reweight_factor(la, 0.0), return_var=True)
    Hs, _ = ed.sparse_hamiltonian(op, N)
    psi = np.exp(la - la.max())
    hp = Hs @ psi
    e_ref = psi @ hp / (psi @ psi)
    v_ref = hp @ hp / (psi @ psi) - e_ref ** 2
    assert abs(m - e_ref) < 1e-11 * abs(e_ref)
    assert abs(v - v_ref) < 1e-9 * max(1.0, v_ref)
    assert abs(oop.expectation(E) - e_ref) > 1e-3  # the weights matter here
    assert abs(oop.expectation(E, osm.reweight_factor(la, 0.0)) - m) < 1e-14 * abs(m)


def test_compiled_tables_equal_literal_application(rng):
    for op, N, nup in [(oop.heisenberg([4, 4], [1, 0.5], [1, 2], True), 16, 8), (oop.ising([10], 0.3, 2.0), 10, None),
                       (oop.heisenberg([6], [1.0, 0.2], [1, 2]), 6, None)]:
        ch = oop.compile_op_list(op, N)
        s = rand_spins(rng, 40, N, nup)
        assert np.array_equal(oop.compiled_diag(s, ch), oop.apply_diag(s, op))
        (sc, H), = oop.apply_off_diag(s, op).values()
        seg, sconn, hc = oop.get_conn(sc, H)
        seg2, sconn2, hc2, _ = oop.compiled_conn(s, ch)
        assert np.array_equal(seg, seg2) and np.array_equal(sconn, sconn2) and np.array_equal(hc, hc2)


def test_heisenberg_term_structure():
    """sigma.sigma on an antiparallel bond: diagonal -1 and exchange amplitude 2 (x sign)."""
    op = oop.heisenberg([2], msr=False, boundary=0)
    s = np.array([[1, -1]], np.int8)
    assert oop.apply_diag(s, op)[0] == -1.0
    (sc, H), = oop.apply_off_diag(s, op).values()
    seg, sconn, hc = oop.get_conn(sc, H)
    assert sconn.tolist() == [[-1, 1]] and hc.tolist() == [2.0]
    op_msr = oop.heisenberg([2], msr=True, boundary=0)
    (sc, H), = oop.apply_off_diag(s, op_msr).values()
    assert oop.get_conn(sc, H)[2].tolist() == [-2.0]


# ------------------------------------------------------------------ RBM
def test_rbm_logpsi_is_log_prod_cosh(rng):
    W = rng.standard_normal((7, 10)) * 0.3
    b = rng.standard_normal(7) * 0.3
    s = rand_spins(rng, 20, 10)
    direct = np.log(np.prod(np.cosh(s @ W.T + b), axis=1))
    assert np.allclose(om.rbm_logpsi(W, b, s)[1], direct, rtol=1e-13, atol=1e-13)
    # large arguments stay finite (log-domain)
    big = om.logcosh(np.array([800.0, -800.0]))
    assert np.allclose(big, 800.0 - np.log(2.0))


def test_rbm_jacobian_finite_differences(rng):
    M, N = 5, 6
    theta = rng.standard_normal(M * N + M) * 0.3
    s = rand_spins(rng, 4, N)
    O = om.rbm_jacobian(*om.rbm_unpack(theta, M, N), s)
    eps = 1e-6
    for k in range(theta.size):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += eps
        tm[k] -= eps
        fd = (om.rbm_logpsi(*om.rbm_unpack(tp, M, N), s)[1] - om.rbm_logpsi(*om.rbm_unpack(tm, M, N), s)[1]) / (2 * eps)
        assert np.allclose(O[:, k], fd, atol=1e-8)


def test_rbm_local_update_equals_recompute(rng):
    W = rng.standard_normal((9, 12)) * 0.3
    b = rng.standard_normal(9) * 0.3
    s = rand_spins(rng, 10, 12)
    s2 = s.copy()
    s2[:, [2, 7]] *= -1
    h = om.rbm_h(W, b, s)
    assert np.allclose(om.rbm_ref_update(W, h, s2, s), om.rbm_h(W, b, s2), atol=1e-12)


# ------------------------------------------------------------------ ResConv
def _torch_resconv_logpsi(spec, theta, s, approximate=True):
    """Independent float64 ResConv in torch (circular pad + conv2d, autograd)."""
    import torch
    import torch.nn.functional as F

    params = spec.unpack(theta)
    x = torch.as_tensor(s, dtype=torch.float64).reshape((-1, 1) + spec.shape)
    tensors = []
    for W, B in params:
        tensors.append((torch.tensor(W, requires_grad=True), torch.tensor(B, requires_grad=True)))

    def conv(x, W, B):
        return F.conv2d(F.pad(x, (1, 1, 1, 1), mode="circular"), W) + B[None, :, None, None]

    if approximate == "relu":
        g = F.relu
    else:
        g = (lambda z: F.gelu(z, approximate="tanh")) if approximate else F.gelu
    for i in range(spec.nblocks):
        (W1, B1), (W2, B2) = tensors[2 * i], tensors[2 * i + 1]
        r = x
        y = x / np.sqrt(i + 1)
        y = y / np.sqrt(2.0) if i == 0 else g(y)
        z = conv(g(conv(y, W1, B1)), W2, B2)
        if z.shape[1] > r.shape[1]:
            r = r.repeat_interleave(z.shape[1] // r.shape[1], dim=1)
        x = z + r
    h = (x / np.sqrt(spec.nblocks + 1)).reshape(x.shape[0], -1)
    la = torch.logsumexp(h, dim=1) - np.log(h.shape[1])
    return la, tensors


@pytest.mark.parametrize("approximate", [True, False, "relu"])
def test_resconv_forward_and_jacobian_match_torch_autograd(approximate, rng):
    import torch

    spec = om.ResConvSpec(3, 4, (4, 5))
    theta = rng.standard_normal(spec.nparams) * 0.4
    s = rand_spins(rng, 3, 20)
    la = om.resconv_forward(spec, theta, s, approximate)[1]
    O = om.resconv_jacobian(spec, theta, s, approximate)
    for n in range(3):
        lt, tens = _torch_resconv_logpsi(spec, theta, s[n: n + 1], approximate)
        assert abs(lt.item() - la[n]) < 1e-12
        lt.sum().backward()
        g = []
        for (cin, cout, hb), (W, B) in zip(spec.layers, tens):
            g.append(W.grad.numpy().ravel())
            if hb:
                g.append(B.grad.numpy().ravel())
        assert np.allclose(O[n], np.concatenate(g), atol=1e-11)


def test_resconv_translation_invariance(rng):
    spec = om.ResConvSpec(2, 4, (5, 5))
    theta = rng.standard_normal(spec.nparams) * 0.5
    s = rand_spins(rng, 4, 25)
    sh = np.roll(s.reshape(4, 5, 5), shift=(2, 3), axis=(1, 2)).reshape(4, 25)
    assert np.allclose(om.resconv_forward(spec, theta, s)[1], om.resconv_forward(spec, theta, sh)[1], atol=1e-12)


def test_resconv_head_closed_form():
    """All weights and biases zero: every block adds 0, so h = s / sqrt(nb + 1)
    broadcast over channels and log psi = LSE(h) - log(C N)."""
    spec = om.ResConvSpec(3, 4, (3, 3))
    s = np.array([[1, -1, 1, 1, 1, -1, -1, 1, 1]], np.int8)
    la = om.resconv_forward(spec, np.zeros(spec.nparams), s)[1][0]
    h = s[0] / 2.0
    assert abs(la - (np.log(4 * np.exp(h).sum()) - np.log(36))) < 1e-13


def test_resconv_nparams_reference_counts():
    assert om.ResConvSpec(8, 16, (16, 16)).nparams == 34944
    assert om.ResConvSpec(16, 32, (20, 20)).nparams == 286976


# ------------------------------------------------------------------ sampler
def _exact_probs(logpsi_fn, basis):
    la = logpsi_fn(basis)[1]
    p = np.exp(2 * (la - la.max()))
    return p / p.sum()


def _int(s):
    return ((s > 0).astype(np.int64) * (1 << np.arange(s.shape[1] - 1, -1, -1))).sum(1)


@pytest.mark.parametrize("kind", ["flip", "exchange"])
def test_metropolis_stationary_distribution(kind, rng):
    # a flat enough |psi|^2 that 60 proposals per chain mix between basins
    N = 6
    W = rng.standard_normal((4, N)) * 0.25
    b = rng.standard_normal(4) * 0.2
    f = lambda x: om.rbm_logpsi(W, b, x)
    if kind == "flip":
        basis, _ = ed.basis_states(N)
        s0 = rand_spins(rng, 4000, N)
        kd, nbr = osm.LOCALFLIP, None
    else:
        basis, _ = ed.basis_states(N, nup=3)
        s0 = rand_spins(rng, 4000, N, nup=3)
        kd, nbr = osm.EXCHANGE, olat.site_neighbor_table([N], 1)
    p = _exact_probs(f, basis)
    s, *_ = osm.sweep(f, s0, seed=3, nsteps=60, kind=kd, nbr=nbr)
    idx = {int(v): k for k, v in enumerate(_int(basis))}
    counts = np.bincount([idx[int(v)] for v in _int(s)], minlength=basis.shape[0])
    expected = p * s.shape[0]
    chi2 = np.sum((counts - expected) ** 2 / expected)
    dof = basis.shape[0] - 1
    assert chi2 < dof + 6 * np.sqrt(2 * dof), chi2


def test_exchange_proposal_structure():
    s = np.array([[1, -1, 1, -1, -1, 1]], np.int8)
    nbr = olat.site_neighbor_table([6], 1)
    # k-th hopping (+1) site: hopping sites are 0, 2, 5; w0 chosen for k = 1 -> site 2
    w = np.array([[int(1.5 / 3 * 2 ** 32), 0, 0, 0]], np.uint32)
    out = osm.propose(osm.EXCHANGE, s, w, nbr, hopping=1)
    # slot 0 of site 2 is neighbour 1 (spin -1): swap
    assert out.tolist() == [[1, 1, -1, -1, -1, 1]]


# ------------------------------------------------------------------ SR solvers
def _problem(rng, ns, np_):
    O = rng.standard_normal((ns, np_)) + 0.5
    E = rng.standard_normal(ns)
    Ebar, _, _ = osr.get_Ebar(E)
    Obar, _ = osr.get_Obar(O)
    return Obar, Ebar


def test_push_through_identity_minsr_equals_sr(rng):
    A, b = _problem(rng, 40, 90)
    x1 = osr.minnorm_shift_eig(A, b, rshift=0.0, ashift=1e-3)
    x2 = osr.lstsq_shift_eig(A, b, rshift=0.0, ashift=1e-3)
    assert np.allclose(x1, x2, rtol=1e-9, atol=1e-12)
    ref = np.linalg.solve(A.T @ A + 1e-3 * np.eye(90), A.T @ b)
    assert np.allclose(x1, ref, rtol=1e-8, atol=1e-11)


def test_pinv_solvers_give_minimum_norm_least_squares(rng):
    for ns, np_ in ((30, 80), (90, 20)):
        A, b = _problem(rng, ns, np_)
        x = osr.auto_pinv_eig(A, b, rtol=1e-10)
        ref = np.linalg.lstsq(A, b, rcond=None)[0]
        assert np.allclose(x, ref, rtol=1e-6, atol=1e-9)


def test_sum_without_noise_closed_forms():
    """SNR truncation (solver.py:114-125) on columns whose statistics are
    known: constant columns have infinite SNR (factor 1), +-a columns sum to
    zero, and mu +- sigma columns of length m have snr = |mu| sqrt(m) / sigma,
    so tol = snr halves the sum and twice that divides it by 65; at or below
    tol 1e-6 nothing changes."""
    m = 16
    pm = np.tile([1.0, -1.0], m // 2)
    c = np.stack([np.full(m, 0.3), 2.5 * pm, 1.0 + 2.0 * pm, -3.0 + 0.5 * pm], axis=1)
    assert np.array_equal(osr.sum_without_noise(c, 0.0), c.sum(0))
    assert np.array_equal(osr.sum_without_noise(c, 1e-6), c.sum(0))
    tol = 1.0 * np.sqrt(m) / 2.0  # the snr of column 2: mu 1, sigma 2
    x = osr.sum_without_noise(c, tol)
    assert abs(x[0] - 0.3 * m) < 1e-12 and x[1] == 0.0
    assert abs(x[2] - m / 2.0) < 1e-12
    q3 = tol / (3.0 * np.sqrt(m) / 0.5)
    assert abs(x[3] - (-3.0 * m) / (1 + q3 ** 6)) < 1e-12
    x2 = osr.sum_without_noise(c, 2 * tol)
    assert abs(x2[2] - m / 65.0) < 1e-12
    # invariant under the sign and the scale of a column
    assert np.allclose(osr.sum_without_noise(-7.0 * c, tol), -7.0 * x, rtol=1e-13, atol=0)


def test_snr_truncation_keeps_the_push_through_identity(rng):
    """With A = P Sigma Q^T, the SR branch's c_sk = (A V)_sk b_s is sigma_k
    times the MinSR branch's c_ts = U_ts b_t for the same singular pair, and
    the SNR is scale-free: both branches truncate the same components by the
    same factors, so minnorm_pinv_eig == lstsq_pinv_eig for any tol_snr.  A
    sum over the wrong index, or a transposed U, breaks the equality."""
    for ns, np_ in ((40, 90), (90, 30)):
        A, b = _problem(rng, ns, np_)
        for tol in (0.0, 0.3, 1.0, 3.0):
            x1 = osr.minnorm_pinv_eig(A, b, rtol=1e-10, tol_snr=tol)
            x2 = osr.lstsq_pinv_eig(A, b, rtol=1e-10, tol_snr=tol)
            assert np.linalg.norm(x1 - x2) <= 1e-8 * np.linalg.norm(x1), (ns, np_, tol)
        # a tolerance above every component's SNR leaves almost nothing
        x0 = osr.auto_pinv_eig(A, b, rtol=1e-10)
        xs = osr.auto_pinv_eig(A, b, rtol=1e-10, tol_snr=1e4)
        assert np.linalg.norm(xs) < 1e-6 * np.linalg.norm(x0)


def test_minsr_pinv_eig_solves_indefinite_and_matches_minnorm(rng):
    """minsr_pinv_eig (solver.py:262-294): for a well-conditioned indefinite
    T and a negligible cutoff it is T^-1 b (the soft inverse keeps the sign of
    each eigenvalue); on T = A A^T, A^T x is minnorm_pinv_eig(A, b) with the
    same SNR tolerance."""
    n = 60
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    d = np.concatenate([rng.uniform(0.5, 2.0, n // 2), -rng.uniform(0.5, 2.0, n - n // 2)])
    T = (Q * d) @ Q.T
    T = 0.5 * (T + T.T)
    b = rng.standard_normal(n)
    x = osr.minsr_pinv_eig(T, b, rtol=1e-14)
    assert np.linalg.norm(x - np.linalg.solve(T, b)) < 1e-11 * np.linalg.norm(x)
    A, bb = _problem(rng, 40, 90)
    for tol in (0.0, 0.7):
        y = osr.minsr_pinv_eig(A @ A.T, bb, rtol=1e-10, tol_snr=tol)
        ref = osr.minnorm_pinv_eig(A, bb, rtol=1e-10, tol_snr=tol)
        assert np.linalg.norm(A.T @ y - ref) < 1e-9 * np.linalg.norm(ref)


def test_eigs_inv_soft_cutoff():
    v = np.array([1e-20, 1e-6, 1.0, 0.0, -2.0])
    inv = osr.eigs_inv(v, rtol=1e-6)
    assert inv[3] == 0.0
    assert abs(inv[2] - 1.0) < 1e-12 and abs(inv[4] + 0.5) < 1e-12
    assert inv[0] < 1e-20  # far below the cut: suppressed, not amplified
    assert abs(inv[1] - 1e6 / (1 + (2e-6 / 1e-6) ** 6)) < 1e-3


def test_ebar_obar_centering(rng):
    O = rng.standard_normal((50, 7))
    E = rng.standard_normal(50) + 3
    Ebar, Em, Ev = osr.get_Ebar(E)
    Obar, Om = osr.get_Obar(O)
    assert abs(Ebar.sum()) < 1e-12 and np.allclose(Obar.sum(0), 0, atol=1e-12)
    assert abs(Em - E.mean()) < 1e-13 and abs(Ev - E.var()) < 1e-12
    assert np.allclose(np.linalg.norm(Ebar) ** 2, E.var(), rtol=1e-12)


def test_update_skips_nonfinite():
    th = np.ones(4, np.float32)
    assert osr.update(th, np.array([1, np.nan, 0, 0])) is None
    assert np.array_equal(osr.update(th, np.array([0.5, 0, 0, 1.0])), np.array([0.5, 1, 1, 0], np.float32))


@pytest.mark.parametrize("shape,nb,C", [((6, 6), 3, 16), ((5, 4), 2, 32)])
def test_cpu_resconv_f32_matches_float64_restatement(shape, nb, C, rng):
    """The FP32 torch-CPU forward (the CPU baseline's) against the float64
    restatement, which the torch float64 autograd check above pins: agreement
    to FP32 rounding; a dropped block, a missing bias or an unflipped
    circular shift would move log psi by far more."""
    from oracle import cpu_resconv

    spec = om.ResConvSpec(nb, C, shape)
    theta = spec.he_normal(rng).astype(np.float64)
    for (cin, cout, hb), (w, bo) in zip(spec.layers, spec.offsets):
        if hb:
            theta[bo: bo + cout] = 0.1 * rng.standard_normal(cout)
    theta = theta.astype(np.float32).astype(np.float64)
    s = rand_spins(rng, 12, shape[0] * shape[1])
    ref = om.resconv_forward(spec, theta, s)[1]
    got = cpu_resconv.resconv_forward_f32(spec, theta, s)
    assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)) < 2e-5
