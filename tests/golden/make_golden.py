"""Writes tests/golden/golden_small.npz from the CPU oracle.

The reference ships no fixtures or test vectors for this path (it has no test
suite), so the committed vectors are the oracle's own outputs on tiny seeded
problems; the oracle itself is pinned by tests/test_oracle.py (known answers,
closed forms, brute force).  Regenerate with:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ed, models as om, operator as oop, philox, sampler as osm, sr as osr  # noqa: E402
from oracle import lattice as olat  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_small.npz")


def generate():
    g = {}
    rng = np.random.default_rng(20240601)
    # lattice / operator: 4x4 J1-J2 with Marshall signs
    Ho = oop.heisenberg([4, 4], [1.0, 0.5], [1, 2], True)
    s = -np.ones((8, 16), np.int8)
    s[:, :8] = 1
    for r in s:
        rng.shuffle(r)
    g["j1j2_spins"] = s
    g["j1j2_diag"] = oop.apply_diag(s, Ho)
    (sc, H), = oop.apply_off_diag(s, Ho).values()
    seg, sconn, hc = oop.get_conn(sc, H)
    g["j1j2_conn_segment"], g["j1j2_conn_spins"], g["j1j2_conn_h"] = seg, sconn, hc
    # RBM float64 on the same spins
    M = 24
    W = rng.standard_normal((M, 16)) * 0.15
    b = rng.standard_normal(M) * 0.15
    g["rbm_theta"] = om.rbm_pack(W, b)
    f = lambda x: om.rbm_logpsi(W, b, x)
    g["rbm_logabs"] = f(s)[1]
    g["rbm_eloc"] = oop.oloc(s, f, Ho)
    g["rbm_O"] = om.rbm_jacobian(W, b, s)
    # ResConv(2, 16) on 4x4
    spec = om.ResConvSpec(2, 16, (4, 4))
    theta = spec.he_normal(rng).astype(np.float64)
    for (cin, cout, hb), (w, bo) in zip(spec.layers, spec.offsets):
        if hb:
            theta[bo: bo + cout] = 0.1 * rng.standard_normal(cout)
    theta = theta.astype(np.float32).astype(np.float64)
    g["resconv_theta"] = theta.astype(np.float32)
    fr = lambda x: om.resconv_forward(spec, theta, x)
    g["resconv_logabs"] = fr(s)[1]
    g["resconv_eloc"] = oop.oloc(s, fr, Ho)
    g["resconv_O"] = om.resconv_jacobian(spec, theta, s[:4]).astype(np.float32)  # first 4 samples
    # Metropolis: float64 RBM on a 6x6 lattice, LocalFlip, 32 chains x 40 steps
    W6 = rng.standard_normal((20, 36)) * 0.2
    b6 = rng.standard_normal(20) * 0.2
    g["mc_theta"] = om.rbm_pack(W6, b6)
    s0 = philox.init_spins(11, 32, 36)
    words = philox.sweep_words(12, 40, 32)
    sf, laf, nacc, near = osm.metropolis(lambda x: om.rbm_logpsi(W6, b6, x), s0, words, osm.LOCALFLIP)
    g["mc_s0"], g["mc_seed"], g["mc_s"], g["mc_logabs"], g["mc_nacc"] = s0, np.int64(12), sf, laf, nacc
    # exchange variant on the same lattice, Sz = 0
    s0x = philox.init_spins(13, 32, 36, nup=18)
    wx = philox.sweep_words(14, 40, 32)
    sx, lax, naccx, _ = osm.metropolis(lambda x: om.rbm_logpsi(W6, b6, x), s0x, wx, osm.EXCHANGE,
                                       olat.site_neighbor_table([6, 6], 1), 1)
    g["mcx_s0"], g["mcx_seed"], g["mcx_s"], g["mcx_logabs"], g["mcx_nacc"] = s0x, np.int64(14), sx, lax, naccx
    # SR / MinSR steps on the RBM Jacobian and local energies
    O, E = g["rbm_O"], g["rbm_eloc"]
    g["sr_chol"] = osr.sr_step(O, E, lambda A, v: osr.auto_shift_eig(A, v, rshift=0.0, ashift=1e-3))[0]
    g["sr_eigh"] = osr.sr_step(O, E, lambda A, v: osr.auto_pinv_eig(A, v, rtol=1e-10))[0]
    # known answers
    g["ed_j1j2_4x4"] = ed.ground_energy(Ho, 16, nup=8)
    g["philox_words"] = philox.sweep_words(99, 3, 5, step0=7)
    return g


if __name__ == "__main__":
    np.savez_compressed(OUT, **generate())
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
