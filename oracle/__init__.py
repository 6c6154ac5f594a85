"""CPU FP64 oracle for the quantax VMC step — TEST INFRASTRUCTURE ONLY.

This package restates, in numpy/scipy, the algorithm of the reference
(ChenAo-Phys/quantax) for the hot path: Metropolis sampling (LocalFlip,
SpinExchange), RBM_Dense and ResConv wavefunctions, Heisenberg/Ising local
energies over connected configurations, the per-sample Jacobian O_k and the
SR/MinSR solve.  Every function cites the reference file:line it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker.  The product package
``quantax_amd`` never imports anything from here; it fails loudly when the
HIP library is missing instead of falling back to this code.

Pinning (how each part is checked independently of retyping its formula):

* ``lattice``      — brute-force bond counts and explicit small lattices.
* ``operator``     — exact diagonalisation against closed forms (TFIM chain)
                     and the reference's published ED energies
                     (``tutorials/exact_diag.ipynb:191-192``,
                     ``tutorials/J1J2.ipynb:374``); brute-force
                     <psi|H|psi>/<psi|psi> equals the |psi|^2-weighted mean of
                     E_loc on N <= 12.
* ``philox``       — Random123 known-answer vectors for Philox4x32-10.
* ``models.rbm_*`` — closed forms (log prod cosh, tanh Jacobian) checked
                     against finite differences and direct products.
* ``models.resconv_*`` — translation invariance, the LSE head closed form,
                     finite-difference and torch-autograd (float64) checks of
                     the backward pass, for the tanh and erf GELU and the
                     opt-in ReLU variant.
* ``cpu_resconv``  — the FP32 torch-CPU forward (the CPU baseline) against the
                     float64 restatement above, to FP32 rounding.
* ``sampler``      — exact stationary distribution (chi^2 on N <= 8) and
                     the detailed structure of proposals.
* ``sr``           — scipy cho_solve / eigh / lstsq minimum-norm solutions,
                     the push-through identity SR == MinSR for a shift > 0;
                     ``sum_without_noise`` by columns of known statistics
                     (constant, +-a, mu +- sigma: snr = |mu| sqrt(m) / sigma);
                     the SNR-truncated minnorm and lstsq branches equal for
                     any tol_snr (their components differ by the singular
                     values, which the SNR does not see); ``minsr_pinv_eig``
                     = T^-1 b on well-conditioned indefinite T and
                     A^T minsr_pinv_eig(A A^T, b) = minnorm_pinv_eig(A, b).
* ``sampler.reweight_factor`` — the closed form |psi|^(2-n) / mean (constant 1
                     at n = 2); ``operator.expectation`` — the dense <H> and
                     <H^2> - <H>^2 on N <= 12.

Parity unpinned (no reference output exists for it, JAX is absent): the
threefry streams of ``jax.random`` (replaced here by supplied Philox streams),
the JAX parameter initialisers, and ``jax.nn.gelu``'s default form (the tanh
approximation is assumed, the erf form is available and tested as a variant).
"""
