"""The solver factories' branches and options on the device against the
oracle (needs an MI355X).

* ``minnorm_*`` / ``lstsq_*`` take their branch whatever the shape
  (``optimizer/solver.py:50-77, 128-164``): both shapes, both branches, the
  shifted Cholesky and the soft pseudo-inverse (partial fractions and dsyevd).
* ``tol_snr > 0`` (``_sum_without_noise``, ``solver.py:114-125``) on the
  MinSR and SR branches, with and without reweighting, and across the row
  blocks in which the SR branch forms Obar.
* ``minsr_pinv_eig`` (``solver.py:262-294``) on indefinite and Gram matrices.
* the soft pseudo-inverse's Chebyshev route, forced on small cases and by
  size on larger ones.

Tolerance: FP64 solves 1e-9 relative (summation order against LAPACK; the
matrices here have condition numbers below 1e5 where the soft cutoff does
not act), the split-bf16 Gram 1e-4.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import sr as osr

pytestmark = pytest.mark.gpu

import quantax_amd as qtx  # noqa: E402
from quantax_amd import _lib  # noqa: E402

DEV = "cuda"


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _problem(seed, ns, np_, odt=np.float64, weights=False):
    rng = np.random.default_rng(seed)
    O = (rng.standard_normal((ns, np_)) + 0.5).astype(odt)
    e = rng.standard_normal(ns) - 2.0
    r = rng.uniform(0.5, 1.5, ns) if weights else None
    if r is not None:
        r = r / r.mean()
    return O, e, r


def _device_inputs(O, e, r):
    Od = to_dev(O)
    ns, np_ = O.shape
    od = _lib.QTX_F32 if O.dtype == np.float32 else _lib.QTX_F64
    mean = torch.empty(np_, dtype=torch.float64, device=DEV)
    _lib.call("qtx_colmean", _lib.stream(), od, Od.data_ptr(), ns, np_, mean.data_ptr())
    rd = None if r is None else to_dev(r.astype(np.float64))
    ed = to_dev(e.astype(np.float64))
    eb = torch.empty(ns, dtype=torch.float64, device=DEV)
    out2 = torch.empty(2, dtype=torch.float64, device=DEV)
    _lib.call("qtx_ebar", _lib.stream(), ed.data_ptr(), _lib.ptr(rd), ns, eb.data_ptr(), out2.data_ptr())
    return Od, mean, eb, rd


def _oracle(O, e, r):
    Obar, _ = osr.get_Obar(O.astype(np.float64), r)
    Ebar, _, _ = osr.get_Ebar(e, r)
    return Obar, Ebar


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("shape", [(120, 300), (300, 120)])
@pytest.mark.parametrize("odt", [np.float64, np.float32])
def test_forced_branches_match_oracle(shape, odt):
    ns, np_ = shape
    O, e, _ = _problem(1, ns, np_, odt)
    Od, mean, eb, _ = _device_inputs(O, e, None)
    A, b = _oracle(O, e, None)
    cases = [
        (qtx.optimizer.minnorm_shift_eig(rshift=0.0, ashift=1e-3),
         osr.minnorm_shift_eig(A, b, rshift=0.0, ashift=1e-3), "chol_shift"),
        (qtx.optimizer.lstsq_shift_eig(rshift=0.0, ashift=1e-3),
         osr.lstsq_shift_eig(A, b, rshift=0.0, ashift=1e-3), "chol_shift"),
        (qtx.optimizer.minnorm_pinv_eig(rtol=1e-4), osr.minnorm_pinv_eig(A, b, rtol=1e-4), "partial_fractions"),
        (qtx.optimizer.lstsq_pinv_eig(rtol=1e-4), osr.lstsq_pinv_eig(A, b, rtol=1e-4), "partial_fractions"),
        # small cutoff: the shifted Cholesky route on the shape's own branch;
        # the other branch's Gram is rank-deficient, so dsyevd
        (qtx.optimizer.minnorm_pinv_eig(rtol=1e-10), osr.minnorm_pinv_eig(A, b, rtol=1e-10),
         "shifted_cholesky" if ns < np_ else "syevd"),
        (qtx.optimizer.lstsq_pinv_eig(rtol=1e-10), osr.lstsq_pinv_eig(A, b, rtol=1e-10),
         "syevd" if ns < np_ else "shifted_cholesky"),
    ]
    for solver, ref, path in cases:
        x = qtx.optimizer.solve_step(Od, mean, eb, solver).cpu().numpy()
        info = qtx.optimizer.last_solve_info()
        err = _rel(x, ref)
        print(f"{shape} {odt.__name__} {solver.branch} mode {solver.mode} rtol {solver.rtol}: {info['path']} "
              f"err {err:.2e}")
        assert info["path"] == path, info
        assert err < 1e-9, err


@pytest.mark.parametrize("shape", [(120, 300), (300, 120)])
@pytest.mark.parametrize("weights", [False, True])
def test_chebyshev_route_matches_oracle(shape, weights):
    """The Chebyshev route (QTX_SOLVE_CHEB forces it below its size
    threshold) on both branches and shapes, the rank-deficient Grams
    included, with and without reweighting, at cutoffs across the route's
    range (rtol 1e-2 .. 1e-5), against the oracle's eigh form."""
    ns, np_ = shape
    O, e, r = _problem(3, ns, np_, np.float64, weights)
    Od, mean, eb, rd = _device_inputs(O, e, r)
    A, b = _oracle(O, e, r)
    for branch in ("minnorm", "lstsq"):
        for rtol in (1e-2, 1e-4, 1e-5):
            solver = getattr(qtx.optimizer, f"{branch}_pinv_eig")(rtol=rtol)
            solver.mode |= _lib.QTX_SOLVE_CHEB
            ref = getattr(osr, f"{branch}_pinv_eig")(A, b, rtol=rtol)
            x = qtx.optimizer.solve_step(Od, mean, eb, solver, rd).cpu().numpy()
            info = qtx.optimizer.last_solve_info()
            err = _rel(x, ref)
            print(f"{shape} weights {weights} {branch} rtol {rtol}: {info['path']} err {err:.2e}")
            assert info["path"] == "chebyshev", info
            assert err < 1e-9, err


def test_chebyshev_route_by_size_several_blocks():
    """From n = 2048 on EIGH_PINV takes the Chebyshev route by itself: 2500
    samples (three 1024-wide diagonal blocks of the factorisation, the last
    partial; a triangular inverse of the same size), against the oracle."""
    O, e, _ = _problem(4, 2500, 4000, np.float32)
    O[:, :2000] *= 0.05  # spread the spectrum
    Od, mean, eb, _ = _device_inputs(O, e, None)
    A, b = _oracle(O, e, None)
    for rtol in (1e-4, 1e-2):
        x = qtx.optimizer.solve_step(Od, mean, eb, qtx.optimizer.auto_pinv_eig(rtol=rtol, gram="fp64"))
        info = qtx.optimizer.last_solve_info()
        err = _rel(x.cpu().numpy(), osr.auto_pinv_eig(A, b, rtol=rtol))
        print(f"2500 x 4000 rtol {rtol}: {info['path']} err {err:.2e}")
        assert info["path"] == "chebyshev", info
        assert err < 1e-9, err


def test_chebyshev_route_split_bf16_gram_large():
    """8192 x 17000 with the split-bf16 Gram, whose rounding leaves T with
    eigenvalues slightly below zero: the series covers them down to -c / 6
    and otherwise (its last term grows) the solve takes the partial fractions
    (fallback "chebyshev_interval").  Either way the result equals the
    dsyevd route's on the same Gram (the dsyevd route is pinned to the
    oracle in the smaller cases)."""
    ns, np_ = 8192, 17000
    g = torch.Generator(device=DEV).manual_seed(5)
    O = torch.randn(ns, np_, device=DEV, dtype=torch.float32, generator=g)
    O *= torch.logspace(0, -3, np_, device=DEV, dtype=torch.float32)[None, :]
    mean = O.mean(0).double()
    eb = torch.randn(ns, device=DEV, dtype=torch.float64, generator=g)
    eb -= eb.mean()
    for rtol in (1e-4, 1e-5):
        x = qtx.optimizer.solve_step(O, mean, eb, qtx.optimizer.auto_pinv_eig(rtol=rtol, gram="bf16x3"))
        info = qtx.optimizer.last_solve_info()
        direct = qtx.optimizer.Solver(_lib.QTX_SOLVE_EIGH_PINV_DIRECT, rtol=rtol, gram="bf16x3")
        xd = qtx.optimizer.solve_step(O, mean, eb, direct)
        err = (torch.linalg.norm(x - xd) / torch.linalg.norm(xd)).item()
        print(f"8192 x 17000 split-bf16 rtol {rtol}: {info['path']} ({info['fallback']}) err vs dsyevd {err:.2e}")
        assert info["path"] in ("chebyshev", "partial_fractions"), info
        assert info["path"] == "chebyshev" or info["fallback"] == "chebyshev_interval", info
        assert err < 1e-10, err


def test_forced_minnorm_split_bf16_gram_with_more_samples():
    """The split-bf16 Gram on the MinSR branch with ns > np.  T = Obar Obar^T
    (600 x 600) has rank 199 here, so the Gram's rounding (~1e-6 of its
    largest entries) is amplified by lambda_max / shift in the null space:
    2e-3 relative at shift 1e-3 on an MI355X, hence shift 0.1 for the 1e-4
    bound of the split-bf16 paths (the FP64 Gram matches to 1e-12 at any
    shift above)."""
    O, e, _ = _problem(2, 600, 200, np.float32)
    Od, mean, eb, _ = _device_inputs(O, e, None)
    A, b = _oracle(O, e, None)
    x = qtx.optimizer.solve_step(Od, mean, eb, qtx.optimizer.minnorm_shift_eig(0.0, 0.1, gram="bf16x3"))
    err = _rel(x.cpu().numpy(), osr.minnorm_shift_eig(A, b, rshift=0.0, ashift=0.1))
    print(f"600 x 200 split-bf16 MinSR shift 0.1: err {err:.2e}")
    assert err < 1e-4, err


@pytest.mark.parametrize("shape", [(120, 300), (300, 120)])
@pytest.mark.parametrize("weights", [False, True])
def test_snr_truncation_matches_oracle(shape, weights):
    ns, np_ = shape
    O, e, r = _problem(3, ns, np_, np.float64, weights)
    Od, mean, eb, rd = _device_inputs(O, e, r)
    A, b = _oracle(O, e, r)
    for tol in (0.5, 2.0):
        for branch in ("auto", "minnorm", "lstsq"):
            fac = {"auto": qtx.optimizer.auto_pinv_eig, "minnorm": qtx.optimizer.minnorm_pinv_eig,
                   "lstsq": qtx.optimizer.lstsq_pinv_eig}[branch]
            ofac = {"auto": osr.auto_pinv_eig, "minnorm": osr.minnorm_pinv_eig, "lstsq": osr.lstsq_pinv_eig}[branch]
            x = qtx.optimizer.solve_step(Od, mean, eb, fac(rtol=1e-6, tol_snr=tol), rd).cpu().numpy()
            info = qtx.optimizer.last_solve_info()
            ref = ofac(A, b, rtol=1e-6, tol_snr=tol)
            err = _rel(x, ref)
            print(f"{shape} weights {weights} tol_snr {tol} {branch}: {info['path']}/{info['fallback']} "
                  f"err {err:.2e}, |x| / |x(tol 0)| = "
                  f"{np.linalg.norm(ref) / np.linalg.norm(ofac(A, b, rtol=1e-6)):.3f}")
            assert info["path"] == "syevd" and info["fallback"] == "snr_truncation", info
            assert err < 1e-9, err


def test_snr_truncation_sr_branch_across_obar_row_blocks():
    """140000 x 300 in FP32: the SR branch forms Obar in two row blocks of at
    most 256 MiB, and the SNR sums run over both."""
    ns, np_ = 140000, 300
    O, e, _ = _problem(4, ns, np_, np.float32)
    Od, mean, eb, _ = _device_inputs(O, e, None)
    A, b = _oracle(O, e, None)
    x = qtx.optimizer.solve_step(Od, mean, eb, qtx.optimizer.auto_pinv_eig(rtol=1e-8, tol_snr=1.5)).cpu().numpy()
    ref = osr.lstsq_pinv_eig(A, b, rtol=1e-8, tol_snr=1.5)
    err = _rel(x, ref)
    print(f"140000 x 300 SR tol_snr 1.5: err {err:.2e}")
    assert err < 1e-9, err


def test_minsr_pinv_eig_matches_oracle():
    rng = np.random.default_rng(5)
    n = 257
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    d = np.concatenate([rng.uniform(0.1, 3.0, 150), -rng.uniform(0.1, 3.0, n - 150)])
    T = (Q * d) @ Q.T
    T = 0.5 * (T + T.T)
    b = rng.standard_normal(n)
    A = rng.standard_normal((n, 700)) / 30.0
    G = A @ A.T
    for M in (T, G):
        for rtol, tol in ((1e-12, 0.0), (1e-3, 0.0), (1e-12, 0.8), (1e-3, 3.0)):
            x = qtx.optimizer.minsr_pinv_eig(rtol=rtol, tol_snr=tol)(to_dev(M), to_dev(b)).cpu().numpy()
            ref = osr.minsr_pinv_eig(M, b, rtol=rtol, tol_snr=tol)
            err = _rel(x, ref)
            print(f"minsr_pinv_eig n {n} rtol {rtol} tol_snr {tol}: err {err:.2e}")
            assert err < 1e-9, err


def test_solve_many_mixes_snr_and_plain_tasks():
    """qtx_minsr_solve_many over tasks with and without SNR truncation on
    separate streams gives each task's single-solve result."""
    specs = [((120, 300), 1.0), ((300, 120), 0.0), ((300, 120), 0.7), ((150, 400), 0.0)]
    streams = [torch.cuda.Stream() for _ in specs]
    keep, tasks, refs = [], [], []
    for k, ((ns, np_), tol) in enumerate(specs):
        O, e, _ = _problem(10 + k, ns, np_, np.float32)
        Od, mean, eb, _ = _device_inputs(O, e, None)
        solver = qtx.optimizer.auto_pinv_eig(rtol=1e-6, tol_snr=tol)
        refs.append(qtx.optimizer.solve_step(Od, mean, eb, solver).cpu().numpy())
        ws = torch.empty(solver.ws_bytes(ns, np_), dtype=torch.uint8, device=DEV)
        step = torch.empty(np_, dtype=torch.float64, device=DEV)
        t = _lib.QtxMinsrTask()
        t.stream = streams[k].cuda_stream
        t.o_dtype, t.mode, t.gram = _lib.QTX_F32, solver.cmode, _lib.QTX_GRAM_FP64
        t.O, t.mean, t.r, t.ebar = Od.data_ptr(), mean.data_ptr(), None, eb.data_ptr()
        t.ns, t.np = ns, np_
        t.rshift, t.ashift, t.rtol, t.atol, t.tol_snr = 0.0, 0.0, solver.rtol, 0.0, tol
        t.step, t.ws, t.ws_bytes = step.data_ptr(), ws.data_ptr(), ws.numel()
        keep += [Od, mean, eb, ws, step]
        tasks.append(t)
    torch.cuda.synchronize()
    arr = (_lib.QtxMinsrTask * len(tasks))(*tasks)
    _lib.call("qtx_minsr_solve_many", arr, len(tasks))
    torch.cuda.synchronize()
    for k, t in enumerate(arr):
        assert t.status == 0
        assert t.fallback == (5 if specs[k][1] > 0 else t.fallback)
        x = keep[5 * k + 4].cpu().numpy()
        assert _rel(x, refs[k]) < 1e-13, k


def test_snr_needs_its_workspace_flag():
    """tol_snr > 1e-6 in a task whose mode lacks QTX_SOLVE_SNR is refused
    before anything runs; so are both branch flags at once."""
    O, e, _ = _problem(6, 50, 80)
    Od, mean, eb, _ = _device_inputs(O, e, None)
    lib = _lib.lib()
    mode = _lib.QTX_SOLVE_EIGH_PINV
    ws = torch.empty(int(lib.qtx_minsr_ws_bytes(50, 80, mode, 0)), dtype=torch.uint8, device=DEV)
    step = torch.empty(80, dtype=torch.float64, device=DEV)
    t = _lib.QtxMinsrTask()
    t.stream = torch.cuda.current_stream().cuda_stream
    t.o_dtype, t.mode, t.gram = _lib.QTX_F64, mode, 0
    t.O, t.mean, t.r, t.ebar = Od.data_ptr(), mean.data_ptr(), None, eb.data_ptr()
    t.ns, t.np, t.rtol, t.tol_snr = 50, 80, 1e-6, 1.0
    t.step, t.ws, t.ws_bytes = step.data_ptr(), ws.data_ptr(), ws.numel()
    arr = (_lib.QtxMinsrTask * 1)(t)
    assert lib.qtx_minsr_solve_many(arr, 1) == _lib.QTX_EINVAL
    assert b"QTX_SOLVE_SNR" in lib.qtx_last_error()
    rc = lib.qtx_minsr_solve(C.c_void_p(t.stream), _lib.QTX_F64, C.c_void_p(Od.data_ptr()),
                             C.c_void_p(mean.data_ptr()), None, C.c_void_p(eb.data_ptr()), 50, 80,
                             mode | _lib.QTX_SOLVE_MINNORM | _lib.QTX_SOLVE_LSTSQ, 0, 0.0, 0.0, 1e-6, 0.0,
                             C.c_void_p(step.data_ptr()), C.c_void_p(ws.data_ptr()), ws.numel())
    assert rc == _lib.QTX_EINVAL


@pytest.mark.parametrize("shape,solver", [
    ((2500, 4000), "chol"), ((1500, 3000), "pf"), ((600, 1500), "syevd"), ((3000, 700), "snr"),
    ((1200, 3000), "small")])
def test_caller_scratch_matches_library_scratch(shape, solver):
    """QTX_SOLVE_BLAS_WS: the rocBLAS / rocSOLVER scratch taken from the tail
    of the caller's workspace (sized by qtx_minsr_blas_ws_bytes) gives the
    same update as scratch rocBLAS allocates itself, on the blocked Cholesky
    (n = 2500, three diagonal blocks), the partial fractions, dsyevd, the
    SR branch with SNR truncation and the shifted Cholesky route of the
    reference default cutoff (n = 1200, two diagonal blocks)."""
    ns, np_ = shape
    O, e, _ = _problem(7, ns, np_, np.float32)
    Od, mean, eb, _ = _device_inputs(O, e, None)
    mk = {"chol": lambda: qtx.optimizer.auto_shift_eig(rshift=0.0, ashift=1e-3),
          "pf": lambda: qtx.optimizer.auto_pinv_eig(rtol=1e-4),
          "syevd": lambda: qtx.optimizer.auto_pinv_eig(rtol=1e-10),
          "snr": lambda: qtx.optimizer.auto_pinv_eig(rtol=1e-8, tol_snr=1.0),
          "small": lambda: qtx.optimizer.auto_pinv_eig()}[solver]
    s_lib, s_user = mk(), mk()
    if solver == "syevd":
        s_lib.mode = s_user.mode = _lib.QTX_SOLVE_EIGH_PINV_DIRECT
    s_lib.caller_scratch = False
    own = int(_lib.lib().qtx_minsr_ws_bytes(ns, np_, s_user.cmode, 0))
    total = s_user.ws_bytes(ns, np_)
    print(f"{shape} {solver}: own {own} B, rocBLAS/rocSOLVER scratch {total - own} B")
    x_lib = qtx.optimizer.solve_step(Od, mean, eb, s_lib).cpu().numpy()
    info_lib = qtx.optimizer.last_solve_info()
    x_user = qtx.optimizer.solve_step(Od, mean, eb, s_user).cpu().numpy()
    info_user = qtx.optimizer.last_solve_info()
    assert info_lib["path"] == info_user["path"]
    # rocBLAS may pick other kernels with a user-owned workspace: rounding only
    err = _rel(x_user, x_lib)
    print(f"  caller scratch against library scratch: {err:.2e}")
    assert err < 1e-12, err


def test_caller_scratch_too_small_is_refused():
    """A workspace tail below the rocSOLVER need is a workspace error, not a
    fault: dsyevd at n = 600 with a 256-byte tail (the smallest accepted;
    without the tail the call is refused before anything runs)."""
    ns, np_ = 600, 1500
    O, e, _ = _problem(8, ns, np_, np.float64)
    Od, mean, eb, _ = _device_inputs(O, e, None)
    s = qtx.optimizer.auto_pinv_eig(rtol=1e-10)
    s.mode = _lib.QTX_SOLVE_EIGH_PINV_DIRECT  # dsyevd always
    lib = _lib.lib()
    own = int(lib.qtx_minsr_ws_bytes(ns, np_, s.cmode, 0))
    assert s.ws_bytes(ns, np_) > own  # dsyevd needs scratch
    step = torch.empty(np_, dtype=torch.float64, device=DEV)
    ws0 = torch.empty(own, dtype=torch.uint8, device=DEV)
    rc0 = lib.qtx_minsr_solve(C.c_void_p(torch.cuda.current_stream().cuda_stream), _lib.QTX_F64,
                              C.c_void_p(Od.data_ptr()), C.c_void_p(mean.data_ptr()), None, C.c_void_p(eb.data_ptr()),
                              ns, np_, s.cmode, 0, 0.0, 0.0, 1e-10, 0.0, C.c_void_p(step.data_ptr()),
                              C.c_void_p(ws0.data_ptr()), ws0.numel())
    assert rc0 == _lib.QTX_EWORKSPACE
    ws = torch.empty(own + 256, dtype=torch.uint8, device=DEV)
    rc = lib.qtx_minsr_solve(C.c_void_p(torch.cuda.current_stream().cuda_stream), _lib.QTX_F64,
                             C.c_void_p(Od.data_ptr()), C.c_void_p(mean.data_ptr()), None, C.c_void_p(eb.data_ptr()),
                             ns, np_, s.cmode, 0, 0.0, 0.0, 1e-10, 0.0, C.c_void_p(step.data_ptr()),
                             C.c_void_p(ws.data_ptr()), ws.numel())
    torch.cuda.synchronize()
    assert rc == _lib.QTX_EWORKSPACE, (rc, lib.qtx_last_error())
