"""VMC-step benchmark on one MI355X.

Workload (default): J1-J2 Heisenberg (J2 = 0.5, Marshall sign) on a 16x16
periodic square lattice, Sz = 0; ResConv(8 blocks, 16 channels, 3x3) with
He-normal weights; 16384 chains (one sample per chain, the reference's
semantics) advanced by SpinExchange for 2N = 512 proposals per sweep; MinSR
with a 16384 x 16384 T solved by FP64 eigh (soft pseudo-inverse).  One step =
sweep -> E_loc -> O -> Gram + eigh -> update, i.e. ``Euler.step``.

Modes: "mixed" = split-bf16 MFMA convolutions and split-bf16 MFMA Gram (FP32
accumulation), FP64 centring/eigh/update; "plain" = FP32 MFMA convolutions and
FP64 MFMA Gram.  The JSON line reports the mixed mode as ``value`` and the
plain mode beside it.

    python bench.py [--steps K] [--warmup W] [--mode mixed|plain|both] [--config c3|c2|c1|c4]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
MFMA_PEAK = {"bf16": 2500.0, "fp32": 157.3, "fp64": 78.6}  # dense TFLOP/s


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--mode", default="both", choices=["mixed", "plain", "both"])
    p.add_argument("--config", default="c3", choices=["c3", "c2", "c1", "c4"])
    p.add_argument("--sweep-steps", type=int, default=None,
                   help="proposals per chain per sweep (default 2N; 8 for c4)")
    p.add_argument("--thermal-sweeps", type=int, default=None, help="thermalisation sweeps (default 1; 0 for c4)")
    p.add_argument("--solver", default=None, choices=["eigh", "eigh-direct", "chol"],
                   help="eigh: soft pseudo-inverse without eigenvectors when the cutoff allows (Chebyshev "
                        "series from n = 2048, else partial fractions; shifted Cholesky for small cutoffs); "
                        "eigh-direct: the same through dsyevd; chol: diagonal-shift Cholesky "
                        "(default: the configuration's own, chol for c1/c2, eigh for c3/c4)")
    p.add_argument("--rtol", type=float, default=1e-4, help="soft pseudo-inverse cutoff of the eigh solver")
    p.add_argument("--lr", type=float, default=None, help="learning rate (default: the configuration's own)")
    p.add_argument("--no-ref-solve", action="store_true",
                   help="skip the solve with the reference's default cutoff rtol = 1e-12 (dsyevd route)")
    p.add_argument("--profile-kernels", action="store_true",
                   help="kernel profiling run: one sweep of --sweep-steps proposals (default 16) and E_loc of "
                        "--profile-eloc-samples samples, nothing else; prints one JSON line of HIP-event times")
    p.add_argument("--profile-eloc-samples", type=int, default=2048)
    p.add_argument("--precision-breakdown", action="store_true",
                   help="also attribute the mixed update-vector error to E_loc and to the Gram")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-serve", action="store_true", help="skip the served job stream")
    p.add_argument("--serve-jobs", type=int, default=64)
    p.add_argument("--serve-steps", type=int, default=50, help="MinSR steps per served job")
    p.add_argument("--serve-gap", type=float, default=0.25, help="mean Poisson inter-arrival time (s)")
    p.add_argument("--serve-max-active", type=int, default=12,
                   help="cap on concurrently served jobs (0: no cap); the cap lowers median latency")
    p.add_argument("--serve-seq-steps", type=int, default=0,
                   help="steps per job timed for the sequential baselines (0: every step of every job; "
                        "otherwise the rest is extrapolated at the job's later-step cost)")
    p.add_argument("--activation", default="gelu", choices=["gelu", "relu"],
                   help="ResConv activation: gelu (the reference's tanh-form GELU) or the opt-in ReLU variant")
    p.add_argument("--no-relu-variant", action="store_true",
                   help="c3 with GELU in both modes: skip the same step with the opt-in ReLU activation, "
                        "reported beside the headline as relu_variant")
    p.add_argument("--calib-copy-mb", type=int, default=0,
                   help="run a device-to-device copy of this many MiB first (HBM counter calibration)")
    return p.parse_args()


CONFIGS = {
    # name: (L, lattice kind, model kind, model args, ns, J, n_neighbor)
    "c3": dict(L=16, model="resconv", nblocks=8, channels=16, ns=16384, J=[1.0, 0.5], nn=[1, 2], msr=True,
               solver="eigh", lr=1e-3,
               workload="J1-J2 16x16 (J2=0.5, MSR, Sz=0), ResConv(8 blocks, 16 ch), 16384 chains, MinSR eigh"),
    "c2": dict(L=10, model="rbm", features=400, ns=4096, J=[1.0], nn=[1], msr=True, solver="chol", lr=1e-2,
               workload="Heisenberg 10x10 (MSR, Sz=0), RBM_Dense(alpha=4) float32, 4096 chains, MinSR with "
                        "diagonal shift 1e-4 (Cholesky)"),
    "c1": dict(L=16, model="rbm1d", features=16, ns=1024, solver="chol", lr=1e-2,
               workload="TFIM chain L=16 (h=1), RBM_Dense(16) float64, 1024 chains, SR with diagonal shift 1e-3"),
    # the large case: sampling is cut to 8 proposals per chain and no
    # thermalisation (a full 2N sweep is 6 PFLOP of forwards); E_loc, O, the
    # 32768 x 32768 Gram and the solve run in full
    "c4": dict(L=20, model="resconv", nblocks=16, channels=32, ns=32768, J=[1.0, 0.5], nn=[1, 2], msr=True,
               sweep_steps=8, thermal_sweeps=0, nconn_est=800, solver="eigh", lr=1e-3,
               workload="J1-J2 20x20 (J2=0.5, MSR, Sz=0), ResConv(16 blocks, 32 ch), 32768 chains, MinSR "
                        "(T 32768^2); sampling cut to 8 proposals per chain, no thermalisation"),
}


# HBM traffic of one rc_metropolis_kernel launch at the 16x16 workload (mixed,
# 512 steps, 16384 chains): NOT measured by this run.  It comes from two
# rocprofv3 --pmc runs of this script (FETCH_SIZE, then WRITE_SIZE), each
# scaled by what the counter reads for a 256 MiB device copy; the commands,
# raw counts and the calibration are in the profile named below, and the
# bench line reports the figure with that source.
PMC_C3_MIXED = {"fetch_kib": 68745.5, "write_kib": 4352.0, "fetch_scale": 262144.0 / 131092.8125, "write_scale": 1.0,
                "source": "profiles/rc_metropolis_16x16_c16_nb8_hbm_traffic.md"}


def conv_flops_per_forward(nblocks, C, N):
    """Algorithmic FLOPs of one ResConv forward (multiply-add = 2)."""
    f = 2 * C * 1 * 9 * N  # first conv (Cin = 1)
    f += (2 * nblocks - 1) * 2 * C * C * 9 * N
    return f


class Stage:
    """HIP-event timer on the current torch stream."""

    def __init__(self, torch):
        self.torch = torch
        self.t = {}

    def run(self, name, fn):
        t = self.torch
        a, b = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        a.record()
        out = fn()
        b.record()
        self.t.setdefault(name, []).append((a, b))
        return out

    def ms(self):
        self.torch.cuda.synchronize()
        return {k: float(np.mean([a.elapsed_time(b) for a, b in v])) for k, v in self.t.items()}


def build_problem(cfg, mode, args, qtx):
    L = cfg["L"]
    qtx.set_random_seed(0)
    if cfg["model"] == "rbm1d":
        lat = qtx.sites.Chain(L)
        H = qtx.operator.Ising(h=1.0)
        model = qtx.model.RBM_Dense(cfg["features"], dtype=np.float64, init_std=0.01, seed=0)
        state = qtx.state.Variational(model)
        sampler = qtx.sampler.LocalFlip(state, cfg["ns"], thermal_steps=args.thermal_sweeps * 2 * L, seed=1,
                                        sweep_steps=args.sweep_steps)
        solver = qtx.optimizer.lstsq_shift_eig(rshift=0.0, ashift=1e-3)
        return lat, H, state, sampler, solver
    N = L * L
    lat = qtx.sites.Square(L, Nparticles=(N // 2, N // 2))
    H = qtx.operator.Heisenberg(J=cfg["J"], n_neighbor=cfg["nn"], msr=cfg["msr"])
    gram = "bf16x3" if mode == "mixed" else "fp64"
    if cfg["model"] == "resconv":
        model = qtx.model.ResConv(cfg["nblocks"], cfg["channels"], precision="bf16x3" if mode == "mixed" else "fp32",
                                  seed=0, gelu="relu" if args.activation == "relu" else "tanh")
    else:
        model = qtx.model.RBM_Dense(cfg["features"], dtype=np.float32, init_std=0.01, seed=0)
        gram = "bf16x3" if mode == "mixed" else "fp64"
    state = qtx.state.Variational(model)
    sampler = qtx.sampler.SpinExchange(state, cfg["ns"], thermal_steps=args.thermal_sweeps * 2 * N,
                                       sweep_steps=args.sweep_steps, seed=1)
    if args.solver in ("eigh", "eigh-direct"):
        # soft eigenvalue cutoff at 1e-4 of the largest eigenvalue: the default
        # 1e-12 gives steps of norm ~1e3 on a freshly initialised deep ResConv
        solver = qtx.optimizer.auto_pinv_eig(rtol=args.rtol, gram=gram)
        if args.solver == "eigh-direct":
            solver.mode = qtx._lib.QTX_SOLVE_EIGH_PINV_DIRECT
    else:
        solver = qtx.optimizer.auto_shift_eig(rshift=0.0, ashift=1e-4, gram=gram)
    return lat, H, state, sampler, solver


def run_mode(cfg, mode, args, torch, qtx):
    lat, H, state, sampler, solver = build_problem(cfg, mode, args, qtx)
    opt = qtx.optimizer.SR(state, H, solver=solver)
    stages = Stage(torch)
    lr = args.lr

    last = {}

    def vmc_step():
        samples = stages.run("sweep", sampler.sweep)
        ebar = stages.run("eloc", lambda: opt.get_Ebar(samples))
        O, mean = stages.run("jacobian", lambda: opt.get_Obar(samples))
        step = stages.run("solve", lambda: qtx.optimizer.solve_step(O, mean, ebar, solver, None, opt._ws))
        last["solve_info"] = qtx.optimizer.last_solve_info()
        last.update(samples=samples, O=O, mean=mean, ebar=ebar, step=step, energy=opt.energy,
                    step_norm=float(torch.linalg.norm(step).item()))
        stages.run("update", lambda: state.update(lr * step))
        return opt

    for _ in range(args.warmup):
        vmc_step()
        print(f"[bench] {mode}: warmup step done, E = {opt.energy:.6f}, |step| = {last['step_norm']:.3e}",
              file=sys.stderr, flush=True)
    stages.t.clear()
    sampler.kernel_events.clear()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    energies = []
    kept = []  # each timed step's (O, mean, Ebar, step) for the per-step Gram-precision check below
    per_step = mode == "mixed" and cfg["model"] == "resconv" and cfg["L"] == 16
    for i in range(args.steps):
        vmc_step()
        energies.append(opt.energy)
        if per_step:
            kept.append((last["O"], last["mean"], last["ebar"], last["step"]))
        print(f"[bench] {mode}: timed step {i + 1}/{args.steps}, E = {opt.energy:.6f}, "
              f"|step| = {last['step_norm']:.3e}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kms = [a.elapsed_time(b) for a, b in sampler.kernel_events]
    out = dict(ms_per_step=1e3 * dt / args.steps, steps_per_s=args.steps / dt, stages=stages.ms(),
               metropolis_kernel_ms=float(np.mean(kms)) if kms else None,
               energies=energies, acceptance=sampler.acceptance, nconn=float(H._connectivity.float().mean().item()),
               nparams=state.nparams, solve_info=last.get("solve_info"))
    out["eloc_ms"] = out["stages"]["eloc"]
    if mode == "mixed" and cfg["model"] == "resconv":
        if cfg["L"] == 16:
            # per timed step: the split-bf16 solve against the same route with the
            # FP64 Gram on that step's own O and E_loc (after the timed region)
            ref_solver = qtx.optimizer.Solver(solver.mode, solver.rshift, solver.ashift, solver.rtol, solver.atol,
                                              "fp64")
            errs = []
            for O_k, m_k, e_k, x_k in kept:
                x_ref = qtx.optimizer.solve_step(O_k, m_k, e_k, ref_solver)
                errs.append((torch.linalg.norm(x_k - x_ref) / torch.linalg.norm(x_ref)).item())
            kept.clear()
            out["per_step_gram_err"] = errs
            out["gram"] = gram_rate(last["O"], last["mean"], torch)
            out["precision"] = precision_report(cfg, state, H, solver, last, torch, qtx, args.precision_breakdown)
            if not args.no_ref_solve:
                out["reference_default_solve"] = reference_default_solve(last, solver, torch, qtx)
        else:
            out["gram"] = gram_report(last, solver, torch, qtx, opt._ws)
    return out


def precision_report(cfg, state, H, solver, last, torch, qtx, breakdown=False):
    """Errors of the mixed path against the FP32-network / FP64-Gram path,
    both evaluated on the last timed step's spins at the current parameters."""
    model = state.model
    s = last["samples"].spins
    la_m = model.logpsi(s)
    e_m, _ = model.eloc(H, s, la_m)
    model.set_precision("fp32")
    la_p = model.logpsi(s)
    e_p, _ = model.eloc(H, s, la_p)
    model.set_precision("bf16x3")
    O = model.jacobian(s)
    mean = torch.empty(O.shape[1], dtype=torch.float64, device=O.device)
    from quantax_amd import _lib
    _lib.call("qtx_colmean", _lib.stream(), _lib.QTX_F32, O.data_ptr(), O.shape[0], O.shape[1], mean.data_ptr())

    def ebar_of(e):
        eb = torch.empty_like(e)
        o2 = torch.empty(2, dtype=torch.float64, device=e.device)
        _lib.call("qtx_ebar", _lib.stream(), e.data_ptr(), None, e.shape[0], eb.data_ptr(), o2.data_ptr())
        return eb, o2[0].item()

    eb_m, E_m = ebar_of(e_m)
    eb_p, E_p = ebar_of(e_p)
    ref_mode = qtx._lib.QTX_SOLVE_EIGH_PINV_DIRECT if solver.mode == qtx._lib.QTX_SOLVE_EIGH_PINV else solver.mode
    ref_solver = qtx.optimizer.Solver(ref_mode, solver.rshift, solver.ashift, solver.rtol, solver.atol, "fp64")
    x_m = qtx.optimizer.solve_step(O, mean, eb_m, solver)
    x_p = qtx.optimizer.solve_step(O, mean, eb_p, ref_solver)
    def rel(a):
        return (torch.linalg.norm(a - x_p) / torch.linalg.norm(x_p)).item()

    rep = {"reference": "same spins and parameters: FP32-MFMA network E_loc, FP64-MFMA Gram, FP64 eigh",
           "update_vector_rel_err": rel(x_m), "energy_rel_err": abs(E_m - E_p) / abs(E_p),
           "logpsi_max_abs_diff": (la_m - la_p).abs().max().item()}
    if breakdown:
        rep["err_from_eloc_only"] = rel(qtx.optimizer.solve_step(O, mean, eb_m, ref_solver))
        rep["err_from_gram_only"] = rel(qtx.optimizer.solve_step(O, mean, eb_p, solver))
    return rep


def reference_default_solve(last, solver, torch, qtx):
    """The last step's solve with the reference's default cutoff rtol = 1e-12
    (``optimizer/solver.py:12-21``), below 1e-5 lambda_max: the shifted
    Cholesky route when T allows it, else dsyevd; and the same solve forced
    through dsyevd (EIGH_PINV_DIRECT) for its time and the two results'
    relative difference.  HIP events around each call, the step's own O and
    Ebar, the faster of two calls."""
    def timed(s):
        ms, x = [], None
        for _ in range(2):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            x = qtx.optimizer.solve_step(last["O"], last["mean"], last["ebar"], s)
            b.record()
            torch.cuda.synchronize()
            ms.append(a.elapsed_time(b))
        return min(ms), x, qtx.optimizer.last_solve_info()

    ref = qtx.optimizer.Solver(solver.mode, solver.rshift, solver.ashift, 1e-12, 0.0, solver.gram)
    ms, x, info = timed(ref)
    direct = qtx.optimizer.Solver(qtx._lib.QTX_SOLVE_EIGH_PINV_DIRECT, solver.rshift, solver.ashift, 1e-12, 0.0,
                                  solver.gram)
    ms_d, x_d, _ = timed(direct)
    torch.cuda.empty_cache()
    rep = {"rtol": 1e-12, "gram": solver.gram, "ms": round(ms, 1), "path": info["path"],
           "fallback": info["fallback"], "dsyevd_ms": round(ms_d, 1),
           "rel_diff_vs_dsyevd": (torch.linalg.norm(x - x_d) / torch.linalg.norm(x_d)).item(),
           "note": "not part of the timed step, which uses rtol = "
           f"{solver.rtol:g} (its route in solve_path)"}
    print(f"[bench] reference-default solve (rtol 1e-12): {rep['ms']} ms via {rep['path']}; dsyevd "
          f"{rep['dsyevd_ms']} ms, relative difference {rep['rel_diff_vs_dsyevd']:.2e}", file=sys.stderr, flush=True)
    return rep


def gram_rate(O, mean, torch):
    """MFMA rate of the split-bf16 Gram entry point (centring/split pass + MFMA
    Gram) on a step's O: two launches timed by HIP events, the faster kept."""
    from quantax_amd import _lib

    ns, np_ = O.shape
    g = _lib.QTX_GRAM_BF16X3
    nws = int(_lib.lib().qtx_gram_ws_bytes(ns, np_, g))
    T = torch.empty((ns, ns), dtype=torch.float64, device=O.device)
    ws = torch.empty(nws, dtype=torch.uint8, device=O.device)
    ev = []
    for _ in range(2):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        _lib.call("qtx_gram", _lib.stream(), _lib.QTX_F32, O.data_ptr(), mean.data_ptr(), None, ns, np_, g,
                  T.data_ptr(), ws.data_ptr(), nws)
        b.record()
        ev.append((a, b))
    torch.cuda.synchronize()
    ms = min(a.elapsed_time(b) for a, b in ev)
    del T, ws
    torch.cuda.empty_cache()
    useful = float(ns) * ns * np_  # lower triangle (the kernel computes 128 x 128 tiles on and below the diagonal)
    rep = {"kernel": "qtx_gram (split-bf16: centring/split pass + MFMA Gram)", "ms": round(ms, 2),
           "algorithmic_flops": useful, "achieved_tflops": round(useful / ms / 1e9, 1),
           "bf16_mfma_issue_tflops": round(3 * useful / ms / 1e9, 1),
           "bf16_mfma_utilisation": round(3 * useful / ms / 1e9 / MFMA_PEAK["bf16"], 4),
           "flop_convention": "Ns^2 Np (SYRK-halved, one multiply-add = 2 flops over the lower triangle); "
                              "issue = 3 split products"}
    print(f"[bench] gram: {ms:.1f} ms, {rep['achieved_tflops']} TFLOP/s algorithmic", file=sys.stderr, flush=True)
    return rep


def gram_report(last, solver, torch, qtx, ws_solve):
    """The large case: the Gram's MFMA rate on the last step's O and the
    update-vector error of the mixed solve against the same route with the
    FP64 Gram, both on the last step's O and E_loc (the errors of the network
    precision are measured on the 16x16 workload)."""
    O, mean, eb = last["O"], last["mean"], last["ebar"]
    rep = gram_rate(O, mean, torch)
    ref_solver = qtx.optimizer.Solver(solver.mode, solver.rshift, solver.ashift, solver.rtol, solver.atol, "fp64")
    x_m = qtx.optimizer.solve_step(O, mean, eb, solver, None, ws_solve)
    x_p = qtx.optimizer.solve_step(O, mean, eb, ref_solver)
    torch.cuda.empty_cache()
    rep["update_vector_rel_err"] = (torch.linalg.norm(x_m - x_p) / torch.linalg.norm(x_p)).item()
    rep["reference"] = "same O and E_loc, FP64 MFMA Gram, same soft-inverse route"
    return rep


def run_serving(args, torch):
    """The served job stream (64 J1-J2 RBM jobs, Poisson arrivals) against the
    same jobs run one after another through the single-job path, once with
    the FP64 Gram (the plain path) and once with the same split-bf16 Gram as
    the server, so batching and precision gains are reported apart.  Every
    step of every job is timed unless --serve-seq-steps caps it."""
    from quantax_amd import serve

    specs = serve.c5_stream(args.serve_jobs, mean_interarrival=args.serve_gap, nsteps=args.serve_steps, seed=5)
    log = lambda m: print(f"[bench] serve: {m}", file=sys.stderr, flush=True)
    srv = serve.Server(progress=log, max_active=args.serve_max_active or None)
    rep = srv.run(specs)
    rounds = srv.round_summary()
    srv.close()
    log(f"served: {rep.summary()}, rounds: {rounds}")
    nseq = args.serve_seq_steps or None
    seq = {}
    for gram in ("fp64", "bf16x3"):
        q = serve.run_sequential(specs, gram=gram, nsteps=nseq, progress=log)
        seq[gram] = q.summary()
        log(f"sequential ({gram} Gram): {seq[gram]}")
    how = ("every step of every job timed" if nseq is None else
           f"{nseq} steps per job timed, the rest extrapolated at the job's later-step cost")
    s = rep.summary()
    return {
        "workload": (f"{args.serve_jobs} J1-J2 RBM_Dense jobs (L in 6..16, alpha in 1,2,4, ns in 1024..8192, "
                     f"J2 in [0, 0.6]), Poisson arrivals, mean gap {args.serve_gap} s, {args.serve_steps} MinSR "
                     f"steps each (split-bf16 Gram, soft pseudo-inverse rtol 1e-4, lr 1e-2), at most "
                     f"{args.serve_max_active or 'unlimited'} jobs in flight"),
        "served": s,
        "rounds": {**rounds, "note": "per round: host time enqueueing the three grouped launches and the per-job "
                                     "Ebar/Omean/update launches; host time inside qtx_minsr_solve_many "
                                     "(its enqueues and the readbacks it waits on); device time of the grouped "
                                     "launches (HIP events)"},
        "sequential_plain": {**seq["fp64"], "note": f"single-job path, FP64 Gram, {how}; latency = the job's own run"},
        "sequential_mixed": {**seq["bf16x3"], "note": f"single-job path, split-bf16 Gram (as served), {how}"},
        "served_over_sequential": round(s["steps_per_s"] / seq["fp64"]["steps_per_s"], 3),
        "served_over_sequential_mixed": round(s["steps_per_s"] / seq["bf16x3"]["steps_per_s"], 3),
    }


def cpu_baseline(cfg, seconds, sweep_steps):
    """CPU baseline on the host cores, rank 0: the oracle's FP32 ResConv
    forward through torch's CPU convolution (oracle/cpu_resconv.py, the kind
    of compiled CPU convolution a JAX-CPU run of the reference uses), timed on
    a bounded sample of the workload's configurations and converted to VMC
    steps/s by the forwards one step needs (sweep proposals + connected
    configurations; the forwards are >99 % of a step's arithmetic).  The
    float64 numpy restatement is timed beside it as a secondary figure."""
    import oracle.models as om
    from oracle import cpu_resconv
    import torch

    cores = len(os.sched_getaffinity(0))
    rng = np.random.default_rng(0)
    L = cfg["L"]
    if cfg["model"] != "resconv":
        return None
    N = L * L
    spec = om.ResConvSpec(cfg["nblocks"], cfg["channels"], (L, L))
    theta = spec.he_normal(rng).astype(np.float64)
    ns = cfg["ns"]
    # sweep proposals + connected configurations per sample (~550 at 16x16, ~800 at 20x20)
    fwd_per_step = ns * sweep_steps + ns * cfg.get("nconn_est", 550)

    def rate(fn, batch, budget):
        s = (2 * rng.integers(0, 2, size=(batch, N)) - 1).astype(np.int8)
        fn(s)  # warm-up (allocator, thread pool)
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget:
            fn(s)
            n += batch
        return n, time.perf_counter() - t0

    threads = torch.get_num_threads()
    torch.set_num_threads(cores)
    try:
        n, dt = rate(lambda s: cpu_resconv.resconv_forward_f32(spec, theta, s), 256, seconds)
    finally:
        torch.set_num_threads(threads)
    fwd = n / dt
    n64, dt64 = rate(lambda s: om.resconv_forward(spec, theta, s), 16, max(3.0, seconds / 3))
    fwd64 = n64 / dt64
    flops = conv_flops_per_forward(cfg["nblocks"], cfg["channels"], N)
    return dict(value=fwd / fwd_per_step, unit="VMC steps/s (extrapolated)", cores=cores, kind="port",
                sample=f"{n} ResConv forwards in FP32 through torch's CPU conv2d (circular padding, "
                       f"oracle/cpu_resconv.py) in {dt:.1f} s on {cores} threads: {fwd:.0f} forwards/s = "
                       f"{fwd * flops / 1e9:.1f} GFLOP/s, scaled by {fwd_per_step} forwards per VMC step",
                secondary={"kind": "float64 numpy oracle (oracle/models.py)", "forwards_per_s": round(fwd64, 1),
                           "value": fwd64 / fwd_per_step, "gflops": round(fwd64 * flops / 1e9, 2)})


def dtype_label(cfg, mode):
    """Arithmetic types of the step: network, Gram, solve."""
    if cfg["model"] == "rbm1d":
        return "fp64"  # float64 RBM; the SR branch (ns >= np) forms S in FP64
    net = "bf16x3" if (mode == "mixed" and cfg["model"] == "resconv") else "fp32"
    gram = "bf16x3" if mode == "mixed" else "fp64"
    return f"{net} network + {gram} Gram + fp64 solve"


def profile_kernels(cfg, args, torch, qtx):
    """For rocprofv3 runs: build the configuration's mixed (or --mode plain)
    ResConv problem with no thermalisation, run one sweep of --sweep-steps
    proposals (16 by default) over all chains and E_loc of the first
    --profile-eloc-samples samples; print the HIP-event times."""
    mode = "plain" if args.mode == "plain" else "mixed"
    args.thermal_sweeps = 0
    if args.sweep_steps is None:
        args.sweep_steps = 16
    lat, H, state, sampler, solver = build_problem(cfg, mode, args, qtx)
    samples = sampler.sweep()
    torch.cuda.synchronize()
    ms_sweep = sampler.kernel_events[-1][0].elapsed_time(sampler.kernel_events[-1][1])
    n = min(args.profile_eloc_samples, cfg["ns"])
    s, la = samples.spins[:n].contiguous(), samples.psi.logabs[:n].contiguous()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    e, nc = state.model.eloc(H, s, la)
    b.record()
    torch.cuda.synchronize()
    N = cfg["L"] ** 2
    fwd = conv_flops_per_forward(cfg["nblocks"], cfg["channels"], N)
    ms_e = a.elapsed_time(b)
    nconn = float(nc.float().mean().item())
    print(json.dumps({"profile": args.config, "mode": mode, "sweep_steps": args.sweep_steps, "chains": cfg["ns"],
                      "metropolis_ms": round(ms_sweep, 3),
                      "metropolis_tflops": round(cfg["ns"] * args.sweep_steps * fwd / ms_sweep / 1e9, 1),
                      "eloc_samples": n, "eloc_ms": round(ms_e, 3), "mean_nconn": round(nconn, 1),
                      "eloc_tflops": round(n * nconn * fwd / ms_e / 1e9, 1)}))


def relu_variant(cfg, args, torch, qtx, flops_per_sweep):
    """The same VMC step, both modes, with the opt-in ReLU activation (the
    reference's network uses GELU); timed like the headline."""
    va = argparse.Namespace(**vars(args))
    va.activation = "relu"
    va.no_ref_solve = True
    res = {}
    for m in ("mixed", "plain"):
        print(f"[bench] building {args.config} {m} (ReLU variant)", file=sys.stderr, flush=True)
        res[m] = run_mode(cfg, m, va, torch, qtx)
    r, p = res["mixed"], res["plain"]
    ach = flops_per_sweep / (r["metropolis_kernel_ms"] / 1e3) / 1e12
    out = {"activation": "relu", "value": round(r["steps_per_s"], 5), "ms_per_step": round(r["ms_per_step"], 2),
           "stages_ms": {k: round(v, 2) for k, v in r["stages"].items()},
           "plain": {"value": round(p["steps_per_s"], 5), "ms_per_step": round(p["ms_per_step"], 2),
                     "stages_ms": {k: round(v, 2) for k, v in p["stages"].items()}},
           "mixed_over_plain": round(r["steps_per_s"] / p["steps_per_s"], 3),
           "metropolis_kernel": {"avg_launch_ms": round(r["metropolis_kernel_ms"], 3),
                                 "achieved_tflops": round(ach, 2), "frac": round(ach / MFMA_PEAK["bf16"], 4)}}
    if "precision" in r:
        out["precision"] = {k: r["precision"][k] for k in ("update_vector_rel_err", "energy_rel_err",
                                                           "logpsi_max_abs_diff") if k in r["precision"]}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import torch

    import quantax_amd as qtx
    from quantax_amd import _lib

    _lib.device()
    if args.calib_copy_mb > 0:
        n = args.calib_copy_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] calibration copy of {args.calib_copy_mb} MiB done", file=sys.stderr, flush=True)
        del src, dst
    cfg = CONFIGS[args.config]
    if args.solver is None:
        args.solver = cfg["solver"]
    if args.lr is None:
        args.lr = cfg["lr"]
    if args.profile_kernels:
        profile_kernels(cfg, args, torch, qtx)
        return
    if args.sweep_steps is None:
        args.sweep_steps = cfg.get("sweep_steps")
    if args.thermal_sweeps is None:
        args.thermal_sweeps = cfg.get("thermal_sweeps", 1)
    modes = ["mixed", "plain"] if args.mode == "both" else [args.mode]
    res = {}
    for m in modes:
        print(f"[bench] building {args.config} {m}", file=sys.stderr, flush=True)
        res[m] = run_mode(cfg, m, args, torch, qtx)
        r = res[m]
        print(f"[bench] {m}: {r['ms_per_step']:.1f} ms/step, stages {r['stages']}", file=sys.stderr, flush=True)
    main_mode = modes[0]
    r = res[main_mode]
    N = cfg["L"] ** 2 if cfg["model"] != "rbm1d" else cfg["L"]
    sweep_steps = args.sweep_steps or 2 * N
    roof = None
    if cfg["model"] == "resconv" and args.config == "c4":
        # sampling is cut here, so the dominant kernel is the fused E_loc (one
        # forward per valid connected configuration); HIP events on its stream
        fwd = conv_flops_per_forward(cfg["nblocks"], cfg["channels"], N)
        flops = cfg["ns"] * r["nconn"] * fwd
        ach = flops / (r["eloc_ms"] / 1e3) / 1e12
        peak = MFMA_PEAK["bf16"] if main_mode == "mixed" else MFMA_PEAK["fp32"]
        roof = {"bound": "mfma", "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(ach / peak, 4), "traffic": None, "kernel": "rc_eloc_kernel",
                "flops_per_launch": flops, "avg_launch_ms": round(r["eloc_ms"], 3)}
    elif cfg["model"] == "resconv":
        fwd = conv_flops_per_forward(cfg["nblocks"], cfg["channels"], N)
        flops = cfg["ns"] * sweep_steps * fwd  # one Metropolis launch = one sweep
        t = r["metropolis_kernel_ms"] / 1e3  # HIP events around the kernel launch, same stream
        ach = flops / t / 1e12
        peak = MFMA_PEAK["bf16"] if main_mode == "mixed" else MFMA_PEAK["fp32"]
        traffic = None
        if args.config == "c3" and main_mode == "mixed" and sweep_steps == 512:
            p = PMC_C3_MIXED
            traffic = int(1024 * (p["fetch_kib"] * p["fetch_scale"] + p["write_kib"] * p["write_scale"]))
        roof = {"bound": "mfma", "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(ach / peak, 4), "traffic": traffic,
                "traffic_source": None if traffic is None else
                "not measured in this run: rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE of this kernel, " + PMC_C3_MIXED["source"],
                "hbm_gbps": None if traffic is None else round(traffic / t / 1e9, 3),
                "hbm_peak_gbps": HBM_PEAK_GBS,
                "traffic_algorithmic": cfg["ns"] * sweep_steps * 16 + 2 * cfg["ns"] * N + 24 * cfg["ns"],
                "kernel": "rc_metropolis_kernel", "flops_per_launch": flops,
                "avg_launch_ms": round(r["metropolis_kernel_ms"], 3)}
    out = {
        "metric": "VMC steps/s, 16x16 J1J2 ResConv MinSR (mixed vs FP64); served p50/p99 latency"
        if args.config == "c3" else
        f"VMC steps/s ({args.config})",
        "value": round(r["steps_per_s"], 5),
        "unit": "VMC steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype_label(cfg, main_mode),
        "data": ("synthetic (He-normal ResConv weights, Philox-seeded chains)" if cfg["model"] == "resconv" else
                 "synthetic (RBM_Dense weights N(0, 0.01^2), Philox-seeded chains)"),
        "config": {"workload": cfg["workload"], "mode": main_mode, "nsamples": cfg["ns"],
                   "sweep_steps": sweep_steps, "solver": args.solver, "rtol": args.rtol, "lr": args.lr,
                   "nparams": r["nparams"]},
        "stages_ms": {k: round(v, 2) for k, v in r["stages"].items()},
        "solve_path": r["solve_info"],
        "energy": r["energies"][-1],
        "acceptance": round(r["acceptance"], 4),
        "mean_nconn": round(r["nconn"], 1),
        "roofline": roof,
    }
    if "plain" in res and main_mode != "plain":
        p = res["plain"]
        out["plain"] = {"value": round(p["steps_per_s"], 5), "ms_per_step": round(p["ms_per_step"], 2),
                        "stages_ms": {k: round(v, 2) for k, v in p["stages"].items()}}
        out["mixed_over_plain"] = round(r["steps_per_s"] / p["steps_per_s"], 3)
    if "precision" in r:
        out["precision"] = r["precision"]
        out["precision"]["per_step_update_vector_rel_err_gram"] = r.get("per_step_gram_err")
    if "gram" in r:
        out["gram"] = r["gram"]
    if "reference_default_solve" in r:
        out["reference_default_solve"] = r["reference_default_solve"]
    if cfg["model"] == "resconv":
        # fused E_loc: one forward per valid connected configuration; reads the
        # spins and log psi, writes E_loc and the counts (compute-bound)
        fwd = conv_flops_per_forward(cfg["nblocks"], cfg["channels"], N)
        t_e = r["eloc_ms"] / 1e3
        nbytes = cfg["ns"] * (N + 8 + 8 + 4)
        out["eloc"] = {"kernel": "rc_eloc_kernel", "ms": round(r["eloc_ms"], 2),
                       "achieved_tflops": round(cfg["ns"] * r["nconn"] * fwd / t_e / 1e12, 1),
                       "hbm_algorithmic_bytes": nbytes, "hbm_gbps": round(nbytes / t_e / 1e9, 4)}
    if (args.config == "c3" and args.activation == "gelu" and args.mode == "both" and
            not args.no_relu_variant):
        out["relu_variant"] = relu_variant(cfg, args, torch, qtx, fwd * cfg["ns"] * sweep_steps)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_seconds, sweep_steps)
    if not args.no_serve and args.config == "c3":
        out["serving"] = run_serving(args, torch)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
