"""Serving loop: independent RBM_Dense VMC jobs on one GPU with continuous batching.

The reference runs one lattice per process (``quantax/sites/sites.py:46-51``)
and has no serving layer; this module is the request loop around the per-job
VMC step of ``optimizer/driver.py:30-39`` (sweep, E_loc, Jacobian, MinSR
solve, update), for a stream of jobs that differ in lattice size, coupling,
RBM width and sample count.

* Admission: a job is admitted (FIFO) once it has arrived and its device
  footprint (O, the solver workspace, chains) fits the HBM budget next to the
  jobs already running; finished jobs release their memory at once.
* Continuous batching: every round advances every admitted job by at least
  one VMC step; a job whose step costs 1/k of the largest active job's takes
  min(k, pacing) steps, in sub-rounds that group only the jobs with steps
  left in the round, so small jobs are not held in lock-step with the
  largest.  Sampling, E_loc and the Jacobian of all jobs are grouped launches
  (``qtx_rbm_group_*``: one persistent launch per register class), so a job
  that arrives while others run joins the next round's launches.  The per-job
  SR solves (Gram, soft inverse, update) then run on per-job streams, all
  driven from one host thread by ``qtx_minsr_solve_many``, which enqueues
  each solve phase for every job before it reads any result back, so the
  jobs' solver kernels overlap on the device.
* Each job keeps its own Philox streams and the grouped kernels run the
  single-job per-chain code, so a job's trajectory is the same alone and in
  any mix (``test_gpu_serve.py`` checks it).
* Latency of a job = completion time - arrival time; the report gives p50 and
  p99 beside the aggregate VMC steps/s.

``run_sequential`` is the baseline the stream is compared with: the same
jobs one after another through the single-job entry points (``qtx_rbm_*``)
with the plain FP64 Gram.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from . import operator as qop
from . import sites as qsites
from .model import RBM_Dense
from .optimizer import auto_pinv_eig, last_solve_info, _GRAM

_SEED_INIT = 0x5DEECE66D  # initial-spin key offset used by the single-job samplers
_NSTREAMS = 8  # job solves run on a fixed pool of streams (job id mod 8)
_streams: List[torch.cuda.Stream] = []


def _job_stream(job_id: int) -> torch.cuda.Stream:
    if not _streams:
        _streams.extend(torch.cuda.Stream(device=_lib.device()) for _ in range(_NSTREAMS))  # default priority
    return _streams[job_id % _NSTREAMS]


@dataclass
class JobSpec:
    """One VMC job: J1-J2 Heisenberg (Marshall sign, Sz = 0) on an L x L
    periodic square lattice, RBM_Dense with alpha N hidden units (float32,
    N(0, init_std) weights), ``ns`` samples = chains, ``nsteps`` MinSR steps."""
    job_id: int
    L: int
    J2: float
    alpha: int
    ns: int
    nsteps: int = 50
    seed: int = 0
    arrival: float = 0.0
    lr: float = 1e-2
    rtol: float = 1e-4
    init_std: float = 0.01

    @property
    def N(self) -> int:
        return self.L * self.L

    @property
    def M(self) -> int:
        return self.alpha * self.N

    @property
    def nparams(self) -> int:
        return self.M * self.N + self.M


def c5_stream(njobs: int = 64, mean_interarrival: float = 0.25, nsteps: int = 50, seed: int = 0,
              L_choices: Sequence[int] = (6, 8, 10, 12, 14, 16), alpha_choices: Sequence[int] = (1, 2, 4),
              ns_choices: Sequence[int] = (1024, 2048, 4096, 8192), J2_range=(0.0, 0.6)) -> List[JobSpec]:
    """The served workload: Poisson arrivals (exponential gaps), L, alpha and
    the sample count uniform over their choices, J2 uniform in its range."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for j in range(njobs):
        t += float(rng.exponential(mean_interarrival))
        out.append(JobSpec(job_id=j, L=int(rng.choice(L_choices)), J2=float(rng.uniform(*J2_range)),
                           alpha=int(rng.choice(alpha_choices)), ns=int(rng.choice(ns_choices)), nsteps=nsteps,
                           seed=int(rng.integers(1, 2**62)), arrival=t))
    return out


def footprint_bytes(spec: JobSpec, gram: str) -> int:
    """Device bytes a job holds while admitted (O, solver workspace, chains)."""
    ns, npar, N = spec.ns, spec.nparams, spec.N
    ws = auto_pinv_eig(rtol=spec.rtol, gram=gram).ws_bytes(ns, npar)  # with the rocSOLVER scratch
    o_bytes = ns * npar * 4
    small = ns * (N + 8 * 4 + 4 * 2) + npar * (4 * 2 + 8 * 2) + (1 << 20)
    return o_bytes + ws + small


class JobRun:
    """Device state of one admitted job."""

    def __init__(self, spec: JobSpec, gram: str = "bf16x3"):
        self.spec = spec
        self.gram = gram
        N = spec.N
        dev = _lib.device()
        # lattice, Hamiltonian and model are built with this job's lattice current;
        # afterwards the job only uses its own device copies
        lat = qsites.Square(spec.L, Nparticles=(N // 2, N - N // 2))
        self._H = qop.Heisenberg(J=[1.0, spec.J2], n_neighbor=[1, 2], msr=True)
        self.ham = self._H.device_tables()
        tab = lat.neighbor_table(1)
        self.nbr = torch.from_numpy(np.ascontiguousarray(tab)).to(dev)
        self.width = int(tab.shape[1])
        self.hopping = 1 if 2 * (N // 2) <= N else -1
        self.model = RBM_Dense(spec.M, dtype=np.float32, init_std=spec.init_std, seed=spec.seed)
        ns = spec.ns
        self.s = torch.empty((ns, N), dtype=torch.int8, device=dev)
        _lib.call("qtx_init_spins", _lib.stream(), spec.seed ^ _SEED_INIT, ns, N, N // 2, self.s.data_ptr())
        self.logabs = torch.zeros(ns, dtype=torch.float64, device=dev)
        self.n_accept = torch.zeros(ns, dtype=torch.int32, device=dev)
        self.n_near = torch.zeros(ns, dtype=torch.int32, device=dev)
        self.eloc = torch.empty(ns, dtype=torch.float64, device=dev)
        self.O = torch.empty((ns, spec.nparams), dtype=torch.float32, device=dev)
        self.mean = torch.empty(spec.nparams, dtype=torch.float64, device=dev)
        self.ebar = torch.empty(ns, dtype=torch.float64, device=dev)
        self.out2 = torch.empty(2, dtype=torch.float64, device=dev)
        self.step = torch.empty(spec.nparams, dtype=torch.float64, device=dev)
        self.applied = torch.empty(1, dtype=torch.int32, device=dev)
        self.energy = torch.full((spec.nsteps,), float("nan"), dtype=torch.float64, device=dev)
        self.solver = auto_pinv_eig(rtol=spec.rtol, gram=gram)
        nws = self.solver.ws_bytes(ns, spec.nparams)
        self.ws = torch.empty(max(nws, 1), dtype=torch.uint8, device=dev)
        self.solve_paths: List[str] = []  # soft-inverse route taken at each step
        self.sweep_counter = 0       # Philox step counter of the sweep stream
        self.thermalised = False
        self.steps_done = 0
        self.stream = _job_stream(spec.job_id)
        self.done_event = torch.cuda.Event()

    # ------------------------------------------------------------ sampling stage
    def sweep_steps(self) -> int:
        """Metropolis steps of this job's next sweep: 2N, plus 20N of
        thermalisation before the first (metropolis.py:125-132)."""
        N = self.spec.N
        return 2 * N + (0 if self.thermalised else 20 * N)

    def job_struct(self, nsteps: int) -> _lib.QtxRbmJob:
        j = _lib.QtxRbmJob()
        j.packed = self.model.packed.data_ptr()
        j.M, j.N = self.spec.M, self.spec.N
        j.kind = _lib.QTX_EXCHANGE
        j.width = self.width
        j.nbr = self.nbr.data_ptr()
        j.hopping = self.hopping
        j.nsteps = nsteps
        j.seed = self.spec.seed
        j.step0 = self.sweep_counter
        j.ns = self.spec.ns
        j.s = self.s.data_ptr()
        j.logabs = self.logabs.data_ptr()
        j.n_accept = self.n_accept.data_ptr()
        j.n_near = self.n_near.data_ptr()
        j.reweight = 2.0
        j.tie_tol = 1e-9
        j.H = self.ham
        j.eloc = self.eloc.data_ptr()
        j.nconn = None
        j.O = self.O.data_ptr()
        return j

    def advance_counters(self, nsteps: int) -> None:
        self.sweep_counter += nsteps
        self.thermalised = True

    def sample_single(self) -> None:
        """The same sweep, E_loc and Jacobian through the single-job entry points."""
        ns, N, M = self.spec.ns, self.spec.N, self.spec.M
        nsteps = self.sweep_steps()
        words = torch.empty((nsteps, ns, 4), dtype=torch.int32, device=self.s.device)
        _lib.call("qtx_philox_fill", _lib.stream(), self.spec.seed, _lib.TAG_SWEEP, self.sweep_counter, nsteps, 0, ns,
                  words.data_ptr())
        _lib.call("qtx_rbm_metropolis", _lib.stream(), _lib.QTX_F32, self.model.packed.data_ptr(), M, N,
                  _lib.QTX_EXCHANGE, self.nbr.data_ptr(), self.width, self.hopping, words.data_ptr(), nsteps, ns, 2.0,
                  1e-9, self.s.data_ptr(), self.logabs.data_ptr(), self.n_accept.data_ptr(), self.n_near.data_ptr())
        self.advance_counters(nsteps)
        _lib.call("qtx_rbm_eloc", _lib.stream(), _lib.QTX_F32, self.model.packed.data_ptr(), M, N, self.ham,
                  self.s.data_ptr(), ns, self.eloc.data_ptr(), None)
        _lib.call("qtx_rbm_jacobian", _lib.stream(), _lib.QTX_F32, self.model.packed.data_ptr(), M, N,
                  self.s.data_ptr(), ns, _lib.QTX_F32, self.O.data_ptr())

    # ------------------------------------------------------------ SR stage
    def sr_pre(self) -> None:
        """Ebar and Omean on the current stream."""
        sp, st = self.spec, _lib.stream()
        _lib.call("qtx_ebar", st, self.eloc.data_ptr(), None, sp.ns, self.ebar.data_ptr(), self.out2.data_ptr())
        _lib.call("qtx_colmean", st, _lib.QTX_F32, self.O.data_ptr(), sp.ns, sp.nparams, self.mean.data_ptr())

    def sr_task(self) -> _lib.QtxMinsrTask:
        """The MinSR/SR solve of this step as a qtx_minsr_solve_many task on ``self.stream``."""
        sp, s = self.spec, self.solver
        t = _lib.QtxMinsrTask()
        t.stream = self.stream.cuda_stream
        t.o_dtype, t.mode, t.gram = _lib.QTX_F32, s.cmode, _GRAM[s.gram]
        t.O, t.mean, t.r, t.ebar = self.O.data_ptr(), self.mean.data_ptr(), None, self.ebar.data_ptr()
        t.ns, t.np = sp.ns, sp.nparams
        t.rshift, t.ashift, t.rtol, t.atol, t.tol_snr = s.rshift, s.ashift, s.rtol, s.atol, s.tol_snr
        t.step, t.ws, t.ws_bytes = self.step.data_ptr(), self.ws.data_ptr(), self.ws.numel()
        return t

    def sr_post(self, path: str) -> None:
        """The guarded update and this step's energy, on the current stream."""
        sp = self.spec
        self.solve_paths.append(path)
        _lib.call("qtx_apply_update", _lib.stream(), _lib.QTX_F32, self.model.theta.data_ptr(), self.step.data_ptr(),
                  sp.lr, sp.nparams, self.applied.data_ptr())
        self.model.repack()
        self.energy[self.steps_done: self.steps_done + 1].copy_(self.out2[0:1])
        self.steps_done += 1

    def sr_update(self) -> None:
        """Ebar, Omean, MinSR/SR solve and the guarded update on the current
        stream (the single-job route)."""
        sp, s = self.spec, self.solver
        self.sr_pre()
        _lib.call("qtx_minsr_solve", _lib.stream(), _lib.QTX_F32, self.O.data_ptr(), self.mean.data_ptr(), None,
                  self.ebar.data_ptr(), sp.ns, sp.nparams, s.cmode, _GRAM[s.gram], s.rshift, s.ashift, s.rtol, s.atol,
                  self.step.data_ptr(), self.ws.data_ptr(), self.ws.numel())
        info = last_solve_info()
        self.sr_post(info["path"] if info["fallback"] is None else f"{info['path']}:{info['fallback']}")

    def result(self) -> dict:
        return {"job_id": self.spec.job_id, "energy": self.energy.cpu().numpy(),
                "theta": self.model.theta.detach().cpu().numpy(), "solve_paths": list(self.solve_paths)}


def admit(pending: List[JobSpec], now: float, used: int, budget: int, cost) -> List[JobSpec]:
    """FIFO admission: pop from the head of ``pending`` (sorted by arrival)
    every job that has arrived and fits next to ``used`` bytes; stop at the
    first that does not (no overtaking, so a large job is not starved)."""
    out = []
    while pending and pending[0].arrival <= now and used + cost(pending[0]) <= budget:
        sp = pending.pop(0)
        used += cost(sp)
        out.append(sp)
    return out


@dataclass
class ServeReport:
    latencies: List[float] = field(default_factory=list)
    makespan: float = 0.0
    total_steps: int = 0
    rounds: int = 0
    max_active: int = 0
    measured_s: Optional[float] = None
    results: dict = field(default_factory=dict)

    def summary(self) -> dict:
        lat = np.asarray(self.latencies)
        return {"jobs": len(lat), "vmc_steps": self.total_steps, "makespan_s": round(self.makespan, 3),
                "steps_per_s": round(self.total_steps / self.makespan, 3) if self.makespan > 0 else None,
                "latency_p50_s": round(float(np.percentile(lat, 50)), 3) if lat.size else None,
                "latency_p99_s": round(float(np.percentile(lat, 99)), 3) if lat.size else None,
                "rounds": self.rounds, "max_active": self.max_active,
                **({"measured_s": round(self.measured_s, 3)} if self.measured_s is not None else {})}


class Server:
    """Continuous-batching VMC server on the current GPU."""

    def __init__(self, hbm_budget: Optional[int] = None, gram: str = "bf16x3", keep_results: bool = False,
                 progress=None, max_active: Optional[int] = None, pacing: int = 8):
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = int(0.7 * free)
        self.budget = int(hbm_budget)
        self.gram = gram
        self.max_active = max_active  # optional cap on concurrent jobs (lower median latency)
        # per-job pacing: a job whose step costs 1/k of the largest active
        # job's runs min(k, pacing) steps per round (sub-rounds of the smaller
        # jobs only), so small jobs are not held in lock-step with the largest
        self.pacing = max(1, int(pacing))
        self.keep_results = keep_results
        self.progress = progress
        self._ws = None
        # per round: host seconds spent enqueueing the grouped launches and the
        # per-job pre/post work, host seconds inside qtx_minsr_solve_many
        # (enqueue + its readbacks), round wall seconds, and HIP events around
        # the three grouped launches
        self.round_log = []

    def close(self) -> None:
        self._ws = None

    def _group_ws(self, njobs: int) -> torch.Tensor:
        need = max(int(_lib.lib().qtx_rbm_group_ws_bytes(njobs)), 1)
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(max(need, 64 * 1024), dtype=torch.uint8, device=_lib.device())
        return self._ws

    def _grouped(self, name: str, jobs: List[JobRun], nsteps: Optional[List[int]] = None) -> None:
        arr = (_lib.QtxRbmJob * len(jobs))()
        for i, j in enumerate(jobs):
            arr[i] = j.job_struct(0 if nsteps is None else nsteps[i])
        ws = self._group_ws(len(jobs))
        _lib.call(name, _lib.stream(), arr, len(jobs), ws.data_ptr(), ws.numel())

    def round(self, active: List[JobRun]) -> None:
        """One VMC step of every active job: grouped sampling, E_loc and
        Jacobian on the current stream, then every job's solve on its own
        stream through one qtx_minsr_solve_many call."""
        t0 = time.perf_counter()
        ga, gb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ga.record()
        nst = [j.sweep_steps() for j in active]
        self._grouped("qtx_rbm_group_metropolis", active, nst)
        for j, n in zip(active, nst):
            j.advance_counters(n)
        self._grouped("qtx_rbm_group_eloc", active)
        self._grouped("qtx_rbm_group_jacobian", active)
        gb.record()
        ready = torch.cuda.Event()
        ready.record()
        main = torch.cuda.current_stream()
        for j in active:
            with torch.cuda.stream(j.stream):
                j.stream.wait_event(ready)
                j.sr_pre()
        tasks = (_lib.QtxMinsrTask * len(active))(*[j.sr_task() for j in active])
        t1 = time.perf_counter()
        rc = _lib.lib().qtx_minsr_solve_many(tasks, len(active))
        t2 = time.perf_counter()
        if rc != 0:
            raise _lib.QtxError(f"qtx_minsr_solve_many failed ({rc}): "
                                f"{_lib.lib().qtx_last_error().decode(errors='replace')}")
        names = {0: "chol_shift", 1: "partial_fractions", 2: "syevd", 3: "shifted_cholesky", 4: "chebyshev"}
        reasons = {1: "small_cutoff", 2: "nonfinite", 3: "nonpositive_lambda_max", 4: "cholesky_failed",
                   5: "snr_truncation", 6: "shifted_cholesky_failed", 7: "shifted_cholesky_not_converged",
               8: "chebyshev_interval"}
        for j, tk in zip(active, tasks):
            path = names.get(tk.path, "none") + (f":{reasons[tk.fallback]}" if tk.fallback else "")
            with torch.cuda.stream(j.stream):
                j.sr_post(path)
                j.done_event.record(j.stream)
        for j in active:
            main.wait_event(j.done_event)
        t3 = time.perf_counter()
        self.round_log.append(((t1 - t0) + (t3 - t2), t2 - t1, t3 - t0, ga, gb, len(active)))

    def round_summary(self) -> dict:
        """Mean per-round host and device times of the rounds run so far."""
        if not self.round_log:
            return {}
        torch.cuda.synchronize()
        enq = np.mean([r[0] for r in self.round_log])
        solve = np.mean([r[1] for r in self.round_log])
        wall = np.mean([r[2] for r in self.round_log])
        grouped = np.mean([r[3].elapsed_time(r[4]) for r in self.round_log])
        return {"rounds": len(self.round_log), "mean_jobs_per_round": round(float(np.mean([r[5] for r in self.round_log])), 2),
                "round_wall_ms": round(1e3 * wall, 2),
                "host_enqueue_ms": round(1e3 * enq, 2),
                "solve_many_host_ms": round(1e3 * solve, 2),
                "grouped_launches_gpu_ms": round(grouped, 2)}

    @staticmethod
    def step_cost(sp: JobSpec) -> float:
        """Relative cost of one VMC step: sampling and E_loc ~ ns N^2 M (2N
        proposals and ~N connected configurations per sample, O(M) each), the
        Jacobian ns N M, the MinSR Gram ns^2 Np (the largest term here)."""
        return float(sp.ns) * sp.N * sp.N * sp.M + float(sp.ns) ** 2 * sp.nparams

    def paces(self, specs: Sequence[JobSpec]) -> List[int]:
        """Steps per round of each job: floor(largest step cost / its own), in 1..pacing."""
        cmax = max(self.step_cost(sp) for sp in specs)
        return [max(1, min(self.pacing, int(cmax / self.step_cost(sp)))) for sp in specs]

    def run(self, specs: Sequence[JobSpec]) -> ServeReport:
        rep = ServeReport()
        pending = sorted(specs, key=lambda s: (s.arrival, s.job_id))
        for s in pending:
            if footprint_bytes(s, self.gram) > self.budget:
                raise ValueError(f"job {s.job_id} needs {footprint_bytes(s, self.gram)} bytes, "
                                 f"more than the HBM budget {self.budget}")
        active: List[JobRun] = []
        used = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        while pending or active:
            now = time.perf_counter() - t0
            room = len(pending) if self.max_active is None else max(0, self.max_active - len(active))
            for sp in admit(pending[:room], now, used, self.budget, lambda s: footprint_bytes(s, self.gram)):
                pending.remove(sp)
                used += footprint_bytes(sp, self.gram)
                active.append(JobRun(sp, self.gram))
            if not active:
                time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
                continue
            rep.max_active = max(rep.max_active, len(active))
            paces = self.paces([j.spec for j in active])
            for r in range(max(paces)):
                sub = [j for j, k in zip(active, paces) if k > r and j.steps_done < j.spec.nsteps]
                if not sub:
                    break
                self.round(sub)
                rep.rounds += 1
                finished = [j for j in sub if j.steps_done >= j.spec.nsteps]
                if finished:
                    torch.cuda.synchronize()
                    tdone = time.perf_counter() - t0
                    for j in finished:
                        rep.latencies.append(tdone - j.spec.arrival)
                        rep.total_steps += j.spec.nsteps
                        used -= footprint_bytes(j.spec, self.gram)
                        if self.keep_results:
                            rep.results[j.spec.job_id] = j.result()
                    if self.progress:
                        self.progress(f"{len(rep.latencies)} jobs done, {len(active) - len(finished)} active, "
                                      f"{len(pending)} waiting, t = {tdone:.1f} s")
            active = [j for j in active if j.steps_done < j.spec.nsteps]
        torch.cuda.synchronize()
        rep.makespan = time.perf_counter() - t0
        return rep


def run_sequential(specs: Sequence[JobSpec], gram: str = "fp64", nsteps: Optional[int] = None,
                   keep_results: bool = False, progress=None) -> ServeReport:
    """The jobs one after another through the single-job entry points (no
    grouping, no overlap).  With ``nsteps`` below a job's step count the job
    runs that many steps and its full run is extrapolated as (first step,
    which carries the thermalisation) + (steps - 1) x (mean later step): the
    per-step cost of a job does not change over its run.  ``makespan`` is the
    extrapolated total, ``measured_s`` what actually ran."""
    rep = ServeReport()
    measured = 0.0
    for k, sp in enumerate(specs):
        job = JobRun(sp, gram)
        n = sp.nsteps if nsteps is None else max(2, min(nsteps, sp.nsteps))
        times = []
        for _ in range(n):
            torch.cuda.synchronize()
            ts = time.perf_counter()
            job.sample_single()
            job.sr_update()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - ts)
        measured += sum(times)
        full = times[0] + (sp.nsteps - 1) * float(np.mean(times[1:])) if n < sp.nsteps else sum(times)
        rep.latencies.append(full)
        rep.makespan += full
        rep.total_steps += sp.nsteps
        if keep_results:
            rep.results[sp.job_id] = job.result()
        del job
        if progress and (k + 1) % 8 == 0:
            progress(f"sequential: {k + 1}/{len(specs)} jobs, {measured:.1f} s measured")
    rep.measured_s = measured
    return rep
