"""Serving loop on the GPU: the grouped launches reproduce the single-job
path, and every job's result is the same alone and in a mix (including a job
that arrives while others run)."""
import numpy as np
import pytest
import torch

from quantax_amd import serve

pytestmark = pytest.mark.gpu


def _specs():
    # different lattice sizes, couplings, widths (register classes R = 1, 2, 4)
    # and sample counts; the third job arrives later
    return [
        serve.JobSpec(job_id=0, L=4, J2=0.3, alpha=2, ns=96, nsteps=3, seed=101, arrival=0.0),
        serve.JobSpec(job_id=1, L=6, J2=0.5, alpha=1, ns=160, nsteps=4, seed=202, arrival=0.0),
        serve.JobSpec(job_id=2, L=6, J2=0.1, alpha=4, ns=64, nsteps=2, seed=303, arrival=0.0),
    ]


def test_grouped_equals_single_job_path():
    """One job through qtx_rbm_group_* equals the same job through qtx_rbm_*:
    same chains (bit-identical spins), same E_loc and O."""
    sp = serve.JobSpec(job_id=0, L=6, J2=0.4, alpha=2, ns=128, nsteps=2, seed=77)
    a = serve.JobRun(sp, gram="fp64")
    b = serve.JobRun(sp, gram="fp64")
    srv = serve.Server()
    for _ in range(2):
        a.sample_single()
        nst = [b.sweep_steps()]
        srv._grouped("qtx_rbm_group_metropolis", [b], nst)
        b.advance_counters(nst[0])
        srv._grouped("qtx_rbm_group_eloc", [b])
        srv._grouped("qtx_rbm_group_jacobian", [b])
        torch.cuda.synchronize()
        assert torch.equal(a.s, b.s)
        assert torch.equal(a.n_accept, b.n_accept)
        assert torch.equal(a.eloc, b.eloc)
        assert torch.equal(a.O, b.O)
        a.sr_update()
        b.sr_update()
        torch.cuda.synchronize()
        assert torch.allclose(a.model.theta, b.model.theta, rtol=1e-6, atol=1e-9)
    srv.close()


def _run(specs, **kw):
    srv = serve.Server(keep_results=True, **kw)
    try:
        rep = srv.run(specs)
    finally:
        srv.close()
    return rep


def _check_same(r1, r2):
    e1, e2 = r1["energy"], r2["energy"]
    assert np.all(np.isfinite(e1))
    assert np.allclose(e1, e2, rtol=1e-10, atol=0), (e1, e2)
    assert np.allclose(r1["theta"], r2["theta"], rtol=1e-5, atol=1e-7)


def test_alone_equals_mix():
    specs = _specs()
    mix = _run(specs)
    assert mix.total_steps == sum(s.nsteps for s in specs)
    assert len(mix.latencies) == 3 and mix.max_active == 3
    for sp in specs:
        alone = _run([serve.JobSpec(**{**sp.__dict__, "arrival": 0.0})])
        _check_same(alone.results[sp.job_id], mix.results[sp.job_id])


def test_late_arrival_joins_running_jobs():
    specs = _specs()
    specs[2].arrival = 0.3  # arrives while the first two are running
    late = _run(specs)
    ref = _run([serve.JobSpec(**{**specs[2].__dict__, "arrival": 0.0})])
    _check_same(ref.results[2], late.results[2])
    s = late.summary()
    assert s["jobs"] == 3 and s["latency_p99_s"] >= s["latency_p50_s"] > 0


def test_budget_limits_concurrency():
    specs = _specs()
    need = max(serve.footprint_bytes(s, "bf16x3") for s in specs)
    rep = _run(specs, hbm_budget=need)  # room for one job at a time
    assert rep.max_active == 1 and len(rep.latencies) == 3


def test_solve_many_equals_single_solves(rng):
    """qtx_minsr_solve_many on uneven tasks (MinSR and SR branches, three
    modes) gives each task's qtx_minsr_solve result."""
    from quantax_amd import _lib
    dev = torch.device("cuda")
    shapes = [(300, 1000, _lib.QTX_SOLVE_EIGH_PINV), (128, 2000, _lib.QTX_SOLVE_CHOL_SHIFT),
              (500, 200, _lib.QTX_SOLVE_EIGH_PINV), (200, 700, _lib.QTX_SOLVE_EIGH_PINV_DIRECT),
              (1024, 3000, _lib.QTX_SOLVE_EIGH_PINV), (900, 2500, _lib.QTX_SOLVE_CHOL_SHIFT),
              (1500, 4000, _lib.QTX_SOLVE_EIGH_PINV)]
    jobs = []
    for ns, np_, mode in shapes:
        O = torch.from_numpy((rng.standard_normal((ns, np_)) * 0.3 + 0.1).astype(np.float32)).to(dev)
        E = torch.from_numpy(rng.standard_normal(ns) - 2.0).to(dev)
        mean = torch.empty(np_, dtype=torch.float64, device=dev)
        _lib.call("qtx_colmean", _lib.stream(), _lib.QTX_F32, O.data_ptr(), ns, np_, mean.data_ptr())
        eb = torch.empty(ns, dtype=torch.float64, device=dev)
        o2 = torch.empty(2, dtype=torch.float64, device=dev)
        _lib.call("qtx_ebar", _lib.stream(), E.data_ptr(), None, ns, eb.data_ptr(), o2.data_ptr())
        gram = _lib.QTX_GRAM_BF16X3
        nws = int(_lib.lib().qtx_minsr_ws_bytes(ns, np_, mode, gram))
        ws = torch.empty(nws, dtype=torch.uint8, device=dev)
        step1 = torch.empty(np_, dtype=torch.float64, device=dev)
        _lib.call("qtx_minsr_solve", _lib.stream(), _lib.QTX_F32, O.data_ptr(), mean.data_ptr(), None, eb.data_ptr(),
                  ns, np_, mode, gram, 0.0, 1e-3, 1e-3, 0.0, step1.data_ptr(), ws.data_ptr(), nws)
        jobs.append(dict(O=O, mean=mean, eb=eb, ws=ws, ns=ns, np=np_, mode=mode, gram=gram, ref=step1,
                         out=torch.empty(np_, dtype=torch.float64, device=dev), stream=torch.cuda.Stream()))
    torch.cuda.synchronize()
    tasks = (_lib.QtxMinsrTask * len(jobs))()
    for t, j in zip(tasks, jobs):
        t.stream = j["stream"].cuda_stream
        t.o_dtype, t.mode, t.gram = _lib.QTX_F32, j["mode"], j["gram"]
        t.O, t.mean, t.r, t.ebar = j["O"].data_ptr(), j["mean"].data_ptr(), None, j["eb"].data_ptr()
        t.ns, t.np = j["ns"], j["np"]
        t.rshift, t.ashift, t.rtol, t.atol = 0.0, 1e-3, 1e-3, 0.0
        t.step, t.ws, t.ws_bytes = j["out"].data_ptr(), j["ws"].data_ptr(), j["ws"].numel()
    assert _lib.lib().qtx_minsr_solve_many(tasks, len(jobs)) == 0
    torch.cuda.synchronize()
    assert [t.path for t in tasks] == [1, 0, 1, 2, 1, 0, 1]
    assert all(t.status == 0 and t.fallback == 0 for t in tasks)
    for j in jobs:
        err = (torch.linalg.norm(j["out"] - j["ref"]) / torch.linalg.norm(j["ref"])).item()
        assert err < 1e-12, err


@pytest.mark.parametrize("ns,path", [(2000, 1), (2048, 4)])
def test_solve_many_overlapped_factorisations_stay_exact(ns, path, rng):
    """Twelve soft-inverse solves (a load where GPU time is shared)
    overlapping on their streams, three times over: the partial fractions
    (n = 2000) and the Chebyshev route (n = 2048 on).  No factorisation may
    fail, and each result equals the task's own solve alone.  A panel kernel
    that stored its factored diagonal tile while late workgroups of the same
    launch still read it failed here in about 4 of 10 solves."""
    from quantax_amd import _lib
    dev = torch.device("cuda")
    np_ = 5000
    gram, mode = _lib.QTX_GRAM_BF16X3, _lib.QTX_SOLVE_EIGH_PINV
    nws = int(_lib.lib().qtx_minsr_ws_bytes(ns, np_, mode, gram))
    jobs = []
    for k in range(12):
        O = torch.from_numpy((rng.standard_normal((ns, np_)) * 0.3 + 0.1).astype(np.float32)).to(dev)
        O[:, : np_ // 2] *= 0.05
        E = torch.from_numpy(rng.standard_normal(ns) - 2.0).to(dev)
        mean = torch.empty(np_, dtype=torch.float64, device=dev)
        _lib.call("qtx_colmean", _lib.stream(), _lib.QTX_F32, O.data_ptr(), ns, np_, mean.data_ptr())
        eb = torch.empty(ns, dtype=torch.float64, device=dev)
        o2 = torch.empty(2, dtype=torch.float64, device=dev)
        _lib.call("qtx_ebar", _lib.stream(), E.data_ptr(), None, ns, eb.data_ptr(), o2.data_ptr())
        ws = torch.empty(nws, dtype=torch.uint8, device=dev)
        ref = torch.empty(np_, dtype=torch.float64, device=dev)
        _lib.call("qtx_minsr_solve", _lib.stream(), _lib.QTX_F32, O.data_ptr(), mean.data_ptr(), None, eb.data_ptr(),
                  ns, np_, mode, gram, 0.0, 0.0, 1e-4, 0.0, ref.data_ptr(), ws.data_ptr(), nws)
        jobs.append(dict(O=O, mean=mean, eb=eb, ws=ws, ref=ref, out=torch.empty_like(ref),
                         stream=torch.cuda.Stream()))
    torch.cuda.synchronize()
    for rep in range(3):
        tasks = (_lib.QtxMinsrTask * len(jobs))()
        for t, j in zip(tasks, jobs):
            t.stream = j["stream"].cuda_stream
            t.o_dtype, t.mode, t.gram = _lib.QTX_F32, mode, gram
            t.O, t.mean, t.r, t.ebar = j["O"].data_ptr(), j["mean"].data_ptr(), None, j["eb"].data_ptr()
            t.ns, t.np = ns, np_
            t.rshift, t.ashift, t.rtol, t.atol = 0.0, 0.0, 1e-4, 0.0
            t.step, t.ws, t.ws_bytes = j["out"].data_ptr(), j["ws"].data_ptr(), j["ws"].numel()
        assert _lib.lib().qtx_minsr_solve_many(tasks, len(jobs)) == 0
        torch.cuda.synchronize()
        assert all(t.status == 0 and t.path == path and t.fallback == 0 for t in tasks), \
            [(t.path, t.fallback) for t in tasks]
        for j in jobs:
            err = (torch.linalg.norm(j["out"] - j["ref"]) / torch.linalg.norm(j["ref"])).item()
            assert err < 1e-9, err
