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
    srv = serve.Server(workers=1)
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
    mix = _run(specs, workers=3)
    assert mix.total_steps == sum(s.nsteps for s in specs)
    assert len(mix.latencies) == 3 and mix.max_active == 3
    for sp in specs:
        alone = _run([serve.JobSpec(**{**sp.__dict__, "arrival": 0.0})], workers=1)
        _check_same(alone.results[sp.job_id], mix.results[sp.job_id])


def test_late_arrival_joins_running_jobs():
    specs = _specs()
    specs[2].arrival = 0.3  # arrives while the first two are running
    late = _run(specs, workers=2)
    ref = _run([serve.JobSpec(**{**specs[2].__dict__, "arrival": 0.0})], workers=1)
    _check_same(ref.results[2], late.results[2])
    s = late.summary()
    assert s["jobs"] == 3 and s["latency_p99_s"] >= s["latency_p50_s"] > 0


def test_budget_limits_concurrency():
    specs = _specs()
    need = max(serve.footprint_bytes(s, "bf16x3") for s in specs)
    rep = _run(specs, workers=2, hbm_budget=need)  # room for one job at a time
    assert rep.max_active == 1 and len(rep.latencies) == 3
