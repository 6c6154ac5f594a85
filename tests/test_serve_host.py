"""Host-side parts of the serving loop: the job stream, FIFO admission by
budget and the argument checks of the grouped entry points (no GPU work)."""
import ctypes as C

import numpy as np
import pytest

from quantax_amd import _lib
from quantax_amd.serve import JobSpec, Server, admit, c5_stream


def test_c5_stream_shape():
    a = c5_stream(64, mean_interarrival=0.25, seed=3)
    b = c5_stream(64, mean_interarrival=0.25, seed=3)
    assert [s.__dict__ for s in a] == [s.__dict__ for s in b]
    arr = np.array([s.arrival for s in a])
    assert np.all(np.diff(arr) > 0) and arr[0] > 0
    assert {s.L for s in a} <= {6, 8, 10, 12, 14, 16}
    assert {s.alpha for s in a} <= {1, 2, 4}
    assert {s.ns for s in a} <= {1024, 2048, 4096, 8192}
    assert all(0.0 <= s.J2 <= 0.6 and s.nsteps == 50 for s in a)
    # exponential gaps: the mean is near the requested one
    assert 0.15 < np.mean(np.diff(np.concatenate([[0.0], arr]))) < 0.4


def test_admit_fifo_budget():
    specs = [JobSpec(job_id=i, L=4, J2=0.0, alpha=1, ns=64, arrival=float(i)) for i in range(5)]
    cost = {0: 40, 1: 30, 2: 50, 3: 10, 4: 10}
    pending = list(specs)
    got = admit(pending, now=10.0, used=0, budget=100, cost=lambda s: cost[s.job_id])
    assert [s.job_id for s in got] == [0, 1]  # job 2 does not fit and blocks 3, 4 (no overtaking)
    assert [s.job_id for s in pending] == [2, 3, 4]
    got = admit(pending, now=10.0, used=20, budget=100, cost=lambda s: cost[s.job_id])
    assert [s.job_id for s in got] == [2, 3, 4]
    pending = list(specs)
    got = admit(pending, now=1.5, used=0, budget=1000, cost=lambda s: cost[s.job_id])
    assert [s.job_id for s in got] == [0, 1]  # later jobs have not arrived


def test_pacing_by_step_cost():
    """Steps per round: the largest job takes one, a job costing 1/k of it
    min(k, pacing), equal jobs one each."""
    big = JobSpec(job_id=0, L=16, J2=0.5, alpha=4, ns=8192)
    mid = JobSpec(job_id=1, L=12, J2=0.5, alpha=2, ns=4096)
    small = JobSpec(job_id=2, L=6, J2=0.0, alpha=1, ns=1024)
    srv = Server(hbm_budget=1)
    c = [Server.step_cost(s) for s in (big, mid, small)]
    assert c[0] > c[1] > c[2]
    p = srv.paces([big, mid, small])
    assert p[0] == 1 and p[2] == 8 and p[1] == min(8, int(c[0] / c[1]))
    assert srv.paces([mid, mid]) == [1, 1]
    assert Server(hbm_budget=1, pacing=1).paces([big, mid, small]) == [1, 1, 1]


def test_group_entry_points_reject_bad_jobs():
    lib = _lib.lib()
    arr = (_lib.QtxRbmJob * 2)()
    for j in arr:
        j.packed, j.M, j.N, j.ns, j.s = 1, 8, 4, 16, 1
        j.logabs = j.n_accept = j.n_near = j.eloc = j.O = 1
        j.kind, j.nsteps = _lib.QTX_LOCALFLIP, 4
        j.H.N = 4
    arr[1].M = 0
    need = int(lib.qtx_rbm_group_ws_bytes(2))
    assert need > 0
    assert lib.qtx_rbm_group_metropolis(None, arr, 2, None, need) == -4  # QTX_EWORKSPACE
    rc = lib.qtx_rbm_group_metropolis(None, arr, 2, 1, need)
    assert rc == -1 and b"job 1" in lib.qtx_last_error()
    arr[1].M = 8
    arr[1].kind = _lib.QTX_EXCHANGE  # without a neighbour table
    assert lib.qtx_rbm_group_metropolis(None, arr, 2, 1, need) == -1
    arr[1].kind = _lib.QTX_LOCALFLIP
    arr[0].H.N = 5
    assert lib.qtx_rbm_group_eloc(None, arr, 2, 1, need) == -1
    assert lib.qtx_rbm_group_logpsi(None, arr, 0, None, 0) == 0  # empty group is a no-op
