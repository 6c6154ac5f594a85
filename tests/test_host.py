"""Host-side checks that need no GPU: the C ABI library loads and exports every
declared entry point, the package's lattice/operator tables equal the
oracle's, and the bench contract for more than one GPU."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import lattice as olat
from oracle import operator as oop

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qtx_hip.h")
LIB = os.path.join(ROOT, "quantax_amd", "_lib", "libqtxhip.so")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(qtx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    syms = declared_symbols()
    for name in ["qtx_rbm_metropolis", "qtx_resconv_metropolis", "qtx_rbm_eloc", "qtx_resconv_eloc",
                 "qtx_rbm_jacobian", "qtx_resconv_jacobian", "qtx_minsr_solve", "qtx_philox_fill"]:
        assert name in syms


@pytest.mark.skipif(not os.path.exists(LIB), reason="libqtxhip.so not built (run __graft_entry__.build())")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    from quantax_amd import _lib

    assert set(_lib.PROTOS) == set(declared_symbols())


@pytest.mark.skipif(not os.path.exists(LIB), reason="libqtxhip.so not built")
def test_host_only_entry_points_validate_arguments():
    from quantax_amd import _lib

    d = _lib.QtxResConvDesc(8, 16, 16, 16, 0, 0)
    assert _lib.lib().qtx_resconv_nparams(d) == 34944
    assert _lib.lib().qtx_resconv_packed_bytes(d) > 34944 * 4
    bad = _lib.QtxResConvDesc(8, 24, 16, 16, 0, 0)
    assert _lib.lib().qtx_resconv_packed_bytes(bad) == 0
    assert _lib.lib().qtx_minsr_ws_bytes(100, 50, 0, 0) >= 50 * 50 * 8
    # argument errors are reported without touching a device
    rc = _lib.lib().qtx_rbm_pack(None, 0, None, 4, 4, None)
    assert rc == -1 and b"pointers" in _lib.lib().qtx_last_error()


@pytest.mark.parametrize("extent,ns", [([4, 4], [1, 2]), ([6, 5], [1, 2, 3]), ([16], [1, 2]), ([3, 3], [1])])
def test_package_lattice_equals_oracle(extent, ns):
    import quantax_amd.sites as qs

    lat = qs.Grid(extent)
    got = lat.get_neighbor(ns)
    ref = olat.get_neighbor(extent, ns)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    assert np.array_equal(lat.neighbor_table(ns), olat.site_neighbor_table(extent, ns))


@pytest.mark.parametrize("case", ["j1j2", "heis_chain", "ising", "custom"])
def test_package_operator_tables_equal_oracle(case):
    import quantax_amd.operator as qo
    import quantax_amd.sites as qs

    if case == "j1j2":
        qs.Square(6)
        H = qo.Heisenberg(J=[1, 0.5], n_neighbor=[1, 2], msr=True)
        Ho = oop.heisenberg([6, 6], [1, 0.5], [1, 2], True)
    elif case == "heis_chain":
        qs.Chain(12)
        H = qo.Heisenberg(msr=True)
        Ho = oop.heisenberg([12], msr=True)
    elif case == "ising":
        qs.Chain(9)
        H = qo.Ising(h=0.3, J=-1.5)
        Ho = oop.ising([9], h=0.3, J=-1.5)
    else:
        qs.Chain(5)
        H = 0.5 * qo.sigma_x(1) @ qo.sigma_x(3) - qo.sigma_z(0) + 2 * qo.sigma_p(2) @ qo.sigma_m(4)
        Ho = 0.5 * oop.sigma_x(1) @ oop.sigma_x(3) + (-1) * oop.sigma_z(0) + 2 * oop.sigma_p(2) @ oop.sigma_m(4)
    t = H.tables()
    ch = oop.compile_op_list(Ho, t["N"])
    for k in ["off_a", "off_b", "off_ra", "off_rb", "off_c", "diag_c", "diag_a", "diag_b", "diag_kind"]:
        assert np.array_equal(getattr(ch, k), t[k]), k
    assert [o for o, _ in H.op_list] == [o for o, _ in Ho.op_list]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in r.stdout
