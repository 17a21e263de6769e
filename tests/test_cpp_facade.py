"""The C++ mirror of FEDDLib's FE assembly interface (feddlib_amd/host/FE_gpu.hpp).

tests/cpp/fe_gpu_test.cpp calls FE::assemblyLaplace / assemblyRHS / assemblyLaplaceAssFE /
assemblyNavierStokes / assemblyLinearElasticity / assemblyNonLinearElasticity (+ E field)
/ assemblyLinElasXDim with the reference's arguments and parameter-list keys, and checks
each result against the oracle (graph bit-exact, values to the 1e-12 row rule, vectors to
1e-12 relative); it also checks the reference's exception types and messages.  The
binary is built by tests/cpp/Makefile (from __graft_entry__.build()).
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "fe_gpu_test")


def _binary(build_if_missing):
    if not os.path.exists(BIN):
        if not build_if_missing:
            pytest.fail("tests/cpp/fe_gpu_test is not built (run __graft_entry__.build())")
        subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "cpp")])
    return BIN


def _run(args, build_if_missing):
    p = subprocess.run([_binary(build_if_missing)] + args, capture_output=True, text=True, timeout=300)
    print(p.stdout)
    print(p.stderr)
    return p


def test_facade_error_behaviour_cpu():
    """Exceptions raised before any device work: checkFE, P0, null vector, factory, parameters."""
    p = _run(["--cpu"], True)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "FAIL" not in p.stdout


@pytest.mark.gpu
def test_facade_against_oracle_gpu():
    p = _run([], False)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "FAIL" not in p.stdout and p.stdout.count("PASS") >= 25
