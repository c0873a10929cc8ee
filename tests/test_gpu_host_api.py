"""GPU test of the C++ host mirror of the reference's evaluator API
(mkf_amd/host/mkhip.hpp): tests/cpp/test_host_api runs mkckks / mkbfv
MulRelinNew and the MKF CKKS / BFV pipelines on map-based ciphertexts and key
sets, compares every output bit for bit with the CPU oracle, and checks the
reference's error behaviour (a panic on a missing key, an error value for MKF)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_host_api")


def test_cpp_host_api_matches_oracle():
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/test_host_api is not built (run __graft_entry__.build())")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all host-API checks passed" in r.stdout
