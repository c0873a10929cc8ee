"""The C ABI library loads, exports every symbol include/mkhip.h declares, and
rejects invalid parameters before touching a device (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "mkhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mkhip_\w+)\s*\(", text)))


def test_header_and_python_binding_agree():
    import mkf_amd

    assert _declared() == sorted(mkf_amd.EXPORTS)


def test_library_exports_every_declared_symbol():
    import mkf_amd

    lib = mkf_amd.lib()
    for name in _declared():
        assert hasattr(lib, name), name


def _create(logN, Q, P, gamma):
    import mkf_amd

    h = ctypes.c_void_p()
    q = np.asarray(Q, np.uint64)
    p = np.asarray(P or [0], np.uint64)
    return mkf_amd.lib().mkhip_ctx_create(ctypes.byref(h), logN, mkf_amd._p(q), len(Q), mkf_amd._p(p), len(P),
                                          gamma, 0, None, 0)


def test_invalid_parameters_are_rejected_on_the_host():
    from params import config

    logN, Q, P = config("c1")
    assert _create(12, Q, P, 1) == -2            # unsupported N
    assert _create(logN, Q, P, 2) == -3          # K = 1 with gamma = 2: alpha = 0 (mkrlwe/params.go:63-71)
    assert _create(logN, [Q[0] + 2], P, 1) == -2  # not a prime / not 1 mod 2N
    assert _create(logN, Q, P, 0) == -1


def test_status_strings():
    import mkf_amd

    for code in mkf_amd.STATUS:
        assert mkf_amd.lib().mkhip_status_string(code)


def test_cpp_host_mirror_compiles():
    """The C++ mirror of the reference's evaluator API is self-contained C++17 over the C ABI."""
    import shutil
    import subprocess

    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    hdr = os.path.join(ROOT, "mkf_amd", "host", "mkhip.hpp")
    r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-x", "c++", hdr],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
