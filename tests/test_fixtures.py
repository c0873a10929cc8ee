"""The fixture reader / writer (mkf_amd/fixtures.py) on the CPU: directories in
the reference's export format round-trip, and damaged payloads are refused."""
import os

import numpy as np
import pytest

import fixture_gen


def test_kklss_ckks_fixture_round_trip(tmp_path):
    from mkf_amd import fixtures

    d = str(tmp_path / "kklss_ckks")
    exp = fixture_gen.kklss_ckks(d)
    fx = fixtures.load(d)
    assert fx["scheme"] == "KKLSS-CKKS" and fx["parties"] == 2
    assert fx["in_level"] == fx["out_level"] + 1
    assert np.array_equal(fx["expected"], exp)
    assert fx["keys"].shape[:3] == (2, 3, fx["crs_u"].shape[0])


def test_mkf_fixture_round_trip(tmp_path):
    from mkf_amd import fixtures

    d = str(tmp_path / "mkf_bfv")
    exp = fixture_gen.mkf(d, bfv=True)
    fx = fixtures.load(d)
    assert fx["scheme"] == "MKF-BFV" and fx["t"] == 65537
    assert np.array_equal(fx["expected"], exp)
    assert len(fx["fold_keys"]) == 2


def test_damaged_payload_is_refused(tmp_path):
    from mkf_amd import fixtures

    d = str(tmp_path / "bad")
    fixture_gen.kklss_ckks(d)
    path = os.path.join(d, "expected_ciphertext_q.bin")
    raw = bytearray(open(path, "rb").read())
    raw[17] ^= 1
    open(path, "wb").write(bytes(raw))
    with pytest.raises(fixtures.FixtureError):
        fixtures.load(d)
    open(path, "wb").write(bytes(raw[:-8]))  # truncated
    with pytest.raises(fixtures.FixtureError):
        fixtures.load(d, verify=False)
