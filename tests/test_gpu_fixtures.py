"""Fixture directories in the reference's export format, run through the HIP
library by mkf_amd.fixtures.check: all four schemes bit-exact."""
import pytest

import fixture_gen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scheme", ["kklss_ckks", "kklss_bfv", "mkf_ckks", "mkf_bfv"])
def test_fixture_bit_exact(tmp_path, scheme):
    from mkf_amd import fixtures

    d = str(tmp_path / scheme)
    {"kklss_ckks": lambda: fixture_gen.kklss_ckks(d), "kklss_bfv": lambda: fixture_gen.kklss_bfv(d),
     "mkf_ckks": lambda: fixture_gen.mkf(d, bfv=False), "mkf_bfv": lambda: fixture_gen.mkf(d, bfv=True)}[scheme]()
    rule, diffs = fixtures.check(d)
    assert rule == 0, diffs

