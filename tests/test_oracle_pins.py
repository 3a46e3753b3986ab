"""The oracle's Rtrsm (every branch), Rgetrs, Rpotrs, Rgemv (strided and negative
increments), Rsyrk (both triangles, both forms), upper Rpotrf and iRamax against
element-by-element definitions in exact rational posit arithmetic (tests/oracle_pins.py).
CPU only."""
import pytest

from oracle_pins import CHECKS


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_oracle_matches_exact_definition(oracle, name):
    fails = CHECKS[name](oracle)
    assert not fails, "\n".join(fails[:10])
