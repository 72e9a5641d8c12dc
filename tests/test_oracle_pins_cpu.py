"""The oracle's flow functions against mathematics and physics (the checks
live in tests/pins.py so that the mutant test can run them too)."""
from __future__ import annotations

import pytest

from tests import pins


@pytest.mark.parametrize("check", pins.ALL_CHECKS, ids=lambda c: c.__name__)
def test_oracle_pin(check, oracle, adi, thd):
    check(oracle, adi, thd)
