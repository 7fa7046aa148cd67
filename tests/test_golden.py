"""Committed oracle vectors (tests/golden/oracle_golden.npz, written by
tests/golden/make_golden.py) are reproduced by the oracle."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

GOLD = np.load(os.path.join(HERE, "golden", "oracle_golden.npz"))


@pytest.mark.parametrize("name", sorted(make_golden.CASES))
def test_oracle_reproduces_golden(name):
  got = make_golden.compute(name)
  for k, v in got.items():
    np.testing.assert_allclose(v, GOLD[k], rtol=1e-12, atol=1e-13, err_msg=k)
