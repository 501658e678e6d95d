"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle) against the oracle as built now:
bit-exact, since the strict oracle is deterministic.  A change to the oracle
that moves any number fails here before it can move the GPU parity checks."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)

import make_golden  # noqa: E402


@pytest.mark.parametrize("name", [n for n, _ in make_golden.CASES])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_fixture_matches_oracle(name, dt):
    path = os.path.join(HERE, f"{name}_{np.dtype(dt).name}.npz")
    stored = np.load(path)
    f = dict(make_golden.CASES)[name]
    fresh = f(dt, np.random.default_rng([ord(c) for c in name + np.dtype(dt).name]))
    assert sorted(stored.files) == sorted(fresh)
    for k in stored.files:
        a, b = stored[k], np.asarray(fresh[k])
        assert a.dtype == b.dtype and a.shape == b.shape and a.tobytes() == b.tobytes(), (name, k)
