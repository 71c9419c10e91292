"""Committed golden fixtures (tests/golden/*.npz, written by make_golden.py
from the oracle): the oracle must keep reproducing them bit for bit, and so
must the GPU implementation."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import cases  # noqa: E402


def _compare(got, want, name):
    for key in want.files:
        a, b = np.asarray(got[key]), want[key]
        assert a.shape == b.shape, (name, key, a.shape, b.shape)
        if a.dtype.kind == "f":
            assert np.array_equal(a, b, equal_nan=True), (name, key, float(np.nanmax(np.abs(a - b))))
        else:
            assert np.array_equal(a, b), (name, key)


@pytest.mark.parametrize("name", list(cases.CASES))
def test_oracle_reproduces_fixture(name):
    want = np.load(cases.fixture_path(name))
    _compare(cases.run_oracle(name), want, name)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(cases.CASES))
def test_gpu_reproduces_fixture(name):
    want = np.load(cases.fixture_path(name))
    _compare(cases.run_gpu(name), want, name)
