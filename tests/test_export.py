"""The PLY exporter's number formatting (host entry of the device formatter)
against Python's own "{:.7f}" of float(v), the formatting the reference's
export_ply uses (karman_vortex.py:205-211): random bit patterns (every
exponent, subnormals, NaN, infinities), values spread over 15 decades, and
exact ties at the seventh decimal (k / 2^7 and (k + 1/2) 1e-7)."""
import ctypes
import os

import numpy as np
import pytest


def _fmt():
    import _native as N
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("libdimcv.so not built")
    L = N.lib()
    buf = ctypes.create_string_buffer(64)

    def f(v):
        n = L.dimcv_format_f7(float(v), buf)
        assert n == len(buf.value)
        return buf.value.decode()
    return f


def test_float_format_matches_python():
    f = _fmt()
    rng = np.random.default_rng(0)
    special = np.float32([0.0, -0.0, np.inf, -np.inf, np.nan, -np.nan, 1.0, -1.0, 0.5, 5e-8, -5e-8, 1.5e-7,
                          2.5e-7, 1e-45, -1e-45, 3.4028235e38, -3.4028235e38, 1.17549435e-38, 0.99999995,
                          123456.789, 16777216.0, 16777217.0, 1e20, 2.0 ** 63, 2.0 ** 64, -9.9999995e-8])
    vals = [special,
            rng.integers(0, 2 ** 32, 100_000, dtype=np.uint64).astype(np.uint32).view(np.float32),
            (rng.normal(size=60_000) * 10.0 ** rng.integers(-9, 6, 60_000)).astype(np.float32),
            np.float32(np.arange(-4000, 4000) / 128.0),
            np.float32((np.arange(0, 20000) + 0.5) * 1e-7)]
    bad = [(v, f(v), f"{float(v):.7f}") for a in vals for v in a if f(v) != f"{float(v):.7f}"]
    assert not bad, bad[:10]
