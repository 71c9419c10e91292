"""The vector-ALU issue-rate calibration entry (dimcv_valu_issue_probe),
called through the C ABI: it reports a shader clock and per-SIMD issue rates
that a CDNA SIMD can reach (at most two vector instructions per quad-cycle,
and a scalar-operand stream no faster than an all-register one), and it
rejects a scratch buffer below its stated size."""
import ctypes

import pytest

import _native as N

pytestmark = pytest.mark.gpu


def test_issue_probe_reports_plausible_clock_and_rates():
    import torch
    L = N.lib()
    need = L.dimcv_valu_issue_probe_bytes()
    assert need > 256
    buf = torch.empty(int(need), dtype=torch.uint8, device="cuda")
    out = (ctypes.c_double * 4)()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.dimcv_valu_issue_probe(ctypes.c_void_p(buf.data_ptr()), int(need) - 1, out, stream) == -3
    assert "needed" in L.dimcv_last_error().decode()
    N.check(L.dimcv_valu_issue_probe(ctypes.c_void_p(buf.data_ptr()), int(need), out, stream),
            "dimcv_valu_issue_probe")
    clock_v, rate_v, clock_s, rate_s = list(out)
    print(f"clock {clock_v:.3f} / {clock_s:.3f} GHz, issue {rate_v:.3f} / {rate_s:.3f} per quad-cycle")
    for clock in (clock_v, clock_s):
        assert 0.5 < clock < 3.0
    assert 0.5 < rate_v <= 2.05
    assert 0.5 < rate_s <= rate_v * 1.05
