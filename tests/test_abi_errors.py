"""The C ABI's argument checks: entry points given null or out-of-range
arguments return their negative status with a message from
dimcv_last_error(), before any device work (so these run without a GPU).
The Python layer turns a status into RuntimeError (_native.check)."""
import ctypes
import os

import numpy as np
import pytest

ERR_ARG, ERR_HIP, ERR_WORKSPACE, ERR_KEY, ERR_NBR, ERR_RANGE = -1, -2, -3, -4, -5, -6


@pytest.fixture(scope="module")
def L():
    import _native as N
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("libdimcv.so not built")
    return N.lib()


def last_error(L):
    return L.dimcv_last_error().decode()


def params(n_max=1000):
    import _native as N
    p = N.Params()
    p.n_max = n_max
    p.nbr_kmax = 64
    p.grid_num[0] = p.grid_num[1] = p.grid_num[2] = 4
    p.n_cells = 64
    return p


def test_null_arguments_are_rejected(L):
    import _native as N
    p = params()
    st = N.State()
    cases = [
        ("dimcv_step_begin", (None, None, None, None, 0, None)),
        ("dimcv_substep", (None, None, None, 0, 0, None)),
        ("dimcv_batch_substep", (None, None, None, None, None)),
        ("dimcv_batch_graph_step", (None, 0, None, None)),
        ("dimcv_segment_velocity", (None, 0, ctypes.c_float(0.0), None, None, None, None, None)),
        ("dimcv_segment_rk4", (None, None, None, None, None)),
        ("dimcv_rigid_moments", (None, None, 0, 0, None, None, 0, None, None)),
        ("dimcv_rigid_apply", (None, None, 0, None, 0, None)),
        ("dimcv_cdf_sample_noalloc", (None, None, None, 0, 0, None)),
        ("dimcv_valu_issue_probe", (None, 0, None, None)),
    ]
    for name, args in cases:
        rc = getattr(L, name)(*args)
        assert rc == ERR_ARG, (name, rc)
        assert last_error(L), name


def test_segment_forms_and_outputs_are_checked(L):
    import _native as N
    sg = N.Segments()
    sg.n = 0
    out = (ctypes.c_float * 3)()
    # form 2 does not exist; form 1 needs the stage inputs; outputs are required
    assert L.dimcv_segment_velocity(ctypes.byref(sg), 2, ctypes.c_float(0.0), None, None, out, out, None) == ERR_ARG
    assert "form" in last_error(L)
    assert L.dimcv_segment_velocity(ctypes.byref(sg), 1, ctypes.c_float(0.5), None, None, out, out, None) == ERR_ARG
    assert L.dimcv_segment_velocity(ctypes.byref(sg), 0, ctypes.c_float(0.0), None, None, None, None, None) == ERR_ARG
    # a negative pool size
    sg.n = -1
    assert L.dimcv_segment_velocity(ctypes.byref(sg), 0, ctypes.c_float(0.0), None, None, out, out, None) == ERR_ARG


def test_sampler_size_limits(L):
    import _native as N
    p = params(n_max=1000)
    st = N.State()
    # more samples than slots
    rc = L.dimcv_cdf_sample_noalloc(ctypes.byref(st), ctypes.byref(p), None, 1001, 10, None)
    assert rc == ERR_RANGE and "samples" in last_error(L)
    # a fluid bound past n_max
    rc = L.dimcv_cdf_sample_noalloc(ctypes.byref(st), ctypes.byref(p), None, 10, 1001, None)
    assert rc == ERR_RANGE
    # past the no-alloc scratch (2^20 fluid particles)
    big = params(n_max=1 << 22)
    rc = L.dimcv_cdf_sample_noalloc(ctypes.byref(st), ctypes.byref(big), None, 10, (1 << 20) + 1, None)
    assert rc == ERR_RANGE and "scratch" in last_error(L)


def test_shape_matching_mode_is_checked(L):
    import _native as N
    p = params()
    st = N.State()
    out = (ctypes.c_double * 9)()
    assert L.dimcv_rigid_moments(ctypes.byref(st), ctypes.byref(p), 1, 2, None, None, 10, out, None) == ERR_ARG
    # mode 1 needs the centre and the rest centre
    assert L.dimcv_rigid_moments(ctypes.byref(st), ctypes.byref(p), 1, 1, None, None, 10, out, None) == ERR_ARG


def test_policy_and_size_queries(L):
    assert L.dimcv_far_split_count(-1) == -1
    assert L.dimcv_far_split_count(1 << 20) == 2
    counts = (ctypes.c_int32 * 2)(100, 200)
    assert L.dimcv_batch_workspace_bytes(0, counts, 1000) == -1
    assert L.dimcv_batch_workspace_bytes(2, None, 1000) == -1
    assert L.dimcv_batch_workspace_bytes(2, counts, 1000) > 0


def test_python_layer_raises_runtime_error(L):
    import _native as N
    with pytest.raises(RuntimeError, match="null"):
        N.check(L.dimcv_rigid_apply(None, None, 0, None, 0, None), "dimcv_rigid_apply")
