"""The C ABI: libdimcv.so loads, exports every entry point include/dimcv.h
declares, and the ctypes mirrors of dimcv_params / dimcv_state have the
exact C layout (checked against a gcc-compiled probe of the header).
No compute call is made (no GPU is needed)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "dimcv.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dimcv_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_entry():
    import _native as N
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("libdimcv.so not built")
    lib = ctypes.CDLL(N.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_library_version_and_workspace_query():
    import _native as N
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("libdimcv.so not built")
    L = N.lib()
    assert L.dimcv_abi_version() == 1
    small = L.dimcv_workspace_bytes(1000, 100, 1)
    big = L.dimcv_workspace_bytes(1_048_576, 131_072, 1_048_576)
    assert 0 < small < big
    assert big % 256 == 0
    # keys/values 16 B, CDF 16 B, near field 16 B, (x, m_V) and (x, w)
    # records 32 B, kappa 4 B, sample table 32 B per particle, plus at most
    # 2^21 split partials of 32 B
    assert 116 * 1_048_576 <= big <= 116 * 1_048_576 + 32 * 2 ** 21 + (1 << 20)


def _probe_layout():
    code = r'''
#include <stdio.h>
#include <stddef.h>
#include "dimcv.h"
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  printf("sizeof.params %zu\n", sizeof(dimcv_params));
  printf("sizeof.state %zu\n", sizeof(dimcv_state));
  printf("sizeof.scene %zu\n", sizeof(dimcv_scene));
  O(dimcv_params, n_max) O(dimcv_params, grid_size) O(dimcv_params, gravity)
  O(dimcv_params, has_vortex_domain) O(dimcv_params, vp_factor) O(dimcv_params, bounce_hi)
  O(dimcv_params, far_mode) O(dimcv_params, m_V0) O(dimcv_params, vp_factor_f64)
  O(dimcv_state, grid_ids) O(dimcv_state, grid_ids_buffer) O(dimcv_state, num_particles_each_emitter)
  O(dimcv_state, n_emitters) O(dimcv_state, fluid_list) O(dimcv_state, workspace)
  O(dimcv_state, workspace_bytes) O(dimcv_state, scene_id) O(dimcv_state, n_scenes)
  O(dimcv_state, scene_counters) O(dimcv_state, block_scene) O(dimcv_state, batch_workspace_bytes)
  O(dimcv_state, particle_id_buffer)
  O(dimcv_scene, cell_base) O(dimcv_scene, domain_start) O(dimcv_scene, visc_coef)
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write(code)
        subprocess.check_call(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), c, "-o", exe])
        out = subprocess.check_output([exe], text=True)
    return dict((k, int(v)) for k, v in (ln.split() for ln in out.strip().splitlines()))


def test_ctypes_structs_match_c_layout():
    import _native as N
    lay = _probe_layout()
    assert ctypes.sizeof(N.Params) == lay["sizeof.params"]
    assert ctypes.sizeof(N.State) == lay["sizeof.state"]
    assert ctypes.sizeof(N.Scene) == lay["sizeof.scene"]
    for key, off in lay.items():
        if key.startswith("sizeof"):
            continue
        tname, field = key.split(".")
        cls = {"dimcv_params": N.Params, "dimcv_state": N.State, "dimcv_scene": N.Scene}[tname]
        assert getattr(cls, field).offset == off, key


def test_far_split_policy_fits_the_partial_capacity():
    """The MC far field's source-split count (shared by the single-scene and
    batched steps) never asks for more partials than either reserves:
    splits * targets <= min(16 targets, 2^21) and 1 <= splits <= 16."""
    import _native as N
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("libdimcv.so not built")
    L = N.lib()
    ts = list(range(1, 5000, 37)) + list(range((1 << 19) - 300, (1 << 21) + 300, 997)) + \
        [525_285, 700_000, 1 << 19, 1 << 20, 1 << 21, (1 << 21) + 1, 3_000_000, 10_000_000]
    for t in sorted(ts):
        s = L.dimcv_far_split_count(t)
        assert 1 <= s <= 16, (t, s)
        assert s == 1 or s * t <= min(16 * t, 1 << 21), (t, s)
    assert L.dimcv_far_split_count(1 << 20) == 2        # the two halves at 1M targets
    assert L.dimcv_far_split_count(-1) == -1
