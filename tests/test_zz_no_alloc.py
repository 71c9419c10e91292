"""dimcv_cdf_sample_noalloc: the inverse-CDF sampler with its scratch in a
fixed device array of the library instead of the caller's workspace.

Called through the C ABI on a state whose workspace pointer is NULL, it must
give bitwise the sample indices, sample count and CDF total of
dimcv_cdf_sample (which uses the workspace) on the same inputs: twice in a
row, the second time on a smaller system (the scratch keeps the first call's
contents), and it must refuse a fluid bound above its stated maximum.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(nx, ny, nz, seed):
    import scenes
    ps, solver, info = scenes.build_ring_system(nx, ny, nz, n_rings=2, seed=seed, mc_samples=4096,
                                                far_field="fast")
    rng = np.random.default_rng(seed)
    kd = rng.random(ps.particle_max_num, dtype=np.float32) * np.float32(5.0)
    kd[rng.random(kd.size) < 0.3] = 0.0
    solver.kernel_density.from_numpy(kd)
    return ps, solver, rng


def _run(ps, solver, name, u, m, workspace=True):
    import torch
    import _native as N
    tu = torch.as_tensor(u, dtype=torch.float64).to(ps.device)
    solver.sample_idx.fill(-7)
    solver.num_samples[None] = 0
    ps._dscal.zero_()
    st = type(ps.state()).from_buffer_copy(ps.state())
    if not workspace:
        st.workspace = None
        st.workspace_bytes = 0
    rc = getattr(N.lib(), name)(ctypes.byref(st), ctypes.byref(ps.params), tu.data_ptr(), m, ps._nfl,
                                ps.stream())
    N.check(rc, name)
    torch.cuda.synchronize()
    return (solver.sample_idx.to_numpy().copy(), int(solver.num_samples[None]),
            float(ps._dscal[N.D_CDF_TOTAL].item()))


def test_noalloc_cdf_sample_matches_workspace_entry():
    import _native as N
    for shape, m, seed in (((64, 64, 16), 4096, 6), ((12, 12, 10), 300, 7)):   # second call smaller
        ps, solver, rng = _system(*shape, seed)
        u = rng.random(m)
        want = _run(ps, solver, "dimcv_cdf_sample", u, m)
        got = _run(ps, solver, "dimcv_cdf_sample_noalloc", u, m, workspace=False)
        assert want[1] == m and want[2] > 0.0
        assert np.array_equal(got[0], want[0]), shape
        assert got[1:] == want[1:], shape
        print(f"{ps._nfl} fluid, {m} samples: no-alloc entry bitwise equal to the workspace entry", flush=True)
    # a fluid bound above the stated maximum (on a system with room for it)
    # is refused before any launch
    import scenes
    from config_builder import SimConfig
    from particle_system import ParticleSystem
    scene = scenes.ring_scene_config(12, 12, 10)
    scene["Configuration"]["reservedCapacity"] = (1 << 20) + 4096
    big = ParticleSystem(SimConfig(scene=scene))
    big.build_solver()
    st = type(big.state()).from_buffer_copy(big.state())
    st.workspace = None
    st.workspace_bytes = 0
    rc = N.lib().dimcv_cdf_sample_noalloc(ctypes.byref(st), ctypes.byref(big.params), None, 0,
                                          (1 << 20) + 1, big.stream())
    assert rc == -6                                                   # DIMCV_ERR_RANGE
    assert b"scratch maximum" in N.lib().dimcv_last_error()
