"""The Kármán scene at its configured size (von_karman_vortex_street.json:
r = 0.0045, 223 x 56 x 56 cells, 1,751,080 solid particles of which
1,201,233 are carved phantoms parked in cell 0, a 3,570-particle emitter
layer every 2 steps).

The first 32 steps are compared with the oracle bit for bit (fluid reaches
the vortex domain at x = 0.2 at step 26, so the Bernoulli sample set, AKDE
and vort2vel act in the last steps); a 200-step run, the scene's length, is
then held to the invariants the reference's bookkeeping fixes.
"""
import copy
import json
import os

import numpy as np
import pytest

import harness as H

pytestmark = pytest.mark.gpu

SCENE = os.path.join(H.HERE, "..", "dimcv-sph_amd", "data", "scenes", "von_karman_vortex_street.json")


def _scene():
    with open(SCENE) as f:
        return json.load(f)


def test_configured_karman_scene_bitexact_first_steps():
    ps, solver, ref = H.make_karman_pair(_scene())
    assert ps.particle_max_num == ref.particle_max_num == 2_751_080
    assert ps.particle_num[None] == ref.particle_num == 1_751_080
    assert int(np.prod(ps.grid_num)) == ref.ncells == 699_328
    rng = np.random.default_rng(200)
    full_at = {0, 7, 15, 23, 27, 31}
    for k in range(32):
        u1 = rng.random(ps.particle_max_num, dtype=np.float32)
        u2 = rng.random(ps.particle_max_num, dtype=np.float32)
        solver.step(u1, u2)
        ref.step(u1, u2)
        assert solver.last_iterations_v == ref.last_kv and solver.last_iterations_p == ref.last_kp, k
        assert solver.num_samples[None] == ref.num_samples, k
        if k in full_at:
            bad = H.mismatches(ps, solver, ref)
        else:
            bad = H.mismatches(ps, solver, ref, ps_fields=["x", "v", "is_sample"], solver_fields=[])
        assert not bad, f"step {k}: {bad}"
        print(f"step {k}: {ref.particle_num} particles, {ref.fluid_particle_num} fluid, "
              f"{ref.num_samples} samples, bit-exact", flush=True)
    assert ref.num_samples > 0


def test_configured_karman_scene_200_steps_invariants():
    from config_builder import SimConfig
    from particle_system import ParticleSystem
    ps = ParticleSystem(SimConfig(scene=copy.deepcopy(_scene())))
    solver = ps.build_solver()
    solver.initialize()
    solids = 1_751_080
    emitted = 0
    max_samples = 0
    for k in range(200):
        if k % solver.emit_interval == 0:
            emitted += 3570
        solver.step()
        pn = ps.particle_num[None]
        fluid = ps.fluid_particle_num[None]
        assert pn == solids + fluid, k                       # particle_system.py:467-468
        assert 0 < fluid <= emitted, k                        # nothing created but by the emitter
        obj = ps.num_particles_each_emitter.to_numpy()
        assert obj[0] == fluid, k       # all fluid is the emitter's (object 0): its count (:414-427)
        ns = solver.num_samples[None]
        assert 0 <= ns <= fluid, k
        max_samples = max(max_samples, ns)
        if k % 20 == 19 or k == 199:
            x = ps.x.to_numpy()[:pn]
            v = ps.v.to_numpy()[:pn]
            assert np.isfinite(x).all() and np.isfinite(v).all(), k
            mat = ps.material.to_numpy()[:pn]
            assert int((mat == 1).sum()) == fluid and int((mat == 0).sum()) == solids, k
            print(f"step {k}: {fluid} fluid, {ns} samples, k_v={solver.last_iterations_v}, "
                  f"k_p={solver.last_iterations_p}", flush=True)
    assert emitted == 100 * 3570
    assert max_samples > 1000
