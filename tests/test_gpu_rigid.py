"""F4: dynamic rigid blocks.  A fluid block thrown at 15 m/s onto a dynamic
rigid block (density 500) resting above a static floor: the rigid particles'
boundary volumes are recomputed every step (sph_base.py:105-112), their
acceleration is gravity plus the pressure update's reactions gathered from
their fluid neighbours (dimcv_sph.py:536-549), and they advect with it
(:210-218).  Every field of every particle is compared with the oracle bit
for bit after every step, in the Bernoulli (per-pass substep) and the
fixed-M (one-call step) paths."""
import copy

import numpy as np
import pytest

import harness as H

pytestmark = pytest.mark.gpu


def rigid_drop_scene(mc=None):
    def blk(oid, start, end, vel, dens, dyn, color):
        return {"objectId": oid, "start": start, "end": end, "translation": [0, 0, 0], "scale": [1, 1, 1],
                "velocity": vel, "density": dens, "color": color, "isDynamic": dyn}
    c = {"domainStart": [0, 0, 0], "domainEnd": [0.6, 0.6, 0.6], "particleRadius": 0.01, "density0": 1000,
         "simulationMethod": 0, "gravitation": [0, 0, -9.8], "timeStepSize": 5e-4, "bounceBackBoundary": False,
         "volumeScale": 0.2, "GenProbPara": 5, "DelProbPara": 3.5, "reservedCapacity": 64}
    if mc:
        c["mcSamples"] = mc
        c["farField"] = "exact"
    return {"Configuration": c,
            "FluidBlocks": [blk(0, [0.1, 0.1, 0.215], [0.5, 0.5, 0.4], [0.0, 0.0, -15.0], 1000.0, None,
                                [50, 100, 200])],
            "RigidBlocks": [blk(3, [0.15, 0.15, 0.12], [0.45, 0.45, 0.19], [0.0, 0.0, 0.0], 500.0, True,
                                [200, 50, 50]),
                            blk(1, [0.05, 0.05, 0.05], [0.55, 0.55, 0.09], [0.0, 0.0, 0.0], 1000.0, False,
                                [201, 112, 100])]}


@pytest.mark.parametrize("mc", [None, 64], ids=["bernoulli-substep", "fixed-m-step"])
def test_dynamic_rigid_block_bitexact(mc):
    ps, solver, ref = H.make_karman_pair(rigid_drop_scene(mc))
    assert ps.has_dynamic_rigid
    rng = np.random.default_rng(4)
    reacted = 0
    for k in range(10):
        u1 = rng.random(ps.particle_max_num, dtype=np.float32)
        u2 = rng.random(ps.particle_max_num, dtype=np.float32)
        u = rng.random(64)
        if mc:
            solver.step(u_cdf=u)
            ref.step(u_cdf=u)
        else:
            solver.step(u1, u2)
            ref.step(u1, u2)
        bad = H.mismatches(ps, solver, ref)
        assert not bad, f"step {k}: {bad}"
        n = ref.particle_num
        rig = (ref.field("material")[:n] == 0) & (ref.field("is_dynamic")[:n] == 1)
        acc = ref.field("acceleration")[:n][rig]
        reacted += int((np.abs(acc[:, 2] + np.float32(9.8)) > 1e-3).sum())
    assert reacted > 100    # the reactions acted on many rigid particles
