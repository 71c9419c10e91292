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


def test_shape_matching_recovers_the_rigid_transform():
    """compute_com / solve_constraints (sph_base.py:183-222) on a cubic
    dynamic block: moved to c + Rot S q (q = x_0 - rest_cm, S a symmetric
    stretch), its centre of mass is c (the lattice's first moment vanishes),
    A = sum m p q^T = Rot S Q with Q isotropic, so the polar rotation is Rot
    and every particle returns to c + Rot q.  The rest centres are indexed by
    object id over rigid + fluid objects (particle_system.py:126-132), so
    the block takes id 1 after the fluid block's 0."""
    from config_builder import SimConfig
    from particle_system import ParticleSystem
    scene = rigid_drop_scene()
    scene["RigidBlocks"] = [{"objectId": 1, "start": [0.2, 0.2, 0.2], "end": [0.3, 0.3, 0.3],
                             "translation": [0, 0, 0], "scale": [1, 1, 1], "velocity": [0, 0, 0],
                             "density": 500.0, "color": [200, 50, 50], "isDynamic": True}]
    ps = ParticleSystem(SimConfig(scene=scene))
    solver = ps.build_solver()
    solver.initialize()
    solver.compute_rigid_rest_cm(1)
    n = ps.particle_num[None]
    obj = ps.object_id.to_numpy()[:n] == 1
    assert obj.sum() == 125
    x0 = ps.x_0.to_numpy()[:n][obj].astype(np.float64)
    rest = ps.rigid_rest_cm[1].astype(np.float64)
    np.testing.assert_allclose(rest, x0.mean(axis=0), rtol=1e-6)
    np.testing.assert_allclose(rest, [0.24, 0.24, 0.24], rtol=1e-5)
    th = 0.3
    rot = np.array([[np.cos(th), -np.sin(th), 0.0], [np.sin(th), np.cos(th), 0.0], [0.0, 0.0, 1.0]])
    stretch = np.diag([1.08, 0.95, 1.02])
    c = np.array([0.31, 0.27, 0.22])
    q = x0 - rest
    x = ps.x.to_numpy()
    x[:n][obj] = (c + q @ (rot @ stretch).T).astype(np.float32)
    ps.x.from_numpy(x)
    np.testing.assert_allclose(solver.compute_com(1), c, rtol=2e-6)
    R = solver.solve_constraints(1)
    np.testing.assert_allclose(R, rot, atol=2e-6)
    got = ps.x.to_numpy()[:n][obj].astype(np.float64)
    np.testing.assert_allclose(got, c + q @ rot.T, atol=3e-6)
    # the other particles are untouched
    np.testing.assert_array_equal(ps.x.to_numpy()[:n][~obj], x[:n][~obj])
