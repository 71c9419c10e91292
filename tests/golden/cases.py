"""Golden cases: small seeded scenes run for a few steps.

Each case is defined once here and run by the oracle (to write / check the
committed fixtures) and by the GPU implementation (to check against them).
The reference ships no golden vectors for this path; these fixtures pin the
oracle against regressions and give the GPU tests a stored expectation.
"""
from __future__ import annotations

import copy
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = {
    # 16 x 16 x 8 jittered lattice, 2 vortex rings, M = 64 inverse-CDF samples
    "rings16": dict(kind="rings", dims=(16, 16, 8), n_rings=2, seed=21, m=64, steps=3),
    # 10 x 10 x 10 perfect lattice with a linear shear, M = 16
    "shear10": dict(kind="shear", n=10, m=16, steps=2),
    # the mini Karman scene (emitter, walls, carved cylinder, vortex domain)
    "mini_karman": dict(kind="karman", scene="mini_karman.json", steps=34),
}

FIELDS = ["x", "v", "density", "density_adv", "dfsph_factor", "vorticity", "is_sample", "life_time",
          "material", "grid_ids"]
SOLVER_FIELDS = ["kernel_density", "d_vort_smoothed", "sample_idx"]


def shear_scene(n: int, m: int) -> tuple[dict, np.ndarray, np.ndarray]:
    r = 0.01
    d = 2 * r
    cfg = {"domainStart": [0, 0, 0], "domainEnd": [(n + 2) * d, (n + 2) * d, (n + 2) * d],
           "particleRadius": r, "density0": 1000, "simulationMethod": 0,
           "gravitation": [0, 0, 0], "timeStepSize": 5e-4, "volumeScale": 0.2,
           "GenProbPara": 5, "DelProbPara": 3.5, "reservedCapacity": n ** 3,
           "mcSamples": m, "farField": "exact"}
    g = (np.arange(n) + 1.5) * d
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    vel = np.zeros_like(pos)
    vel[:, 0] = 2.0 * (pos[:, 1] - pos[:, 1].mean())
    vel[:, 2] = 0.5 * np.sin(6.0 * pos[:, 0])
    return {"Configuration": cfg}, pos, vel.astype(np.float32)


def ring_inputs(case):
    import scenes as SH
    nx, ny, nz = case["dims"]
    rng = np.random.default_rng(case["seed"])
    pos = SH.jittered_lattice(nx, ny, nz, 0.01, rng)
    rings = SH.random_rings(case["n_rings"], np.array([nx, ny, nz]) * 0.02, rng, margin=0.05)
    return pos, rings


def uniforms(case, step, n_max):
    rng = np.random.default_rng(1000 + 17 * step)
    if case["kind"] == "karman":
        return dict(u1=rng.random(n_max, dtype=np.float32), u2=rng.random(n_max, dtype=np.float32))
    return dict(u_cdf=rng.random(case["m"]))


def add_block(sim, pos, vel):
    n = pos.shape[0]
    sim.add_particles(0, n, pos, vel, np.full(n, 1000.0, np.float32), np.zeros(n, np.float32),
                      np.ones(n, np.int32), np.ones(n, np.int32),
                      np.tile(np.array([[50, 100, 200]], np.int32), (n, 1)))


def run_oracle(name: str) -> dict:
    from oracle.refsim import RefSim, ring_velocity
    case = CASES[name]
    if case["kind"] == "karman":
        scene = json.load(open(os.path.join(HERE, case["scene"])))
        sim = RefSim(copy.deepcopy(scene))
    elif case["kind"] == "shear":
        scene, pos, vel = shear_scene(case["n"], case["m"])
        sim = RefSim(copy.deepcopy(scene))
        add_block(sim, pos, vel)
    else:
        import scenes as SH
        pos, rings = ring_inputs(case)
        nx, ny, nz = case["dims"]
        scene = SH.ring_scene_config(nx, ny, nz, mc_samples=case["m"], far_field="exact", seed=case["seed"])
        vel = ring_velocity(pos, *rings, sigma=0.04, K=64)
        sim = RefSim(copy.deepcopy(scene))
        add_block(sim, pos, vel)
    sim.initialize()
    for k in range(case["steps"]):
        sim.step(**uniforms(case, k, sim.particle_max_num))
    pn = sim.particle_num
    out = {"particle_num": np.array(pn), "kv": np.array(sim.last_kv), "kp": np.array(sim.last_kp),
           "num_samples": np.array(sim.num_samples),
           "grid_particles_num": sim.grid_particles_num().copy()}
    for f in FIELDS:
        out[f] = sim.field(f)[:pn].copy()
    for f in SOLVER_FIELDS:
        out[f] = sim.field({"kinematic_vort_num": "kvn"}.get(f, f)).copy()
    return out


def run_gpu(name: str) -> dict:
    import torch
    from config_builder import SimConfig
    from particle_system import ParticleSystem
    case = CASES[name]
    if case["kind"] == "karman":
        scene = json.load(open(os.path.join(HERE, case["scene"])))
        ps = ParticleSystem(SimConfig(scene=copy.deepcopy(scene)))
    elif case["kind"] == "shear":
        scene, pos, vel = shear_scene(case["n"], case["m"])
        ps = ParticleSystem(SimConfig(scene=copy.deepcopy(scene)))
        add_block(ps, pos, vel)
    else:
        import scenes as SH
        from oracle.refsim import ring_velocity
        pos, rings = ring_inputs(case)
        nx, ny, nz = case["dims"]
        scene = SH.ring_scene_config(nx, ny, nz, mc_samples=case["m"], far_field="exact", seed=case["seed"])
        vel = ring_velocity(pos, *rings, sigma=0.04, K=64)   # identical initial input
        ps = ParticleSystem(SimConfig(scene=copy.deepcopy(scene)))
        add_block(ps, pos, vel)
    solver = ps.build_solver()
    solver.initialize()
    for k in range(case["steps"]):
        solver.step(**uniforms(case, k, ps.particle_max_num))
    torch.cuda.synchronize()
    pn = ps.particle_num[None]
    out = {"particle_num": np.array(pn), "kv": np.array(solver.last_iterations_v),
           "kp": np.array(solver.last_iterations_p), "num_samples": np.array(solver.num_samples[None]),
           "grid_particles_num": ps.grid_particles_num.to_numpy()}
    for f in FIELDS:
        a = getattr(ps, f).to_numpy()[:pn]
        out[f] = a.astype(np.uint8) if a.dtype == bool else a
    for f in SOLVER_FIELDS:
        out[f] = getattr(solver, f).to_numpy()
    return out


def fixture_path(name: str) -> str:
    return os.path.join(HERE, f"{name}.npz")
