"""Shared test harness: build the GPU implementation and the oracle on the
same scene and drive both with identical host uniforms."""
from __future__ import annotations

import copy
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

PS_FIELDS = ["object_id", "x", "x_0", "v", "acceleration", "m_V", "m", "density", "pressure",
             "material", "color", "is_dynamic", "vorticity", "vorticity_vis", "dfsph_factor",
             "density_adv", "is_active", "is_sample", "life_time", "grid_ids"]
# solver fields whose stale entries persist across steps (compared over capacity)
SOLVER_FULL = [("d_vort_smoothed", "d_vort_smoothed"), ("kinematic_vort_num", "kvn"),
               ("kernel_density", "kernel_density"), ("sample_idx", "sample_idx"),
               ("d_vort", "d_vort"), ("v_star", "v_star"), ("grad_v", "grad_v"), ("x_temp", "x_temp")]


def load_scene(name: str) -> dict:
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def make_karman_pair(scene: dict):
    from config_builder import SimConfig
    from oracle.refsim import RefSim
    from particle_system import ParticleSystem
    ps = ParticleSystem(SimConfig(scene=copy.deepcopy(scene)))
    solver = ps.build_solver()
    solver.initialize()
    ref = RefSim(copy.deepcopy(scene))
    ref.initialize()
    return ps, solver, ref


def make_ring_pair(nx, ny, nz, n_rings=2, seed=3, mc_samples=128, far_field="exact", **kw):
    import scenes
    from oracle.refsim import RefSim
    ps, solver, info = scenes.build_ring_system(nx, ny, nz, n_rings=n_rings, seed=seed,
                                                mc_samples=mc_samples, far_field=far_field, **kw)
    ref = RefSim(copy.deepcopy(info["scene"]))
    n = info["pos"].shape[0]
    ref.add_particles(0, n, info["pos"], info["vel"], np.full(n, 1000.0, np.float32),
                      np.zeros(n, np.float32), np.ones(n, np.int32), np.ones(n, np.int32),
                      np.tile(np.array([[50, 100, 200]], np.int32), (n, 1)))
    ref.initialize()
    return ps, solver, ref, info


def gpu_field(ps, solver, name):
    f = getattr(ps, name, None)
    if f is None:
        f = getattr(solver, name)
    a = f.to_numpy()
    if a.dtype == bool:
        a = a.astype(np.uint8)
    return a.reshape(a.shape[0], -1) if a.ndim > 1 else a


def ref_field(ref, name):
    a = ref.field(name)
    return a.reshape(a.shape[0], -1) if a.ndim > 1 else a


def mismatches(ps, solver, ref, ps_fields=PS_FIELDS, solver_fields=SOLVER_FULL):
    """Names of fields that are not bit-identical (ps fields over [0, particle_num))."""
    pn = ref.particle_num
    bad = []
    if ps.particle_num[None] != pn:
        bad.append(f"particle_num {ps.particle_num[None]} != {pn}")
        return bad
    for name in ps_fields:
        a = gpu_field(ps, solver, name)[:pn]
        b = ref_field(ref, name)[:pn]
        if not np.array_equal(a, b, equal_nan=True) if a.dtype.kind == "f" else not np.array_equal(a, b):
            d = np.abs(a.astype(np.float64) - b.astype(np.float64))
            bad.append(f"{name}: {int((d > 0).sum())} differ, max {np.nanmax(d):.3e}")
    for gname, rname in solver_fields:
        a = gpu_field(ps, solver, gname)
        b = ref_field(ref, rname)
        if not np.array_equal(a, b, equal_nan=True) if a.dtype.kind == "f" else not np.array_equal(a, b):
            d = np.abs(a.astype(np.float64) - b.astype(np.float64))
            bad.append(f"{gname}: {int((d > 0).sum())} differ, max {np.nanmax(d):.3e}")
    gp = ps.grid_particles_num.to_numpy()
    if not np.array_equal(gp, ref.grid_particles_num()):
        bad.append("grid_particles_num")
    return bad


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))
