"""Synthetic vortex-ring scenes for the 1M-particle, precision, batching and
differentiation workloads.

A scene is a jittered cubic lattice of fluid particles (spacing 2r, jitter
uniform +-0.1 r, seeded) in the box [0, n_x 2r] x [0, n_y 2r] x [0, n_z 2r],
with the initial velocity induced by seeded Lamb-Oseen vortex rings:
centres uniform in the box shrunk by ``margin``, axes uniform on the sphere,
radii U[0.1, 0.3], circulations N(1, 0.2), core radius sigma (default the
support radius 4r).  Ring velocity: each ring is discretised into K points
with line elements dl = (2 pi R / K) t, u(x) = sum G/(4 pi) dl x (x - p) /
|x - p|^3 (1 - exp(-|x - p|^2 / sigma^2)), evaluated in FP64 on the GPU
(``dimcv_ring_velocity``).

The scene uses the reference's own configuration keys (no vortex domain, no
transit area, gravity 0) plus ``mcSamples`` for fixed-M importance sampling.
"""
from __future__ import annotations

import numpy as np
import torch

import _native as N


def ring_scene_config(nx: int, ny: int, nz: int, r: float = 0.01, dt: float = 2e-4,
                      mc_samples: int | None = 4096, far_field: str = "fast", viscosity: float = 1e-3,
                      seed: int = 2, periodic: bool = False) -> dict:
    d = 2 * r
    cfg = {
        "domainStart": [0.0, 0.0, 0.0],
        "domainEnd": [nx * d, ny * d, nz * d],
        "particleRadius": r,
        "density0": 1000,
        "simulationMethod": 0,
        "gravitation": [0.0, 0.0, 0.0],
        "timeStepSize": dt,
        "volumeScale": 0.2,
        "GenProbPara": 5,
        "DelProbPara": 3.5,
        "reservedCapacity": nx * ny * nz,
        "bounceBackBoundary": False,
        "farField": far_field,
        "viscosity": viscosity,
        "rngSeed": seed,
    }
    if periodic:
        cfg["periodic"] = True
    if mc_samples:
        cfg["mcSamples"] = int(mc_samples)
    return {"Configuration": cfg}


def jittered_lattice(nx: int, ny: int, nz: int, r: float, rng: np.random.Generator) -> np.ndarray:
    """Positions (i + 1/2) 2r + U(-0.1 r, 0.1 r) per axis, 'ij' order, f32."""
    d = 2 * r
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    base = np.stack([ii, jj, kk], axis=-1).reshape(-1, 3).astype(np.float64)
    pos = (base + 0.5) * d + rng.uniform(-0.1 * r, 0.1 * r, size=base.shape)
    return pos.astype(np.float32)


def random_rings(n_rings: int, box, rng: np.random.Generator, margin: float = 0.3,
                 gamma_mean: float = 1.0, gamma_std: float = 0.2):
    box = np.asarray(box, dtype=np.float64)
    lo = np.minimum(margin, 0.5 * box)
    hi = np.maximum(box - margin, 0.5 * box)
    centres = rng.uniform(lo, hi, size=(n_rings, 3))
    axes = rng.normal(size=(n_rings, 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    radii = rng.uniform(0.1, 0.3, size=n_rings)
    gammas = rng.normal(gamma_mean, gamma_std, size=n_rings)
    return centres, axes, radii, gammas


def ring_points(centres, axes, radii, gammas, K: int = 256) -> np.ndarray:
    """(n_rings K, 7) table: point, line element, G / (4 pi)."""
    rows = []
    for c, a, R, G in zip(centres, axes, radii, gammas):
        a = np.asarray(a, np.float64) / np.linalg.norm(a)
        t = np.array([1.0, 0.0, 0.0]) if abs(a[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
        e1 = t - np.dot(t, a) * a
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(a, e1)
        th = 2.0 * np.pi * (np.arange(K) + 0.5) / K
        cs, sn = np.cos(th)[:, None], np.sin(th)[:, None]
        pts = np.asarray(c)[None, :] + R * (cs * e1 + sn * e2)
        dl = (2.0 * np.pi * R / K) * (-sn * e1 + cs * e2)
        g = np.full((K, 1), G / (4.0 * np.pi))
        rows.append(np.concatenate([pts, dl, g], axis=1))
    return np.ascontiguousarray(np.concatenate(rows, axis=0), dtype=np.float64)


def ring_velocity(x: torch.Tensor, table: np.ndarray, sigma: float) -> torch.Tensor:
    """FP64 ring-induced velocity at device positions x (n, 3) -> f32 (n, 3)."""
    pts = torch.as_tensor(table, dtype=torch.float64, device=x.device)
    vel = torch.zeros_like(x)
    rc = N.lib().dimcv_ring_velocity(pts.shape[0], pts.data_ptr(), 1.0 / (sigma * sigma), x.shape[0],
                                     x.data_ptr(), vel.data_ptr(), torch.cuda.current_stream().cuda_stream)
    N.check(rc, "dimcv_ring_velocity")
    return vel


def make_ring_scene(nx: int, ny: int, nz: int, n_rings: int = 8, r: float = 0.01, seed: int = 2,
                    mc_samples: int | None = 4096, far_field: str = "fast", dt: float = 2e-4,
                    viscosity: float = 1e-3, sigma: float | None = None, K: int = 256,
                    periodic: bool = False, trim: float = 0.0, device: str = "cuda") -> dict:
    """Configuration, positions and initial velocities of one ring scene.

    ``trim`` > 0 moves the box's upper x face ``trim`` particle diameters
    inwards, so that the outermost lattice layer starts outside the box and
    is deactivated by the first grid build (update_activity).
    """
    rng = np.random.default_rng(seed)
    pos = jittered_lattice(nx, ny, nz, r, rng)
    box = np.array([nx, ny, nz], dtype=np.float64) * 2 * r
    rings = random_rings(n_rings, box, rng)
    sigma = 4 * r if sigma is None else sigma
    table = ring_points(*rings, K=K)
    scene = ring_scene_config(nx, ny, nz, r=r, dt=dt, mc_samples=mc_samples, far_field=far_field,
                              viscosity=viscosity, seed=seed, periodic=periodic)
    if trim:
        scene["Configuration"]["domainEnd"][0] = (nx - trim) * 2 * r
    xt = torch.as_tensor(pos, device=device)
    vel = ring_velocity(xt, table, sigma).cpu().numpy()
    return {"pos": pos, "vel": vel, "rings": rings, "table": table, "sigma": sigma, "scene": scene}


def add_ring_particles(ps, pos: np.ndarray, vel: np.ndarray) -> int:
    """Append a ring scene's fluid particles (object 0, rest density, colour)."""
    n = pos.shape[0]
    return ps.add_particles(0, n, pos, vel, np.full(n, 1000.0, np.float32), np.zeros(n, np.float32),
                            np.ones(n, np.int32), np.ones(n, np.int32),
                            np.tile(np.array([[50, 100, 200]], np.int32), (n, 1)))


def build_ring_system(nx: int, ny: int, nz: int, n_rings: int = 8, r: float = 0.01, seed: int = 2,
                      mc_samples: int | None = 4096, far_field: str = "fast", dt: float = 2e-4,
                      viscosity: float = 1e-3, sigma: float | None = None, K: int = 256,
                      periodic: bool = False, trim: float = 0.0, device: str = "cuda"):
    """ParticleSystem + KarmanVortexSolver for a ring-lattice scene, initialised.

    Returns (ps, solver, info) where info holds the positions, ring table and
    initial velocities (host copies) for feeding another implementation.
    """
    from config_builder import SimConfig
    from particle_system import ParticleSystem

    info = make_ring_scene(nx, ny, nz, n_rings=n_rings, r=r, seed=seed, mc_samples=mc_samples,
                           far_field=far_field, dt=dt, viscosity=viscosity, sigma=sigma, K=K,
                           periodic=periodic, trim=trim, device=device)
    ps = ParticleSystem(SimConfig(scene=info["scene"]), device=device)
    add_ring_particles(ps, info["pos"], info["vel"])
    solver = ps.build_solver()
    solver.initialize()
    return ps, solver, info
