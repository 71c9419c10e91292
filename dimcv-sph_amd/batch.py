"""Batched independent scenes (build-defined; the reference runs one scene per
process and has no batching).

``SceneBatch`` advances many independent vortex-ring scenes of uneven sizes
with one set of launches per pass (``dimcv_batch_*`` in libdimcv): the scenes
share one particle store, one sort and one neighbour build, and every per-scene
quantity -- kden_sum, the Jacobi residuals and convergence, the CDF and the
far-field sum -- is evaluated over blocks aligned to that scene's own fluid
list, so each scene's state after every step is bitwise the state of the same
scene run alone (``scenes.build_ring_system`` + ``KarmanVortexSolver.step``).

Scenes share particle radius, rest density, time step and far-field mode;
each has its own box, viscosity, ring set, seed and sample count M (default
2% of its particles).  Per-scene uniforms come from ``numpy`` generators
seeded with the scene seed, as the solo solver's ``rngSeed`` generator does.

A step runs eagerly (``step``) or from a HIP graph (``capture`` then
``step_graph``): the Jacobi solves converge on the device (a persistent
launch tests every scene and iterates the ones that have not converged), so
the whole step is captured as one graph per parity of the sorted-buffer swap
and a step is one graph launch plus one status readback.
``autotune`` picks the far-field persistent grid size and the neighbour-build
workgroup size; both only change launch shapes, never results.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import threading
import time
from dataclasses import dataclass

import numpy as np
import torch

import _native as N
from _field import Field
from scenes import add_ring_particles, make_ring_scene, ring_scene_config


@dataclass
class RingSpec:
    """One ring scene of a batch: lattice nx x ny x nz, rings, seed, nu, M."""
    nx: int
    ny: int
    nz: int
    n_rings: int = 4
    seed: int = 0
    viscosity: float = 1e-3
    mc_samples: int = 0     # 0: 2% of the particles
    trim: float = 0.0       # see scenes.make_ring_scene

    @property
    def n(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def m(self) -> int:
        return int(self.mc_samples) if self.mc_samples else max(1, int(round(0.02 * self.n)))


def random_specs(n_scenes: int, seed: int = 7, n_lo: int = 8192, n_hi: int = 131072,
                 nu_lo: float = 1e-4, nu_hi: float = 1e-2, rings=(2, 6)) -> list[RingSpec]:
    """Seeded scenes: particle counts and viscosities log-uniform, 2-6 rings.

    A scene of target count N is an s x s x ceil(N / s^2) lattice with
    s = round(N^(1/3)).
    """
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_scenes):
        target = math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))
        side = max(3, int(round(target ** (1.0 / 3.0))))
        nx = max(3, int(math.ceil(target / (side * side))))
        nu = math.exp(rng.uniform(math.log(nu_lo), math.log(nu_hi)))
        nr = int(rng.integers(rings[0], rings[1] + 1))
        out.append(RingSpec(nx, side, side, n_rings=nr, seed=int(seed * 1000 + k), viscosity=float(nu)))
    return out


def solo_system(spec: RingSpec, far_field: str = "fast", r: float = 0.01, dt: float = 2e-4,
                device: str = "cuda"):
    """The scene of ``spec`` alone: (ps, solver), initialised."""
    from scenes import build_ring_system
    ps, solver, _ = build_ring_system(spec.nx, spec.ny, spec.nz, n_rings=spec.n_rings, r=r, seed=spec.seed,
                                      mc_samples=spec.m, far_field=far_field, dt=dt,
                                      viscosity=spec.viscosity, trim=spec.trim, device=device)
    return ps, solver


class SceneBatch:

    def __init__(self, specs, far_field: str = "fast", r: float = 0.01, dt: float = 2e-4,
                 device: str = "cuda"):
        from config_builder import SimConfig
        from particle_system import ParticleSystem

        self.specs = list(specs)
        S = len(self.specs)
        if S == 0:
            raise ValueError("a batch needs at least one scene")
        self.n_scenes = S
        self.far_field = far_field
        infos = [make_ring_scene(sp.nx, sp.ny, sp.nz, n_rings=sp.n_rings, r=r, seed=sp.seed,
                                 mc_samples=sp.m, far_field=far_field, dt=dt, viscosity=sp.viscosity,
                                 trim=sp.trim, device=device) for sp in self.specs]
        counts = [sp.n for sp in self.specs]
        if max(counts) > 4096 * 4096:
            raise ValueError("a batched scene holds at most 4096^2 particles (one-workgroup tile scan)")
        total = int(sum(counts))
        ext = np.max([np.array(inf["scene"]["Configuration"]["domainEnd"], np.float64) for inf in infos], axis=0)
        cfg = ring_scene_config(1, 1, 1, r=r, dt=dt, mc_samples=self.specs[0].m, far_field=far_field)
        cfg["Configuration"]["domainEnd"] = [float(e) for e in ext]
        cfg["Configuration"]["reservedCapacity"] = total + 256 * S
        ps = ParticleSystem(SimConfig(scene=cfg), device=device)
        # the batched activity kernel keeps no per-emitter counts
        if ps.fluid_emitters:
            raise ValueError("batched scenes have no fluid emitters")
        self.ps = ps
        dev = ps.device

        # scene table: own grid (particle_system.py:52-56), cell base, samples
        h = ps.grid_size
        table = (N.Scene * S)()
        cell_base, sample_base = 0, 0
        self.cell_base, self.sample_base = [], []
        for s, (sp, inf) in enumerate(zip(self.specs, infos)):
            c = inf["scene"]["Configuration"]
            start = np.array(c["domainStart"], np.float64)
            end = np.array(c["domainEnd"], np.float64)
            if np.any(start != 0):
                raise ValueError("batched scenes start at the origin (pos_to_index ignores domain_start)")
            gn = np.ceil((end - start) / h).astype(int)
            e = table[s]
            for d in range(3):
                e.grid_num[d] = int(gn[d])
                e.domain_start[d] = np.float32(start[d])
                e.domain_end[d] = np.float32(end[d])
            e.cell_base = cell_base
            e.n_cells = int(np.prod(gn))
            e.m = sp.m
            e.sample_base = sample_base
            e.visc_coef = np.float32(2 * (ps.dim + 2) * sp.viscosity)
            self.cell_base.append(cell_base)
            self.sample_base.append(sample_base)
            cell_base += e.n_cells
            sample_base += sp.m
        if cell_base >= 2 ** 31 - 1:
            raise ValueError("too many cells in the batch")
        self.total_cells = cell_base
        self.m_total = sample_base
        raw = np.frombuffer(bytes(table), dtype=np.uint8)
        self._scenes = torch.as_tensor(raw.copy(), device=dev)

        # the batch's cell table spans every scene's cells
        ps.grid_particles_num = Field(cell_base, 0, torch.int32, dev)
        ps.grid_particles_num_temp = Field(cell_base, 0, torch.int32, dev)
        ps.params.n_cells = cell_base
        ps.invalidate_state()

        self.solver = ps.build_solver()
        n = int(ps.particle_max_num)
        self.scene_id = Field(n, 0, torch.int32, dev)
        self.scene_id_buffer = Field(n, 0, torch.int32, dev)
        self.scene_counters = torch.zeros((S, N.SC_COUNT), dtype=torch.int32, device=dev)
        self.scene_dscal = torch.zeros((S, N.SD_COUNT), dtype=torch.float64, device=dev)
        self.block_scene = torch.zeros(n // 256 + 1, dtype=torch.int32, device=dev)
        cnt = np.array(counts, np.int32)
        wsb = N.lib().dimcv_batch_workspace_bytes(S, cnt.ctypes.data, n)
        if wsb < 0:
            raise RuntimeError("dimcv_batch_workspace_bytes failed")
        self._bws = torch.zeros(int(wsb), dtype=torch.uint8, device=dev)

        self.slot_offset = []
        off = 0
        for s, inf in enumerate(infos):
            self.slot_offset.append(off)
            k = add_ring_particles(ps, inf["pos"], inf["vel"])
            if k != inf["pos"].shape[0]:
                raise RuntimeError("batch capacity exceeded")
            self.scene_id.t[off:off + k] = s
            off += k
        self.rngs = [np.random.default_rng(sp.seed) for sp in self.specs]
        self._st = None
        self._graph = None
        self._stream = None
        self._parity = 0
        self.last_iterations = np.zeros((S, 2), np.int32)
        self.build_grid()

    # ------------------------------------------------------------------ ABI --
    def state(self) -> N.State:
        if self._st is not None and self.ps._state is not None:
            return self._st
        st = N.State.from_buffer_copy(self.ps.state())
        st.scene_id = self.scene_id.ptr()
        st.scene_id_buffer = self.scene_id_buffer.ptr()
        st.scenes = self._scenes.data_ptr()
        st.n_scenes = self.n_scenes
        st.scene_counters = self.scene_counters.data_ptr()
        st.scene_dscal = self.scene_dscal.data_ptr()
        st.block_scene = self.block_scene.data_ptr()
        st.batch_workspace = self._bws.data_ptr()
        st.batch_workspace_bytes = self._bws.numel()
        self._st = st
        return st

    def _call(self, name: str, *args):
        fn = getattr(N.lib(), name)
        N.check(fn(ctypes.byref(self.state()), ctypes.byref(self.ps.params), *args), name)

    def _stream_handle(self) -> int:
        return torch.cuda.current_stream().cuda_stream

    def _swap_sorted(self):
        for name in N.SORTED_FIELDS:
            getattr(self.ps, name).swap(getattr(self.ps, name + "_buffer"))
        self.scene_id.swap(self.scene_id_buffer)
        self.ps.invalidate_state()
        self._st = None

    def status(self) -> np.ndarray:
        hc = self.ps._host_counters
        self._call("dimcv_batch_status", hc.ctypes.data, self._stream_handle())
        self.ps._pn = int(hc[N.C_PARTICLE_NUM])
        self.ps._nfl = int(hc[N.C_FLUID_LIST])
        return hc

    def build_grid(self):
        """initialize_particle_system of every scene + neighbour lists."""
        self._call("dimcv_batch_grid_build", self._stream_handle())
        self._swap_sorted()
        self._call("dimcv_batch_neighbor_build", self._stream_handle())
        self.status()

    # ------------------------------------------------------------------ step --
    def uniforms(self) -> np.ndarray:
        """This step's inverse-CDF uniforms of every scene (scene s at sample_base[s])."""
        return np.concatenate([g.random(sp.m) for g, sp in zip(self.rngs, self.specs)])

    def _opts(self, u_ptr: int) -> N.StepOpts:
        sv = self.solver
        o = N.StepOpts()
        o.u_cdf, o.m = u_ptr, self.m_total
        o.max_iter_v = sv.m_max_iterations_v
        o.max_iter_p = sv.m_max_iterations
        o.eta_v = sv.inv_dt * sv.max_error_V * 0.01 * sv.density_0
        o.eta_p = sv.max_error * 0.01 * sv.density_0
        o.inv_dt = np.float32(sv.inv_dt)
        o.dt = np.float32(sv.dt[None])
        o.inv_dt2 = np.float32(sv.inv_dt2)
        o.density0 = np.float32(sv.density_0)
        return o

    def step(self, u=None):
        """One step of every scene (KarmanVortexSolver.step per scene), eagerly."""
        if u is None:
            u = self.uniforms()
        tu = torch.as_tensor(np.ascontiguousarray(u, dtype=np.float64)).to(self.ps.device)
        self.build_grid()
        o = self._opts(tu.data_ptr())
        iters = np.zeros((self.n_scenes, 2), np.int32)
        self._call("dimcv_batch_substep", ctypes.byref(o), iters.ctypes.data, self._stream_handle())
        self.last_iterations = iters
        del tu

    # ---------------------------------------------------------------- graphs --
    def capture(self):
        """Capture the step graphs for the current buffers (replaces old ones)."""
        self.release_graph()
        if self._stream is None:
            self._stream = torch.cuda.Stream(device=self.ps.device)
        self._u_dev = torch.zeros(self.m_total, dtype=torch.float64, device=self.ps.device)
        self._gopts = self._opts(self._u_dev.data_ptr())
        torch.cuda.synchronize()
        even = N.State.from_buffer_copy(self.state())
        odd = N.State.from_buffer_copy(even)
        for name in N.SORTED_FIELDS + ["scene_id"]:
            a, b = getattr(odd, name), getattr(odd, name + "_buffer")
            setattr(odd, name, b)
            setattr(odd, name + "_buffer", a)
        handle = ctypes.c_void_p()
        rc = N.lib().dimcv_batch_graph_create(ctypes.byref(even), ctypes.byref(odd), ctypes.byref(self.ps.params),
                                              ctypes.byref(self._gopts), self._stream.cuda_stream,
                                              ctypes.byref(handle))
        N.check(rc, "dimcv_batch_graph_create")
        self._graph = handle
        self._parity = 0

    def release_graph(self):
        if self._graph is not None:
            torch.cuda.synchronize()
            N.lib().dimcv_batch_graph_destroy(self._graph)
            self._graph = None

    def step_graph(self, u=None):
        """One step of every scene from the captured graphs."""
        if self._graph is None:
            raise RuntimeError("capture() first")
        if u is None:
            u = self.uniforms()
        host = torch.from_numpy(np.ascontiguousarray(u, dtype=np.float64))
        iters = np.zeros((self.n_scenes, 2), np.int32)
        with torch.cuda.stream(self._stream):
            self._u_dev.copy_(host)
            rc = N.lib().dimcv_batch_graph_step(self._graph, self._parity, iters.ctypes.data,
                                                self._stream.cuda_stream)
        N.check(rc, "dimcv_batch_graph_step")
        self._swap_sorted()
        self._parity ^= 1
        self.last_iterations = iters

    def __del__(self):
        try:
            self.release_graph()
        except Exception:
            pass

    # ---------------------------------------------------------- inspection --
    def scene_counters_host(self) -> np.ndarray:
        return self.scene_counters.cpu().numpy()

    def scene_field(self, name: str, s: int) -> np.ndarray:
        """Field ``name`` (ps or solver field) of scene s's active slots."""
        sc = self.scene_counters_host()[s]
        a, k = int(sc[N.SC_SLOT_START]), int(sc[N.SC_SLOTS])
        f = getattr(self.ps, name, None)
        if f is None:
            f = getattr(self.solver, name)
        return f.to_numpy()[a:a + k]

    def scene_samples(self, s: int) -> np.ndarray:
        """Scene s's sample slots, relative to the scene's first slot."""
        sc = self.scene_counters_host()[s]
        m = int(sc[N.SC_NUM_SAMPLES])
        b = self.sample_base[s]
        return self.solver.sample_idx.to_numpy()[b:b + m] - int(sc[N.SC_SLOT_START])

    # ------------------------------------------------------ state snapshot --
    def _snapshot_tensors(self):
        ps, sv = self.ps, self.solver
        out = [f.t for f in vars(ps).values() if isinstance(f, Field)]
        out += [f.t for f in vars(sv).values() if isinstance(f, Field)]
        out += [self.scene_id.t, self.scene_id_buffer.t, self.scene_counters, self.scene_dscal,
                self.block_scene, ps._counters, ps._dscal]
        return out

    def snapshot(self):
        return [(t, t.clone()) for t in self._snapshot_tensors()] + [("rng", [g.bit_generator.state for g in self.rngs])]

    def restore(self, snap):
        torch.cuda.synchronize()
        for t, c in snap[:-1]:
            t.copy_(c)
        for g, st in zip(self.rngs, snap[-1][1]):
            g.bit_generator.state = st
        torch.cuda.synchronize()
        self.status()


# ---------------------------------------------------------------- autotuner --
_TUNE_LOCK = threading.Lock()
_TUNE_CACHE: dict = {}


def set_launch(far_workgroups: int, nbr_block: int) -> None:
    N.check(N.lib().dimcv_batch_set_launch(int(far_workgroups), int(nbr_block)), "dimcv_batch_set_launch")


def autotune(batch: SceneBatch, far_workgroups=(512, 1024, 2048, 4096), nbr_blocks=(64, 128, 256),
             steps: int = 3, cache_path: str | None = None) -> dict:
    """Choose the far-field persistent grid and the neighbour-build workgroup
    size by timing graph steps (coordinate search: grid first, then block).

    Launch shapes never change results, and the batch state (fields and
    generators) is restored afterwards.  The choice is cached per device and
    scene-size signature (in memory, and in ``cache_path`` as JSON if given).
    """
    key = "|".join([torch.cuda.get_device_name(batch.ps.device), batch.far_field,
                    ",".join(str(sp.n) for sp in sorted(batch.specs, key=lambda q: q.n))])
    with _TUNE_LOCK:
        if not _TUNE_CACHE and cache_path and os.path.exists(cache_path):
            with open(cache_path) as fh:
                _TUNE_CACHE.update(json.load(fh))
        if key in _TUNE_CACHE:
            cfg = _TUNE_CACHE[key]
            set_launch(cfg["far_workgroups"], cfg["nbr_block"])
            return dict(cfg, cached=True)
        snap = batch.snapshot()
        u = np.full(batch.m_total, 0.5)

        def timed(wg, nb):
            set_launch(wg, nb)
            batch.capture()
            batch.step_graph(u)      # warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                batch.step_graph(u)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / steps
            batch.release_graph()
            batch.restore(snap)
            if batch._parity:        # restored buffers are the even ones again
                batch._swap_sorted()
                batch._parity = 0
            return ms

        trials = {}
        nb0 = 256
        for wg in far_workgroups:
            trials[(wg, nb0)] = timed(wg, nb0)
        wg_best = min((k for k in trials if k[1] == nb0), key=trials.get)[0]
        for nb in nbr_blocks:
            if (wg_best, nb) not in trials:
                trials[(wg_best, nb)] = timed(wg_best, nb)
        best = min(trials, key=trials.get)
        cfg = {"far_workgroups": best[0], "nbr_block": best[1], "ms_per_step": round(trials[best], 4),
               "trials": {f"{k[0]}x{k[1]}": round(v, 4) for k, v in sorted(trials.items())}}
        set_launch(best[0], best[1])
        _TUNE_CACHE[key] = cfg
        if cache_path:
            with open(cache_path, "w") as fh:
                json.dump(_TUNE_CACHE, fh, indent=1)
        return dict(cfg, cached=False)
