"""Particle data and uniform grid (drop-in for particle_system.py of the reference).

``ParticleSystem`` keeps the reference's constructor, attribute names and
methods (particle_system.py:12-307, 309-314, 334-347, 555-560, 641-696).
Fields live in HBM as PyTorch tensors (``_field.Field``); every kernel runs in
libdimcv.so (gfx950 HIP) through the C ABI in include/dimcv.h.

Differences of substance from the reference (all deliberate):
  * the grid build (update_activity, partition, cell keys, prefix sum,
    counting sort) is one call, ``dimcv_grid_build``, producing the serial
    reference slot order with a stable radix sort;
  * it also builds each fluid particle's neighbour list (the set and order
    ``for_all_neighbors`` visits), reused by every neighbour pass of the step;
  * an active particle whose flat cell index (formed as the reference forms
    it, so one exactly on the upper y or z face lands in the next row's cell)
    falls outside the cell table raises instead of writing out of bounds;
    slots at or beyond ``particle_num`` hold unspecified values
    (the reference fills them with -1/0/255 sentinels; nothing reads them);
  * 3-D only; mesh rigid bodies (trimesh) and dynamic rigid blocks raise.
"""
from __future__ import annotations

from functools import reduce

import numpy as np
import torch

import _native as N
from _field import CounterField, DeviceScalarField, Field


class ParticleSystem:

    def __init__(self, config, GGUI=False, device: str = "cuda"):
        self.cfg = config
        self.GGUI = GGUI
        self.device = torch.device(device)
        if self.device.type != "cuda" or not torch.cuda.is_available():
            raise RuntimeError("ParticleSystem needs a ROCm GPU (libdimcv runs on gfx950 only)")

        self.domain_start = np.array(self.cfg.get_cfg("domainStart"))
        self.domain_end = np.array(self.cfg.get_cfg("domainEnd"))
        self.domain_size = self.domain_end - self.domain_start
        self.window_res = self.cfg.get_cfg("windowRes")
        self.dim = len(self.domain_size)
        assert self.dim > 1
        if self.dim != 3:
            raise NotImplementedError("only 3-D scenes are supported")
        self.simulation_method = self.cfg.get_cfg("simulationMethod")

        self.material_solid = 0
        self.material_fluid = 1

        self.particle_radius = self.cfg.get_cfg("particleRadius")
        self.particle_diameter = 2 * self.particle_radius
        self.support_radius = self.particle_radius * 4.0
        self.inv_support_radius = 1.0 / self.support_radius
        self.m_V0 = 0.8 * self.particle_diameter ** self.dim

        self.grid_size = self.support_radius
        self.inv_grid_size = 1.0 / self.grid_size
        self.grid_num = np.ceil(self.domain_size / self.grid_size).astype(int)
        self.flattened_grid_num = int(reduce(lambda a, b: a * b, self.grid_num))
        self.padding = self.grid_size

        self.object_collection = dict()
        self.object_id_rigid_body = set()
        self.obj_id_emitters = set()
        self.obj_id_fluid_blocks = set()
        self.obj_id_rigid_blocks = set()

        fluid_blocks = self.cfg.get_fluid_blocks()
        fluid_particle_num = 0
        for fluid in fluid_blocks:
            n = self.compute_cube_particle_num(fluid["start"], fluid["end"])
            fluid["particleNum"] = n
            self.object_collection[fluid["objectId"]] = fluid
            fluid_particle_num += n

        rigid_blocks = self.cfg.get_rigid_blocks()
        rigid_particle_num = 0
        # dynamic rigid blocks: moving boundary volumes every step and the
        # pressure update's reactions on them (dimcv_sph.py:536-549)
        self.has_dynamic_rigid = any(bool(r.get("isDynamic")) for r in rigid_blocks)
        for rigid in rigid_blocks:
            n = self.compute_cube_particle_num(rigid["start"], rigid["end"])
            rigid["particleNum"] = n
            self.object_collection[rigid["objectId"]] = rigid
            self.obj_id_rigid_blocks.add(rigid["objectId"])
            rigid_particle_num += n
        if self.cfg.get_rigid_bodies():
            raise NotImplementedError("mesh rigid bodies (trimesh voxelisation) are not supported")

        self.fluid_emitters = self.cfg.get_fluid_emitters()
        self.num_emitters = 0
        for emitter in self.fluid_emitters:
            obj_id = emitter["objectId"]
            self.obj_id_emitters.add(obj_id)
            self.num_emitters += 1
            emitter["particleNum"] = 0
            self.object_collection[obj_id] = emitter

        self.particle_max_num = fluid_particle_num + rigid_particle_num
        self.num_rigid_bodies = len(rigid_blocks)
        if self.num_rigid_bodies > 0:
            # rest centres of mass of the shape matching (particle_system.py:126-132), host side
            self.rigid_rest_cm = np.zeros((self.num_rigid_bodies + len(fluid_blocks) + len(self.fluid_emitters), 3),
                                          dtype=np.float32)
        self.reserved_capacity = self.cfg.get_cfg("reservedCapacity")
        if self.reserved_capacity is not None:
            self.particle_max_num += self.reserved_capacity
        n = int(self.particle_max_num)
        dev = self.device

        # device scalars (ABI counters / FP64 scalars)
        self._counters = torch.zeros(N.C_COUNT, dtype=torch.int32, device=dev)
        self._dscal = torch.zeros(N.D_COUNT, dtype=torch.float64, device=dev)
        self.particle_num = CounterField(self._counters, N.C_PARTICLE_NUM)
        self.fluid_particle_num = CounterField(self._counters, N.C_FLUID_NUM)
        self.solid_particle_num = CounterField(self._counters, N.C_SOLID_NUM)
        self.num_particles_each_emitter = Field(max(self.num_emitters, 1), 0, torch.int32, dev)

        # grid
        nc = self.flattened_grid_num
        self.grid_particles_num = Field(nc, 0, torch.int32, dev)
        self.grid_particles_num_temp = Field(nc, 0, torch.int32, dev)

        # particle fields (particle_system.py:143-174) and sort buffers (:176-220)
        F32, I32, U8 = torch.float32, torch.int32, torch.uint8
        spec = {
            "object_id": (0, I32, False), "x": (3, F32, False), "x_0": (3, F32, False),
            "v": (3, F32, False), "acceleration": (3, F32, False), "m_V": (0, F32, False),
            "m": (0, F32, False), "density": (0, F32, False), "pressure": (0, F32, False),
            "material": (0, I32, False), "color": (3, I32, False), "is_dynamic": (0, U8, True),
            "vorticity": (3, F32, False), "vorticity_vis": (3, F32, False),
            "dfsph_factor": (0, F32, False), "density_adv": (0, F32, False),
            "is_active": (0, U8, True), "is_sample": (0, U8, True), "life_time": (0, F32, False),
            "grid_ids": (0, I32, False), "particle_id": (0, I32, False),
        }
        for name, (c, dt, b) in spec.items():
            setattr(self, name, Field(n, c, dt, dev, as_bool=b))
        for name in N.SORTED_FIELDS:
            c, dt, b = spec[name]
            setattr(self, name + "_buffer", Field(n, c, dt, dev, as_bool=b))
        self.grid_ids_new = Field(n, 0, I32, dev)

        # build-owned: fluid list, neighbour lists, workspace
        self.nbr_kmax = int(self.cfg.get_cfg("neighborListWidth") or 128)
        self.fluid_list = Field(n, 0, I32, dev)
        self.nbr_count = Field(n, 0, I32, dev)
        self.nbr = torch.zeros((self.nbr_kmax, n), dtype=torch.int32, device=dev)
        ws = N.lib().dimcv_workspace_bytes(n, nc, n)
        self._workspace = torch.zeros(int(ws), dtype=torch.uint8, device=dev)
        self._host_counters = np.zeros(N.C_COUNT, dtype=np.int32)
        self._host_counters_ptr = self._host_counters.ctypes.data
        self._pn = 0          # exact particle_num, mirrored on the host
        self._n_inserted = 0  # particles inserted so far (next particle_id)
        self._nfl = 0         # exact fluid-list length after the last grid build
        self._nbr_valid = False
        self.solver = None
        self._state = None
        # the grid build swaps the sorted fields with their buffers every step,
        # so the state alternates between two pointer sets: one cached per parity
        self._state_cache = {}
        self._parity = 0
        self._swap_pairs = None  # (field, buffer) of every sorted field

        self.params = N.Params()
        self._init_params()

        self.x_vis_buffer = None
        self.cnt = 0

        for fluid in fluid_blocks:
            obj_id = fluid["objectId"]
            self.obj_id_fluid_blocks.add(obj_id)
            offset = np.array(fluid["translation"])
            start = np.array(fluid["start"]) + offset
            end = np.array(fluid["end"]) + offset
            self.add_cube(object_id=obj_id, lower_corner=start,
                          cube_size=(end - start) * np.array(fluid["scale"]),
                          velocity=fluid["velocity"], density=fluid["density"], is_dynamic=1,
                          color=fluid["color"], material=1)
        for rigid in rigid_blocks:
            obj_id = rigid["objectId"]
            offset = np.array(rigid["translation"])
            start = np.array(rigid["start"]) + offset
            end = np.array(rigid["end"]) + offset
            self.add_cube(object_id=obj_id, lower_corner=start,
                          cube_size=(end - start) * np.array(rigid["scale"]),
                          velocity=rigid["velocity"], density=rigid["density"],
                          is_dynamic=rigid["isDynamic"], color=rigid["color"], material=0)
        self.solid_particle_num[None] = rigid_particle_num

    # ------------------------------------------------------------------ ABI --
    def _init_params(self):
        p = self.params
        f32 = np.float32
        for d in range(3):
            p.grid_num[d] = int(self.grid_num[d])
            p.domain_start[d] = f32(self.domain_start[d])
            p.domain_end[d] = f32(self.domain_end[d])
            p.box_size[d] = f32(self.domain_size[d])
            p.bounce_hi[d] = f32(self.domain_size[d] - self.padding)
        p.n_cells = int(self.flattened_grid_num)
        p.n_max = int(self.particle_max_num)
        p.nbr_kmax = self.nbr_kmax
        p.periodic = 1 if self.cfg.get_cfg("periodic") else 0
        if p.periodic and (np.any(self.domain_start != 0) or np.any(self.grid_num < 3)):
            raise ValueError("a periodic box must start at the origin and span at least 3 cells per axis")
        p.grid_size = f32(self.grid_size)
        p.support_radius = f32(self.support_radius)
        p.inv_support_radius = f32(self.inv_support_radius)
        inv3 = f32(self.inv_support_radius ** self.dim)
        p.kernel_k = f32(2.5465) * inv3
        p.kernel_dk = (f32(6.0) * f32(2.5465)) * inv3
        p.m_eps = f32(1e-5)
        p.eps = f32(1e-6)
        p.inv_eps = f32(1.0 / 1e-6)
        p.bounce_padding = f32(self.padding)
        p.m_V0 = f32(self.m_V0)
        p.dyn_rigid = 1 if self.has_dynamic_rigid else 0
        p.far_r2_max_rejected = float(far_threshold(f32(self.support_radius)))
        p.nbr_r2_limit = float(near_threshold(f32(self.support_radius)))

    def state(self) -> N.State:
        """The dimcv_state of the current buffers (cached; rebuilt after a swap)."""
        if self._state is not None:
            return self._state
        s = self._state_cache.get(self._parity)
        if s is not None:
            self._state = s
            return s
        s = N.State()
        for name in N.STATE_POINTERS:
            if name in ("d_vort", "d_vort_smoothed", "v_star", "grad_v", "x_temp", "sample_idx",
                        "kinematic_vort_num", "kernel_density"):
                f = getattr(self.solver, name, None) if self.solver is not None else None
                setattr(s, name, f.ptr() if f is not None else None)
            else:
                setattr(s, name, getattr(self, name).ptr())
        s.n_emitters = self.num_emitters
        s.fluid_list = self.fluid_list.ptr()
        s.nbr_count = self.nbr_count.ptr()
        s.nbr = self.nbr.data_ptr()
        vv = getattr(self.solver, "v_vort", None) if self.solver is not None else None
        s.v_vort = vv.ptr() if vv is not None else None
        s.particle_id = self.particle_id.ptr()
        s.particle_id_buffer = self.particle_id_buffer.ptr()
        s.counters = self._counters.data_ptr()
        s.dscal = self._dscal.data_ptr()
        s.workspace = self._workspace.data_ptr()
        s.workspace_bytes = self._workspace.numel()
        self._state = s
        self._state_cache[self._parity] = s
        return s

    def invalidate_state(self):
        """Forget the cached states (after any change of a field's storage)."""
        self._state = None
        self._state_cache.clear()
        self._swap_pairs = None

    def swap_sorted_fields(self):
        """Exchange every sorted field with its buffer (the grid build's
        gather writes the buffers) and flip the cached state's parity."""
        pairs = self._swap_pairs
        if pairs is None:
            pairs = self._swap_pairs = [(getattr(self, f), getattr(self, f + "_buffer")) for f in N.SORTED_FIELDS]
        for a, b in pairs:
            a.t, b.t = b.t, a.t
        self._parity ^= 1
        self._state = None

    def stream(self) -> int:
        """The current HIP stream of this system's device (raw handle)."""
        return torch._C._cuda_getCurrentRawStream(self.device.index or 0)

    def call(self, name: str, *args):
        st = self.state()
        fn = getattr(N.lib(), name)
        N.check(fn(ctypes_ref(st), ctypes_ref(self.params), *args), name)

    # ------------------------------------------------------------- insertion --
    def build_solver(self):
        if self.simulation_method == 0:
            from karman_vortex import KarmanVortexSolver
            return KarmanVortexSolver(self)
        raise NotImplementedError(f"Solver type {self.simulation_method} has not been implemented.")

    def add_particles(self, object_id: int, new_particles_num: int, new_particles_positions,
                      new_particles_velocity, new_particle_density, new_particle_pressure,
                      new_particles_material, new_particles_is_dynamic, new_particles_color):
        """particle_system.py:334-380: append at particle_num, drop beyond capacity."""
        n = int(new_particles_num)
        if n <= 0:
            return 0
        dev = self.device

        def up(a, dt, shape):
            t = torch.as_tensor(np.ascontiguousarray(a), dtype=dt).reshape(shape)
            return t.to(dev)

        F32, I32 = torch.float32, torch.int32
        pos = up(new_particles_positions, F32, (n, 3))
        vel = up(new_particles_velocity, F32, (n, 3))
        den = up(new_particle_density, F32, (n,))
        pre = up(new_particle_pressure, F32, (n,))
        mat = up(new_particles_material, I32, (n,))
        dyn = up(new_particles_is_dynamic, I32, (n,))
        col = up(new_particles_color, I32, (n, 3))
        dropped = np.zeros(1, np.int32)
        self.call("dimcv_add_particles", int(object_id), n, pos.data_ptr(), vel.data_ptr(),
                  den.data_ptr(), pre.data_ptr(), mat.data_ptr(), dyn.data_ptr(), col.data_ptr(),
                  dropped.ctypes.data, self.stream())
        added = n - int(dropped[0])
        self._note_appended(added)
        return added

    def appended(self, n: int) -> int:
        """Host bookkeeping after n particles were appended on the device by
        an asynchronous entry (dimcv_emit): the slots written are those below
        n_max (the host count is exact, so no read-back is needed)."""
        added = max(0, min(int(n), int(self.particle_max_num) - self._pn))
        self._note_appended(added)
        return added

    def _note_appended(self, added: int) -> None:
        # build-owned identity: insertion index of each particle, carried
        # through every sort (particle_id)
        self.particle_id.t[self._pn:self._pn + added] = torch.arange(
            self._n_inserted, self._n_inserted + added, dtype=torch.int32, device=self.device)
        self._n_inserted += added
        self._pn += added
        self._nbr_valid = False

    def compute_cube_particle_num(self, start, end):
        num_dim = [np.arange(start[i], end[i], self.particle_diameter) for i in range(self.dim)]
        return reduce(lambda a, b: a * b, [len(x) for x in num_dim])

    def cube_lattice(self, lower_corner, cube_size) -> np.ndarray:
        """The positions add_cube places (particle_system.py:658-672): numpy
        arange per axis in f64, 'ij' meshgrid, f32, C order (n, 3)."""
        num_dim = [np.arange(lower_corner[i], lower_corner[i] + cube_size[i], self.particle_diameter)
                   for i in range(self.dim)]
        pos = np.array(np.meshgrid(*num_dim, sparse=False, indexing="ij"), dtype=np.float32)
        pos = pos.reshape(-1, reduce(lambda a, b: a * b, list(pos.shape[1:]))).transpose()
        return np.ascontiguousarray(pos)

    def add_cube(self, object_id, lower_corner, cube_size, material, is_dynamic, color=(0, 0, 0),
                 density=None, pressure=None, velocity=None):
        """particle_system.py:647-696 (numpy lattice in 'ij' order, f64 arange -> f32)."""
        pos = self.cube_lattice(lower_corner, cube_size)
        n = pos.shape[0]
        if velocity is None:
            vel = np.full_like(pos, 0, dtype=np.float32)
        else:
            vel = np.array([velocity for _ in range(n)], dtype=np.float32)
        mat = np.full(n, material, dtype=np.int32)
        dyn = np.full(n, is_dynamic, dtype=np.int32)
        col = np.stack([np.full(n, c, dtype=np.int32) for c in color], axis=1)
        den = np.full(n, density if density is not None else 1000.0, dtype=np.float32)
        pre = np.full(n, pressure if pressure is not None else 0.0, dtype=np.float32)
        self.add_particles(object_id, n, pos, vel, den, pre, mat, dyn, col)

    # ------------------------------------------------------------------ grid --
    def initialize_particle_system(self):
        """particle_system.py:555-560 + the step's neighbour lists."""
        hc = self._host_counters
        self.call("dimcv_grid_build", int(self._pn), self._host_counters_ptr, self.stream())
        self.swap_sorted_fields()
        self._pn = int(hc[N.C_PARTICLE_NUM])
        self._nfl = int(hc[N.C_FLUID_LIST])
        self.build_neighbor_lists()

    def build_neighbor_lists(self):
        self.call("dimcv_neighbor_build", self._nfl, self.stream())
        self._nbr_valid = True

    def ensure_neighbor_lists(self):
        if not self._nbr_valid:
            self.build_neighbor_lists()

    def update_activity(self):
        raise NotImplementedError("update_activity runs inside initialize_particle_system()")

    def neighbor_lists(self):
        """Host copy of the neighbour lists: list of ascending slot arrays per fluid slot."""
        cnt = self.nbr_count.t[: self._nfl].cpu().numpy()
        nb = self.nbr[:, : self._nfl].cpu().numpy().view(np.uint32)
        fl = self.fluid_list.t[: self._nfl].cpu().numpy()
        return {int(fl[f]): (nb[: min(cnt[f], self.nbr_kmax), f] & 0x7FFFFFFF).astype(np.int64)
                for f in range(self._nfl)}


def ctypes_ref(obj):
    import ctypes
    return ctypes.byref(obj)


def far_threshold(h: np.float32) -> np.float32:
    """Largest f32 t with sqrt_rn(t) <= h, so that sqrt(r2) > h  <=>  r2 > t."""
    h = np.float32(h)
    t = np.float32(h * h)
    inf = np.float32(np.inf)
    while np.sqrt(np.nextafter(t, inf, dtype=np.float32)) <= h:
        t = np.nextafter(t, inf, dtype=np.float32)
    while np.sqrt(t) > h:
        t = np.nextafter(t, np.float32(0), dtype=np.float32)
    return t


def near_threshold(h: np.float32) -> np.float32:
    """Smallest f32 t with sqrt_rn(t) >= h, so that sqrt(r2) < h  <=>  r2 < t."""
    h = np.float32(h)
    t = np.float32(h * h)
    zero = np.float32(0)
    while np.sqrt(np.nextafter(t, zero, dtype=np.float32)) >= h:
        t = np.nextafter(t, zero, dtype=np.float32)
    while np.sqrt(t) < h:
        t = np.nextafter(t, np.float32(np.inf), dtype=np.float32)
    return t
