"""Oracle (test infrastructure only): Python orchestration of the CPU restatement.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, as the checker.  The product never does.

``RefSim`` restates the host-side logic of the reference classes:

* ``ParticleSystem.__init__`` / ``add_cube`` / ``compute_cube_particle_num``
  (particle_system.py:12-307, 641-696) -- numpy lattice generation;
* ``KarmanVortexSolver.__init__`` / ``step`` / ``substep``
  (karman_vortex.py:10-36, 216-242);
* ``DIMCVSPHSolver.divergence_solve`` / ``pressure_solve`` / ``dimcv`` /
  ``emit_particle`` (dimcv_sph.py:376-393, 597-647, 803-819);
* ``SPHBase.initialize`` (sph_base.py:80-85);

and calls the C++ restatement of every Taichi kernel (refsim.cpp) in the same
order.  ``ti.random`` is replaced by host-fed uniforms (``u1``/``u2`` per slot
for the Bernoulli sample update, ``u_cdf`` for the build-defined inverse-CDF
sampler).
"""
from __future__ import annotations

import ctypes
import os
from functools import reduce

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OrConfig(ctypes.Structure):
    _fields_ = [
        ("domain_start", ctypes.c_double * 3),
        ("domain_end", ctypes.c_double * 3),
        ("particle_radius", ctypes.c_double),
        ("density0", ctypes.c_double),
        ("gravitation", ctypes.c_double * 3),
        ("time_step", ctypes.c_double),
        ("volume_scale", ctypes.c_double),
        ("gen_a", ctypes.c_double),
        ("del_b", ctypes.c_double),
        ("vortex_start", ctypes.c_double * 3),
        ("vortex_end", ctypes.c_double * 3),
        ("transit", ctypes.c_double * 2),
        ("viscosity", ctypes.c_double),
        ("particle_max_num", ctypes.c_int64),
        ("has_vortex_domain", ctypes.c_int32),
        ("has_transit", ctypes.c_int32),
        ("num_emitters", ctypes.c_int32),
        ("periodic", ctypes.c_int32),
    ]


def lib():
    """Load (building if needed) oracle/liboracle.so."""
    global _LIB
    if _LIB is not None:
        return _LIB
    # DIMCV_ORACLE_LIB: a copy built elsewhere (oracle/mutation_check.py
    # points the pins at deliberately broken builds this way)
    path = os.environ.get("DIMCV_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        import subprocess
        subprocess.check_call(["make", "-s", "-C", _HERE])
    L = ctypes.CDLL(path)
    P, I32, I64, F32, F64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_double
    sig = {
        "or_new": (P, [P]), "or_free": (None, [P]), "or_ptr": (P, [P, ctypes.c_char_p]),
        "or_info": (None, [P, P]),
        "or_add_particles": (I32, [P, I32, I32, P, P, P, P, P, P, P]),
        "or_init_cylinder": (None, [P, I32, F32, F32, F32]),
        "or_multiply_time_step": (None, [P, I32, F32]),
        "or_compute_density_error": (F64, [P, F32]),
        "or_compute_kinematic_vorticity_num_sample": (None, [P, P, P, I32]),
        "or_cdf_sample": (I32, [P, P, I32]),
        "or_vort2vel": (None, [P, P]),
        "or_vort2vel_targets": (None, [P, P, I64, P, P, P]),
        "or_direct_bs_f64": (None, [P, P]),
        "or_ring_velocity": (None, [I32, P, P, P, P, F64, I32, I64, P, P]),
        "or_canon_reduce": (F64, [P, I64]),
        "or_canon_scan": (None, [P, P, I64]),
        "or_expf": (F32, [F32]), "or_logf": (F32, [F32]), "or_powf": (F32, [F32, F32]),
        "or_W": (F32, [P, F32]), "or_gradW": (None, [P, P, P]),
        "or_neighbors": (I32, [P, I64, P, I32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    for name in ["or_update_activity", "or_particle_partition", "or_update_grid_id",
                 "or_prefix_sum", "or_counting_sort", "or_initialize_particle_system",
                 "or_compute_static_boundary_volume", "or_compute_moving_boundary_volume",
                 "or_compute_densities", "or_compute_DFSPH_factor", "or_compute_density_change",
                 "or_divergence_solver_iteration_kernel", "or_compute_non_pressure_forces",
                 "or_advect", "or_predict_velocity", "or_compute_v_star", "or_compute_vorticity",
                 "or_compute_grad_v", "or_compute_d_vorticity", "or_compute_all_d_vort_smoothed",
                 "or_record_vort_part_sample_idx", "or_compute_kernel_density",
                 "or_update_particle_life_time", "or_compute_density_adv",
                 "or_pressure_solve_iteration_kernel", "or_copy_x_temp",
                 "or_compute_vorticity_vis", "or_enforce_boundary_3d"]:
        fn = getattr(L, name)
        fn.restype = None
        fn.argtypes = [P]
    _LIB = L
    return L


def _ptr(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def canon_reduce(a) -> float:
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().or_canon_reduce(_ptr(a), a.size)


def canon_scan(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.empty_like(a)
    lib().or_canon_scan(_ptr(a), _ptr(out), a.size)
    return out


def ring_velocity(x: np.ndarray, centres, axes, radii, gammas, sigma: float, K: int = 256):
    """FP64 vortex-ring velocity at f32 positions x (n,3) -> f32 (n,3)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    c = np.ascontiguousarray(centres, dtype=np.float64)
    a = np.ascontiguousarray(axes, dtype=np.float64)
    R = np.ascontiguousarray(radii, dtype=np.float64)
    G = np.ascontiguousarray(gammas, dtype=np.float64)
    out = np.zeros_like(x)
    lib().or_ring_velocity(len(R), _ptr(c), _ptr(a), _ptr(R), _ptr(G), float(sigma), int(K),
                           x.shape[0], _ptr(x), _ptr(out))
    return out


# Field name -> (dtype, components) for numpy views.
_FIELDS = {
    "object_id": (np.int32, 1), "material": (np.int32, 1), "color": (np.int32, 3),
    "grid_ids": (np.int32, 1), "x": (np.float32, 3), "x_0": (np.float32, 3),
    "v": (np.float32, 3), "acceleration": (np.float32, 3), "m_V": (np.float32, 1),
    "m": (np.float32, 1), "density": (np.float32, 1), "pressure": (np.float32, 1),
    "is_dynamic": (np.uint8, 1), "is_active": (np.uint8, 1), "is_sample": (np.uint8, 1),
    "vorticity": (np.float32, 3), "vorticity_vis": (np.float32, 3),
    "dfsph_factor": (np.float32, 1), "density_adv": (np.float32, 1),
    "life_time": (np.float32, 1), "d_vort": (np.float32, 3),
    "d_vort_smoothed": (np.float32, 3), "v_star": (np.float32, 3),
    "grad_v": (np.float32, 9), "x_temp": (np.float32, 3), "kvn": (np.float32, 1),
    "kernel_density": (np.float32, 1), "sample_idx": (np.int32, 1),
}


class RefSim:
    """CPU restatement of ParticleSystem + KarmanVortexSolver for one scene.

    ``scene`` is the parsed scene JSON (dict with a ``Configuration`` block).
    """

    def __init__(self, scene: dict):
        L = lib()
        self.scene = scene
        cfg = scene["Configuration"]
        self.cfg = cfg
        get = cfg.get
        self.domain_start = np.array(get("domainStart"))
        self.domain_end = np.array(get("domainEnd"))
        self.domain_size = self.domain_end - self.domain_start
        self.dim = len(self.domain_size)
        assert self.dim == 3
        self.particle_radius = get("particleRadius")
        self.particle_diameter = 2 * self.particle_radius
        self.support_radius = self.particle_radius * 4.0

        fluid_blocks = scene.get("FluidBlocks", [])
        rigid_blocks = scene.get("RigidBlocks", [])
        self.fluid_emitters = scene.get("FluidEmitters", [])
        fluid_num = sum(self.compute_cube_particle_num(b["start"], b["end"]) for b in fluid_blocks)
        rigid_num = sum(self.compute_cube_particle_num(b["start"], b["end"]) for b in rigid_blocks)
        if scene.get("RigidBodies"):
            raise NotImplementedError("mesh rigid bodies are out of scope")
        self.particle_max_num = fluid_num + rigid_num + (get("reservedCapacity") or 0)
        self.obj_emitters = [e["objectId"] for e in self.fluid_emitters]
        self.emitter_counts = {e["objectId"]: 0 for e in self.fluid_emitters}

        c = OrConfig()
        for d in range(3):
            c.domain_start[d] = float(self.domain_start[d])
            c.domain_end[d] = float(self.domain_end[d])
            c.gravitation[d] = float(get("gravitation")[d])
        c.particle_radius = float(self.particle_radius)
        c.density0 = float(get("density0"))
        c.time_step = float(get("timeStepSize"))
        c.volume_scale = float(get("volumeScale") if get("volumeScale") is not None else 0.0)
        c.gen_a = float(get("GenProbPara") or 0.0)
        c.del_b = float(get("DelProbPara") or 0.0)
        vs, ve = get("vortexEnforcingDomainStart"), get("vortexEnforcingDomainEnd")
        c.has_vortex_domain = 1 if vs is not None else 0
        if vs is not None:
            for d in range(3):
                c.vortex_start[d] = float(vs[d])
                c.vortex_end[d] = float(ve[d])
        ta = get("divergenceTransitArea")
        c.has_transit = 1 if ta is not None else 0
        if ta is not None:
            c.transit[0], c.transit[1] = float(ta[0]), float(ta[1])
        c.viscosity = float(get("viscosity") or 1e-3)
        c.particle_max_num = int(self.particle_max_num)
        c.num_emitters = len(self.fluid_emitters)
        c.periodic = 1 if get("periodic") else 0
        self._c = c
        self.sim = L.or_new(ctypes.byref(c))
        info = np.zeros(5, np.int64)
        L.or_info(self.sim, _ptr(info))
        self.grid_num = info[1:4].copy()
        self.ncells = int(info[4])

        for b in fluid_blocks:
            off = np.array(b["translation"])
            start = np.array(b["start"]) + off
            end = np.array(b["end"]) + off
            self.add_cube(b["objectId"], start, (end - start) * np.array(b["scale"]), material=1,
                          is_dynamic=1, color=b["color"], density=b["density"], velocity=b["velocity"])
        for b in rigid_blocks:
            off = np.array(b["translation"])
            start = np.array(b["start"]) + off
            end = np.array(b["end"]) + off
            self.add_cube(b["objectId"], start, (end - start) * np.array(b["scale"]), material=0,
                          is_dynamic=int(b["isDynamic"]), color=b["color"], density=b["density"],
                          velocity=b["velocity"])
        self.i32("solid_particle_num")[0] = rigid_num

        # KarmanVortexSolver.__init__ (karman_vortex.py:10-17) + DIMCVSPHSolver.__init__
        self.dt = np.float32(get("timeStepSize"))
        self.inv_dt = 1.0 / float(self.dt)
        self.inv_dt2 = 1.0 / (float(self.dt) * float(self.dt))
        self.density_0 = get("density0")
        self.cnt = 0
        self.emit_interval = 1
        for e in self.fluid_emitters:
            init_v = np.array(e["velocity"])
            self.emit_interval = int(np.ceil(self.particle_diameter / (np.linalg.norm(init_v) * float(self.dt))))
            break
        self.num_samples_cfg = get("numSamples")
        self.mc_samples = get("mcSamples")
        L.or_init_cylinder(self.sim, 2, np.float32(0.65), np.float32(0.5), np.float32(0.1))
        self.last_kv = 0
        self.last_kp = 0
        # dimcv_sph.py:21-26 (max_error_V, max_error); tests may tighten them
        self.max_error_V = 0.1
        self.max_error = 0.05

    # -- field access ------------------------------------------------------
    def field(self, name) -> np.ndarray:
        dt, comp = _FIELDS[name]
        p = lib().or_ptr(self.sim, name.encode())
        n = self.particle_max_num * comp
        arr = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dt))), (n,))
        return arr.reshape(-1, comp) if comp > 1 else arr

    def i32(self, name) -> np.ndarray:
        p = lib().or_ptr(self.sim, name.encode())
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_int32)), (1,))

    def emitter_field(self) -> np.ndarray:
        """ps.num_particles_each_emitter (one entry per emitter, by object id)."""
        p = lib().or_ptr(self.sim, b"num_particles_each_emitter")
        n = max(1, len(self.fluid_emitters))
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_int32)), (n,))

    def grid_particles_num(self) -> np.ndarray:
        p = lib().or_ptr(self.sim, b"grid_particles_num")
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_int32)), (self.ncells,))

    @property
    def particle_num(self) -> int:
        return int(self.i32("particle_num")[0])

    @property
    def fluid_particle_num(self) -> int:
        return int(self.i32("fluid_particle_num")[0])

    @property
    def num_samples(self) -> int:
        return int(self.i32("num_samples")[0])

    @property
    def kden_sum(self) -> float:
        p = lib().or_ptr(self.sim, b"kden_sum")
        return float(ctypes.cast(p, ctypes.POINTER(ctypes.c_double))[0])

    def __del__(self):
        if getattr(self, "sim", None) is not None and _LIB is not None:
            _LIB.or_free(self.sim)
            self.sim = None

    # -- particle insertion (particle_system.py:334-347, 641-696) ----------
    def compute_cube_particle_num(self, start, end):
        num_dim = [np.arange(start[i], end[i], self.particle_diameter) for i in range(3)]
        return reduce(lambda a, b: a * b, [len(n) for n in num_dim])

    def add_particles(self, object_id, n, pos, vel, density, pressure, material, is_dynamic, color):
        arrs = [np.ascontiguousarray(pos, np.float32), np.ascontiguousarray(vel, np.float32),
                np.ascontiguousarray(density, np.float32), np.ascontiguousarray(pressure, np.float32),
                np.ascontiguousarray(material, np.int32), np.ascontiguousarray(is_dynamic, np.int32),
                np.ascontiguousarray(color, np.int32)]
        return lib().or_add_particles(self.sim, int(object_id), int(n), *[_ptr(a) for a in arrs])

    def add_cube(self, object_id, lower_corner, cube_size, material, is_dynamic, color=(0, 0, 0),
                 density=None, pressure=None, velocity=None):
        num_dim = [np.arange(lower_corner[i], lower_corner[i] + cube_size[i], self.particle_diameter)
                   for i in range(3)]
        n = reduce(lambda a, b: a * b, [len(d) for d in num_dim])
        pos = np.array(np.meshgrid(*num_dim, sparse=False, indexing="ij"), dtype=np.float32)
        pos = pos.reshape(3, -1).transpose()
        vel = (np.zeros_like(pos) if velocity is None
               else np.array([velocity for _ in range(n)], dtype=np.float32).reshape(n, 3))
        dens = np.full(n, density if density is not None else 1000.0, np.float32)
        pres = np.full(n, pressure if pressure is not None else 0.0, np.float32)
        col = np.stack([np.full(n, c, np.int32) for c in color], axis=1)
        self.add_particles(object_id, n, pos, vel, dens, pres, np.full(n, material, np.int32),
                           np.full(n, is_dynamic, np.int32), col)

    # -- solver (sph_base.py:80-85; karman_vortex.py:216-242) -------------
    def initialize_particle_system(self):
        """particle_system.py:555-560; an out-of-grid cell key raises."""
        L, s = lib(), self.sim
        L.or_update_activity(s)
        L.or_particle_partition(s)
        L.or_update_grid_id(s)
        if self.i32("err_key_range")[0]:
            raise RuntimeError("particle outside the cell grid (cell key out of range)")
        L.or_prefix_sum(s)
        L.or_counting_sort(s)

    def initialize(self):
        L = lib()
        self.initialize_particle_system()
        L.or_compute_static_boundary_volume(self.sim)
        L.or_compute_moving_boundary_volume(self.sim)

    def emit_particle(self):
        for e in self.fluid_emitters:
            obj_id = e["objectId"]
            center = np.array(e["squareCenter"])
            size = np.array(e["squareSize"])
            init_v = np.array(e["velocity"])
            lower = center - 0.5 * size
            cube = size + self.particle_radius * np.abs(init_v) / np.linalg.norm(init_v)
            n_layer = self.compute_cube_particle_num(lower, lower + cube)
            self.emitter_counts[obj_id] += n_layer
            # dump_num_particles_each_emitters_ti2np / np2ti around the
            # emission (karman_vortex.py:217-220): the field gains the layer
            self.emitter_field()[obj_id] += n_layer
            self.add_cube(obj_id, lower, cube, material=1, is_dynamic=1, color=e["color"],
                          density=e["density"], velocity=init_v)

    def step(self, u1=None, u2=None, u_cdf=None):
        """One KarmanVortexSolver.step; uniforms are host-fed (see module doc)."""
        L = lib()
        if self.cnt % self.emit_interval == 0:
            self.emit_particle()
        self.cnt += 1
        self.initialize_particle_system()
        L.or_compute_moving_boundary_volume(self.sim)
        self.substep(u1, u2, u_cdf)
        # karman_vortex.py:226-230 (3-D scenes: enforce_boundary_3D on fluid)
        if self.cfg.get("bounceBackBoundary"):
            L.or_enforce_boundary_3d(self.sim)

    def substep(self, u1=None, u2=None, u_cdf=None):
        L, s = lib(), self.sim
        L.or_compute_densities(s)
        L.or_compute_DFSPH_factor(s)
        self.divergence_solve()
        L.or_compute_non_pressure_forces(s)
        self.dimcv(u1, u2, u_cdf)
        L.or_predict_velocity(s)
        self.pressure_solve()
        L.or_copy_x_temp(s)
        L.or_compute_vorticity_vis(s)
        L.or_advect(s)

    def divergence_solve(self):
        L, s = lib(), self.sim
        L.or_compute_density_change(s)
        L.or_multiply_time_step(s, 0, np.float32(self.inv_dt))
        it = 0
        while it < 1 or it < 100:
            L.or_divergence_solver_iteration_kernel(s)
            L.or_compute_density_change(s)
            err = L.or_compute_density_error(s, np.float32(0.0)) / self.fluid_particle_num
            eta = self.inv_dt * self.max_error_V * 0.01 * self.density_0
            it += 1
            if err <= eta:
                break
        self.last_kv = it
        L.or_multiply_time_step(s, 0, self.dt)

    def pressure_solve(self):
        L, s = lib(), self.sim
        L.or_compute_density_adv(s)
        L.or_multiply_time_step(s, 0, np.float32(self.inv_dt2))
        it = 0
        while it < 1 or it < 100:
            L.or_pressure_solve_iteration_kernel(s)
            L.or_compute_density_adv(s)
            err = L.or_compute_density_error(s, np.float32(self.density_0)) / self.fluid_particle_num
            eta = self.max_error * 0.01 * self.density_0
            it += 1
            if err <= eta:
                break
        self.last_kp = it

    def dimcv(self, u1=None, u2=None, u_cdf=None, v_vort_out=None):
        L, s = lib(), self.sim
        L.or_compute_v_star(s)
        L.or_compute_vorticity(s)
        L.or_compute_grad_v(s)
        L.or_compute_d_vorticity(s)
        L.or_compute_all_d_vort_smoothed(s)
        if self.mc_samples:
            L.or_compute_kinematic_vorticity_num_sample(s, None, None, 0)
            L.or_compute_kernel_density(s)
            self.cdf_sample(u_cdf, int(self.mc_samples))
        else:
            a1 = np.ascontiguousarray(u1, np.float32)
            a2 = np.ascontiguousarray(u2, np.float32)
            assert a1.size >= self.particle_num and a2.size >= self.particle_num
            L.or_compute_kinematic_vorticity_num_sample(s, _ptr(a1), _ptr(a2), 1)
            L.or_record_vort_part_sample_idx(s)
            L.or_compute_kernel_density(s)
        L.or_vort2vel(s, None if v_vort_out is None else _ptr(v_vort_out))
        L.or_update_particle_life_time(s)

    def cdf_sample(self, u, m: int):
        """Build-defined inverse-CDF sampler (refsim.cpp or_cdf_sample)."""
        u = np.ascontiguousarray(u, np.float64)
        assert u.size >= m
        if lib().or_cdf_sample(self.sim, _ptr(u), int(m)) != 0:
            raise ValueError(f"{m} samples exceed particle_max_num {self.particle_max_num}")

    def neighbors(self, p_i: int) -> np.ndarray:
        """Slots for_all_neighbors visits for p_i, in visiting order."""
        buf = np.zeros(1024, np.int32)
        k = lib().or_neighbors(self.sim, int(p_i), _ptr(buf), buf.size)
        assert k <= buf.size
        return buf[:k].astype(np.int64)

    def vort2vel_targets(self, slots):
        """(v_vort, far part, accepted count) of vort2vel at `slots`, v unchanged."""
        slots = np.ascontiguousarray(slots, np.int64)
        n = slots.size
        vv = np.zeros((n, 3), np.float32)
        far = np.zeros((n, 3), np.float32)
        added = np.zeros(n, np.int32)
        lib().or_vort2vel_targets(self.sim, _ptr(slots), n, _ptr(vv), _ptr(far), _ptr(added))
        return vv, far, added

    def direct_bs_f64(self) -> np.ndarray:
        out = np.zeros((self.particle_max_num, 3), np.float64)
        lib().or_direct_bs_f64(self.sim, _ptr(out))
        return out


def export_ply(ref: RefSim, cnt: int, ply_path) -> None:
    """karman_vortex.py:122-214 restated on the oracle's fields: the ASCII PLY
    of the active particles below particle_num, each row formatted with the
    reference's own f-strings."""
    N = ref.particle_num
    if N == 0:
        return
    x = ref.field("x_temp")[:N]
    vorticity = ref.field("vorticity")[:N]
    color = ref.field("color")[:N]
    obj_id = ref.field("object_id")[:N]
    is_active = ref.field("is_active")[:N].astype(bool)
    x, vorticity, color, obj_id = x[is_active], vorticity[is_active], color[is_active], obj_id[is_active]
    if x.shape[0] == 0:
        return
    num_vertices = x.shape[0]
    particle_type = np.full(num_vertices, 4, dtype=np.uint8)
    sc = ref.scene

    def mask_from_ids(ids_set):
        ids = list(ids_set)
        if len(ids) == 0:
            return np.zeros_like(obj_id, dtype=bool)
        return np.isin(obj_id, np.array(ids, dtype=obj_id.dtype))

    particle_type[mask_from_ids({b["objectId"] for b in sc.get("FluidBlocks", [])})] = 0
    particle_type[mask_from_ids({e["objectId"] for e in sc.get("FluidEmitters", [])})] = 1
    particle_type[mask_from_ids({b["objectId"] for b in sc.get("RigidBlocks", [])})] = 2
    particle_type[mask_from_ids({b["objectId"] for b in sc.get("RigidBodies", [])})] = 3
    ply_path.mkdir(parents=True, exist_ok=True)
    with open(ply_path / f"frame_{cnt:04}.ply", "w") as f:
        f.write("ply\n")
        f.write("format ascii 1.0\n")
        f.write(f"element vertex {num_vertices}\n")
        for nm in ("x", "y", "z", "vort_x", "vort_y", "vort_z"):
            f.write(f"property float {nm}\n")
        for nm in ("red", "green", "blue"):
            f.write(f"property uchar {nm}\n")
        f.write("property int object_id\n")
        f.write("property uchar particle_type\n")
        f.write("end_header\n")
        for i in range(num_vertices):
            px, py, pz = x[i]
            vx, vy, vz = vorticity[i]
            r, g, b = color[i]
            f.write(f"{px:.7f} {py:.7f} {pz:.7f} "
                    f"{vx:.7f} {vy:.7f} {vz:.7f} "
                    f"{int(r)} {int(g)} {int(b)} {int(obj_id[i])} {int(particle_type[i])}\n")


def export_png(ref: RefSim, cnt: int, image_path) -> None:
    """karman_vortex.py:38-120 restated on the oracle's fields."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    from matplotlib.colors import Normalize
    N = ref.particle_num
    material = ref.field("material")[:N]
    obj_id = ref.field("object_id")[:N]
    fluid_mask = material == 1
    solid_mask = obj_id == 2
    x = ref.field("x_temp")[:N]
    vort = ref.field("vorticity_vis")[:N][:, 2]
    fluid_x, fluid_vort, solid_x = x[fluid_mask], vort[fluid_mask], x[solid_mask]
    fig = plt.figure(figsize=(10, 4), dpi=200)
    ax = fig.add_subplot(111, projection="3d")
    ax.view_init(elev=30, azim=-60)
    ax.scatter(fluid_x[:, 0], fluid_x[:, 1], fluid_x[:, 2], c=fluid_vort, cmap="coolwarm", s=0.2,
               norm=Normalize(vmin=-40, vmax=40), edgecolors="none", alpha=0.6)
    ax.scatter(solid_x[:, 0], solid_x[:, 1], solid_x[:, 2], color="#00C853", s=0.5, edgecolors="none")
    ax.set_xlim(0, 4)
    ax.set_ylim(0, 1)
    ax.set_zlim(0, 1)
    ax.set_box_aspect((4, 1, 1))
    ax.set_axis_off()
    plt.subplots_adjust(left=0, right=1, top=1, bottom=0)
    plt.savefig(image_path / f"vorticity_{cnt:04}.png", bbox_inches="tight", pad_inches=0,
                transparent=True, dpi=400)
    plt.cla()
    plt.close("all")


class OrSegConst(ctypes.Structure):
    _fields_ = [("reg_radius", ctypes.c_float), ("dt", ctypes.c_float), ("dt6", ctypes.c_float),
                ("inv4pi", ctypes.c_float), ("gamma_decay", ctypes.c_float), ("u_inf", ctypes.c_float * 3)]


def seg_const(reg_radius=0.01, dt=0.002, gamma_decay=1.0, u_inf=(0.0, 0.0, 0.0)) -> OrSegConst:
    """SegmentSolver constants as the kernels see them (segment_solver.py:9-20):
    f32 of the Python values, dt/6 and 1/(4 pi) formed in f64 first."""
    c = OrSegConst()
    c.reg_radius = np.float32(reg_radius)
    c.dt = np.float32(dt)
    c.dt6 = np.float32(dt / 6.0)
    c.inv4pi = np.float32(1.0 / (4.0 * np.pi))
    c.gamma_decay = np.float32(gamma_decay)
    for d in range(3):
        c.u_inf[d] = np.float32(u_inf[d])
    return c


def _seg_lib():
    L = lib()
    if not getattr(L, "_seg_sigs", False):
        P, I32, F32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_float
        L.or_segment_velocity.restype = None
        L.or_segment_velocity.argtypes = [I32, P, P, P, P, P, I32, F32, P, P, P, P]
        L.or_segment_rk4.restype = None
        L.or_segment_rk4.argtypes = [I32, P, P, P, P, P, P, P, P, P]
        L._seg_sigs = True
    return L


def segment_velocity(xm, xp, gamma, active, const, form, factor=0.0, km=None, kp=None, out_m=None, out_p=None):
    """Endpoint velocities (form 0, segment_solver.py:378-418) or velocities
    at x + factor dt k (form 1, :468-521); out arrays updated in place for
    active segments (copies of zeros when not given)."""
    n = len(gamma)
    a = [np.ascontiguousarray(v, np.float32) for v in (xm, xp, gamma)]
    act = np.ascontiguousarray(active, np.int32)
    km = np.zeros((n, 3), np.float32) if km is None else np.ascontiguousarray(km, np.float32)
    kp = np.zeros((n, 3), np.float32) if kp is None else np.ascontiguousarray(kp, np.float32)
    out_m = np.zeros((n, 3), np.float32) if out_m is None else out_m
    out_p = np.zeros((n, 3), np.float32) if out_p is None else out_p
    _seg_lib().or_segment_velocity(n, _ptr(a[0]), _ptr(a[1]), _ptr(a[2]), _ptr(act), ctypes.byref(const), int(form),
                                   float(factor), _ptr(km), _ptr(kp), _ptr(out_m), _ptr(out_p))
    return out_m, out_p


def segment_rk4(xm, xp, gamma, active, age, const, vm, vp):
    """advect_segments_rk4 (segment_solver.py:420-447, 523-542) in place on
    xm, xp, gamma, age (f32 numpy arrays)."""
    n = len(gamma)
    k = np.zeros(24 * n, np.float32)
    vm = np.ascontiguousarray(vm, np.float32)
    vp = np.ascontiguousarray(vp, np.float32)
    act = np.ascontiguousarray(active, np.int32)
    _seg_lib().or_segment_rk4(n, _ptr(xm), _ptr(xp), _ptr(gamma), _ptr(act), _ptr(age), ctypes.byref(const),
                              _ptr(vm), _ptr(vp), _ptr(k))
