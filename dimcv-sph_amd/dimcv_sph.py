"""DIMCV-SPH solver (drop-in for dimcv_sph.py of the reference).

``DIMCVSPHSolver`` keeps the reference's constructor, fields and method names
(dimcv_sph.py:9-819).  Each former ``@ti.kernel`` method launches the matching
libdimcv entry point; the host-side logic (Jacobi loops with their
convergence tests, emission, the ``dimcv()`` pass order) is the reference's.

Build-defined behaviour, beyond the reference (each optional, see
config_builder.py for the keys):
  * ``ti.random`` is replaced by host-fed uniforms (``u1``/``u2`` per slot)
    drawn from a seeded numpy generator, or passed explicitly to ``step``;
  * with ``mcSamples`` = M the Bernoulli birth/death sample set is replaced by
    M inverse-CDF draws from the normalised AKDE density (FP64 CDF);
  * residual sums and ``kden_sum`` are FP64 (fixed reduction order);
  * ``farField`` selects the far-field evaluation of vort2vel: "exact"
    (reference expression, bit-exact), "fast" or "kahan".
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

import _native as N
from _field import CounterField, DeviceScalarField, Field
from sph_base import SPHBase


class DIMCVSPHSolver(SPHBase):

    def __init__(self, particle_system):
        super().__init__(particle_system)
        ps = self.ps
        self.surface_tension = 0
        self.dt[None] = ps.cfg.get_cfg("timeStepSize")

        self.eps = 1e-6
        self.inv_eps = 1.0 / self.eps
        self.volume_scale = ps.cfg.get_cfg("volumeScale")
        if self.volume_scale is not None:
            self.vp_volume = self.volume_scale * ps.m_V0

        self.m_max_iterations_v = 100
        self.m_max_iterations = 100
        self.m_eps = 1e-5
        self.max_error_V = 0.1
        self.max_error = 0.05
        self.inv_dt = 1.0 / self.dt[None]
        self.inv_dt2 = 1.0 / (self.dt[None] * self.dt[None])

        n = int(ps.particle_max_num)
        dev = ps.device
        self.d_vort = Field(n, 3, torch.float32, dev)
        self.d_vort_smoothed = Field(n, 3, torch.float32, dev)
        self.v_star = Field(n, 3, torch.float32, dev)
        self.grad_v = Field(n, 9, torch.float32, dev, view_shape=(3, 3))
        self.x_temp = Field(n, 3, torch.float32, dev)
        self.num_samples = CounterField(ps._counters, N.C_NUM_SAMPLES)
        self.sample_idx = Field(n, 0, torch.int32, dev)
        self.kinematic_vort_num = Field(n, 0, torch.float32, dev)
        self.kernel_density = Field(n, 0, torch.float32, dev)
        self.kden_sum = DeviceScalarField(ps._dscal, N.D_KDEN_SUM)
        self.v_vort = Field(n, 3, torch.float32, dev)   # recovered velocity (analysis)
        self.vortex_enforcing_domain_start = ps.cfg.get_cfg("vortexEnforcingDomainStart")
        self.vortex_enforcing_domain_end = ps.cfg.get_cfg("vortexEnforcingDomainEnd")
        self.a = ps.cfg.get_cfg("GenProbPara")
        self.b = ps.cfg.get_cfg("DelProbPara")
        num_samples = ps.cfg.get_cfg("numSamples")
        if num_samples:
            self.num_samples[None] = num_samples
        self.divergence_transit_area = ps.cfg.get_cfg("divergenceTransitArea")

        self.cnt = 0
        self.emit_interval = 1
        for emitter in ps.fluid_emitters:
            init_v = np.array(emitter["velocity"])
            self.emit_interval = np.ceil(
                ps.particle_diameter / (np.linalg.norm(init_v) * self.dt[None])).astype(int)
            break
        # emission on the device (dimcv_emit): each emitter's layer, built
        # once by the reference's add_cube lattice, stays on the device; an
        # emission is one kernel per emitter with no host copy or sync, and
        # the emitter counts live on the device.  emit_on_device = False
        # restores the host path (add_cube per emission, counts mirrored).
        self.emit_on_device = True
        self._emitters = None
        self._emitter_layers = []

        # build extensions
        self.mc_samples = ps.cfg.get_cfg("mcSamples")
        far = ps.cfg.get_cfg("farField") or ("fast" if self.mc_samples else "exact")
        self.far_mode = {"exact": N.FAR_EXACT, "fast": N.FAR_FAST, "kahan": N.FAR_KAHAN}[far]
        seed = ps.cfg.get_cfg("rngSeed")
        self.rng = np.random.default_rng(0 if seed is None else seed)
        self._staged = {}             # pinned host + device buffers of the per-step uniforms
        self._upload_stream = None    # side stream (default priority) and event of those uploads
        self._upload_event = None
        self._uniforms = (None, None, None)
        self.last_iterations_v = 0
        self.last_iterations_p = 0
        self.time_far_field = False   # record HIP events around the far-field kernel
        self.last_far_ms = None
        # pipelined fixed-M steps (no counterpart in the reference): step()
        # returns once the step is enqueued; its counts and iteration numbers
        # reach the host, and a failed step raises, at the next step() or at
        # flush().  Host-side counts (ps._pn, last_iterations_*) lag by one
        # step until then; device-side reads (fields, counters) are current.
        self.async_steps = False
        self._pending = None
        self._step_parity = 0

        p = ps.params
        f32 = np.float32
        p.dt = f32(self.dt[None])
        p.inv_dt = f32(self.inv_dt)
        if self.vortex_enforcing_domain_start is not None:
            p.has_vortex_domain = 1
            for d in range(3):
                p.vortex_start[d] = f32(self.vortex_enforcing_domain_start[d])
                p.vortex_end[d] = f32(self.vortex_enforcing_domain_end[d])
        if self.divergence_transit_area is not None:
            p.has_transit = 1
            p.transit_a = f32(self.divergence_transit_area[0])
            p.transit_ba = f32(self.divergence_transit_area[1] - self.divergence_transit_area[0])
        if self.volume_scale is not None:
            p.vp_factor = f32(0.1592 * self.vp_volume)
            p.vp_factor_f64 = 0.1592 * self.vp_volume
        p.neg_gen_a = f32(-(self.a if self.a is not None else 0))
        p.neg_del_b = f32(-(self.b if self.b is not None else 0))
        p.far_mode = self.far_mode
        ps.solver = self
        ps.invalidate_state()

    # ------------------------------------------------------------ helpers --
    def _fluid(self, name, *args):
        self.ps.ensure_neighbor_lists()
        self.ps.call(name, *args, self.ps._nfl, self.ps.stream())

    def _slots(self, name, *args):
        self.ps.call(name, *args, self.ps._pn, self.ps.stream())

    def _fluid_count(self) -> int:
        return int(self.ps._host_counters[N.C_FLUID_NUM])

    def is_in_vortex_domain(self, pos) -> bool:
        if self.vortex_enforcing_domain_start is None:
            return True
        for i in range(self.ps.dim):
            if pos[i] < np.float32(self.vortex_enforcing_domain_start[i]) or \
                    pos[i] > np.float32(self.vortex_enforcing_domain_end[i]):
                return False
        return True

    # --------------------------------------------------- DFSPH (dimcv_sph.py) --
    def compute_DFSPH_factor(self):
        self._fluid("dimcv_compute_DFSPH_factor")

    def compute_densities(self):
        self._fluid("dimcv_compute_densities")

    def compute_non_pressure_forces(self):
        self.ps.ensure_neighbor_lists()
        self.ps.call("dimcv_compute_non_pressure_forces", self.ps._pn, self.ps._nfl, self.ps.stream())

    def advect(self):
        self._slots("dimcv_advect")
        self.ps._nbr_valid = False

    def predict_velocity(self):
        self._slots("dimcv_predict_velocity")

    def compute_density_change(self):
        self._fluid("dimcv_compute_density_change")

    def multiply_time_step(self, field, time_step: float):
        self.ps.call("dimcv_multiply_time_step", field.ptr(), float(time_step), self.ps._nfl,
                     self.ps.stream())

    def compute_density_error(self, offset: float) -> float:
        out = np.zeros(1, np.float64)
        self.ps.call("dimcv_compute_density_error", float(offset), self.ps._nfl, out.ctypes.data,
                     self.ps.stream())
        return float(out[0])

    def divergence_solver_iteration_kernel(self):
        self._fluid("dimcv_divergence_solver_iteration")

    def divergence_solver_iteration(self):
        self.divergence_solver_iteration_kernel()
        self.compute_density_change()
        density_err = self.compute_density_error(0.0)
        return density_err / self._fluid_count()

    def divergence_solve(self):
        self.compute_density_change()
        self.multiply_time_step(self.ps.dfsph_factor, self.inv_dt)
        m_iterations_v = 0
        avg_density_err = 0.0
        while m_iterations_v < 1 or m_iterations_v < self.m_max_iterations_v:
            avg_density_err = self.divergence_solver_iteration()
            eta = self.inv_dt * self.max_error_V * 0.01 * self.density_0
            m_iterations_v += 1
            if avg_density_err <= eta:
                break
        self.last_iterations_v = m_iterations_v
        self.multiply_time_step(self.ps.dfsph_factor, self.dt[None])

    def compute_density_adv(self):
        self._fluid("dimcv_compute_density_adv")

    def pressure_solve_iteration_kernel(self):
        self._fluid("dimcv_pressure_solve_iteration")

    def pressure_solve_iteration(self):
        self.pressure_solve_iteration_kernel()
        self.compute_density_adv()
        density_err = self.compute_density_error(self.density_0)
        return density_err / self._fluid_count()

    def pressure_solve(self):
        inv_dt2 = self.inv_dt2
        self.compute_density_adv()
        self.multiply_time_step(self.ps.dfsph_factor, inv_dt2)
        m_iterations = 0
        avg_density_err = 0.0
        while m_iterations < 1 or m_iterations < self.m_max_iterations:
            avg_density_err = self.pressure_solve_iteration()
            eta = self.max_error * 0.01 * self.density_0
            m_iterations += 1
            if avg_density_err <= eta:
                break
        self.last_iterations_p = m_iterations

    # ------------------------------------------------------- emission --
    def _emitter_table(self):
        """dimcv_emitter entries (device lattices made on first use)."""
        if self._emitters is None:
            ps = self.ps
            table = (N.Emitter * max(len(ps.fluid_emitters), 1))()
            for k, emitter in enumerate(ps.fluid_emitters):
                square_center = np.array(emitter["squareCenter"])
                square_size = np.array(emitter["squareSize"])
                init_v = np.array(emitter["velocity"])
                lower = square_center - 0.5 * square_size
                cube = square_size + ps.particle_radius * np.abs(init_v) / np.linalg.norm(init_v)
                pos = ps.cube_lattice(lower, cube)
                dev = torch.as_tensor(pos, device=ps.device).contiguous()
                self._emitter_layers.append(dev)
                e = table[k]
                e.pos, e.n, e.object_id = dev.data_ptr(), pos.shape[0], int(emitter["objectId"])
                for d in range(3):
                    e.velocity[d] = np.float32(init_v[d])
                    e.color[d] = int(emitter["color"][d])
                e.density = np.float32(emitter["density"])
            self._emitters = table
        return self._emitters

    def emit_particle(self):
        ps = self.ps
        if self.emit_on_device:
            table = self._emitter_table()
            ne = len(ps.fluid_emitters)
            if ne == 0:
                return
            ps.call("dimcv_emit", table, ne, ps.stream())
            for k in range(ne):
                ps.appended(int(table[k].n))
            return
        for emitter in ps.fluid_emitters:
            obj_id = emitter["objectId"]
            square_center = np.array(emitter["squareCenter"])
            square_size = np.array(emitter["squareSize"])
            init_v = np.array(emitter["velocity"])
            lower_corner = square_center - 0.5 * square_size
            cube_size = square_size + ps.particle_radius * np.abs(init_v) / np.linalg.norm(init_v)
            ps.object_collection[obj_id]["particleNum"] += ps.compute_cube_particle_num(
                lower_corner, lower_corner + cube_size)
            ps.add_cube(object_id=obj_id, lower_corner=lower_corner, cube_size=cube_size,
                        velocity=init_v, density=emitter["density"], is_dynamic=1,
                        color=emitter["color"], material=1)

    def dump_num_particles_each_emitters_ti2np(self):
        """Host mirror of the emitter counts (object_collection particleNum)
        from the device (synchronises)."""
        if not self.ps.obj_id_emitters:
            return
        counts = self.ps.num_particles_each_emitter.to_numpy()
        for obj_id in self.ps.obj_id_emitters:
            if 0 <= obj_id < counts.shape[0]:
                self.ps.object_collection[obj_id]["particleNum"] = int(counts[obj_id])

    def dump_num_particles_each_emitters_np2ti(self):
        if not self.ps.obj_id_emitters:
            return
        counts = self.ps.num_particles_each_emitter.to_numpy()
        for obj_id in self.ps.obj_id_emitters:
            if 0 <= obj_id < counts.shape[0]:
                counts[obj_id] = self.ps.object_collection[obj_id]["particleNum"]
        self.ps.num_particles_each_emitter.from_numpy(counts)

    # ------------------------------------------------------ DIMCV passes --
    def compute_v_star(self):
        self._fluid("dimcv_compute_v_star")

    def compute_grad_v(self):
        self._fluid("dimcv_compute_grad_v")

    def compute_vorticity(self):
        self._fluid("dimcv_compute_vorticity")

    def compute_d_vorticity(self):
        self._fluid("dimcv_compute_d_vorticity")

    def compute_all_d_vort_smoothed(self):
        self._fluid("dimcv_compute_all_d_vort_smoothed")

    def _upload(self, a, dtype):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(self.ps.device)

    def _staged_buffers(self, key, n, dtype):
        """Pinned host and device buffers kept for `key`, with room for n
        values: (host tensor, device tensor, numpy view of the host one)."""
        bufs = self._staged.get(key)
        if bufs is None or bufs[0].numel() < n or bufs[0].dtype != dtype:
            host = torch.empty(max(n, 1), dtype=dtype, pin_memory=True)
            dev = torch.empty(max(n, 1), dtype=dtype, device=self.ps.device)
            bufs = self._staged[key] = (host, dev, host.numpy())
        return bufs

    def _cdf_key(self):
        """Staging key of the step's inverse-CDF uniforms: with async_steps two
        buffer pairs by step parity, so that a step's upload never overwrites
        the uniforms of the step still running (whose buffers were last used
        by the step before it, already waited for)."""
        return f"u_cdf{self._step_parity}" if self.async_steps else "u_cdf"

    def _upload_staged(self, key, a, dtype):
        """Copy `a` into a device buffer kept for `key`, through a pinned host
        buffer (an asynchronous copy on the current stream).  The host buffer
        is reused by the next call on that key, which comes after the step's
        own synchronisations have retired this copy."""
        a = np.ascontiguousarray(a)
        n = a.size
        host, dev, host_np = self._staged_buffers(key, n, dtype)
        host_np[:n] = a
        if n == dev.numel():
            dev.copy_(host, non_blocking=True)
        else:
            dev[:n].copy_(host[:n], non_blocking=True)
        return dev

    def compute_kinematic_vorticity_num_sample(self, u1=None, u2=None):
        """KVN + Bernoulli sample update; u1/u2 replace the two ti.random() draws."""
        if self.mc_samples:
            self._fluid("dimcv_compute_kvn_sample", None, None, 0)
            return
        pn = self.ps._pn
        if u1 is None:
            u1 = self.rng.random(pn, dtype=np.float32)
        if u2 is None:
            u2 = self.rng.random(pn, dtype=np.float32)
        assert len(u1) >= pn and len(u2) >= pn
        t1 = self._upload(u1, torch.float32)
        t2 = self._upload(u2, torch.float32)
        self._fluid("dimcv_compute_kvn_sample", t1.data_ptr(), t2.data_ptr(), 1)

    def record_vort_part_sample_idx(self):
        self._slots("dimcv_record_sample_idx")

    def compute_kernel_density(self):
        self._fluid("dimcv_compute_kernel_density")

    def importance_sample(self, u=None, m=None):
        """Build-defined: M inverse-CDF draws from kden / kden_sum (FP64 CDF)."""
        m = int(self.mc_samples if m is None else m)
        if u is None:
            u = self.rng.random(m)
        assert len(u) >= m
        tu = self._upload(u, torch.float64)
        self.ps.call("dimcv_cdf_sample", tu.data_ptr(), m, self.ps._nfl, self.ps.stream())

    def vort2vel(self):
        self.ps.ensure_neighbor_lists()
        self.ps.call("dimcv_vort2vel_near", self.ps._nfl, self.ps.stream())
        self.vort2vel_far()

    def vort2vel_far(self):
        """Far-field half of vort2vel (dimcv_sph.py:693-719): the O(N M) sum."""
        self.ps.call("dimcv_vort2vel_far", self.ps._nfl, int(self.ps.particle_max_num), self.ps.stream())

    def update_particle_life_time(self):
        self._fluid("dimcv_update_particle_life_time")

    def compute_kernel_sum(self):
        raise NotImplementedError("compute_kernel_sum is unused by the reference step")

    def compute_vorticity_vis(self):
        self._fluid("dimcv_compute_vorticity_vis")

    def copy_x_temp(self):
        self._slots("dimcv_copy_x_temp")

    def init(self):
        # compute_v_star + compute_vorticity + compute_grad_v as one neighbour pass
        self._fluid("dimcv_dimcv_init")

    def update_sample(self, u1=None, u2=None):
        self.compute_kinematic_vorticity_num_sample(u1, u2)
        if not self.mc_samples:
            self.record_vort_part_sample_idx()

    def dimcv(self):
        u1, u2, u_cdf = self._uniforms
        self.init()
        self.compute_d_vorticity()
        self.compute_all_d_vort_smoothed()
        self.update_sample(u1, u2)
        self.compute_kernel_density()
        if self.mc_samples:
            self.importance_sample(u_cdf)
        self.vort2vel()
        self.update_particle_life_time()

    # ------------------------------------------------- native substep --
    def _step_opts(self, u1, u2, u_cdf, defer_cdf=False):
        """dimcv_step_opts of this step (uniforms uploaded; fluid_count unset).
        With defer_cdf the inverse-CDF uniforms are neither drawn nor
        uploaded: u_cdf points at the device buffer that _upload_staged
        ("u_cdf") fills later."""
        ps = self.ps
        o = N.StepOpts()
        if self.mc_samples:
            m = int(self.mc_samples)
            if defer_cdf:
                t = self._staged_buffers(self._cdf_key(), m, torch.float64)[1]
            else:
                if u_cdf is None:
                    u_cdf = self.rng.random(m)
                t = self._upload_staged("u_cdf", np.asarray(u_cdf)[:m], torch.float64)
            o.u_cdf, o.m = t.data_ptr(), m
        else:
            pn = ps._pn
            if u1 is None:
                u1 = self.rng.random(pn, dtype=np.float32)
            if u2 is None:
                u2 = self.rng.random(pn, dtype=np.float32)
            assert len(u1) >= pn and len(u2) >= pn
            t1 = self._upload_staged("u1", np.asarray(u1)[:pn], torch.float32)
            t2 = self._upload_staged("u2", np.asarray(u2)[:pn], torch.float32)
            o.u1, o.u2, o.m = t1.data_ptr(), t2.data_ptr(), 0
        o.max_iter_v = self.m_max_iterations_v
        o.max_iter_p = self.m_max_iterations
        # jacobiChunk: iterations per host convergence check; 0 checks first
        # after the previous step's iteration count; -1 (the default) lets the
        # one-call fixed-M step (dimcv_step) converge on the device with no
        # host check inside the step (dimcv_substep treats it as 0)
        chunk = ps.cfg.get_cfg("jacobiChunk")
        o.chunk = int(chunk) if chunk is not None else -1
        o.out_kv = max(1, int(self.last_iterations_v or 1))
        o.out_kp = max(1, int(self.last_iterations_p or 1))
        o.eta_v = self.inv_dt * self.max_error_V * 0.01 * self.density_0
        o.eta_p = self.max_error * 0.01 * self.density_0
        o.inv_dt = np.float32(self.inv_dt)
        o.dt = np.float32(self.dt[None])
        o.inv_dt2 = np.float32(self.inv_dt2)
        o.density0 = np.float32(self.density_0)
        o.bounce_back = 0
        o.time_far = 1 if self.time_far_field else 0
        return o

    def _step_outputs(self, o):
        self.last_iterations_v = int(o.out_kv)
        self.last_iterations_p = int(o.out_kp)
        self.last_far_ms = float(o.out_far_ms) if self.time_far_field else None
        self.ps._nbr_valid = False

    def native_substep(self, u1=None, u2=None, u_cdf=None):
        """substep() issued by libdimcv in one call (dimcv_substep): same kernels
        in the same order; the Jacobi loops converge on the device and the host
        checks every ``jacobiChunk`` iterations."""
        ps = self.ps
        self.flush()
        o = self._step_opts(u1, u2, u_cdf)
        o.fluid_count = self._fluid_count()
        if o.fluid_count <= 0:
            raise ZeroDivisionError("no fluid particles (fluid_particle_num is 0)")
        ps.ensure_neighbor_lists()
        ps.call("dimcv_substep", ctypes.byref(o), ps._pn, ps._nfl, ps.stream())
        self._step_outputs(o)

    def native_step(self, u_cdf=None):
        """initialize_particle_system() + substep() in two library calls
        (dimcv_step_begin / dimcv_step_finish): the grid build, the neighbour
        lists and the substep run back to back on the device, bounded by the
        particle count before the build; the counts come back with the
        pressure solve's last check.  The step's inverse-CDF uniforms are
        drawn and uploaded between the two calls, while the device runs the
        build and the divergence solve.  Fixed-M sampling only (the Bernoulli
        uniforms are drawn per particle after the build)."""
        ps = self.ps
        m = int(self.mc_samples)
        if not self.async_steps:
            self.flush()
        self._step_parity ^= 1
        o = self._step_opts(None, None, None, defer_cdf=True)
        pre = ps.state()
        ps.swap_sorted_fields()
        post = ps.state()
        hc = ps._host_counters
        lib = N.lib()
        stream = ps.stream()
        args = (ctypes.byref(pre), ctypes.byref(post), ctypes.byref(ps.params), ctypes.byref(o), ps._pn)
        N.check(lib.dimcv_step_begin(*args, stream), "dimcv_step_begin")
        u = self.rng.random(m) if u_cdf is None else np.asarray(u_cdf)[:m]
        # the copy runs on a side stream, overlapping the device's build and
        # divergence solve; the step's stream waits for it before the sampling
        if self._upload_stream is None:
            self._upload_stream = torch.cuda.Stream(device=ps.device)
            self._upload_event = torch.cuda.Event()
        main = torch.cuda.current_stream(ps.device)
        with torch.cuda.stream(self._upload_stream):
            t = self._upload_staged(self._cdf_key(), u, torch.float64)
            self._upload_event.record(self._upload_stream)
        main.wait_event(self._upload_event)
        assert t.data_ptr() == o.u_cdf
        if self.async_steps:
            # the previous step's counts (its errors are raised here), then
            # this step's rest without a wait: the device already runs this
            # step's build while the host reads them
            self.flush()
            N.check(lib.dimcv_step_finish_async(*args, stream), "dimcv_step_finish_async")
            self._pending = o
            return
        rc = lib.dimcv_step_finish(*args, ps._host_counters_ptr, stream)
        ps._pn = int(hc[N.C_PARTICLE_NUM])
        ps._nfl = int(hc[N.C_FLUID_LIST])
        N.check(rc, "dimcv_step_finish")
        self._step_outputs(o)

    def flush(self):
        """Complete the outstanding asynchronous step (async_steps): its counts,
        iteration counts and far-field time become visible on the host, and a
        failed step raises here."""
        o, self._pending = self._pending, None
        if o is None:
            return
        ps = self.ps
        hc = ps._host_counters
        rc = N.lib().dimcv_step_wait(ctypes.byref(ps.params), ctypes.byref(o), ps._host_counters_ptr)
        ps._pn = int(hc[N.C_PARTICLE_NUM])
        ps._nfl = int(hc[N.C_FLUID_LIST])
        self._step_outputs(o)
        N.check(rc, "dimcv_step_wait")

    def check_async_steps(self):
        """async_steps needs the one-call fixed-M step converging on the
        device, and a particle count that only the step itself changes (its
        launch bound is the count before the pending step)."""
        ps = self.ps
        if not self.mc_samples or ps.cfg.get_cfg("jacobiChunk") is not None or ps.num_emitters > 0 \
                or ps.params.dyn_rigid:
            raise RuntimeError("async_steps needs fixed-M sampling (mcSamples), no jacobiChunk, no emitters "
                               "and no dynamic rigid particles")

    # --------------------------------------------------- comparators --
    def direct_biot_savart(self, precision: str = "f32"):
        """Build-defined direct O(N^2) sum of the quantity vort2vel estimates.

        precision "f32" / "kahan" write ``v_vort`` (returned as numpy);
        "f64" returns an FP64 (n_max, 3) array.  ``v`` is not modified.
        """
        code = {"f32": 0, "kahan": 1, "f64": 2}[precision]
        self.ps.ensure_neighbor_lists()
        out = None
        if code == 2:
            out = torch.zeros((int(self.ps.particle_max_num), 3), dtype=torch.float64, device=self.ps.device)
        self.ps.call("dimcv_direct_bs", code, out.data_ptr() if out is not None else None,
                     self.ps._nfl, self.ps.stream())
        return out.cpu().numpy() if out is not None else self.v_vort.to_numpy()
