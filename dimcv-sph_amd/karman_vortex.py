"""Kármán vortex scene solver (drop-in for karman_vortex.py of the reference).

``KarmanVortexSolver`` keeps the cylinder carving (karman_vortex.py:10-36),
the step/substep order (:216-242) and the exporters (:38-214).  ``step``
issues the substep through one native call (``dimcv_substep``) unless
``use_native_substep`` is False or a subclass overrides ``substep``; with
fixed-M sampling the grid build and the neighbour lists join that call
(``dimcv_step``).  ``step``
accepts optional host uniforms so that two implementations can be driven with
identical random inputs (``u1``/``u2``: per-slot f32 arrays for the Bernoulli
sample update; ``u_cdf``: f64 array of M uniforms for inverse-CDF sampling).
"""
from __future__ import annotations

import numpy as np
import torch

from dimcv_sph import DIMCVSPHSolver


class KarmanVortexSolver(DIMCVSPHSolver):

    def __init__(self, particle_system):
        super().__init__(particle_system)
        self.circle_pos = np.array([0.65, 0.5, 0.5], dtype=np.float32)
        self.circle_vis = np.array([self.circle_pos[0] * 0.25, self.circle_pos[1], self.circle_pos[2]],
                                   dtype=np.float32)
        self.circle_radius = 0.1
        # step() issues the substep through libdimcv in one call unless a
        # subclass overrides substep() (then the Python pass sequence runs)
        self.use_native_substep = True
        self.init_cylinder()

    def init_cylinder(self):
        self.ps.call("dimcv_init_cylinder", 2, np.float32(self.circle_pos[0]), np.float32(self.circle_pos[1]),
                     np.float32(self.circle_radius), self.ps.stream())
        self.ps._nbr_valid = False

    def export_png(self, cnt, image_path):
        """karman_vortex.py:38-120: the fluid (coloured by vorticity_vis z) and
        the cylinder (object 2) as a 3-D scatter.  The rows are selected on
        the device; only the selected positions and values are copied."""
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
        from matplotlib.colors import Normalize
        N = self.ps.particle_num[None]
        material = self.ps.material.t[:N]
        obj_id = self.ps.object_id.t[:N]
        fluid_mask = material == self.ps.material_fluid
        solid_mask = obj_id == 2
        x = self.x_temp.t[:N]
        fluid_x = x[fluid_mask].cpu().numpy()
        fluid_vort = self.ps.vorticity_vis.t[:N, 2][fluid_mask].cpu().numpy()
        solid_x = x[solid_mask].cpu().numpy()
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

    def _particle_type_table(self):
        """Object id -> particle_type code (karman_vortex.py:163-176): 4, then
        fluid block 0, emitter 1, rigid block 2, rigid body 3, later ones
        overriding earlier ones as the reference's mask assignments do."""
        ps = self.ps
        groups = (ps.obj_id_fluid_blocks, ps.obj_id_emitters, ps.obj_id_rigid_blocks, ps.object_id_rigid_body)
        ids = [int(o) for g in groups for o in g]
        n = max(ids) + 1 if ids else 0
        table = np.full(max(n, 1), 4, np.int32)
        for code, g in enumerate(groups):
            for o in g:
                if int(o) >= 0:
                    table[int(o)] = code
        return torch.as_tensor(table, device=ps.device), n

    def export_ply(self, cnt, ply_path):
        """ASCII PLY with the reference schema (karman_vortex.py:122-214).  The
        data rows are formatted on the device (dimcv_export_ply_rows) from the
        particle buffers; only the finished text is copied to the host."""
        ps = self.ps
        N = ps.particle_num[None]
        if N == 0:
            return
        table, n_types = self._particle_type_table()
        rows = np.zeros(1, np.int64)
        nbytes = np.zeros(1, np.int64)
        ps.call("dimcv_export_ply_size", int(N), table.data_ptr(), n_types, rows.ctypes.data,
                nbytes.ctypes.data, ps.stream())
        if rows[0] == 0:
            return
        out = torch.empty(int(nbytes[0]), dtype=torch.uint8, device=ps.device)
        ps.call("dimcv_export_ply_rows", int(N), table.data_ptr(), n_types, out.data_ptr(), int(nbytes[0]),
                ps.stream())
        body = out.cpu().numpy().tobytes()
        ply_path.mkdir(parents=True, exist_ok=True)
        header = ["ply", "format ascii 1.0", f"element vertex {int(rows[0])}"]
        header += [f"property float {nm}" for nm in ("x", "y", "z", "vort_x", "vort_y", "vort_z")]
        header += [f"property uchar {nm}" for nm in ("red", "green", "blue")]
        header += ["property int object_id", "property uchar particle_type", "end_header"]
        with open(ply_path / f"frame_{cnt:04}.ply", "wb") as f:
            f.write(("\n".join(header) + "\n").encode())
            f.write(body)

    def step(self, u1=None, u2=None, u_cdf=None):
        if self.async_steps:
            self.check_async_steps()
        if self.emit_on_device:
            # emission and its counts stay on the device (dimcv_emit)
            if self.cnt % self.emit_interval == 0:
                self.emit_particle()
        else:
            self.dump_num_particles_each_emitters_ti2np()
            if self.cnt % self.emit_interval == 0:
                self.emit_particle()
                self.dump_num_particles_each_emitters_np2ti()
        self.cnt += 1
        native = self.use_native_substep and type(self).substep is KarmanVortexSolver.substep
        if native and self.mc_samples and type(self).compute_moving_boundary_volume is \
                DIMCVSPHSolver.compute_moving_boundary_volume:
            # fixed-M sampling: grid build, lists and substep in one library call
            self.native_step(u_cdf)
        elif native:
            self.flush()
            self.ps.initialize_particle_system()
            self.compute_moving_boundary_volume()
            self.native_substep(u1, u2, u_cdf)
        else:
            self.flush()
            self.ps.initialize_particle_system()
            self.compute_moving_boundary_volume()
            self._uniforms = (u1, u2, u_cdf)
            try:
                self.substep()
            finally:
                self._uniforms = (None, None, None)
        if self.ps.cfg.get_cfg("bounceBackBoundary"):
            self.enforce_boundary_3D(self.ps.material_fluid)

    def substep(self):
        self.compute_densities()
        self.compute_DFSPH_factor()
        self.divergence_solve()
        self.compute_non_pressure_forces()
        self.dimcv()
        self.predict_velocity()
        self.pressure_solve()
        self.copy_x_temp()
        self.compute_vorticity_vis()
        self.advect()
