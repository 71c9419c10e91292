"""SPH base solver (drop-in for sph_base.py of the reference).

Keeps ``SPHBase.__init__`` (sph_base.py:7-20), ``initialize`` (:80-85),
``compute_static_boundary_volume`` / ``compute_moving_boundary_volume``
(:91-112), ``enforce_boundary_3D`` (:149-179) and ``step`` (:224-232).  The
cubic spline kernel and its gradient (:23-68) are device functions of
libdimcv (csrc/common.h); ``cubic_kernel`` / ``cubic_kernel_derivative``
here evaluate the same f32 expressions on the host for inspection.
Shape matching of dynamic rigid bodies (:87-89, :182-222) is not supported.
"""
from __future__ import annotations

import ctypes

import numpy as np

import _native as N
from _field import HostScalarField


class SPHBase:
    def __init__(self, particle_system):
        self.ps = particle_system
        self.g = np.array(self.ps.cfg.get_cfg("gravitation"))
        vis = self.ps.cfg.get_cfg("viscosity")
        self.viscosity = 1e-3 if vis is None else vis
        self.density_0 = self.ps.cfg.get_cfg("density0")
        if self.density_0 is None:
            self.density_0 = 1000.0
        self.dt = HostScalarField(1e-4, on_change=self._dt_changed)
        p = self.ps.params
        f32 = np.float32
        for d in range(3):
            p.gravity[d] = f32(self.g[d])
        p.density0 = f32(self.density_0)
        p.visc_coef = f32(2 * (self.ps.dim + 2) * self.viscosity)
        p.visc_h2 = f32(0.01 * self.ps.support_radius ** 2)

    def _dt_changed(self, v):
        self.ps.params.dt = np.float32(v)
        self.ps.params.inv_dt = np.float32(1.0 / float(np.float32(v)))

    # host evaluation of sph_base.py:23-68 (f32, same expression order)
    def cubic_kernel(self, r_norm):
        p = self.ps.params
        f32 = np.float32
        q = f32(r_norm) * f32(p.inv_support_radius)
        if q <= 1.0:
            if q <= 0.5:
                q2 = q * q
                q3 = q2 * q
                return f32(p.kernel_k) * (f32(6.0) * q3 - f32(6.0) * q2 + f32(1.0))
            t = f32(1.0) - q
            return f32(p.kernel_k) * f32(2.0) * ((t * t) * t)
        return f32(0.0)

    def cubic_kernel_derivative(self, r):
        p = self.ps.params
        f32 = np.float32
        r = np.asarray(r, dtype=np.float32)
        rn = np.sqrt(f32(r[0] * r[0] + r[1] * r[1]) + f32(r[2] * r[2]))
        q = rn * f32(p.inv_support_radius)
        res = np.zeros(3, np.float32)
        if rn > f32(1e-5) and q <= 1.0:
            gq = (r / rn) * f32(p.inv_support_radius)
            if q <= 0.5:
                res = (f32(p.kernel_dk) * q * (f32(3.0) * q - f32(2.0))) * gq
            else:
                fct = f32(1.0) - q
                res = (f32(p.kernel_dk) * (-fct * fct)) * gq
        return res.astype(np.float32)

    def initialize(self):
        self.ps.initialize_particle_system()
        self.compute_static_boundary_volume()
        self.compute_moving_boundary_volume()

    def compute_static_boundary_volume(self):
        self.ps.call("dimcv_compute_static_boundary_volume", self.ps._pn, self.ps.stream())

    def compute_moving_boundary_volume(self):
        # sph_base.py:105-112, over the dynamic rigid particles (none: no work)
        if self.ps.has_dynamic_rigid:
            self.ps.call("dimcv_compute_moving_boundary_volume", self.ps._pn, self.ps.stream())

    # ------------------------------------------------- rigid shape matching --
    def compute_com(self, object_id):
        """sph_base.py:183-194: centre of mass of object_id's dynamic rigid
        particles, m = m_V0 density; FP64 sums in one fixed order (the
        reference's Taichi reduction order is unspecified), rounded to f32."""
        out = (ctypes.c_double * 9)()
        self.ps.call("dimcv_rigid_moments", int(object_id), 0, None, None, self.ps._pn, out, self.ps.stream())
        with np.errstate(divide="ignore", invalid="ignore"):
            return (np.array(out[1:4]) / out[0]).astype(np.float32)   # 0/0 -> nan, as the reference

    def compute_com_kernel(self, object_id: int):
        return self.compute_com(object_id)

    def compute_rigid_rest_cm(self, object_id: int):
        """sph_base.py:83-85 (rest centre of mass, from the current positions)."""
        self.ps.rigid_rest_cm[object_id] = self.compute_com(object_id)

    def solve_constraints(self, object_id: int):
        """sph_base.py:200-222: shape matching of object_id.  A = sum m p q^T
        (FP64, see compute_com); R = the rotation of A's polar decomposition
        (SVD with both factors turned into rotations, so det R = +1; the
        identity when every entry is below 1e-6); every particle moves to
        cm + R (x_0 - rest_cm).  Returns R (f32)."""
        cm = self.compute_com(object_id)
        rest = np.asarray(self.ps.rigid_rest_cm[object_id], dtype=np.float32)
        out = (ctypes.c_double * 9)()
        c3, r3 = (ctypes.c_float * 3)(*cm), (ctypes.c_float * 3)(*rest)
        self.ps.call("dimcv_rigid_moments", int(object_id), 1, c3, r3, self.ps._pn, out, self.ps.stream())
        A = np.array(out[:9], dtype=np.float64).reshape(3, 3)
        U, sig, Vt = np.linalg.svd(A)
        if np.linalg.det(U) < 0:
            U[:, 2] = -U[:, 2]
        if np.linalg.det(Vt) < 0:
            Vt[2, :] = -Vt[2, :]
        R = (U @ Vt).astype(np.float32)
        if np.all(np.abs(R) < 1e-6):
            R = np.eye(3, dtype=np.float32)
        T = (ctypes.c_float * 15)(*np.concatenate([cm, R.ravel(), rest]).astype(np.float32))
        self.ps.call("dimcv_rigid_apply", int(object_id), T, self.ps._pn, self.ps.stream())
        self.ps._nbr_valid = False
        return R

    def substep(self):
        pass

    def enforce_boundary_3D(self, particle_type: int):
        if particle_type != self.ps.material_fluid:
            raise NotImplementedError("enforce_boundary_3D is provided for fluid particles")
        self.ps.call("dimcv_enforce_boundary_3d", self.ps._pn, self.ps.stream())
        self.ps._nbr_valid = False

    def enforce_boundary_2D(self, particle_type: int):
        raise NotImplementedError("only 3-D scenes are supported")

    def step(self):
        self.ps.initialize_particle_system()
        self.compute_moving_boundary_volume()
        self.substep()
        if self.ps.cfg.get_cfg("bounceBackBoundary"):
            self.enforce_boundary_3D(self.ps.material_fluid)
