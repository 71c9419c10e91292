"""Oracle (test infrastructure only): FP64 frozen-structure rollout for
central finite differences of the rollout adjoint (frozen.cpp).

Only tests/ import this module, as the checker.  ``FrozenRollout`` takes the
tapes of a recorded FP32 forward (the product's ``adjoint.record_step``
output, copied to host memory) and evaluates the same rollout in FP64 with
every discrete decision frozen: the function whose gradient the adjoint
computes.  The update gate |v_vort| < |v| is re-derived here from the
recorded FP32 values with the reference's f32 norm.
"""
from __future__ import annotations

import ctypes

import numpy as np

from oracle.refsim import lib as _oracle_lib

F32 = np.float32


class FrozenParams(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("dt", "rho0", "mV", "m", "h", "kernel_k", "kernel_dk", "visc_coef",
                                                "visc_h2", "eps", "vp_factor", "inv_dt", "inv_dt2")] + \
               [("far_t", ctypes.c_float), ("m_eps", ctypes.c_float), ("rho0_f", ctypes.c_float)]


class FrozenStep(ctypes.Structure):
    P = ctypes.c_void_p
    _fields_ = [("n", ctypes.c_int32), ("kv", ctypes.c_int32), ("kp", ctypes.c_int32), ("M", ctypes.c_int32),
                ("ids", P), ("nbr_off", P), ("nbr", P), ("x32", P), ("rho32", P), ("alpha_div32", P),
                ("alpha_p32", P), ("b32", P), ("da32", P), ("gate", P), ("samples", P)]


class FrozenDebug(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("rho", "alpha0", "b0", "acc", "w", "kden", "vvort", "da0")]


def _f32norm(a):
    a = a.astype(F32)
    return np.sqrt((a[:, 0] * a[:, 0] + a[:, 1] * a[:, 1]) + a[:, 2] * a[:, 2])


class FrozenRollout:
    """FP64 frozen rollout built from host copies of the forward's tapes."""

    def __init__(self, tapes, params: dict):
        self.L = _oracle_lib()
        fn = self.L.or_frozen_rollout
        fn.restype = None
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        self.P = FrozenParams(**params)
        self.keep = []
        self.steps = (FrozenStep * len(tapes))()
        for k, T in enumerate(tapes):
            self._fill(self.steps[k], T)
        self.nsteps = len(tapes)

    def _arr(self, a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        self.keep.append(a)
        return a.ctypes.data

    def _fill(self, S, T):
        n = T["n"]
        cnt = np.minimum(T["nbr_count"][:n], T["nbr"].shape[0])
        nb = T["nbr"][:, :n].view(np.uint32)
        off = np.zeros(n + 1, np.int32)
        lists = []
        for f in range(n):
            e = nb[: cnt[f], f]
            e = e[(e >> 31) == 1] & 0x7FFFFFFF
            lists.append(e.astype(np.int32))
            off[f + 1] = off[f] + e.shape[0]
        S.n, S.kv, S.kp, S.M = n, T["kv"], T["kp"], T["M"]
        S.ids = self._arr(T["ids"], np.int32)
        S.nbr_off = self._arr(off, np.int32)
        S.nbr = self._arr(np.concatenate(lists) if lists else np.zeros(1, np.int32), np.int32)
        S.x32 = self._arr(T["x"], F32)
        S.rho32 = self._arr(T["rho"], F32)
        S.alpha_div32 = self._arr(T["alpha_div"], F32)
        S.alpha_p32 = self._arr(T["alpha_p"], F32)
        S.b32 = self._arr(np.stack(T["b"]), F32)
        S.da32 = self._arr(np.stack(T["da"]), F32)
        S.gate = self._arr((_f32norm(T["v_vort"]) < _f32norm(T["v_div"])).astype(np.uint8), np.uint8)
        S.samples = self._arr(T["sample_idx"][: T["M"]], np.int32)

    def run(self, x0: np.ndarray, v0: np.ndarray, debug: bool = False):
        """Final velocities (id order); with debug also the last step's
        per-pass values (slot order of that step)."""
        n = x0.shape[0]
        x0 = np.ascontiguousarray(x0, np.float64)
        v0 = np.ascontiguousarray(v0, np.float64)
        v = np.zeros_like(v0)
        dbg, out = None, None
        if debug:
            m = self.steps[self.nsteps - 1].n
            out = {k: np.zeros((m, 3) if k in ("acc", "w", "vvort") else m, np.float64) for k, _ in FrozenDebug._fields_}
            dbg = FrozenDebug(**{k: a.ctypes.data for k, a in out.items()})
        self.L.or_frozen_rollout(ctypes.byref(self.P), self.nsteps, self.steps, n, x0.ctypes.data, v0.ctypes.data,
                                 v.ctypes.data, None, ctypes.byref(dbg) if dbg is not None else None)
        return (v, out) if debug else v


def record_ref_tapes(ref, u_cdf_list, x0_insertion: np.ndarray):
    """Tapes of ``len(u_cdf_list)`` steps of the f32 oracle ``ref`` (a RefSim
    of a fluid-only ring scene in fixed-M mode), in the layout
    ``FrozenRollout`` reads.  Particle ids are recovered from the sorted
    initial positions x_0 (unique on a jittered lattice)."""
    from oracle.refsim import lib

    L, s = lib(), ref.sim
    key = lambda a: np.ascontiguousarray(a, np.float32).view(np.dtype((np.void, 12))).ravel()
    index = {k.tobytes(): i for i, k in enumerate(key(x0_insertion))}
    f = lambda name, n: ref.field(name)[:n].copy()
    tapes = []
    for u in u_cdf_list:
        ref.cnt += 1
        ref.initialize_particle_system()
        L.or_compute_moving_boundary_volume(s)
        n = ref.particle_num
        T = {"n": n, "ids": np.array([index[k.tobytes()] for k in key(f("x_0", n))], np.int32)}
        lists = [ref.neighbors(i) for i in range(n)]
        kmax = max(1, max(len(q) for q in lists))
        nbr = np.zeros((kmax, n), np.uint32)
        mat = ref.field("material")
        for i, q in enumerate(lists):
            nbr[: len(q), i] = q.astype(np.uint32) | np.where(mat[q] == 1, 0x80000000, 0).astype(np.uint32)
        T["nbr"], T["nbr_count"] = nbr.view(np.int32), np.array([len(q) for q in lists], np.int32)
        T["x"] = f("x", n)
        L.or_compute_densities(s)
        L.or_compute_DFSPH_factor(s)
        T["rho"] = f("density", n)
        # divergence_solve (refsim.py) with the per-iteration record
        L.or_compute_density_change(s)
        T["b"] = [f("density_adv", n)]
        L.or_multiply_time_step(s, 0, np.float32(ref.inv_dt))
        T["alpha_div"] = f("dfsph_factor", n)
        it = 0
        while it < 1 or it < 100:
            L.or_divergence_solver_iteration_kernel(s)
            L.or_compute_density_change(s)
            T["b"].append(f("density_adv", n))
            err = L.or_compute_density_error(s, np.float32(0.0)) / ref.fluid_particle_num
            it += 1
            if err <= ref.inv_dt * 0.1 * 0.01 * ref.density_0:
                break
        T["kv"] = it
        L.or_multiply_time_step(s, 0, ref.dt)
        L.or_compute_non_pressure_forces(s)
        T["v_div"] = f("v", n)
        T["a"] = f("acceleration", n)
        vv = np.zeros((ref.particle_max_num, 3), np.float32)
        ref.dimcv(None, None, u, v_vort_out=vv)
        T["ws"], T["kden"] = f("d_vort_smoothed", n), f("kernel_density", n)
        T["v_vort"] = vv[:n].copy()
        T["M"] = ref.num_samples
        T["sample_idx"] = f("sample_idx", ref.particle_max_num)
        L.or_predict_velocity(s)
        L.or_compute_density_adv(s)
        T["da"] = [f("density_adv", n)]
        L.or_multiply_time_step(s, 0, np.float32(ref.inv_dt2))
        T["alpha_p"] = f("dfsph_factor", n)
        it = 0
        while it < 1 or it < 100:
            L.or_pressure_solve_iteration_kernel(s)
            L.or_compute_density_adv(s)
            T["da"].append(f("density_adv", n))
            err = L.or_compute_density_error(s, np.float32(ref.density_0)) / ref.fluid_particle_num
            it += 1
            if err <= 0.05 * 0.01 * ref.density_0:
                break
        T["kp"] = it
        L.or_copy_x_temp(s)
        L.or_compute_vorticity_vis(s)
        L.or_advect(s)
        tapes.append(T)
    return tapes


def params_of(radius: float, dt: float, viscosity: float, volume_scale: float = 0.2, density0: float = 1000.0):
    """FrozenParams of a scene, from the same f32 constants the step uses
    (particle_system.py:41-46, sph_base.py:14-35, dimcv_sph.py:9-28)."""
    f = np.float32
    sr = 4.0 * radius
    h = f(sr)
    inv3 = f((1.0 / sr) ** 3)
    m_V0_d = 0.8 * (2 * radius) ** 3
    m_V0 = f(m_V0_d)
    far_t = np.float32(h * h)
    inf = np.float32(np.inf)
    while np.sqrt(np.nextafter(far_t, inf, dtype=np.float32)) <= h:
        far_t = np.nextafter(far_t, inf, dtype=np.float32)
    while np.sqrt(far_t) > h:
        far_t = np.nextafter(far_t, np.float32(0), dtype=np.float32)
    return dict(dt=float(f(dt)), rho0=float(density0), mV=float(m_V0), m=float(f(m_V0 * f(density0))),
                h=float(h), kernel_k=float(f(2.5465) * inv3), kernel_dk=float((f(6.0) * f(2.5465)) * inv3),
                visc_coef=float(f(2 * (3 + 2) * viscosity)), visc_h2=float(f(0.01 * sr ** 2)),
                eps=float(f(1e-6)), vp_factor=float(f(0.1592 * volume_scale * m_V0_d)),
                inv_dt=float(f(1.0 / float(f(dt)))), inv_dt2=float(f(1.0 / (float(f(dt)) * float(f(dt))))),
                far_t=float(far_t), m_eps=float(f(1e-5)), rho0_f=float(density0))
